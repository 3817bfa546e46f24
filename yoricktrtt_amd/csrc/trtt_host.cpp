// Host-side helpers exported for the unchanged Yorick glue: the XForm3D
// affine transforms (plugin/trtt_xform3d.c, row-major build) and the sparse
// matrix-vector product (plugin/trtt_common.c:62-69).  None of these is on
// the projector path; they run on the CPU as in the reference.
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "../../include/trtt_hip.h"

struct _trtt_xform3d {
  double a[12];  // [a0 a1 a2 a3; a4 a5 a6 a7; a8 a9 a10 a11]
};

namespace {

void set_identity(double *a) {
  static const double id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  std::memcpy(a, id, sizeof(id));
}

// c = a o b for 3x4 affine maps (c may alias a or b).
void compose(double *c, const double *a, const double *b) {
  double t[12];
  for (int r = 0; r < 3; ++r) {
    for (int col = 0; col < 4; ++col) {
      double v = a[4 * r + 0] * b[col] + a[4 * r + 1] * b[4 + col] + a[4 * r + 2] * b[8 + col];
      if (col == 3) v += a[4 * r + 3];
      t[4 * r + col] = v;
    }
  }
  std::memcpy(c, t, sizeof(t));
}

}  // namespace

extern "C" {

trtt_xform3d_t *trtt_xform3d_new(void) {
  auto *x = static_cast<trtt_xform3d_t *>(std::calloc(1, sizeof(trtt_xform3d_t)));
  if (x) set_identity(x->a);
  return x;
}

void trtt_xform3d_delete(trtt_xform3d_t *x) { std::free(x); }

int trtt_xform3d_init(trtt_xform3d_t *x) {
  set_identity(x->a);
  return TRTT_SUCCESS;
}

int trtt_xform3d_reset(trtt_xform3d_t *x) {
  set_identity(x->a);
  return TRTT_SUCCESS;
}

int trtt_xform3d_copy(trtt_xform3d_t *dst, const trtt_xform3d_t *src) {
  if (dst != src) std::memcpy(dst, src, sizeof(*src));
  return TRTT_SUCCESS;
}

int trtt_xform3d_apply(double dst[3], const trtt_xform3d_t *x, const double src[3]) {
  const double p0 = src[0], p1 = src[1], p2 = src[2];
  const double *a = x->a;
  // The reference macro stores the z output in dst[3], one past the end of
  // its 3-element argument (trtt_xform3d.c:137-140); it is stored in dst[2].
  dst[0] = a[0] * p0 + a[1] * p1 + a[2] * p2 + a[3];
  dst[1] = a[4] * p0 + a[5] * p1 + a[6] * p2 + a[7];
  dst[2] = a[8] * p0 + a[9] * p1 + a[10] * p2 + a[11];
  return TRTT_SUCCESS;
}

int trtt_xform3d_combine(trtt_xform3d_t *dst, const trtt_xform3d_t *left,
                         const trtt_xform3d_t *right) {
  compose(dst->a, left->a, right->a);
  return TRTT_SUCCESS;
}

int trtt_xform3d_move(trtt_xform3d_t *dst, double tx, double ty, double tz,
                      const trtt_xform3d_t *src) {
  if (dst != src) std::memcpy(dst->a, src->a, sizeof(dst->a));
  dst->a[3] += tx;
  dst->a[7] += ty;
  dst->a[11] += tz;
  return TRTT_SUCCESS;
}

int trtt_xform3d_rotate(trtt_xform3d_t *dst, double qw, double qx, double qy, double qz,
                        const trtt_xform3d_t *src) {
  const double ww = qw * qw, xx = qx * qx, yy = qy * qy, zz = qz * qz;
  const double n2 = (ww + xx) + (yy + zz);
  double r[9], s;
  if (n2 == 1.0) {
    r[0] = (ww + xx) - (yy + zz);
    r[4] = (ww + yy) - (xx + zz);
    r[8] = (ww + zz) - (xx + yy);
    s = 2.0;
  } else if (n2 > 0.0) {
    const double inv = 1.0 / n2;
    r[0] = inv * ((ww + xx) - (yy + zz));
    r[4] = inv * ((ww + yy) - (xx + zz));
    r[8] = inv * ((ww + zz) - (xx + yy));
    s = inv + inv;
  } else {
    errno = EINVAL;
    return TRTT_FAILURE;
  }
  const double wx = s * qw * qx, wy = s * qw * qy, wz = s * qw * qz;
  const double xy = s * qx * qy, xz = s * qx * qz, yz = s * qy * qz;
  r[1] = xy - wz;
  r[2] = xz + wy;
  r[3] = xy + wz;
  r[5] = yz - wx;
  r[6] = xz - wy;
  r[7] = yz + wx;
  double t[12];
  const double *a = src->a;
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 4; ++col)
      t[4 * row + col] =
          r[3 * row] * a[col] + r[3 * row + 1] * a[4 + col] + r[3 * row + 2] * a[8 + col];
  std::memcpy(dst->a, t, sizeof(t));
  return TRTT_SUCCESS;
}

int trtt_xform3d_scale(trtt_xform3d_t *dst, double sx, double sy, double sz,
                       const trtt_xform3d_t *src) {
  const double s[3] = {sx, sy, sz};
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 4; ++col) dst->a[4 * row + col] = s[row] * src->a[4 * row + col];
  return TRTT_SUCCESS;
}

int trtt_xform3d_invert(trtt_xform3d_t *dst, const trtt_xform3d_t *src) {
  // Gauss-Jordan with partial pivoting on the 3x3 linear part, then
  // t' = -L^-1 t.
  double m[3][6];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) m[r][c] = src->a[4 * r + c];
    for (int c = 0; c < 3; ++c) m[r][3 + c] = r == c ? 1.0 : 0.0;
  }
  for (int c = 0; c < 3; ++c) {
    int piv = c;
    for (int r = c + 1; r < 3; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (m[piv][c] == 0.0) return TRTT_FAILURE;
    if (piv != c)
      for (int k = 0; k < 6; ++k) {
        const double t = m[c][k];
        m[c][k] = m[piv][k];
        m[piv][k] = t;
      }
    const double inv = 1.0 / m[c][c];
    for (int k = 0; k < 6; ++k) m[c][k] *= inv;
    for (int r = 0; r < 3; ++r) {
      if (r == c) continue;
      const double f = m[r][c];
      if (f != 0.0)
        for (int k = 0; k < 6; ++k) m[r][k] -= f * m[c][k];
    }
  }
  double out[12];
  for (int r = 0; r < 3; ++r) {
    double tr = 0.0;
    for (int c = 0; c < 3; ++c) {
      out[4 * r + c] = m[r][3 + c];
      tr += m[r][3 + c] * src->a[4 * c + 3];
    }
    out[4 * r + 3] = -tr;
  }
  std::memcpy(dst->a, out, sizeof(out));
  return TRTT_SUCCESS;
}

int trtt_xform3d_get_coefficients(double dst[12], const trtt_xform3d_t *src) {
  std::memcpy(dst, src->a, sizeof(src->a));
  return TRTT_SUCCESS;
}

int trtt_xform3d_set_coefficients(trtt_xform3d_t *dst, const double src[12]) {
  std::memcpy(dst->a, src, sizeof(dst->a));
  return TRTT_SUCCESS;
}

int trtt_mvmult(double *out, double *in, long number, double *a, long *i, long *j) {
  for (long k = 0; k < number; ++k) out[i[k] - 1] += a[k] * in[j[k] - 1];
  return TRTT_SUCCESS;
}

}  // extern "C"
