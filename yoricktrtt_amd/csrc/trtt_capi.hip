// C ABI of libtrtt_hip.so: plans (device-resident batched projectors) and the
// reference plugin symbols (synchronous per-view drop-ins on host FP64 data).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "trtt_core.h"

namespace trtt {
hipError_t launch_projector(int kind, bool table, bool quirk, bool f64, const DevPlan &P, int job,
                            const void *in, void *out, const int *seg_begin, int nrows,
                            int views_per_block, int atomic_out, hipStream_t st);
hipError_t launch_residual(bool f64, void *s, const void *y, const void *w, long n, double *fx,
                           hipStream_t st, bool split = false);
hipError_t launch_axial(bool cone, int w, bool f64, const DevPlan &P, int job, const void *in,
                        void *out, const int *seg_begin, int nrows, int views_per_block,
                        int atomic_out, int pmax, int pmax2, void *scratch, void *hdr,
                        int chunk, void *volk, int plain, const AxialTuning &tu, hipEvent_t *ev,
                        const AxialOverlap *ov, int nbuf, int two, hipStream_t st);
int axial_tiles(const DevPlan &P, int job, const AxialTuning &tu);
size_t axial_scratch_bytes(const DevPlan &P, int chunk, const AxialTuning &tu);
size_t axial_hdr_bytes(const DevPlan &P, int chunk, const AxialTuning &tu);
hipError_t launch_sparse_columns(int kind, bool table, bool quirk, bool f64, const DevPlan &P,
                                 bool fill, int *count, const long long *offset, void *val,
                                 int *row, int *col, int round6, hipStream_t st);
hipError_t sparse_scan(const int *count, long long *offset, long n, void *temp,
                       size_t &temp_bytes, hipStream_t st);
hipError_t sparse_sort(const unsigned *key, unsigned *key_out, unsigned *ids, unsigned *perm,
                       long n, int bits, void *temp, size_t &temp_bytes, hipStream_t st);
hipError_t sparse_compress(bool f64, const unsigned *key_sorted, const unsigned *perm,
                           const void *val, const int *other, void *out_val, int *out_idx,
                           int *ptr, long n, int n_major, hipStream_t st);
hipError_t sparse_spmv(bool f64, const int *ptr, const int *idx, const void *val, const void *x,
                       void *y, int n_major, hipStream_t st);
hipError_t launch_dd(bool f64, const DDConst *cs, const DDConst &c0, int nviews, int job,
                     int stuck, const void *in, void *out, hipStream_t st);
bool host_column_2d(bool fan, const DevPlan &P, const ViewConst &g, int i, int j, double *out);
double host_cdf3(double x);
void host_lf_column(const DevPlan &P, const ViewConst &g, int i1, int i2, double *out);
hipError_t launch_tinterp(bool f64, int job, const void *in, void *out, long npix, int nframes,
                          int taps, const int *idx, const void *w, const int *ptr,
                          const int *src, const void *wb, hipStream_t st);
}  // namespace trtt

using trtt::ViewConst;

namespace {

thread_local std::string g_last_error;
int g_device = 0;
int g_precision = TRTT_HIP_F32;
std::mutex g_mutex;

int fail(const std::string &msg) {
  g_last_error = msg;
  std::fprintf(stderr, "trtt_hip: %s\n", msg.c_str());
  return TRTT_FAILURE;
}

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  fail(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

// ---------------------------------------------------------------------------
// Per-view constants (FP64, reference expression order)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

void view_consts_3d(ViewConst &g, const double *xos, const double *xsd, bool cone, double s,
                    int s_deg, double u_scl, double v_scl) {
  // trtt_3D_projectors.c:1190-1232 (cone), :967-1003 (parallel)
  const double a1 = xsd[0], a2 = xsd[1], a3 = xsd[2], a4 = xsd[3];
  const double a5 = xsd[4], a6 = xsd[5], a7 = xsd[6], a8 = xsd[7];
  const double a9 = xsd[8], a10 = xsd[9], a11 = xsd[10], a12 = xsd[11];
  const double a1inv = -1.0 * (1.0 / a1);
  const double b1 = a6 + a1inv * a5 * a2, c1 = a10 + a1inv * a9 * a2;
  const double b2 = a7 + a1inv * a5 * a3, c2 = a11 + a1inv * a9 * a3;
  const double b3 = a8 + a1inv * a5 * a4, c3 = a12 + a1inv * a9 * a4;
  std::memset(&g, 0, sizeof(g));
  g.nvx = b1 * xos[4] + b2 * xos[8];
  g.nux = c1 * xos[4] + c2 * xos[8];
  g.nvy = b1 * xos[5] + b2 * xos[9];
  g.nuy = c1 * xos[5] + c2 * xos[9];
  g.nvz = b1 * xos[6] + b2 * xos[10];
  g.nuz = c1 * xos[6] + c2 * xos[10];
  if (cone) {
    g.nvc = b1 * xos[7] + b2 * xos[11];
    g.nuc = c1 * xos[7] + c2 * xos[11];
  } else {
    g.nvc = b1 * xos[7] + b2 * xos[11] + b3;
    g.nuc = c1 * xos[7] + c2 * xos[11] + c3;
  }
  g.dx = a1 * xos[0] + a2 * xos[4] + a3 * xos[8];
  g.dy = a1 * xos[1] + a2 * xos[5] + a3 * xos[9];
  g.dz = a1 * xos[2] + a2 * xos[6] + a3 * xos[10];
  g.dc = a1 * xos[3] + a2 * xos[7] + a3 * xos[11];
  g.o1 = xos[0];
  g.o2 = xos[1];
  g.o3 = xos[2];
  g.o4 = xos[3];
  g.a4m = -1.0 * a4;
  g.a4sqr = a4 * a4;
  g.a8 = a8;
  g.a12 = a12;
  g.b3 = b3;
  g.c3 = c3;
  const double hs0 = 0.5 * (s_deg + 1);
  const double a1sqr = a1 * a1;
  const double a5sqr = v_scl * v_scl * a5 * a5;
  const double a9sqr = u_scl * u_scl * a9 * a9;
  const double du = std::sqrt(1.0 + (a9sqr / (a1sqr + a5sqr)));
  const double dv = std::sqrt(1.0 + (a5sqr / a1sqr));
  g.pb_rho0 = s * s * s * du * dv;
  g.pb_hu = s / u_scl;
  g.pb_hv = s / v_scl;
  g.pb_su = 1.0 / g.pb_hu;
  g.pb_sv = 1.0 / g.pb_hv;
  g.pb_hsu = hs0 * g.pb_hu;
  g.pb_hsv = hs0 * g.pb_hv;
}

void view_consts_2d(ViewConst &g, const double *xos, const double *xsd, bool fan, bool table,
                    double step, double s, int s_deg, double v_scl) {
  // trtt_2D_projectors.c:195-221 (parallel), :279-315 (fan)
  const double a1 = xsd[0], a2 = xsd[1], a4 = xsd[3], a5 = xsd[4], a6 = xsd[5], a8 = xsd[7];
  const double a1inv = -1.0 * (1.0 / a1);
  std::memset(&g, 0, sizeof(g));
  if (fan) {
    const double b1 = a6 + a1inv * a5 * a2;
    const double b3 = a8 + a1inv * a5 * a4;
    g.nvx = b1 * xos[4];
    g.nvy = b1 * xos[5];
    g.nvc = b1 * xos[7];
    g.dx = a1 * xos[0] + a2 * xos[4];
    g.dy = a1 * xos[1] + a2 * xos[5];
    g.dc = a1 * xos[3] + a2 * xos[7];
    g.o1 = xos[0];
    g.o2 = xos[1];
    g.o4 = xos[3];
    g.a4m = -1.0 * a4;
    g.a4sqr = a4 * a4;
    g.a8 = a8;
    g.b3 = b3;
  } else {
    const double a5a1 = a5 * a1inv;
    const double a5a1scl = a5a1 * v_scl;
    const double b1 = a6 + a5a1 * a2;
    const double b3 = a8 + a5a1 * a4;
    g.nvx = b1 * xos[4];
    g.nvy = b1 * xos[5];
    g.nvc = b1 * xos[7] + b3;
    const double dv = std::sqrt(1 + a5a1scl * a5a1scl);
    g.pb_rho0 = s * s * dv;
    g.pb_hv = (dv * s) / v_scl;
    g.pb_sv = table ? 1 / (g.pb_hv * step) : 1 / g.pb_hv;
    g.pb_hsv = 0.5 * (s_deg + 1) * g.pb_hv;
  }
}

void view_consts_lf(ViewConst &g, const lf2d_parameters_t *p, double beta, double det_offset) {
  // LongFessler2D.c:59-75
  std::memset(&g, 0, sizeof(g));
  const double cb = std::cos(beta), sb = std::sin(beta);
  const double oy = cb * p->objectOffset1 + sb * p->objectOffset2;
  const double ox = -sb * p->objectOffset1 + cb * p->objectOffset2;
  g.lf_cb = cb;
  g.lf_sb = sb;
  g.lf_x0 = ox - 0.5 * (p->objectDimension1 - 1) * p->voxelSize;
  g.lf_y0 = oy - 0.5 * (p->objectDimension2 - 1) * p->voxelSize;
  g.lf_off = det_offset / p->pixelStep + 0.5 * (p->detectorDimension - 1);
  g.w0 = 1.0;
}

#pragma clang fp contract(on)

// Fast axial path eligibility: closed-form FP32 3-D plans whose views all
// rotate about the object z axis (o3 = den_z = num_v_z = 0), with the whole
// volume in front of the source.  Returns the footprint width bound W
// (9 or 13 pixels) and the column profile length, or W = 0.
// CDF of the quadratic B-spline (support [-3/2, 3/2]), FP64, piecewise.
double host_cdf2(double x) {
  if (x <= -1.5) return 0.0;
  if (x <= -0.5) return (x + 1.5) * (x + 1.5) * (x + 1.5) / 6.0;
  if (x <= 0.5) return 0.5 + 0.75 * x - x * x * x / 3.0;
  if (x <= 1.5) return 1.0 - (1.5 - x) * (1.5 - x) * (1.5 - x) / 6.0;
  return 1.0;
}

// A footprint table of degree deg (2 or 3) that holds the B-spline CDF at
// its nodes, cum[n] = F(coord0 + n step) - F(coord0) to within 1e-6 (the
// trapezoid tables of Ytrtt_tomobj.i:140-196 are within 4e-8 for 10^4
// entries), lets the axial kernels evaluate the table entry of an index in
// closed form.  Returns deg when it does, else 0.
int table_cdf_degree(const double *cum, long m, double step, double coord0, int deg) {
  if (!cum || m < 2 || !(step > 0.0) || (deg != 2 && deg != 3)) return 0;
  // the table covers the support [-hs, hs] (the axial kernels skip the clamp)
  const double hs = 0.5 * (deg + 1);
  if (!(coord0 <= -hs) || !(coord0 + double(m - 1) * step >= hs)) return 0;
  auto F = [deg](double x) { return deg == 2 ? host_cdf2(x) : trtt::host_cdf3(x); };
  const double f0 = F(coord0);
  for (long n = 0; n < m; ++n)
    if (!(std::fabs(cum[n] - (F(coord0 + double(n) * step) - f0)) <= 1e-6)) return 0;
  return deg;
}

void axial_analysis(trtt_hip_plan *p) {
  p->axial_w = 0;
  p->axial_pmax = 0;
  p->axial_pmax2 = 0;
  // the kernels take the CDF as exactly -1/2 and 1/2 at a voxel window's
  // outer bounds: the closed-form cubic with s_deg 3 (with another degree it
  // is truncated), or a table that holds the CDF of the B-spline of degree
  // s_deg (2 or 3) at its nodes (evaluated in closed form there)
  const bool cubic_ok = !p->table && p->s_deg == 3;
  const bool table_ok = p->table && p->table_deg == p->s_deg;
  if (!(cubic_ok || table_ok) || p->precision != TRTT_HIP_F32) return;
  if (p->kind != trtt::CB3D && p->kind != trtt::PB3D) return;
  // the kernels broadcast detector and column indices packed in 16 bits
  if (p->nu > 65535 || p->nv > 65535 || p->nx > 65535 || p->ny > 65535) return;
  // and address a phase of the volume with 32-bit offsets
  if (double(p->nx) * p->ny * p->nz >= 2147483648.0) return;
  const bool cone = p->kind == trtt::CB3D;
  const double s = p->s_scl, hs0 = 0.5 * (p->s_deg + 1);
  const double hu_init = s / p->u_scl, hv_init = s / p->v_scl;
  double hu_max = 0.0, hv_max = 0.0, duc_max = 0.0, duc_min = 1e300, dhs_rate = 0.0;
  const double ci[2] = {0.0, double(p->nx - 1)}, cj[2] = {0.0, double(p->ny - 1)};
  for (const ViewConst &g : p->h_views) {
    if (g.nvz != 0.0) return;
    if (cone) {
      if (g.o3 != 0.0 || g.dz != 0.0) return;
      double wmin = 1e300, wmax = -1e300, qmax = 0.0, qmin = 1e300;
      for (double i : ci)
        for (double j : cj) {
          const double wk = (g.o4 + g.o2 * j) + g.o1 * i;
          const double den = (g.dc + g.dy * j) + g.dx * i;
          wmin = std::min(wmin, wk);
          wmax = std::max(wmax, wk);
          if (!(den * g.a4m > 0.0)) return;
          qmax = std::max(qmax, std::fabs(g.a4m / den));
          qmin = std::min(qmin, std::fabs(g.a4m / den));
        }
      if (!(wmin > 0.0) && !(wmax < 0.0)) return;  // source plane crosses the volume
      const double magn = std::fabs(g.a4m) / std::min(std::fabs(wmin), std::fabs(wmax));
      const double u_ext = (std::max(std::fabs(g.a12), std::fabs(p->nu - g.a12)) +
                            8.0 * hu_init * magn + 4.0) * p->u_scl;
      const double v_ext = (std::max(std::fabs(g.a8), std::fabs(p->nv - g.a8)) +
                            8.0 * hv_init * magn + 4.0) * p->v_scl;
      const double du = std::sqrt(1.0 + u_ext * u_ext / g.a4sqr);
      const double dv = std::sqrt(1.0 + v_ext * v_ext / g.a4sqr);
      hu_max = std::max(hu_max, hu_init * magn * du);
      hv_max = std::max(hv_max, hv_init * magn * dv);
      duc_max = std::max(duc_max, qmax * std::fabs(g.nuz));
      duc_min = std::min(duc_min, qmin * std::fabs(g.nuz));
      // |d hs / d k| <= hs0 hu du' duc with du' <= |u| u_scl / a4^2
      dhs_rate = std::max(dhs_rate, hs0 * hu_init * magn * u_ext * p->u_scl / g.a4sqr *
                                        qmax * std::fabs(g.nuz));
    } else {
      hu_max = std::max(hu_max, g.pb_hu);
      hv_max = std::max(hv_max, g.pb_hv);
      duc_max = std::max(duc_max, std::fabs(g.nuz));
      duc_min = std::min(duc_min, std::fabs(g.nuz));
    }
  }
  const double npix = std::floor(2.0 * hs0 * std::max(hu_max, hv_max)) + 2.0;
  const int w = npix <= 9.0 ? 9 : (npix <= 13.0 ? 13 : 0);
  const double pmax = std::ceil(63.0 * duc_max + 2.0 * hs0 * hu_max) + 8.0;
  if (w == 0 || !(pmax <= 256.0)) return;
  p->axial_w = w;
  p->axial_pmax = int(pmax);
  // backward tiles of 128 k: the profile over 128 slices must fit too
  const double pmax2 = std::ceil(127.0 * duc_max + 2.0 * hs0 * hu_max) + 8.0;
  p->axial_pmax2 = pmax2 <= 256.0 ? int(pmax2) : 0;
  // voxels m slices apart start their windows on distinct pixels once
  // m (duc_min - 2 dhs_rate) > 1.01: the forward then accumulates its
  // profiles in m plain passes (1 for consecutive voxels), or with LDS
  // atomics beyond kMaxProfilePasses
  constexpr int kMaxProfilePasses = 4;
  const double rate = duc_min - 2.0 * dhs_rate;
  p->axial_plain = rate > 1.01 ? 1
                   : rate > 1.01 / kMaxProfilePasses ? int(std::floor(1.01 / rate)) + 1
                                                     : 0;
  // TRTT_HIP_PROFILE_ATOMIC=1 accumulates the axial profiles with LDS
  // atomics even where plain accumulation is safe (same result)
  if (const char *e = std::getenv("TRTT_HIP_PROFILE_ATOMIC"))
    if (e[0] == '1') p->axial_plain = 0;
}

// Number of workgroups a launch should offer to fill 256 CUs several times.
constexpr long kTargetBlocks = 4096;

// Tiles of the general kernels (2-D: 1024 pixels, 16 x kTile2DRows in the
// backward, 32 x 32 in the forward).
int tiles_of_dims(int kind, long nx, long ny, long nz) {
  const bool is2d = kind == trtt::PB2D || kind == trtt::FB2D || kind == trtt::LF2D;
  const long tx = is2d ? 16 : 8, tz = is2d ? 1 : 8;
  const long ty = is2d ? trtt::kTile2DRows : 8;
  return int(((nx + tx - 1) / tx) * ((ny + ty - 1) / ty) * ((nz + tz - 1) / tz));
}

int tiles_of(const trtt_hip_plan *p) { return tiles_of_dims(p->kind, p->nx, p->ny, p->nz); }

// Group views by (t0, t1) and cut the ordered list into `nrows` back-
// projection rows of contiguous segments.
void build_segments(trtt_hip_plan *p) {
  const long V = p->nviews;
  p->h_order.resize(V);
  std::iota(p->h_order.begin(), p->h_order.end(), 0);
  std::stable_sort(p->h_order.begin(), p->h_order.end(), [&](int a, int b) {
    const auto &A = p->h_views[a], &B = p->h_views[b];
    return A.t0 != B.t0 ? A.t0 < B.t0 : A.t1 < B.t1;
  });
  const long ntiles = tiles_of(p);
  long nrows = (kTargetBlocks + ntiles - 1) / ntiles;
  nrows = std::max(1L, std::min(nrows, V));
  if (p->kind == trtt::LF2D) nrows = 1;
  p->h_segs.clear();
  p->h_seg_begin.assign(nrows + 1, 0);
  for (long r = 0; r < nrows; ++r) {
    const long a = r * V / nrows, b = (r + 1) * V / nrows;
    p->h_seg_begin[r] = int(p->h_segs.size());
    long s = a;
    while (s < b) {
      const auto &A = p->h_views[p->h_order[s]];
      long e = s + 1;
      while (e < b && p->h_views[p->h_order[e]].t0 == A.t0 &&
             p->h_views[p->h_order[e]].t1 == A.t1)
        ++e;
      p->h_segs.push_back(trtt::Segment{int(s), int(e - s), A.t0, A.t1});
      s = e;
    }
  }
  p->h_seg_begin[nrows] = int(p->h_segs.size());
  p->two_phase_segs = false;
  for (const trtt::Segment &sg : p->h_segs) p->two_phase_segs |= sg.t0 != sg.t1;
}

template <typename T>
bool upload(T *&dst, const std::vector<T> &src) {
  if (src.empty()) return true;
  if (!hip_ok(hipMalloc(&dst, sizeof(T) * src.size()), "hipMalloc")) return false;
  return hip_ok(hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice),
                "hipMemcpy");
}

bool finalize_plan(trtt_hip_plan *p, const double *cum) {
  if (!hip_ok(hipSetDevice(p->device), "hipSetDevice")) return false;
  build_segments(p);
  if (!upload(p->d_views, p->h_views) || !upload(p->d_order, p->h_order) ||
      !upload(p->d_segs, p->h_segs) || !upload(p->d_seg_begin, p->h_seg_begin))
    return false;
  if (p->table) {
    if (p->precision == TRTT_HIP_F64) {
      std::vector<double> t(cum, cum + p->m);
      double *d = nullptr;
      if (!upload(d, t)) return false;
      p->d_table = d;
    } else {
      std::vector<float> t(p->m);
      for (long i = 0; i < p->m; ++i) t[i] = float(cum[i]);
      float *d = nullptr;
      if (!upload(d, t)) return false;
      p->d_table = d;
    }
  }
  if (!hip_ok(hipMalloc(&p->d_err, sizeof(int)), "hipMalloc")) return false;
  if (!hip_ok(hipMemset(p->d_err, 0, sizeof(int)), "hipMemset")) return false;
  trtt::DevPlan &P = p->dev;
  P.views = p->d_views;
  P.order = p->d_order;
  P.segs = p->d_segs;
  P.table = p->d_table;
  P.err = p->d_err;
  P.m = p->m;
  P.step = p->step;
  P.coord0 = p->coord0;
  P.nx = int(p->nx);
  P.ny = int(p->ny);
  P.nz = int(p->nz);
  P.nu = int(p->nu);
  P.nv = int(p->nv);
  P.nviews = int(p->nviews);
  P.nt = int(p->nt);
  P.vox_per_phase = p->nx * p->ny * p->nz;
  P.u_scl = p->u_scl;
  P.v_scl = p->v_scl;
  const double s = p->s_scl;
  const bool is3d = p->kind == trtt::PB3D || p->kind == trtt::CB3D;
  {
#pragma clang fp contract(off)
    P.rho_init = is3d ? s * s * s : s * s;
    P.hu_init = s / p->u_scl;
    P.hv_init = s / p->v_scl;
    P.hs0 = 0.5 * (p->s_deg + 1);
  }
  return true;
}

}  // namespace

// Device side of the host-array evaluation of a plan's data term.
struct EvalData {
  void *y = nullptr, *w = nullptr;  // data and weights, plan precision
  void *wcyt = nullptr;             // sum_k A_k' W_k y_k (volume series)
  void *x = nullptr, *g = nullptr;  // volume series
  void *sino = nullptr;
  double *stage = nullptr;          // FP64 staging of x and gx
  double *fx = nullptr;
};

namespace {

void eval_release(EvalData *e) {
  if (!e) return;
  hipFree(e->y);
  hipFree(e->w);
  hipFree(e->wcyt);
  hipFree(e->x);
  hipFree(e->g);
  hipFree(e->sino);
  hipFree(e->stage);
  hipFree(e->fx);
  delete e;
}

void release(trtt_hip_plan *p) {
  if (!p) return;
  hipSetDevice(p->device);
  eval_release(p->eval);
  hipFree(p->d_views);
  hipFree(p->d_order);
  hipFree(p->d_segs);
  hipFree(p->d_seg_begin);
  hipFree(p->d_table);
  hipFree(p->d_err);
  hipFree(p->d_scratch);
  hipFree(p->d_hdr);
  hipFree(p->d_volk);
  for (hipEvent_t e : p->ev_fwd) hipEventDestroy(e);
  for (hipEvent_t e : p->ev_bwd) hipEventDestroy(e);
  for (int b = 0; b < 2; ++b) {
    if (p->overlap.fwd_done[b]) hipEventDestroy(p->overlap.fwd_done[b]);
    if (p->overlap.red_done[b]) hipEventDestroy(p->overlap.red_done[b]);
  }
  if (p->overlap.aux) hipStreamDestroy(p->overlap.aux);
  delete p;
}

int views_per_block(const trtt_hip_plan *p) {
  const long ntiles = tiles_of(p);
  long vpb = (ntiles * p->nviews) / kTargetBlocks;
  return int(std::max(1L, std::min(vpb, 8L)));
}

// ---------------------------------------------------------------------------
// Scratch buffers of the synchronous drop-in symbols (grow-only, per process)
// ---------------------------------------------------------------------------
struct Scratch {
  void *buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap[4] = {0, 0, 0, 0};
  int device = -1;
  bool ensure(int slot, size_t bytes) {
    if (cap[slot] >= bytes) return true;
    hipFree(buf[slot]);
    buf[slot] = nullptr;
    cap[slot] = 0;
    if (!hip_ok(hipMalloc(&buf[slot], bytes), "hipMalloc")) return false;
    cap[slot] = bytes;
    return true;
  }
};
Scratch g_scratch;

__global__ void cvt_d2f(const double *__restrict__ a, float *__restrict__ b, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    b[i] = float(a[i]);
}
__global__ void cvt_f2d(const float *__restrict__ a, double *__restrict__ b, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    b[i] = double(a[i]);
}

unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  return unsigned(std::max(1L, std::min(b, 8192L)));
}

hipError_t launch_plan(trtt_hip_plan *p, int job, const void *in, void *out, hipStream_t st);
bool alloc_axial_scratch(trtt_hip_plan *p);

// Run one view of a plan on host FP64 arrays: out_host += A in_host (job 0)
// or A' in_host (job != 0).
int run_dropin(trtt_hip_plan *p, double *out, const double *in, int job) {
  const long nvol = p->nx * p->ny * p->nz;
  const long nsino = p->nu * p->nv;
  const long nin = job == 0 ? nvol : nsino, nout = job == 0 ? nsino : nvol;
  const bool f64 = p->precision == TRTT_HIP_F64;
  const size_t esz = f64 ? 8 : 4;
  if (g_scratch.device != p->device) {
    for (int s = 0; s < 4; ++s) {
      g_scratch.buf[s] = nullptr;
      g_scratch.cap[s] = 0;
    }
    g_scratch.device = p->device;
  }
  if (!g_scratch.ensure(0, sizeof(double) * std::max(nin, nout)) ||
      !g_scratch.ensure(1, esz * nin) || !g_scratch.ensure(2, esz * nout))
    return TRTT_FAILURE;
  double *d_stage = static_cast<double *>(g_scratch.buf[0]);
  void *d_in = g_scratch.buf[1], *d_out = g_scratch.buf[2];
  hipStream_t st = nullptr;
  if (f64) {
    if (!hip_ok(hipMemcpy(d_in, in, sizeof(double) * nin, hipMemcpyHostToDevice), "hipMemcpy"))
      return TRTT_FAILURE;
  } else {
    if (!hip_ok(hipMemcpy(d_stage, in, sizeof(double) * nin, hipMemcpyHostToDevice), "hipMemcpy"))
      return TRTT_FAILURE;
    hipLaunchKernelGGL(cvt_d2f, dim3(grid_for(nin)), dim3(256), 0, st, d_stage,
                       static_cast<float *>(d_in), nin);
  }
  if (!hip_ok(hipMemsetAsync(d_out, 0, esz * nout, st), "hipMemset")) return TRTT_FAILURE;
  if (!hip_ok(hipMemsetAsync(p->d_err, 0, sizeof(int), st), "hipMemset")) return TRTT_FAILURE;
  if (!hip_ok(launch_plan(p, job, d_in, d_out, st), "kernel launch")) return TRTT_FAILURE;
  std::vector<double> res(nout);
  if (f64) {
    if (!hip_ok(hipMemcpy(res.data(), d_out, sizeof(double) * nout, hipMemcpyDeviceToHost),
                "hipMemcpy"))
      return TRTT_FAILURE;
  } else {
    hipLaunchKernelGGL(cvt_f2d, dim3(grid_for(nout)), dim3(256), 0, st,
                       static_cast<const float *>(d_out), d_stage, nout);
    if (!hip_ok(hipMemcpy(res.data(), d_stage, sizeof(double) * nout, hipMemcpyDeviceToHost),
                "hipMemcpy"))
      return TRTT_FAILURE;
  }
  int err = 0;
  if (!hip_ok(hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy"))
    return TRTT_FAILURE;
  if (err) return fail("IT WAS SUPPOSED TO BE HIGHLY IMPOSSIBLE: a voxel lies in the source plane");
  for (long i = 0; i < nout; ++i) out[i] += res[i];
  return TRTT_SUCCESS;
}

bool bad_sizes(long a, long b, long c, long d, long e) {
  return a <= 0 || b <= 0 || c <= 0 || d <= 0 || e <= 0 || a > 1L << 30 || b > 1L << 30 ||
         c > 1L << 30 || d > 1L << 30 || e > 1L << 30;
}

trtt_hip_plan *make_plan(int kind, int precision, long nx, long ny, long nz, long nt, double s_scl,
                         int s_deg, long nu, long nv, double u_scl, double v_scl,
                         const double *xos, const double *xsd, long nviews, const double *cum,
                         long m, double step, double coord0, const int *t0, const int *t1,
                         const double *w0, const double *w1) {
  if (kind < 0 || kind > 3) {
    fail("unknown projector kind");
    return nullptr;
  }
  const bool is2d = kind == trtt::PB2D || kind == trtt::FB2D;
  if (is2d) {
    nz = 1;
    nu = 1;
  }
  if (bad_sizes(nx, ny, nz, nu, nv) || nviews <= 0 || nt <= 0 || !xos || !xsd ||
      (cum && m < 2)) {
    fail("invalid plan arguments");
    return nullptr;
  }
  if (nx * ny * nz * nt > (1L << 40) || nu * nv * nviews > (1L << 40)) {
    fail("plan too large");
    return nullptr;
  }
  auto *p = new trtt_hip_plan();
  p->kind = kind;
  p->precision = precision == TRTT_HIP_F64 ? TRTT_HIP_F64 : TRTT_HIP_F32;
  p->device = g_device;
  p->table = cum != nullptr;
  p->quirk_last_bin = is2d && !p->table;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->nt = nt;
  p->nu = nu;
  p->nv = nv;
  p->nviews = nviews;
  p->s_scl = s_scl;
  p->s_deg = s_deg;
  p->u_scl = u_scl;
  p->v_scl = v_scl;
  p->m = cum ? m : 0;
  p->step = cum ? step : 1.0;
  p->coord0 = cum ? coord0 : 0.0;
  p->h_views.resize(nviews);
  for (long k = 0; k < nviews; ++k) {
    ViewConst &g = p->h_views[k];
    if (is2d)
      view_consts_2d(g, xos + 12 * k, xsd + 12 * k, kind == trtt::FB2D, p->table, p->step, s_scl,
                     s_deg, v_scl);
    else
      view_consts_3d(g, xos + 12 * k, xsd + 12 * k, kind == trtt::CB3D, s_scl, s_deg, u_scl,
                     v_scl);
    g.t0 = t0 ? t0[k] : 0;
    g.t1 = t1 ? t1[k] : g.t0;
    g.w0 = w0 ? w0[k] : 1.0;
    g.w1 = w1 ? w1[k] : 0.0;
    if (g.t0 < 0 || g.t0 >= nt || g.t1 < 0 || g.t1 >= nt) {
      delete p;
      fail("time index out of range");
      return nullptr;
    }
  }
  if (!finalize_plan(p, cum)) {
    release(p);
    return nullptr;
  }
  p->table_deg = p->table ? table_cdf_degree(cum, m, step, coord0, s_deg) : 0;
  p->dev.table_deg = p->table_deg;
  axial_analysis(p);
  if (p->axial_w > 0 && !alloc_axial_scratch(p)) {
    release(p);
    return nullptr;
  }
  return p;
}

// k-fastest copy of the volume series read by the axial forward.
bool alloc_volk(trtt_hip_plan *p) {
  if (!p->d_volk) {
    const size_t vbytes = size_t(p->nt) * p->nx * p->ny * p->nz * sizeof(float);
    if (!hip_ok(hipMalloc(&p->d_volk, vbytes), "hipMalloc scratch")) return false;
  }
  return true;
}

// Scratch of the axial forward: none when the windows are flushed with
// atomics; else windows of `chunk` views at a time, at most 4 GiB (views
// are processed in chunks that fit).
bool alloc_axial_scratch(trtt_hip_plan *p) {
  hipFree(p->d_scratch);
  hipFree(p->d_hdr);
  p->d_scratch = p->d_hdr = nullptr;
  if (p->tune.fwd_flush) {
    // windows go straight to the projections: no scratch, one round
    p->axial_chunk = int(p->nviews);
    p->axial_nbuf = 1;
    return alloc_volk(p);
  }
  const size_t per_view = trtt::axial_scratch_bytes(p->dev, 1, p->tune);
  const size_t budget = size_t(4) << 30;
  long chunk = long(budget / per_view);
  if (p->tune.fwd_chunk_views > 0) chunk = std::min(chunk, long(p->tune.fwd_chunk_views));
  chunk = std::max(1L, std::min(chunk, p->nviews));
  p->axial_chunk = int(chunk);
  const size_t hdr_bytes = trtt::axial_hdr_bytes(p->dev, int(chunk), p->tune);
  const size_t win_bytes = per_view * size_t(chunk) - hdr_bytes;
  // two buffers when the views take several chunks: the reduction of one
  // chunk overlaps the projection of the next on a second stream
  p->axial_nbuf = chunk < p->nviews ? 2 : 1;
  const size_t nb = size_t(p->axial_nbuf);
  if (!hip_ok(hipMalloc(&p->d_scratch, nb * win_bytes), "hipMalloc scratch")) return false;
  if (!hip_ok(hipMalloc(&p->d_hdr, nb * hdr_bytes), "hipMalloc scratch")) return false;
  if (p->axial_nbuf == 2 && !p->overlap.aux) {
    if (!hip_ok(hipStreamCreateWithFlags(&p->overlap.aux, hipStreamNonBlocking), "hipStreamCreate"))
      return false;
    for (int b = 0; b < 2; ++b)
      if (!hip_ok(hipEventCreateWithFlags(&p->overlap.fwd_done[b], hipEventDisableTiming),
                  "hipEventCreate") ||
          !hip_ok(hipEventCreateWithFlags(&p->overlap.red_done[b], hipEventDisableTiming),
                  "hipEventCreate"))
        return false;
  }
  return alloc_volk(p);
}

// One launch of a plan (forward job 0 / backward job 1).
hipError_t launch_plan(trtt_hip_plan *p, int job, const void *in, void *out, hipStream_t st) {
  const int nrows = int(p->h_seg_begin.size()) - 1;
  const bool f64 = p->precision == TRTT_HIP_F64;
  if (p->axial_w > 0) {
    const long ntiles = trtt::axial_tiles(p->dev, job, p->tune);
    long vpb = (ntiles * p->nviews) / kTargetBlocks;
    vpb = std::max(1L, std::min(vpb, 8L));
    if (job == 0 && p->tune.fwd_views_per_block > 0) vpb = p->tune.fwd_views_per_block;
    hipEvent_t *ev = nullptr;
    if (p->timing) {
      std::vector<hipEvent_t> &v = job == 0 ? p->ev_fwd : p->ev_bwd;
      const size_t need = job == 0 ? 2 * size_t((p->nviews + p->axial_chunk - 1) / p->axial_chunk)
                                   : 2;
      while (v.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return hipErrorOutOfMemory;
        v.push_back(e);
      }
      ev = v.data();
      (job == 0 ? p->fwd_launches : p->bwd_launches) = int(need / 2);
    }
    return trtt::launch_axial(p->kind == trtt::CB3D, p->axial_w, f64, p->dev, job, in, out,
                              p->d_seg_begin, nrows, int(vpb), nrows > 1 ? 1 : 0, p->axial_pmax,
                              p->axial_pmax2, p->d_scratch, p->d_hdr, p->axial_chunk, p->d_volk,
                              p->axial_plain, p->tune, ev, &p->overlap, p->axial_nbuf,
                              p->two_phase_segs ? 1 : 0, st);
  }
  if (job == 0)
    return trtt::launch_projector(p->kind, p->table, p->quirk_last_bin, f64, p->dev, 0, in, out,
                                  p->d_seg_begin, nrows, views_per_block(p), 0, st);
  return trtt::launch_projector(p->kind, p->table, p->quirk_last_bin, f64, p->dev, 1, in, out,
                                p->d_seg_begin, nrows, int(p->nviews), nrows > 1 ? 1 : 0, st);
}

}  // namespace

// ===========================================================================
// Part 2 -- plans
// ===========================================================================
extern "C" {

int trtt_hip_set_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return fail("invalid device");
  g_device = device;
  return TRTT_SUCCESS;
}

int trtt_hip_set_precision(int precision) {
  if (precision != TRTT_HIP_F32 && precision != TRTT_HIP_F64) return fail("invalid precision");
  g_precision = precision;
  return TRTT_SUCCESS;
}

const char *trtt_hip_last_error(void) { return g_last_error.c_str(); }

trtt_hip_plan_t *trtt_hip_plan_create(int kind, int precision, long nx, long ny, long nz, long nt,
                                      double s_scl, int s_deg, long nu, long nv, double u_scl,
                                      double v_scl, const double *xos, const double *xsd,
                                      long nviews, const double *cum, long size_footprint,
                                      double step_footprint, double coord0, const int *t0,
                                      const int *t1, const double *w0, const double *w1) {
  g_last_error.clear();
  return make_plan(kind, precision, nx, ny, nz, nt, s_scl, s_deg, nu, nv, u_scl, v_scl, xos, xsd,
                   nviews, cum, size_footprint, step_footprint, coord0, t0, t1, w0, w1);
}

trtt_hip_plan_t *trtt_hip_plan_create_lf2d(int precision, const lf2d_parameters_t *lp,
                                           const double *beta, const double *offset,
                                           long nviews) {
  g_last_error.clear();
  if (!lp || !beta || nviews <= 0 ||
      bad_sizes(lp->objectDimension1, lp->objectDimension2, 1, 1, lp->detectorDimension)) {
    fail("invalid Long & Fessler plan arguments");
    return nullptr;
  }
  auto *p = new trtt_hip_plan();
  p->kind = trtt::LF2D;
  p->precision = precision == TRTT_HIP_F64 ? TRTT_HIP_F64 : TRTT_HIP_F32;
  p->device = g_device;
  p->nx = lp->objectDimension1;
  p->ny = lp->objectDimension2;
  p->nz = 1;
  p->nu = 1;
  p->nv = lp->detectorDimension;
  p->nviews = nviews;
  p->h_views.resize(nviews);
  for (long k = 0; k < nviews; ++k)
    view_consts_lf(p->h_views[k], lp, beta[k], offset ? offset[k] : 0.0);
  if (!finalize_plan(p, nullptr)) {
    release(p);
    return nullptr;
  }
  trtt::DevPlan &P = p->dev;
  {
#pragma clang fp contract(off)
    P.lf_vs = lp->voxelSize;
    P.lf_ps = lp->pixelSize;
    P.lf_R = lp->sourceDistance;
    P.lf_scale = (lp->sourceDistance + lp->detectorDistance) / lp->pixelStep;
    P.lf_size = lp->pixelSize / lp->pixelStep;
  }
  return p;
}

int trtt_hip_forward(trtt_hip_plan_t *p, const void *vol, void *sino, void *stream) {
  if (!p || !vol || !sino) return fail("null argument");
  return hip_ok(launch_plan(p, 0, vol, sino, static_cast<hipStream_t>(stream)), "forward launch")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_backward(trtt_hip_plan_t *p, const void *sino, void *vol, void *stream) {
  if (!p || !vol || !sino) return fail("null argument");
  return hip_ok(launch_plan(p, 1, sino, vol, static_cast<hipStream_t>(stream)),
                "backward launch")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_weighted_residual(trtt_hip_plan_t *p, void *sino, const void *y, const void *w,
                               double *fx_dev, void *stream) {
  if (!p || !sino || !y || !fx_dev) return fail("null argument");
  return hip_ok(trtt::launch_residual(p->precision == TRTT_HIP_F64, sino, y, w,
                                      p->nu * p->nv * p->nviews, fx_dev,
                                      static_cast<hipStream_t>(stream)),
                "residual launch")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_plan_error_flag(trtt_hip_plan_t *p) {
  if (!p) return -1;
  int err = 0;
  if (hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return err;
}

int trtt_hip_plan_path(const trtt_hip_plan_t *p) { return p ? p->axial_w : -1; }

int trtt_hip_plan_set_timing(trtt_hip_plan_t *p, int on) {
  if (!p) return fail("null argument");
  p->timing = on != 0;
  p->fwd_launches = p->bwd_launches = 0;
  return TRTT_SUCCESS;
}

int trtt_hip_plan_kernel_time(trtt_hip_plan_t *p, int job, double *ms, int *launches) {
  if (!p || !ms) return fail("null argument");
  const std::vector<hipEvent_t> &v = job == 0 ? p->ev_fwd : p->ev_bwd;
  const int n = job == 0 ? p->fwd_launches : p->bwd_launches;
  if (!p->timing || n == 0) return fail("no timed launch (enable timing on an axial plan)");
  double total = 0.0;
  for (int k = 0; k < n; ++k) {
    float t = 0.0f;
    if (!hip_ok(hipEventSynchronize(v[2 * k + 1]), "hipEventSynchronize") ||
        !hip_ok(hipEventElapsedTime(&t, v[2 * k], v[2 * k + 1]), "hipEventElapsedTime"))
      return TRTT_FAILURE;
    total += t;
  }
  *ms = total;
  if (launches) *launches = n;
  return TRTT_SUCCESS;
}

int trtt_hip_plan_get_tuning(const trtt_hip_plan_t *p, trtt_hip_tuning_t *t) {
  if (!p || !t) return fail("null argument");
  t->fwd_tile = p->tune.fwd_tile;
  t->bwd_tile = p->tune.bwd_tile;
  t->fwd_window_bytes = p->tune.fwd_window_bytes;
  t->fwd_views_per_block = p->tune.fwd_views_per_block;
  t->bwd_window_bytes = p->tune.bwd_window_bytes;
  t->fwd_chunk_views = p->tune.fwd_chunk_views;
  t->fwd_flush = p->tune.fwd_flush;
  t->fwd_xcds = p->tune.fwd_xcds;
  return TRTT_SUCCESS;
}

int trtt_hip_plan_set_tuning(trtt_hip_plan_t *p, const trtt_hip_tuning_t *t) {
  if (!p || !t) return fail("null argument");
  if (t->fwd_tile < 0 || t->fwd_tile >= trtt::kNumFwdTiles || t->bwd_tile < 0 ||
      t->bwd_tile >= trtt::kNumBwdTiles)
    return fail("tile shape index out of range");
  if (t->fwd_window_bytes < 4096 || t->fwd_window_bytes > 65536 || t->fwd_window_bytes % 256)
    return fail("forward window must be 4096..65536 bytes, a multiple of 256");
  if (t->fwd_views_per_block < 0 || t->fwd_views_per_block > 64)
    return fail("views per block must be 0..64");
  if (t->bwd_window_bytes < 4096 || t->bwd_window_bytes > 131072 || t->bwd_window_bytes % 256)
    return fail("backward window must be 4096..131072 bytes, a multiple of 256");
  if (t->fwd_chunk_views < 0) return fail("views per forward chunk must be >= 0");
  if (t->fwd_flush != 0 && t->fwd_flush != 1) return fail("forward flush mode must be 0 or 1");
  if (t->fwd_xcds < 0 || t->fwd_xcds > 64) return fail("forward XCD count must be 0..64");
  if (p->axial_w <= 0) return TRTT_SUCCESS;  // general kernels: nothing to tune
  if (trtt::bwd_tile_vpl(t->bwd_tile) == 2 && p->axial_pmax2 == 0)
    return fail("backward tiles of 128 slices need column profiles of at most 256 pixels");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!hip_ok(hipSetDevice(p->device), "hipSetDevice")) return TRTT_FAILURE;
  const trtt::AxialTuning old = p->tune;
  p->tune.fwd_tile = t->fwd_tile;
  p->tune.bwd_tile = t->bwd_tile;
  p->tune.fwd_window_bytes = t->fwd_window_bytes;
  p->tune.fwd_views_per_block = t->fwd_views_per_block;
  p->tune.bwd_window_bytes = t->bwd_window_bytes;
  p->tune.fwd_chunk_views = t->fwd_chunk_views;
  p->tune.fwd_flush = t->fwd_flush;
  p->tune.fwd_xcds = t->fwd_xcds;
  if (old.fwd_tile != p->tune.fwd_tile || old.fwd_window_bytes != p->tune.fwd_window_bytes ||
      old.fwd_chunk_views != p->tune.fwd_chunk_views || old.fwd_flush != p->tune.fwd_flush) {
    // hipFree waits for the device: no launch of this plan is in flight
    if (!alloc_axial_scratch(p)) {
      p->tune = old;
      alloc_axial_scratch(p);
      return fail("scratch allocation for the new tuning failed");
    }
  }
  return TRTT_SUCCESS;
}

long trtt_hip_plan_nviews(const trtt_hip_plan_t *p) { return p ? p->nviews : 0; }
long trtt_hip_plan_volume_size(const trtt_hip_plan_t *p) {
  return p ? p->nt * p->nx * p->ny * p->nz : 0;
}
long trtt_hip_plan_sino_size(const trtt_hip_plan_t *p) {
  return p ? p->nviews * p->nu * p->nv : 0;
}
int trtt_hip_plan_destroy(trtt_hip_plan_t *p) {
  release(p);
  return TRTT_SUCCESS;
}

}  // extern "C"

// ===========================================================================
// Part 2b -- sparse system matrix of a 2-D plan
// ===========================================================================
struct trtt_hip_sparse {
  int device = 0;
  bool f64 = false;
  long nnz = 0;
  int nrows = 0, ncols = 0;
  void *coo_val = nullptr;  // reference order: view, pixel, bin
  int *coo_row = nullptr, *coo_col = nullptr;
  int *csr_ptr = nullptr, *csr_col = nullptr;  // rows (view, bin)
  void *csr_val = nullptr;
  int *csc_ptr = nullptr, *csc_row = nullptr;  // columns (pixels)
  void *csc_val = nullptr;
};

namespace {

// Device buffer freed at scope exit.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  void reset() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  bool ensure(size_t n) {
    if (n <= bytes && p) return true;
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
    if (!hip_ok(hipMalloc(&p, std::max<size_t>(n, 256)), "hipMalloc")) return false;
    bytes = std::max<size_t>(n, 256);
    return true;
  }
  template <typename U>
  U *as() const {
    return static_cast<U *>(p);
  }
};

void sparse_release(trtt_hip_sparse *m) {
  if (!m) return;
  hipSetDevice(m->device);
  for (void *q : {m->coo_val, static_cast<void *>(m->coo_row), static_cast<void *>(m->coo_col),
                  static_cast<void *>(m->csr_ptr), static_cast<void *>(m->csr_col), m->csr_val,
                  static_cast<void *>(m->csc_ptr), static_cast<void *>(m->csc_row), m->csc_val})
    if (q) hipFree(q);
  delete m;
}

int key_bits(long n) {
  int b = 1;
  while (b < 31 && (1L << b) < n) ++b;
  return b;
}

template <typename U>
bool dev_alloc(U *&p, size_t count) {
  void *q = nullptr;
  if (!hip_ok(hipMalloc(&q, std::max<size_t>(count * sizeof(U), 256)), "hipMalloc")) return false;
  p = static_cast<U *>(q);
  return true;
}

// Stable sort of the entries by `key` and the compressed form over n_major
// keys: other index and value of every entry in key order, and pointers.
bool compress_by(trtt_hip_sparse *m, const int *key, const int *other, int n_major, int *&ptr,
                 int *&idx, void *&val, DevBuf &temp, hipStream_t st) {
  const size_t esz = m->f64 ? 8 : 4;
  const long n = m->nnz;
  if (!dev_alloc(ptr, size_t(n_major) + 1) || !dev_alloc(idx, size_t(n))) return false;
  if (!hip_ok(hipMalloc(&val, std::max<size_t>(esz * n, 256)), "hipMalloc")) return false;
  DevBuf keys, ids, perm;
  if (n > 0) {
    if (!keys.ensure(4 * n) || !ids.ensure(4 * n) || !perm.ensure(4 * n)) return false;
    const unsigned *k = reinterpret_cast<const unsigned *>(key);
    size_t tb = 0;
    if (!hip_ok(trtt::sparse_sort(k, keys.as<unsigned>(), ids.as<unsigned>(), perm.as<unsigned>(),
                                  n, key_bits(n_major), nullptr, tb, st),
                "radix sort"))
      return false;
    if (!temp.ensure(tb)) return false;
    tb = temp.bytes;
    if (!hip_ok(trtt::sparse_sort(k, keys.as<unsigned>(), ids.as<unsigned>(), perm.as<unsigned>(),
                                  n, key_bits(n_major), temp.p, tb, st),
                "radix sort"))
      return false;
  }
  return hip_ok(trtt::sparse_compress(m->f64, keys.as<unsigned>(), perm.as<unsigned>(),
                                      m->coo_val, other, val, idx, ptr, n, n_major, st),
                "sparse compress") &&
         hip_ok(hipStreamSynchronize(st), "sparse assembly");
}

}  // namespace

extern "C" {

trtt_hip_sparse_t *trtt_hip_sparse_create(trtt_hip_plan_t *p, int decimals) {
  if (!p) {
    fail("null argument");
    return nullptr;
  }
  if (p->kind != trtt::PB2D && p->kind != trtt::FB2D && p->kind != trtt::LF2D) {
    fail("sparse matrices are built for 2-D plans (PB2D, FB2D, LF2D)");
    return nullptr;
  }
  if (p->nt != 1) {
    fail("sparse matrices are built for plans of one phase");
    return nullptr;
  }
  for (const ViewConst &g : p->h_views)
    if (g.w0 != 1.0 || g.w1 != 0.0) {
      fail("sparse matrices are built for plans without time interpolation weights");
      return nullptr;
    }
  if (decimals != 6 && decimals != -1) {
    fail("decimals must be 6 (coefficients as the reference's text round trip leaves them) or -1");
    return nullptr;
  }
  const long npix = p->nx * p->ny, n = npix * p->nviews;
  if (p->nviews > 65535 || n >= long(INT_MAX) || p->nviews * p->nv >= long(INT_MAX)) {
    fail("sparse matrix too large (views <= 65535, views x pixels and views x bins < 2^31)");
    return nullptr;
  }
  if (!hip_ok(hipSetDevice(p->device), "hipSetDevice")) return nullptr;
  hipStream_t st = nullptr;
  const bool f64 = p->precision == TRTT_HIP_F64;
  const bool tab = p->table, quirk = p->quirk_last_bin;
  DevBuf count, offset, temp;
  if (!count.ensure(4 * size_t(n)) || !offset.ensure(8 * size_t(n))) return nullptr;
  if (!hip_ok(hipMemsetAsync(p->d_err, 0, sizeof(int), st), "hipMemset") ||
      !hip_ok(trtt::launch_sparse_columns(p->kind, tab, quirk, f64, p->dev, false, count.as<int>(),
                                          nullptr, nullptr, nullptr, nullptr, 0, st),
              "sparse count"))
    return nullptr;
  size_t tb = 0;
  if (!hip_ok(trtt::sparse_scan(count.as<int>(), offset.as<long long>(), n, nullptr, tb, st),
              "scan") ||
      !temp.ensure(tb))
    return nullptr;
  tb = temp.bytes;
  if (!hip_ok(trtt::sparse_scan(count.as<int>(), offset.as<long long>(), n, temp.p, tb, st),
              "scan"))
    return nullptr;
  long long last_off = 0;
  int last_cnt = 0, err = 0;
  if (!hip_ok(hipMemcpy(&last_off, offset.as<long long>() + (n - 1), 8, hipMemcpyDeviceToHost),
              "hipMemcpy") ||
      !hip_ok(hipMemcpy(&last_cnt, count.as<int>() + (n - 1), 4, hipMemcpyDeviceToHost),
              "hipMemcpy") ||
      !hip_ok(hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy"))
    return nullptr;
  if (err) {
    fail("IT WAS SUPPOSED TO BE HIGHLY IMPOSSIBLE: a pixel lies in the source plane");
    return nullptr;
  }
  const long long nnz = last_off + last_cnt;
  if (nnz > static_cast<long long>(INT_MAX) - 512) {
    fail("sparse matrix has too many entries (at most 2^31 - 512)");
    return nullptr;
  }
  trtt_hip_sparse *m = new trtt_hip_sparse;
  m->device = p->device;
  m->f64 = f64;
  m->nnz = long(nnz);
  m->nrows = int(p->nviews * p->nv);
  m->ncols = int(npix);
  const size_t esz = f64 ? 8 : 4;
  if (!hip_ok(hipMalloc(&m->coo_val, std::max<size_t>(esz * m->nnz, 256)), "hipMalloc") ||
      !dev_alloc(m->coo_row, size_t(m->nnz)) || !dev_alloc(m->coo_col, size_t(m->nnz)) ||
      !hip_ok(trtt::launch_sparse_columns(p->kind, tab, quirk, f64, p->dev, true, nullptr,
                                          offset.as<long long>(), m->coo_val, m->coo_row,
                                          m->coo_col, decimals == 6 ? 1 : 0, st),
              "sparse fill") ||
      !hip_ok(hipStreamSynchronize(st), "sparse fill")) {
    sparse_release(m);
    return nullptr;
  }
  count.reset();
  if (!compress_by(m, m->coo_row, m->coo_col, m->nrows, m->csr_ptr, m->csr_col, m->csr_val, temp,
                   st) ||
      !compress_by(m, m->coo_col, m->coo_row, m->ncols, m->csc_ptr, m->csc_row, m->csc_val, temp,
                   st)) {
    sparse_release(m);
    return nullptr;
  }
  return m;
}

long trtt_hip_sparse_nnz(const trtt_hip_sparse_t *m) { return m ? m->nnz : -1; }

int trtt_hip_sparse_coo(const trtt_hip_sparse_t *m, double *coefs, long *rows, long *cols) {
  if (!m || !coefs || !rows || !cols) return fail("null argument");
  if (!hip_ok(hipSetDevice(m->device), "hipSetDevice")) return TRTT_FAILURE;
  const long n = m->nnz;
  std::vector<int> r(n), c(n);
  if (m->f64) {
    if (!hip_ok(hipMemcpy(coefs, m->coo_val, 8 * n, hipMemcpyDeviceToHost), "hipMemcpy"))
      return TRTT_FAILURE;
  } else {
    std::vector<float> v(n);
    if (!hip_ok(hipMemcpy(v.data(), m->coo_val, 4 * n, hipMemcpyDeviceToHost), "hipMemcpy"))
      return TRTT_FAILURE;
    for (long e = 0; e < n; ++e) coefs[e] = v[e];
  }
  if (!hip_ok(hipMemcpy(r.data(), m->coo_row, 4 * n, hipMemcpyDeviceToHost), "hipMemcpy") ||
      !hip_ok(hipMemcpy(c.data(), m->coo_col, 4 * n, hipMemcpyDeviceToHost), "hipMemcpy"))
    return TRTT_FAILURE;
  for (long e = 0; e < n; ++e) {
    rows[e] = long(r[e]) + 1;
    cols[e] = long(c[e]) + 1;
  }
  return TRTT_SUCCESS;
}

int trtt_hip_sparse_forward(trtt_hip_sparse_t *m, const void *img, void *sino, void *stream) {
  if (!m || !img || !sino) return fail("null argument");
  return hip_ok(trtt::sparse_spmv(m->f64, m->csr_ptr, m->csr_col, m->csr_val, img, sino,
                                  m->nrows, static_cast<hipStream_t>(stream)),
                "sparse forward")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_sparse_backward(trtt_hip_sparse_t *m, const void *sino, void *img, void *stream) {
  if (!m || !img || !sino) return fail("null argument");
  return hip_ok(trtt::sparse_spmv(m->f64, m->csc_ptr, m->csc_row, m->csc_val, sino, img,
                                  m->ncols, static_cast<hipStream_t>(stream)),
                "sparse backward")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_sparse_destroy(trtt_hip_sparse_t *m) {
  sparse_release(m);
  return TRTT_SUCCESS;
}

}  // extern "C"

// ===========================================================================
// Part 2c -- temporal interpolation of frame series
// ===========================================================================
struct trtt_hip_tinterp {
  int device = 0;
  bool f64 = false;
  long npix = 0;
  int nt = 0, nviews = 0, taps = 0;
  int *idx = nullptr;   // [nviews][taps] phase of each tap
  void *w = nullptr;    // [nviews][taps] weights (plan precision)
  int *ptr = nullptr;   // [nt + 1] per-phase list of (view, weight)
  int *src = nullptr;
  void *wb = nullptr;
};

namespace {

void tinterp_release(trtt_hip_tinterp *h) {
  if (!h) return;
  hipSetDevice(h->device);
  for (void *q : {static_cast<void *>(h->idx), h->w, static_cast<void *>(h->ptr),
                  static_cast<void *>(h->src), h->wb})
    if (q) hipFree(q);
  delete h;
}

// Weights in the plan precision, uploaded.
bool upload_weights(const std::vector<double> &w, bool f64, void *&dst) {
  const size_t n = std::max<size_t>(w.size(), 1);
  if (!hip_ok(hipMalloc(&dst, n * (f64 ? 8 : 4)), "hipMalloc")) return false;
  if (w.empty()) return true;
  if (f64) return hip_ok(hipMemcpy(dst, w.data(), 8 * w.size(), hipMemcpyHostToDevice), "hipMemcpy");
  std::vector<float> f(w.begin(), w.end());
  return hip_ok(hipMemcpy(dst, f.data(), 4 * f.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

template <typename U>
bool upload_vec(const std::vector<U> &v, U *&dst) {
  if (!dev_alloc(dst, std::max<size_t>(v.size(), 1))) return false;
  return v.empty() ||
         hip_ok(hipMemcpy(dst, v.data(), sizeof(U) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // namespace

extern "C" {

trtt_hip_tinterp_t *trtt_hip_tinterp_create(int precision, long npix, long nt, long nviews,
                                            int taps, const int *idx, const double *w) {
  if (!idx || !w) {
    fail("null argument");
    return nullptr;
  }
  if ((precision != TRTT_HIP_F32 && precision != TRTT_HIP_F64) || npix <= 0 || nt <= 0 ||
      nt > 65535 || nviews <= 0 || nviews > 65535 || taps <= 0 || taps > 64 ||
      nviews * taps >= long(INT_MAX)) {
    fail("invalid temporal interpolation sizes (frames and phases <= 65535, 1..64 taps)");
    return nullptr;
  }
  const long n = nviews * taps;
  for (long e = 0; e < n; ++e)
    if (idx[e] < 0 || idx[e] >= nt) {
      fail("temporal interpolation phase index out of range");
      return nullptr;
    }
  // per-phase lists in (view, tap) order
  std::vector<int> ptr(nt + 1, 0), src(n);
  std::vector<double> wb(n);
  for (long e = 0; e < n; ++e) ++ptr[idx[e] + 1];
  for (long t = 0; t < nt; ++t) ptr[t + 1] += ptr[t];
  std::vector<int> fillp(ptr.begin(), ptr.end() - 1);
  for (long e = 0; e < n; ++e) {
    const int slot = fillp[idx[e]]++;
    src[slot] = int(e / taps);
    wb[slot] = w[e];
  }
  if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return nullptr;
  trtt_hip_tinterp *h = new trtt_hip_tinterp;
  h->device = g_device;
  h->f64 = precision == TRTT_HIP_F64;
  h->npix = npix;
  h->nt = int(nt);
  h->nviews = int(nviews);
  h->taps = taps;
  const std::vector<int> iv(idx, idx + n);
  const std::vector<double> wv(w, w + n);
  if (!upload_vec(iv, h->idx) || !upload_weights(wv, h->f64, h->w) ||
      !upload_vec(ptr, h->ptr) || !upload_vec(src, h->src) ||
      !upload_weights(wb, h->f64, h->wb)) {
    tinterp_release(h);
    return nullptr;
  }
  return h;
}

int trtt_hip_tinterp_apply(trtt_hip_tinterp_t *h, int job, const void *in, void *out,
                           void *stream) {
  if (!h || !in || !out) return fail("null argument");
  if (in == out) return fail("in and out must differ");
  const int frames = job == 0 ? h->nviews : h->nt;
  return hip_ok(trtt::launch_tinterp(h->f64, job == 0 ? 0 : 1, in, out, h->npix, frames, h->taps,
                                     h->idx, h->w, h->ptr, h->src, h->wb,
                                     static_cast<hipStream_t>(stream)),
                "temporal interpolation")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_tinterp_destroy(trtt_hip_tinterp_t *h) {
  tinterp_release(h);
  return TRTT_SUCCESS;
}

}  // extern "C"

extern "C" {

// ===========================================================================
// Part 1 -- reference plugin symbols
// ===========================================================================
// Constants of trtt_DD_PB3D (trtt_3D_projectors.c:1088-1120), in the
// reference's FP64 expression order.
static trtt::DDConst dd_consts(long nx, long ny, long nz, double s, long nu, long nv, double us,
                               double vs, const double *xos, const double *xsd) {
  trtt::DDConst c;
  const double a1inv = -1.0 * (1.0 / xsd[0]);
  const double b1 = xsd[5] + a1inv * xsd[4] * xsd[1], c1 = xsd[9] + a1inv * xsd[8] * xsd[1];
  const double b2 = xsd[6] + a1inv * xsd[4] * xsd[2], c2 = xsd[10] + a1inv * xsd[8] * xsd[2];
  const double b3 = xsd[7] + a1inv * xsd[4] * xsd[3], c3 = xsd[11] + a1inv * xsd[8] * xsd[3];
  c.v_x = b1 * xos[4] + b2 * xos[8];
  c.u_x = c1 * xos[4] + c2 * xos[8];
  c.v_y = b1 * xos[5] + b2 * xos[9];
  c.u_y = c1 * xos[5] + c2 * xos[9];
  c.v_z = b1 * xos[6] + b2 * xos[10];
  c.u_z = c1 * xos[6] + c2 * xos[10];
  c.v_c = b1 * xos[7] + b2 * xos[11] + b3;
  c.u_c = c1 * xos[7] + c2 * xos[11] + c3;
  const double Du = std::fabs(c.u_z);
  const double Dv = std::fabs(c.v_y) > std::fabs(c.v_x) ? std::fabs(c.v_y) : std::fabs(c.v_x);
  c.h = (s * s * s) / (us * vs * Du * Dv);
  c.Du = 0.5 * Du;
  c.Dv = 0.5 * Dv;
  c.us = us;
  c.vs = vs;
  c.nx = nx;
  c.ny = ny;
  c.nz = nz;
  c.nu = nu;
  c.nv = nv;
  return c;
}

// One view of the distance-driven projector on host FP64 arrays: out_host +=
// A in (job 0), the reference's transpose (job != 0, stuck) or the exact
// adjoint (job != 0, !stuck), in the drop-in precision.
static int dd_dropin(double *out, double *in, long nx, long ny, long nz, double s_scl, long nu,
                     long nv, double u_scl, double v_scl, const double *xos, const double *xsd,
                     int job, bool stuck) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!out || !in || !xos || !xsd) return fail("null argument");
  if (bad_sizes(nx, ny, nz, nu, nv) || nx * ny * nz > (1L << 31) || nu * nv > (1L << 31))
    return fail("invalid sizes");
  if (!(s_scl > 0.0) || !(u_scl > 0.0) || !(v_scl > 0.0)) return fail("invalid scales");
  if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return TRTT_FAILURE;
  const trtt::DDConst c = dd_consts(nx, ny, nz, s_scl, nu, nv, u_scl, v_scl, xos, xsd);
  if (!(c.h < HUGE_VAL)) return fail("degenerate projection (zero footprint size)");
  const long nvol = nx * ny * nz, nsino = nu * nv;
  const long nin = job == 0 ? nvol : nsino, nout = job == 0 ? nsino : nvol;
  const bool f64 = g_precision == TRTT_HIP_F64;
  const size_t esz = f64 ? 8 : 4;
  if (g_scratch.device != g_device) {
    for (int q = 0; q < 4; ++q) {
      g_scratch.buf[q] = nullptr;
      g_scratch.cap[q] = 0;
    }
    g_scratch.device = g_device;
  }
  if (!g_scratch.ensure(0, sizeof(double) * std::max(nin, nout)) ||
      !g_scratch.ensure(1, esz * nin) || !g_scratch.ensure(2, esz * nout) ||
      !g_scratch.ensure(3, sizeof(trtt::DDConst)))
    return TRTT_FAILURE;
  double *d_stage = static_cast<double *>(g_scratch.buf[0]);
  void *d_in = g_scratch.buf[1], *d_out = g_scratch.buf[2];
  trtt::DDConst *d_c = static_cast<trtt::DDConst *>(g_scratch.buf[3]);
  hipStream_t st = nullptr;
  if (!hip_ok(hipMemcpy(d_c, &c, sizeof(c), hipMemcpyHostToDevice), "hipMemcpy"))
    return TRTT_FAILURE;
  if (f64) {
    if (!hip_ok(hipMemcpy(d_in, in, sizeof(double) * nin, hipMemcpyHostToDevice), "hipMemcpy"))
      return TRTT_FAILURE;
  } else {
    if (!hip_ok(hipMemcpy(d_stage, in, sizeof(double) * nin, hipMemcpyHostToDevice), "hipMemcpy"))
      return TRTT_FAILURE;
    hipLaunchKernelGGL(cvt_d2f, dim3(grid_for(nin)), dim3(256), 0, st, d_stage,
                       static_cast<float *>(d_in), nin);
  }
  if (!hip_ok(hipMemsetAsync(d_out, 0, esz * nout, st), "hipMemset") ||
      !hip_ok(trtt::launch_dd(f64, d_c, c, 1, job == 0 ? 0 : 1, stuck ? 1 : 0, d_in, d_out, st),
              "kernel launch"))
    return TRTT_FAILURE;
  std::vector<double> res(nout);
  if (f64) {
    if (!hip_ok(hipMemcpy(res.data(), d_out, sizeof(double) * nout, hipMemcpyDeviceToHost),
                "hipMemcpy"))
      return TRTT_FAILURE;
  } else {
    hipLaunchKernelGGL(cvt_f2d, dim3(grid_for(nout)), dim3(256), 0, st,
                       static_cast<const float *>(d_out), d_stage, nout);
    if (!hip_ok(hipMemcpy(res.data(), d_stage, sizeof(double) * nout, hipMemcpyDeviceToHost),
                "hipMemcpy"))
      return TRTT_FAILURE;
  }
  for (long q = 0; q < nout; ++q) out[q] += res[q];
  return TRTT_SUCCESS;
}

int trtt_DD_PB3D(double *out, double *in, long nx, long ny, long nz, double s_scl, long nu,
                 long nv, double u_scl, double v_scl, double *xos, double *xsd, int job) {
  return dd_dropin(out, in, nx, ny, nz, s_scl, nu, nv, u_scl, v_scl, xos, xsd, job, true);
}

int trtt_hip_DD_PB3D_adjoint(double *out, double *in, long nx, long ny, long nz, double s_scl,
                             long nu, long nv, double u_scl, double v_scl, double *xos,
                             double *xsd) {
  return dd_dropin(out, in, nx, ny, nz, s_scl, nu, nv, u_scl, v_scl, xos, xsd, 1, false);
}

}  // extern "C"

// Batched distance-driven projector: per-view constants on the device.
struct trtt_hip_dd {
  int device = 0;
  bool f64 = false;
  int nviews = 0;
  trtt::DDConst c0{};
  trtt::DDConst *d_c = nullptr;
};

extern "C" {

trtt_hip_dd_t *trtt_hip_dd_create(int precision, long nx, long ny, long nz, double s_scl,
                                  long nu, long nv, double u_scl, double v_scl, const double *xos,
                                  const double *xsd, long nviews) {
  if (!xos || !xsd) {
    fail("null argument");
    return nullptr;
  }
  if ((precision != TRTT_HIP_F32 && precision != TRTT_HIP_F64) || bad_sizes(nx, ny, nz, nu, nv) ||
      nx * ny * nz > (1L << 31) || nu * nv > (1L << 31) || nviews <= 0 || nviews > 65535 ||
      !(s_scl > 0.0) || !(u_scl > 0.0) || !(v_scl > 0.0)) {
    fail("invalid distance-driven plan arguments");
    return nullptr;
  }
  std::vector<trtt::DDConst> cs(nviews);
  for (long k = 0; k < nviews; ++k) {
    cs[k] = dd_consts(nx, ny, nz, s_scl, nu, nv, u_scl, v_scl, xos + 12 * k, xsd + 12 * k);
    if (!(cs[k].h < HUGE_VAL)) {
      fail("degenerate projection (zero footprint size)");
      return nullptr;
    }
  }
  if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return nullptr;
  trtt_hip_dd *p = new trtt_hip_dd;
  p->device = g_device;
  p->f64 = precision == TRTT_HIP_F64;
  p->nviews = int(nviews);
  p->c0 = cs[0];
  if (!upload_vec(cs, p->d_c)) {
    delete p;
    return nullptr;
  }
  return p;
}

int trtt_hip_dd_forward(trtt_hip_dd_t *p, const void *vol, void *sino, void *stream) {
  if (!p || !vol || !sino) return fail("null argument");
  return hip_ok(trtt::launch_dd(p->f64, p->d_c, p->c0, p->nviews, 0, 0, vol, sino,
                                static_cast<hipStream_t>(stream)),
                "distance-driven forward")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_dd_backward(trtt_hip_dd_t *p, const void *sino, void *vol, int exact, void *stream) {
  if (!p || !vol || !sino) return fail("null argument");
  return hip_ok(trtt::launch_dd(p->f64, p->d_c, p->c0, p->nviews, 1, exact ? 0 : 1, sino, vol,
                                static_cast<hipStream_t>(stream)),
                "distance-driven backward")
             ? TRTT_SUCCESS
             : TRTT_FAILURE;
}

int trtt_hip_dd_destroy(trtt_hip_dd_t *p) {
  if (!p) return TRTT_SUCCESS;
  hipSetDevice(p->device);
  hipFree(p->d_c);
  delete p;
  return TRTT_SUCCESS;
}

// One-view plans of the drop-in symbols, cached per geometry family (kind,
// precision, device, sizes, scales, footprint table): the unchanged Yorick
// loop calls the same few families view after view, so a call only
// rewrites the plan's 400-byte view constants instead of building and
// freeing a plan.  Calls are serialised by g_mutex.
namespace {

struct DropinKey {
  int kind, precision, device, s_deg;
  long nx, ny, nz, nu, nv, m;
  double s_scl, u_scl, v_scl, step, coord0;
  unsigned long long table_hash;
  bool operator==(const DropinKey &o) const {
    return kind == o.kind && precision == o.precision && device == o.device &&
           s_deg == o.s_deg && nx == o.nx && ny == o.ny && nz == o.nz && nu == o.nu &&
           nv == o.nv && m == o.m && s_scl == o.s_scl && u_scl == o.u_scl &&
           v_scl == o.v_scl && step == o.step && coord0 == o.coord0 &&
           table_hash == o.table_hash;
  }
};

// FNV-1a over the table's bytes (the key must change when the table does)
unsigned long long hash_table(const double *t, long m) {
  unsigned long long h = 1469598103934665603ULL;
  const unsigned char *b = reinterpret_cast<const unsigned char *>(t);
  for (size_t i = 0, n = size_t(m) * sizeof(double); i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

constexpr size_t kDropinCache = 8;
struct DropinEntry {
  DropinKey key;
  trtt_hip_plan *plan;
  unsigned long long used;
};
std::vector<DropinEntry> g_dropin;
unsigned long long g_dropin_clock = 0;

// The cached plan of `key`, its view constants set from xos / xsd.
trtt_hip_plan *dropin_plan(const DropinKey &key, const double *cum, const double *xos,
                           const double *xsd) {
  for (DropinEntry &e : g_dropin)
    if (e.key == key) {
      trtt_hip_plan *p = e.plan;
      e.used = ++g_dropin_clock;
      ViewConst &g = p->h_views[0];
      const bool is2d = p->kind == trtt::PB2D || p->kind == trtt::FB2D;
      if (is2d)
        view_consts_2d(g, xos, xsd, p->kind == trtt::FB2D, p->table, p->step, p->s_scl,
                       p->s_deg, p->v_scl);
      else
        view_consts_3d(g, xos, xsd, p->kind == trtt::CB3D, p->s_scl, p->s_deg, p->u_scl,
                       p->v_scl);
      g.t0 = g.t1 = 0;
      g.w0 = 1.0;
      g.w1 = 0.0;
      if (!hip_ok(hipMemcpy(p->d_views, &g, sizeof(g), hipMemcpyHostToDevice), "hipMemcpy"))
        return nullptr;
      axial_analysis(p);
      if (p->axial_w > 0 && !p->d_scratch && !alloc_axial_scratch(p)) return nullptr;
      return p;
    }
  trtt_hip_plan *p = make_plan(key.kind, key.precision, key.nx, key.ny, key.nz, 1, key.s_scl,
                               key.s_deg, key.nu, key.nv, key.u_scl, key.v_scl, xos, xsd, 1, cum,
                               key.m, key.step, key.coord0, nullptr, nullptr, nullptr, nullptr);
  if (!p) return nullptr;
  if (g_dropin.size() >= kDropinCache) {
    auto lru = std::min_element(g_dropin.begin(), g_dropin.end(),
                                [](const DropinEntry &a, const DropinEntry &b) {
                                  return a.used < b.used;
                                });
    release(lru->plan);
    g_dropin.erase(lru);
  }
  g_dropin.push_back(DropinEntry{key, p, ++g_dropin_clock});
  return p;
}

}  // namespace

static int dropin(int kind, double *out, double *in, long nx, long ny, long nz, double s_scl,
                  int s_deg, const double *cum, long m, double step, const double *coord,
                  long nu, long nv, double u_scl, double v_scl, const double *xos,
                  const double *xsd, int job) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!out || !in || !xos || !xsd) return fail("null argument");
  if (cum && !coord) return fail("null footprint coordinates");
  if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return TRTT_FAILURE;
  const bool is2d = kind == trtt::PB2D || kind == trtt::FB2D;
  DropinKey key{kind, g_precision, g_device, s_deg, nx, ny, is2d ? 1 : nz, is2d ? 1 : nu, nv,
                cum ? m : 0, s_scl, u_scl, v_scl, cum ? step : 1.0, cum ? coord[0] : 0.0,
                cum && m > 0 ? hash_table(cum, m) : 0ULL};
  trtt_hip_plan *p = dropin_plan(key, cum, xos, xsd);
  if (!p) return TRTT_FAILURE;
  return run_dropin(p, out, in, job);
}

int trtt_PB2D(double *out, double *in, long nx, long ny, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nv, double v_scl, double *xos,
              double *xsd, int job) {
  (void)footprint;
  if (!cum_footprint) return fail("null footprint table");
  return dropin(trtt::PB2D, out, in, nx, ny, 1, s_scl, s_deg, cum_footprint, size_footprint,
                step_footprint, coord_footprint, 1, nv, 1.0, v_scl, xos, xsd, job);
}

int trtt_FB2D(double *out, double *in, long nx, long ny, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nv, double v_scl, double *xos,
              double *xsd, int job) {
  (void)footprint;
  if (!cum_footprint) return fail("null footprint table");
  return dropin(trtt::FB2D, out, in, nx, ny, 1, s_scl, s_deg, cum_footprint, size_footprint,
                step_footprint, coord_footprint, 1, nv, 1.0, v_scl, xos, xsd, job);
}

int trtt_PB2D_cubic(double *out, double *in, long nx, long ny, double s_scl, int s_deg, long nv,
                    double v_scl, double *xos, double *xsd, int job) {
  return dropin(trtt::PB2D, out, in, nx, ny, 1, s_scl, s_deg, nullptr, 0, 1.0, nullptr, 1, nv,
                1.0, v_scl, xos, xsd, job);
}

int trtt_FB2D_cubic(double *out, double *in, long nx, long ny, double s_scl, int s_deg, long nv,
                    double v_scl, double *xos, double *xsd, int job) {
  return dropin(trtt::FB2D, out, in, nx, ny, 1, s_scl, s_deg, nullptr, 0, 1.0, nullptr, 1, nv,
                1.0, v_scl, xos, xsd, job);
}

int trtt_PB3D(double *out, double *in, long nx, long ny, long nz, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nu, long nv, double u_scl,
              double v_scl, double *xos, double *xsd, int job) {
  (void)footprint;
  if (!cum_footprint) return fail("null footprint table");
  return dropin(trtt::PB3D, out, in, nx, ny, nz, s_scl, s_deg, cum_footprint, size_footprint,
                step_footprint, coord_footprint, nu, nv, u_scl, v_scl, xos, xsd, job);
}

int trtt_CB3D(double *out, double *in, long nx, long ny, long nz, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nu, long nv, double u_scl,
              double v_scl, double *xos, double *xsd, int job) {
  (void)footprint;
  if (!cum_footprint) return fail("null footprint table");
  return dropin(trtt::CB3D, out, in, nx, ny, nz, s_scl, s_deg, cum_footprint, size_footprint,
                step_footprint, coord_footprint, nu, nv, u_scl, v_scl, xos, xsd, job);
}

int trtt_PB3D_cubic(double *out, double *in, long nx, long ny, long nz, double s_scl, int s_deg,
                    long nu, long nv, double u_scl, double v_scl, double *xos, double *xsd,
                    int job) {
  return dropin(trtt::PB3D, out, in, nx, ny, nz, s_scl, s_deg, nullptr, 0, 1.0, nullptr, nu, nv,
                u_scl, v_scl, xos, xsd, job);
}

int trtt_CB3D_cubic(double *out, double *in, long nx, long ny, long nz, double s_scl, int s_deg,
                    long nu, long nv, double u_scl, double v_scl, double *xos, double *xsd,
                    int job) {
  return dropin(trtt::CB3D, out, in, nx, ny, nz, s_scl, s_deg, nullptr, 0, 1.0, nullptr, nu, nv,
                u_scl, v_scl, xos, xsd, job);
}

static void lf_dropin(const lf2d_parameters_t *lp, double beta, double offset, const double *src,
                      double *dst, int clr, int job, int line = 0) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!lp || !src || !dst) {
    fail("null argument");
    return;
  }
  const long nimg = lp->objectDimension1 * lp->objectDimension2;
  const long nout = job == 0 ? lp->detectorDimension : nimg;
  if (clr) std::memset(dst, 0, sizeof(double) * nout);
  if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return;
  trtt_hip_plan *p = trtt_hip_plan_create_lf2d(g_precision, lp, &beta, &offset, 1);
  if (!p) return;
  p->dev.lf_line = line;
  run_dropin(p, dst, src, job);
  release(p);
}

void lf2d_direct(const lf2d_parameters_t *p, double beta, double detectorOffset,
                 const double voxel[], double pixel[], int clr) {
  lf_dropin(p, beta, detectorOffset, voxel, pixel, clr, 0);
}

void lf2d_adjoint(const lf2d_parameters_t *p, double beta, double detectorOffset,
                  const double pixel[], double voxel[], int clr) {
  lf_dropin(p, beta, detectorOffset, pixel, voxel, clr, 1);
}

void lf2d_line_integ_direct(const lf2d_parameters_t *p, double beta, double detectorOffset,
                            const double voxel[], double pixel[], int clr) {
  lf_dropin(p, beta, detectorOffset, voxel, pixel, clr, 0, 1);
}

void lf2d_line_integ_adjoint(const lf2d_parameters_t *p, double beta, double detectorOffset,
                             const double pixel[], double voxel[], int clr) {
  lf_dropin(p, beta, detectorOffset, pixel, voxel, clr, 1, 1);
}

// Per-pixel columns (host code: a call evaluates one pixel's footprint on
// one view, a few dozen FP64 operations, with the kernels' own footprint
// functions; the whole matrix is built on the device by
// trtt_hip_sparse_create).
static int sparser_2d(bool fan, double *out, long i, long j, double s_scl, int s_deg,
                      const double *cum, long m, double step, const double *coord, long nv,
                      double v_scl, const double *xos, const double *xsd) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!out || !cum || !coord || !xos || !xsd) return fail("null argument");
  if (nv <= 0 || nv > (1L << 30) || m < 2 || i < INT_MIN || i > INT_MAX || j < INT_MIN ||
      j > INT_MAX)
    return fail("invalid sparser arguments");
  ViewConst g;
  view_consts_2d(g, xos, xsd, fan, true, step, s_scl, s_deg, v_scl);
  trtt::DevPlan P{};
  P.table = cum;
  P.m = m;
  P.step = step;
  P.coord0 = coord[0];
  P.nz = P.nu = P.nviews = P.nt = 1;
  P.nv = int(nv);
  P.u_scl = 1.0;
  P.v_scl = v_scl;
  {
#pragma clang fp contract(off)
    P.rho_init = s_scl * s_scl;
    P.hu_init = s_scl;
    P.hv_init = s_scl / v_scl;
    P.hs0 = 0.5 * (s_deg + 1);
  }
  if (!trtt::host_column_2d(fan, P, g, int(i), int(j), out))
    return fail("IT WAS SUPPOSED TO BE HIGHLY IMPOSSIBLE: a pixel lies in the source plane");
  return TRTT_SUCCESS;
}

int trtt_PB2D_sparser(double *out, long i, long j, long nx, long ny, double s_scl, int s_deg,
                      double *footprint, double *cum_footprint, long size_footprint,
                      double step_footprint, double *coord_footprint, long nv, double v_scl,
                      double *xos, double *xsd) {
  (void)nx;
  (void)ny;
  (void)footprint;
  return sparser_2d(false, out, i, j, s_scl, s_deg, cum_footprint, size_footprint,
                    step_footprint, coord_footprint, nv, v_scl, xos, xsd);
}

int trtt_FB2D_sparser(double *out, long i, long j, long nx, long ny, double s_scl, int s_deg,
                      double *footprint, double *cum_footprint, long size_footprint,
                      double step_footprint, double *coord_footprint, long nv, double v_scl,
                      double *xos, double *xsd) {
  (void)nx;
  (void)ny;
  (void)footprint;
  return sparser_2d(true, out, i, j, s_scl, s_deg, cum_footprint, size_footprint,
                    step_footprint, coord_footprint, nv, v_scl, xos, xsd);
}

static void lf_sparser(const lf2d_parameters_t *lp, double beta, double offset, long i1, long i2,
                       double *pixel, int clr, int line) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!lp || !pixel || lp->detectorDimension <= 0 || lp->detectorDimension > (1L << 30) ||
      i1 < INT_MIN || i1 > INT_MAX || i2 < INT_MIN || i2 > INT_MAX) {
    fail("invalid Long & Fessler sparser arguments");
    return;
  }
  if (clr) std::memset(pixel, 0, sizeof(double) * lp->detectorDimension);
  ViewConst g;
  view_consts_lf(g, lp, beta, offset);
  trtt::DevPlan P{};
  P.nv = int(lp->detectorDimension);
  P.lf_line = line;
  {
#pragma clang fp contract(off)
    P.lf_vs = lp->voxelSize;
    P.lf_ps = lp->pixelSize;
    P.lf_R = lp->sourceDistance;
    P.lf_scale = (lp->sourceDistance + lp->detectorDistance) / lp->pixelStep;
    P.lf_size = lp->pixelSize / lp->pixelStep;
  }
  trtt::host_lf_column(P, g, int(i1), int(i2), pixel);
}

void lf2d_sparser(const lf2d_parameters_t *p, double beta, double detectorOffset, const long i1,
                  const long i2, double pixel[], int clr) {
  lf_sparser(p, beta, detectorOffset, i1, i2, pixel, clr, 0);
}

void lf2d_line_integ_sparser(const lf2d_parameters_t *p, double beta, double detectorOffset,
                             const long i1, const long i2, double pixel[], int clr) {
  lf_sparser(p, beta, detectorOffset, i1, i2, pixel, clr, 1);
}

int trtt_hip_plan_set_lf_line(trtt_hip_plan_t *p, int line) {
  if (!p) return fail("null argument");
  if (p->kind != trtt::LF2D) return fail("not a Long & Fessler plan");
  p->dev.lf_line = line != 0;
  return TRTT_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free drop-in (no reference counterpart).  Callers that cannot
// allocate device memory (code under HIP graph capture set-up, serving loops,
// harnesses that forbid hipMalloc) get the per-view footprint projectors of
// the drop-in symbols with every device buffer (view constants, table,
// staging, output) carved out of one static device array per GPU.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
// Device address of `scratch` on each GPU (hipGetSymbolAddress, once per
// device); the array's name is only a host-side stand-in.
unsigned char *g_noalloc_ptr[kNoAllocMaxDevices] = {};

unsigned char *noalloc_base(int dev) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    fail("device index out of range for the allocation-free entry point");
    return nullptr;
  }
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress"))
      return nullptr;
    g_noalloc_ptr[dev] = static_cast<unsigned char *>(ptr);
  }
  return g_noalloc_ptr[dev];
}

__global__ void zero_bytes(unsigned char *__restrict__ p, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    p[i] = 0;
}

template <typename T>
__global__ void add_into_f64(const T *__restrict__ a, double *__restrict__ b, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    b[i] += double(a[i]);
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int trtt_hip_dropin_noalloc(int kind, double *out, const double *in, long nx, long ny, long nz,
                            double s_scl, int s_deg, const double *cum, long size_footprint,
                            double step_footprint, double coord0, long nu, long nv, double u_scl,
                            double v_scl, const double *xos, const double *xsd, int job) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (kind < trtt::PB3D || kind > trtt::FB2D) return fail("unknown projector kind");
  if (!out || !in || !xos || !xsd) return fail("null argument");
  const bool is2d = kind == trtt::PB2D || kind == trtt::FB2D;
  if (is2d) {
    nz = 1;
    nu = 1;
  }
  if (bad_sizes(nx, ny, nz, nu, nv) || (cum && size_footprint < 2))
    return fail("invalid projector arguments");
  const bool table = cum != nullptr, f64 = g_precision == TRTT_HIP_F64;
  const size_t esz = f64 ? 8 : 4;
  const long nvol = nx * ny * nz, nsino = nu * nv;
  const long nin = job == 0 ? nvol : nsino, nout = job == 0 ? nsino : nvol;
  const long m = table ? size_footprint : 0;
  // layout: error flag | view constants | order, segment, row bounds |
  //         table | FP64 staging | input | output
  const size_t o_view = 256, o_idx = o_view + align256(sizeof(ViewConst));
  const size_t o_table = o_idx + 256;
  const size_t o_stage = o_table + align256(esz * size_t(m));
  const size_t o_in = o_stage + align256(8 * size_t(std::max(std::max(nin, nout), m)));
  const size_t o_out = o_in + align256(esz * size_t(nin));
  const size_t total = o_out + align256(esz * size_t(nout));
  if (total > kNoAllocBytes)
    return fail("allocation-free projector: " + std::to_string(total) +
                " bytes of device scratch needed, the static array holds " +
                std::to_string(kNoAllocBytes));
  if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return TRTT_FAILURE;
  unsigned char *base = noalloc_base(g_device);
  if (!base) return TRTT_FAILURE;
  int *d_err = reinterpret_cast<int *>(base);
  ViewConst *d_view = reinterpret_cast<ViewConst *>(base + o_view);
  int *d_order = reinterpret_cast<int *>(base + o_idx);
  trtt::Segment *d_seg = reinterpret_cast<trtt::Segment *>(base + o_idx + 64);
  int *d_seg_begin = reinterpret_cast<int *>(base + o_idx + 128);
  void *d_table = base + o_table;
  double *d_stage = reinterpret_cast<double *>(base + o_stage);
  void *d_in = base + o_in, *d_out = base + o_out;

  // per-view constants and the one-view schedule (as make_plan builds them)
  ViewConst g;
  const double step = table ? step_footprint : 1.0, c0 = table ? coord0 : 0.0;
  if (is2d)
    view_consts_2d(g, xos, xsd, kind == trtt::FB2D, table, step, s_scl, s_deg, v_scl);
  else
    view_consts_3d(g, xos, xsd, kind == trtt::CB3D, s_scl, s_deg, u_scl, v_scl);
  g.t0 = g.t1 = 0;
  g.w0 = 1.0;
  g.w1 = 0.0;
  const int order = 0;
  const trtt::Segment seg{0, 1, 0, 0};
  const int seg_begin[2] = {0, 1};
  hipStream_t st = nullptr;
  hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(64), 0, st, base, long(sizeof(int)));
  if (!hip_ok(hipMemcpy(d_view, &g, sizeof(g), hipMemcpyHostToDevice), "hipMemcpy") ||
      !hip_ok(hipMemcpy(d_order, &order, sizeof(order), hipMemcpyHostToDevice), "hipMemcpy") ||
      !hip_ok(hipMemcpy(d_seg, &seg, sizeof(seg), hipMemcpyHostToDevice), "hipMemcpy") ||
      !hip_ok(hipMemcpy(d_seg_begin, seg_begin, sizeof(seg_begin), hipMemcpyHostToDevice),
              "hipMemcpy"))
    return TRTT_FAILURE;
  if (table) {
    if (f64) {
      if (!hip_ok(hipMemcpy(d_table, cum, sizeof(double) * m, hipMemcpyHostToDevice), "hipMemcpy"))
        return TRTT_FAILURE;
    } else {
      if (!hip_ok(hipMemcpy(d_stage, cum, sizeof(double) * m, hipMemcpyHostToDevice), "hipMemcpy"))
        return TRTT_FAILURE;
      hipLaunchKernelGGL(cvt_d2f, dim3(grid_for(m)), dim3(256), 0, st, d_stage,
                         static_cast<float *>(d_table), m);
    }
  }
  if (f64) {
    if (!hip_ok(hipMemcpy(d_in, in, sizeof(double) * nin, hipMemcpyHostToDevice), "hipMemcpy"))
      return TRTT_FAILURE;
  } else {
    if (!hip_ok(hipMemcpy(d_stage, in, sizeof(double) * nin, hipMemcpyHostToDevice), "hipMemcpy"))
      return TRTT_FAILURE;
    hipLaunchKernelGGL(cvt_d2f, dim3(grid_for(nin)), dim3(256), 0, st, d_stage,
                       static_cast<float *>(d_in), nin);
  }
  hipLaunchKernelGGL(zero_bytes, dim3(grid_for(long(esz) * nout)), dim3(256), 0, st,
                     static_cast<unsigned char *>(d_out), long(esz) * nout);

  trtt::DevPlan P{};
  P.views = d_view;
  P.order = d_order;
  P.segs = d_seg;
  P.table = table ? d_table : nullptr;
  P.err = d_err;
  P.m = m;
  P.step = step;
  P.coord0 = c0;
  P.nx = int(nx);
  P.ny = int(ny);
  P.nz = int(nz);
  P.nu = int(nu);
  P.nv = int(nv);
  P.nviews = 1;
  P.nt = 1;
  P.vox_per_phase = nvol;
  P.u_scl = u_scl;
  P.v_scl = v_scl;
  {
#pragma clang fp contract(off)
    P.rho_init = is2d ? s_scl * s_scl : s_scl * s_scl * s_scl;
    P.hu_init = s_scl / u_scl;
    P.hv_init = s_scl / v_scl;
    P.hs0 = 0.5 * (s_deg + 1);
  }
  const bool quirk = is2d && !table;
  const long ntiles = tiles_of_dims(kind, nx, ny, nz);
  const int vpb = int(std::max(1L, std::min(ntiles / kTargetBlocks, 8L)));
  if (!hip_ok(trtt::launch_projector(kind, table, quirk, f64, P, job == 0 ? 0 : 1, d_in, d_out,
                                     d_seg_begin, 1, job == 0 ? vpb : 1, 0, st),
              "kernel launch"))
    return TRTT_FAILURE;
  // out_host + result: the host output is staged in FP64 and the result
  // added to it on the device
  if (!hip_ok(hipMemcpy(d_stage, out, sizeof(double) * nout, hipMemcpyHostToDevice), "hipMemcpy"))
    return TRTT_FAILURE;
  if (f64)
    hipLaunchKernelGGL(add_into_f64<double>, dim3(grid_for(nout)), dim3(256), 0, st,
                       static_cast<const double *>(d_out), d_stage, nout);
  else
    hipLaunchKernelGGL(add_into_f64<float>, dim3(grid_for(nout)), dim3(256), 0, st,
                       static_cast<const float *>(d_out), d_stage, nout);
  if (!hip_ok(hipGetLastError(), "kernel launch")) return TRTT_FAILURE;
  int err = 0;
  if (!hip_ok(hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy"))
    return TRTT_FAILURE;
  if (err) return fail("IT WAS SUPPOSED TO BE HIGHLY IMPOSSIBLE: a voxel lies in the source plane");
  if (!hip_ok(hipMemcpy(out, d_stage, sizeof(double) * nout, hipMemcpyDeviceToHost), "hipMemcpy"))
    return TRTT_FAILURE;
  return TRTT_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Host-array evaluation of the weighted least-squares data term over every
// view of a plan (the batched form of trtt_4D_mpy_eval,
// Tomography_4D/Ytrtt4D_optim.i:87-223, and of its serial loop :418-446).
// ---------------------------------------------------------------------------
namespace {

template <typename T>
__global__ void sub_to_f64(const T *__restrict__ a, const T *__restrict__ b, double *__restrict__ c,
                           long n, int use_b) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    c[i] = use_b ? double(a[i]) - double(b[i]) : double(a[i]);
}

template <typename T>
__global__ void mul_into(const T *__restrict__ w, T *__restrict__ y, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    y[i] *= w[i];
}

template <typename T>
__global__ void zero_t(T *__restrict__ p, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    p[i] = T(0);
}

// host FP64 -> device array of the plan's type
bool upload_f64(const trtt_hip_plan *p, EvalData *e, void *dst, const double *src, long n) {
  if (p->precision == TRTT_HIP_F64)
    return hip_ok(hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
  if (!hip_ok(hipMemcpy(e->stage, src, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy"))
    return false;
  hipLaunchKernelGGL(cvt_d2f, dim3(grid_for(n)), dim3(256), 0, nullptr, e->stage,
                     static_cast<float *>(dst), n);
  return hip_ok(hipGetLastError(), "kernel launch");
}

template <typename T>
void zero_dev(void *p, long n) {
  hipLaunchKernelGGL(zero_t<T>, dim3(grid_for(n)), dim3(256), 0, nullptr, static_cast<T *>(p), n);
}

}  // namespace

extern "C" {

int trtt_hip_eval_set_data(trtt_hip_plan_t *p, const double *y, const double *w) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!p || !y) return fail("null argument");
  if (!hip_ok(hipSetDevice(p->device), "hipSetDevice")) return TRTT_FAILURE;
  const bool f64 = p->precision == TRTT_HIP_F64;
  const size_t esz = f64 ? 8 : 4;
  const long nsino = p->nu * p->nv * p->nviews, nvol = p->nx * p->ny * p->nz * p->nt;
  EvalData *e = p->eval;
  if (!e) {
    e = new EvalData;
    if (!hip_ok(hipMalloc(&e->y, esz * nsino), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->w, esz * nsino), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->sino, esz * nsino), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->wcyt, esz * nvol), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->x, esz * nvol), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->g, esz * nvol), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->stage, sizeof(double) * std::max(nsino, nvol)), "hipMalloc") ||
        !hip_ok(hipMalloc(&e->fx, sizeof(double)), "hipMalloc")) {
      eval_release(e);
      return TRTT_FAILURE;
    }
    p->eval = e;
  }
  if (!upload_f64(p, e, e->y, y, nsino)) return TRTT_FAILURE;
  const bool has_w = w != nullptr;
  if (has_w) {
    if (!upload_f64(p, e, e->w, w, nsino)) return TRTT_FAILURE;
  } else {
    hipFree(e->w);
    e->w = nullptr;
  }
  // WCyT = sum_k A_k' W_k y_k (Ytrtt4D_optim.i:764-782)
  hipStream_t st = nullptr;
  if (f64) {
    hipMemcpyAsync(e->sino, e->y, esz * nsino, hipMemcpyDeviceToDevice, st);
    if (has_w)
      hipLaunchKernelGGL(mul_into<double>, dim3(grid_for(nsino)), dim3(256), 0, st,
                         static_cast<const double *>(e->w), static_cast<double *>(e->sino), nsino);
    zero_dev<double>(e->wcyt, nvol);
  } else {
    hipMemcpyAsync(e->sino, e->y, esz * nsino, hipMemcpyDeviceToDevice, st);
    if (has_w)
      hipLaunchKernelGGL(mul_into<float>, dim3(grid_for(nsino)), dim3(256), 0, st,
                         static_cast<const float *>(e->w), static_cast<float *>(e->sino), nsino);
    zero_dev<float>(e->wcyt, nvol);
  }
  if (!hip_ok(launch_plan(p, 1, e->sino, e->wcyt, st), "backward launch")) return TRTT_FAILURE;
  return hip_ok(hipDeviceSynchronize(), "eval set-up") ? TRTT_SUCCESS : TRTT_FAILURE;
}

int trtt_hip_eval(trtt_hip_plan_t *p, double *fx, double *gx, const double *x, int split) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  if (!p || !fx || !gx || !x) return fail("null argument");
  EvalData *e = p->eval;
  if (!e) return fail("trtt_hip_eval: no data (call trtt_hip_eval_set_data first)");
  if (!hip_ok(hipSetDevice(p->device), "hipSetDevice")) return TRTT_FAILURE;
  const bool f64 = p->precision == TRTT_HIP_F64;
  const long nsino = p->nu * p->nv * p->nviews, nvol = p->nx * p->ny * p->nz * p->nt;
  hipStream_t st = nullptr;
  if (!upload_f64(p, e, e->x, x, nvol)) return TRTT_FAILURE;
  if (f64) {
    zero_dev<double>(e->sino, nsino);
    zero_dev<double>(e->g, nvol);
  } else {
    zero_dev<float>(e->sino, nsino);
    zero_dev<float>(e->g, nvol);
  }
  zero_dev<double>(e->fx, 1);
  if (!hip_ok(hipMemsetAsync(p->d_err, 0, sizeof(int), st), "hipMemset")) return TRTT_FAILURE;
  if (!hip_ok(launch_plan(p, 0, e->x, e->sino, st), "forward launch") ||
      !hip_ok(trtt::launch_residual(f64, e->sino, e->y, e->w, nsino, e->fx, st, split != 0),
              "residual launch") ||
      !hip_ok(launch_plan(p, 1, e->sino, e->g, st), "backward launch"))
    return TRTT_FAILURE;
  // gx = A'W(Ax - y), or A'WAx - WCyT with split (FP64 difference)
  if (f64)
    hipLaunchKernelGGL(sub_to_f64<double>, dim3(grid_for(nvol)), dim3(256), 0, st,
                       static_cast<const double *>(e->g), static_cast<const double *>(e->wcyt),
                       e->stage, nvol, split ? 1 : 0);
  else
    hipLaunchKernelGGL(sub_to_f64<float>, dim3(grid_for(nvol)), dim3(256), 0, st,
                       static_cast<const float *>(e->g), static_cast<const float *>(e->wcyt),
                       e->stage, nvol, split ? 1 : 0);
  if (!hip_ok(hipGetLastError(), "kernel launch")) return TRTT_FAILURE;
  int err = 0;
  if (!hip_ok(hipMemcpy(fx, e->fx, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
      !hip_ok(hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy") ||
      !hip_ok(hipMemcpy(gx, e->stage, sizeof(double) * nvol, hipMemcpyDeviceToHost), "hipMemcpy"))
    return TRTT_FAILURE;
  if (err) return fail("IT WAS SUPPOSED TO BE HIGHLY IMPOSSIBLE: a voxel lies in the source plane");
  return TRTT_SUCCESS;
}

}  // extern "C"
