// Fast closed-form footprint kernels for "axial" views: geometries whose
// rotation axis is the object z axis and the detector u axis (every circular
// cone- or parallel-beam orbit built by the TRTT system builders).  There the
// transaxial position vc, the magnification and the v footprint of a voxel
// depend on its column (i, j) only, and uc is affine in k
// (trtt_3D_projectors.c:1234-1261 with o3 = den_z = num_v_z = 0).
//
// Mapping: a wave owns one voxel column at a time with its 64 consecutive k
// in the lanes.
//  * The column's v weights are wave-uniform: lane L evaluates the CDF at
//    the L-th v pixel bound, readlane broadcasts the differences.
//  * Forward: the lanes accumulate the column's axial profile
//    P[u] = sum_k x_k rho_k w_k(u) in wave-private LDS (distinct addresses
//    per step, no collisions), then add the outer product P x cv into the
//    wave's detector window in LDS; the window goes to a scratch slot and a
//    reduction kernel sums the slots over each detector block.
//  * Backward: the window is staged from HBM; the lanes first form
//    Q[u] = sum_m cv[m] y[vl+m][u] for the column, then each lane gathers its
//    voxel's axial weights against Q.
// Voxels whose window leaves the column hull take a direct global path.
#include <climits>
#include <type_traits>

#include "trtt_core.h"

namespace trtt {

extern __shared__ __align__(16) unsigned char g_smem[];

namespace axial {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kKT = 64;  // k per tile (one per lane)
// fixed row pitch of the backward's staged window (one row per thread)
constexpr int kPitch = kThreads;
// fixed row pitch of the forward's detector window
constexpr int kFwdPitch = 128;


// F(x) - 1/2 for the cubic B-spline CDF: sign(x) G(min(|x|, 2)) with, on
// [0, 2], G(t) = 1/2 - (2-t)^4/24 + max(1-t, 0)^4/6 (the two pieces
// 2t/3 - t^3/3 + t^4/8 on [0, 1] and 1/2 - (2-t)^4/24 on [1, 2] differ by
// exactly (1-t)^4/6), so no per-piece select is needed.  The 1/24 and 1/6
// are folded into the bases, s = max(2-t, 0)/24^(1/4) and
// r = max(1-t, 0)/6^(1/4), so G = 1/2 - (s^2)^2 + (r^2)^2 is two FMAs on the
// squares.  Both bases lie below 1, so the [0, 1] output clamp of the FMA
// that forms them is the max(., 0), which also makes G = 1/2 for t > 2:
// 7 instructions with the sign transfer (|x| is a source modifier).
constexpr double kC24 = 0.45180100180492244;  // 24^(-1/4)
constexpr double kC6 = 0.6389431042462724;    // 6^(-1/4)

template <typename T>
__device__ __forceinline__ T cdf3c(T x) {
  const T t = fabs(x);
  const T s = fmin(fmax(fma(t, T(-kC24), T(2.0 * kC24)), T(0)), T(1));
  const T r = fmin(fmax(fma(t, T(-kC6), T(kC6)), T(0)), T(1));
  const T s2 = s * s, r2 = r * r;
  return copysign(fma(r2, r2, fma(s2, -s2, T(0.5))), x);
}

// F(x) - 1/2 for the quadratic B-spline CDF, by the same construction: on
// [0, 3/2], G(t) = 1/2 - (3/2-t)^3/6 + max(1/2-t, 0)^3/2 (the pieces
// 3t/4 - t^3/3 on [0, 1/2] and 1/2 - (3/2-t)^3/6 on [1/2, 3/2] differ by
// exactly (1/2-t)^3/2), with s = max(3/2-t, 0)/6^(1/3) and
// r = max(1/2-t, 0)/2^(1/3), both below 1: G = 1/2 - s^2 s + r^2 r.
constexpr double kQ6 = 0.5503212081491045;   // 6^(-1/3)
constexpr double kQ2 = 0.7937005259840998;   // 2^(-1/3)

template <typename T>
__device__ __forceinline__ T cdf2c(T x) {
  const T t = fabs(x);
  const T s = fmin(fmax(fma(t, T(-kQ6), T(1.5 * kQ6)), T(0)), T(1));
  const T r = fmin(fmax(fma(t, T(-kQ2), T(0.5 * kQ2)), T(0)), T(1));
  const T s2 = s * s, r2 = r * r;
  return copysign(fma(r2, r, fma(s2, -s, T(0.5))), x);
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }

template <typename T>
__device__ __forceinline__ T bcast(T v, int lane);
template <>
__device__ __forceinline__ float bcast<float>(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
template <>
__device__ __forceinline__ double bcast<double>(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return __builtin_bit_cast(double, (long long)(unsigned(lo)) | ((long long)hi << 32));
}

__device__ __forceinline__ void clampw(double c, double hs, int n, int &lo, int &hi) {
  const double l = c - hs, r = c + hs;
  const double nm1 = double(n - 1);
  lo = l < 0.0 ? 0 : (l <= nm1 ? int(floor(l)) : n - 1);
  hi = r < 0.0 ? 0 : (r <= nm1 ? int(floor(r)) + 1 : n);
}

// Column constants of one view.
template <typename T>
struct Col {
  double vc;       // transaxial centre
  T xvs;           // -(vc - vl) sv: the v CDF abscissa of the window start (|vc - vl| is a
                   // few pixels, so single precision is exact enough)
  double uc0, duc; // uc(k) = uc0 + duc k
  T rho;           // rho0 / du
  T hu0;           // hu / du
  T inv_a;         // 1 / (a4^2 + vca8^2): du = sqrt(1 + uca12^2 inv_a) (0 for parallel)
  T sv;            // 1 / hv
  int vl, vr;      // v window
  bool ok;
  // tabulated footprint: FP64 table coordinate of the window start and
  // its step per pixel along v; hu / du and 1 / (a4^2 + vca8^2) in FP64
  double xv0, sdv, hu0d, inv_ad;
  // closed-form footprint: 24^(-1/4) / hu0 (the symmetric evaluation)
  T ca;
};

// Tabulated footprint (trtt_PB3D / trtt_CB3D with a table of the cubic
// B-spline CDF, trtt_3D_projectors.c:267-300): the weight of a pixel is the
// difference of the table at the floor() indices of its two bounds, and the
// table at index n holds the cubic CDF at node coord0 + n step (to the
// table's own trapezoid error; the host accepts a table for this path only
// when it is within 1e-6 of that).  The index is taken in FP64 from the
// reference's coordinate expressions, the CDF at the node in T.
// The reference clamps the index to [0, m - 1]; the host accepts a table
// for this path only when it covers the footprint's support [-2, 2], so
// past either end the CDF at the node is the constant the clamped entry
// holds (-1/2 or +1/2 here) and no clamp is needed.
template <int TAB, typename T>
__device__ __forceinline__ T table_cdf(const DevPlan &P, double x) {
  const float node = fmaf(float(floor(x)), float(P.step), float(P.coord0));
  return TAB == 2 ? cdf2c<T>(node) : cdf3c<T>(node);
}

template <bool CONE, int TAB, typename T>
__device__ __forceinline__ void column(const DevPlan &P, const ViewConst &g, int i, int j, Col<T> &c) {
#pragma clang fp contract(off)
  const double di = double(i), dj = double(j);
  if (CONE) {
    const double wk = (g.o4 + g.o2 * dj) + g.o1 * di;
    const double num_v = (g.nvc + g.nvy * dj) + g.nvx * di;
    const double num_u = (g.nuc + g.nuy * dj) + g.nux * di;
    const double den = (g.dc + g.dy * dj) + g.dx * di;
    const double q = g.a4m / den;
    c.ok = wk != 0.0;
    c.vc = q * num_v + g.b3;
    c.uc0 = q * num_u + g.c3;
    c.duc = q * g.nuz;
    const double magn = g.a4m / wk;
    const double vca8 = (c.vc - g.a8) * P.v_scl;
    const double a = g.a4sqr + vca8 * vca8;
    const double dv = sqrt(1.0 + vca8 * vca8 / g.a4sqr);
    const double hv = P.hv_init * magn * dv;
    c.sv = T(1.0 / hv);
    c.rho = T(P.rho_init * magn * magn * dv);
    c.hu0 = T(P.hu_init * magn);
    c.ca = T(kC24 / (P.hu_init * magn));
    c.inv_a = T(1.0 / a);
    clampw(c.vc, P.hs0 * hv, P.nv, c.vl, c.vr);
    c.xvs = -T(c.vc - double(c.vl)) * c.sv;
    if (TAB) {
      // trtt_3D_projectors.c:267-275: scl_norm_v / step and the start
      const double sdv = (1.0 / hv) / P.step;
      c.sdv = sdv;
      c.xv0 = ((double(c.vl) - c.vc) - hv * P.coord0) * sdv;
      c.hu0d = P.hu_init * magn;
      c.inv_ad = 1.0 / a;
    }
  } else {
    c.ok = true;
    c.vc = (g.nvc + g.nvy * dj) + g.nvx * di;
    c.uc0 = (g.nuc + g.nuy * dj) + g.nux * di;
    c.duc = g.nuz;
    c.sv = T(g.pb_sv);
    c.rho = T(g.pb_rho0);
    c.hu0 = T(g.pb_hu);
    c.ca = T(kC24 / g.pb_hu);
    c.inv_a = T(0);
    clampw(c.vc, g.pb_hsv, P.nv, c.vl, c.vr);
    c.xvs = -T(c.vc - double(c.vl)) * c.sv;
    if (TAB) {
      const double sdv = g.pb_sv / P.step;
      c.sdv = sdv;
      c.xv0 = ((double(c.vl) - c.vc) - g.pb_hv * P.coord0) * sdv;
      c.hu0d = g.pb_hu;
      c.inv_ad = 0.0;
    }
  }
}

// Axial footprint of voxel k of a column.  The window bounds are taken
// from the single-precision centre (an error there only moves a bound onto
// a pixel of zero weight); the sub-pixel offset x0 uses the FP64 centre.
template <typename T>
struct Vox {
  double uc;
  T su, rho;
  float hs;      // half support in pixels
  int ul;        // first pixel, before clamping at the left detector edge
  int ur;        // end pixel, clamped at the right detector edge
  int ux;        // end pixel before clamping
  double xu0, sdu;  // tabulated footprint: table coordinate at ul, step per pixel
};

// Pixel window of the support [c - hs, c + hs] on a detector of n pixels:
// first pixel lo (not clamped: it may be negative), end hi clamped at n and
// hx unclamped.  Float-to-int conversion saturates, and both ends are capped
// so that hx + 1 cannot overflow.
__device__ __forceinline__ void clampw_f(float c, float hs, int n, int &lo, int &hi, int &hx) {
  const float l = c - hs, r = c + hs;
  lo = int(floorf(fmaxf(l, -1.0e7f)));
  hx = int(floorf(fminf(r, 1.0e7f))) + 1;
  hi = min(hx, n);
}

// The window meets the detector.
template <typename T>
__device__ __forceinline__ bool on_detector(const Vox<T> &v) {
  return v.ur > max(v.ul, 0);
}

template <int TAB = 0, typename T>
__device__ __forceinline__ void voxel(const DevPlan &P, double a12, const Col<T> &c, int k,
                                      Vox<T> &v) {
  v.uc = c.uc0 + c.duc * double(k);
  const T uca12 = T((v.uc - a12) * P.u_scl);
  const T du = fsqrt(T(1) + uca12 * uca12 * c.inv_a);
  const T hu = c.hu0 * du;
  v.su = frcp(hu);
  v.rho = c.rho * du;
  v.hs = float(T(P.hs0) * hu);
  clampw_f(float(v.uc), v.hs, P.nu, v.ul, v.ur, v.ux);
  if (TAB) {
    // FP64 scale and table coordinate (trtt_3D_projectors.c:1264-1280,
    // 267-275): hu = hu_init magn du, x = ((ul - uc) - hu coord0) / (hu step);
    // the square root and reciprocal start from single precision and take
    // one Newton step each (relative error far below the table's index
    // resolution)
    const double ud = (v.uc - a12) * P.u_scl;
    const double dsq = fma(ud * ud, c.inv_ad, 1.0);
    // 1/sqrt(dsq) from single precision, one Newton step (relative error
    // about 2^-46), then du = dsq / sqrt(dsq)
    double rs = double(__builtin_amdgcn_rsqf(float(dsq)));
    rs = rs * fma(-0.5 * dsq * rs, rs, 1.5);
    const double hud = c.hu0d * (dsq * rs);
    const double den = hud * P.step;
    double r = double(frcp(float(den)));
    r = fma(r, fma(-den, r, 1.0), r);
    v.sdu = r;
    v.xu0 = ((double(v.ul) - v.uc) - hud * P.coord0) * r;
  }
}

// Weights of pixels [ul, ul+NP) for a voxel whose unclamped window
// [ul, ux) has at most NP pixels: NP - 1 CDF evaluations, and the last
// weight is 1/2 - F because the CDF is exactly 1/2 from the footprint's right
// end on.  The window starts (unclamped) at least 2h left of the centre,
// where the CDF is exactly -1/2.  Entries past the window are zero.
template <int NP, int TAB = 0, typename T>
__device__ __forceinline__ void uweights(const DevPlan &P, const Vox<T> &v, T (&w)[NP]) {
  const T x0 = T(double(v.ul) - v.uc) * v.su;
  T prev = T(-0.5);
#pragma unroll
  for (int l = 0; l + 1 < NP; ++l) {
    const T nx = TAB ? table_cdf<TAB, T>(P, fma(double(l + 1), v.sdu, v.xu0))
                     : cdf3c<T>(x0 + T(l + 1) * v.su);
    w[l] = nx - prev;
    prev = nx;
  }
  w[NP - 1] = T(0.5) - prev;
}

// The weights of uweights<np> are exact on the detector when the unclamped
// window fits np pixels, or when pixel ul + np - 1 lies past the right
// detector edge (its weight is then never used).
template <typename T>
__device__ __forceinline__ bool fits(const Vox<T> &v, int np, int nu) {
  return v.ux - v.ul <= np || v.ul + np > nu;
}

// Pixel bound of a column: every voxel window of the column has at most
// ceil(2 hs_max) + 1 pixels, hs_max the larger half support of the two ends
// of the k range (hs is convex in k: uc is affine in k and the distortion
// factor sqrt(1 + uca^2/a) convex in uc).  The margin covers the rounding of
// the window bounds; voxels that still exceed it take the slow path.
__device__ __forceinline__ int pixel_bound(float hs_a, float hs_b) {
  return int(ceilf(2.0f * fmaxf(hs_a, hs_b) + 1.0e-3f)) + 1;
}

// ---------------------------------------------------------------------------
// Symmetric evaluation of the closed-form cubic footprint (no table).
// The pixel bounds are counted outwards from the pixel fl = floor(uc) that
// holds the centre: with f = uc - fl, bound j on the left lies at
// x = -(f + j) su and bound j on the right at x = (1 - f + j) su
// (j = 0, 1, ...; x in units of hu).  On the left half the CDF is
// F(x) = s^4 - r^4 with s = max(2 + x, 0)/24^(1/4), r = max(1 + x, 0)/6^(1/4)
// (the truncated powers (x+2)^4/24 - 4 (x+1)^4/24 of the cubic B-spline);
// on the right half 1 - F(x) has the same form in -x.  With
// A = su/24^(1/4) and A' = su/6^(1/4) the bases of bound j are
// (2/24^(1/4) - A f) - j A and (1/6^(1/4) - A' f) - j A', both below 1, so
// each is one FMA whose [0, 1] output clamp is the max(., 0): six
// instructions per bound, with neither |x| nor a sign transfer.  Past the
// support (j >= hs - f on the left, hs = 2 hu the half support in pixels)
// both forms are exactly 0, so J = floor(hs) bounds per side cover the
// window fl - J - 1 .. fl + J + 1 of 2J + 3 pixels:
//   w(fl - 1 - j) = FL_j - FL_(j+1),  w(fl) = 1 - FL_0 - FR_0,
//   w(fl + 1 + j) = FR_j - FR_(j+1),  FL_(J+1) = FR_(J+1) = 0
// with FL_j the CDF at left bound j and FR_j the tail 1 - F at right
// bound j; the weights telescope to 1 (up to rounding).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double frsq(double x) { return 1.0 / sqrt(x); }

// Centre of voxel kd of a column in FP64: every use (column hull, voxels)
// takes this one expression, so the floors agree.
template <typename T>
__device__ __forceinline__ double col_uc(const Col<T> &c, double kd) {
  return fma(c.duc, kd, c.uc0);
}

template <typename T>
struct Sym {
  T A, A2;            // su / 24^(1/4), su / 6^(1/4)
  T bl, rl, br, rr;   // bases of bound 0: s and r on the left, s and r on the right
  T rho;              // rho0 du
  int p;              // centre pixel fl relative to the column profile start
};

// uscl = u_scl, a12s = a12 u_scl; lo: first pixel of the column profile.  The distortion
// factor takes the centre in single precision (its relative error reaches
// du only through (uc - a12)^2 / a, far below single precision); the
// sub-pixel offset f comes from the FP64 centre.
template <typename T>
__device__ __forceinline__ void sym_voxel(T uscl, T a12s, const Col<T> &c, double kd, int lo,
                                          Sym<T> &s) {
  const double uc = col_uc(c, kd);
  const double fl = floor(uc);
  const T f = T(uc - fl);
  s.p = int(fl) - lo;
  const T uca = fma(T(uc), uscl, -a12s);
  const T dsq = fma(uca * c.inv_a, uca, T(1));
  const T rs = frsq(dsq);  // 1 / du
  s.A = c.ca * rs;
  s.A2 = s.A * T(kC6 / kC24);
  s.rho = c.rho * (dsq * rs);
  const T g = T(1) - f;
  s.bl = fma(-s.A, f, T(2.0 * kC24));
  s.rl = fma(-s.A2, f, T(kC6));
  s.br = fma(-s.A, g, T(2.0 * kC24));
  s.rr = fma(-s.A2, g, T(kC6));
}

template <typename T>
__device__ __forceinline__ T sym_cdf(T bs, T br) {
  const T s = fmin(fmax(bs, T(0)), T(1)), r = fmin(fmax(br, T(0)), T(1));
  const T s2 = s * s, r2 = r * r;
  return fma(s2, s2, -(r2 * r2));
}

// The r term of bound j vanishes once j >= hu: its base is
// (1 - (f + j)/hu)/6^(1/4) on the left and (1 - (1 - f + j)/hu)/6^(1/4) on
// the right, with f in [0, 1).  The column's J = floor(hs_max) with
// hs = 2 hu (the closed-form cubic, the only footprint of this path) gives
// floor(hu_max) = floor(J/2), so every bound j >= J/2 + 1 lies past hu for
// every voxel of the column and takes F = s^4 alone (three instructions).
template <typename T>
__device__ __forceinline__ T sym_cdf_s(T bs) {
  const T s = fmin(fmax(bs, T(0)), T(1));
  const T s2 = s * s;
  return s2 * s2;
}

// CDF values at the bounds: L[j] = F at left bound j, R[j] = 1 - F at
// right bound j.
template <int J, typename T>
__device__ __forceinline__ void sym_bounds(const Sym<T> &s, T (&L)[J + 1], T (&R)[J + 1]) {
  constexpr int JR = J / 2 + 1;  // first bound without the r term
#pragma unroll
  for (int j = 0; j <= J; ++j) {
    const T bl = j == 0 ? s.bl : fma(T(-j), s.A, s.bl);
    const T br = j == 0 ? s.br : fma(T(-j), s.A, s.br);
    if (j < JR) {
      const T rl = j == 0 ? s.rl : fma(T(-j), s.A2, s.rl);
      const T rr = j == 0 ? s.rr : fma(T(-j), s.A2, s.rr);
      L[j] = sym_cdf(bl, rl);
      R[j] = sym_cdf(br, rr);
    } else {
      L[j] = sym_cdf_s(bl);
      R[j] = sym_cdf_s(br);
    }
  }
}

// Weights of the pixels fl - J - 1 + t, t < 2J + 3.
template <int J, typename T>
__device__ __forceinline__ void sym_weights(const Sym<T> &s, T (&w)[2 * J + 3]) {
  T L[J + 1], R[J + 1];
  sym_bounds<J>(s, L, R);
  w[0] = L[J];
#pragma unroll
  for (int j = 0; j < J; ++j) w[J - j] = L[j] - L[j + 1];
  w[J + 1] = (T(1) - L[0]) - R[0];
#pragma unroll
  for (int j = 0; j < J; ++j) w[J + 2 + j] = R[j] - R[j + 1];
  w[2 * J + 2] = R[J];
}

// Lane L receives v from lane L + 1 (DPP wave shift; the last lane's value
// is unused).
__device__ __forceinline__ float from_next_lane(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130 /* wave_shl:1 */,
                                         0xF, 0xF, false));
}

// v weights of the column: lane m evaluates the CDF at the m-th v pixel
// bound, one shifted difference gives the weight of pixel m in lane m, and
// readlane broadcasts the W weights.  (An
// LDS slot per wave read back with broadcast reads instead was slower in
// both kernels: the backward's extra LDS costs a workgroup per CU.)
template <int W, int TAB = 0, typename T>
__device__ __forceinline__ void vweights(const DevPlan &P, const Col<T> &c, int lane, T (&cv)[W]) {
  const T f = TAB ? table_cdf<TAB, T>(P, fma(double(lane), c.sdv, c.xv0))
                  : cdf3c<T>(c.xvs + T(lane) * c.sv);
  // the shift runs on all lanes (a lane reading a disabled lane gets 0);
  // the weights of rows m >= npv are not zeroed: every row loop that
  // takes them (outer products, Q rows) stops at the column's npv rows
  const T fn = from_next_lane(f);
  const T d = fn - f;
#pragma unroll
  for (int m = 0; m < W; ++m) cv[m] = bcast(d, m);
}


// Row loops with a wave-uniform row count N (1..W), dispatched once per
// column so the rows run without per-row branches.
template <int N, typename T>
__device__ __forceinline__ void outer_rows(T *dst, int RW, const T *cv, T pu) {
  T t[N];
#pragma unroll
  for (int m = 0; m < N; ++m) t[m] = dst[m * RW];
#pragma unroll
  for (int m = 0; m < N; ++m) dst[m * RW] = t[m] + cv[m] * pu;
}

template <int N, typename T>
__device__ __forceinline__ T q_rows(const T *src, int RW, const T *cv) {
  T q = T(0);
#pragma unroll
  for (int m = 0; m < N; ++m) q += cv[m] * src[m * RW];
  return q;
}

// Q of both 64-pixel halves of a column span of at most 128 pixels with all
// reads in one batch.
template <int N, typename T>
__device__ __forceinline__ void q_rows2(const T *src, int RW, const T *cv, T &q0, T &q1) {
  T t0[N], t1[N];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    t0[m] = src[m * RW];
    t1[m] = src[m * RW + 64];
  }
  q0 = T(0);
  q1 = T(0);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    q0 += cv[m] * t0[m];
    q1 += cv[m] * t1[m];
  }
}

// q_rows2 with a compile-time row pitch: the row offsets become immediate
// offsets of the LDS reads.
template <int N, int PITCH, typename T>
__device__ __forceinline__ void q_rows2c(const T *src, const T *cv, T &q0, T &q1) {
  T t0[N], t1[N];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    t0[m] = src[m * PITCH];
    t1[m] = src[m * PITCH + 64];
  }
  q0 = T(0);
  q1 = T(0);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    q0 += cv[m] * t0[m];
    q1 += cv[m] * t1[m];
  }
}

template <int W, int PITCH, typename T>
__device__ __forceinline__ void q_n2c(int n, const T *src, const T (&cv)[W], T &q0, T &q1) {
  q0 = q1 = T(0);
  switch (n) {
#define TRTT_Q2C(NN) \
  case NN:           \
    if (NN <= W) q_rows2c<(NN <= W ? NN : 1), PITCH, T>(src, cv, q0, q1); \
    break;
    TRTT_Q2C(1) TRTT_Q2C(2) TRTT_Q2C(3) TRTT_Q2C(4) TRTT_Q2C(5) TRTT_Q2C(6) TRTT_Q2C(7)
    TRTT_Q2C(8) TRTT_Q2C(9) TRTT_Q2C(10) TRTT_Q2C(11) TRTT_Q2C(12) TRTT_Q2C(13)
#undef TRTT_Q2C
    default: break;
  }
}

template <int W, typename T>
__device__ __forceinline__ void q_n2(int n, const T *src, int RW, const T (&cv)[W], T &q0, T &q1) {
  q0 = q1 = T(0);
  switch (n) {
#define TRTT_Q2(NN) \
  case NN:          \
    if (NN <= W) q_rows2<(NN <= W ? NN : 1), T>(src, RW, cv, q0, q1); \
    break;
    TRTT_Q2(1) TRTT_Q2(2) TRTT_Q2(3) TRTT_Q2(4) TRTT_Q2(5) TRTT_Q2(6) TRTT_Q2(7) TRTT_Q2(8)
    TRTT_Q2(9) TRTT_Q2(10) TRTT_Q2(11) TRTT_Q2(12) TRTT_Q2(13)
#undef TRTT_Q2
    default: break;
  }
}

// Both 64-pixel halves of a column span of at most 128 pixels in one batch:
// every read is in flight before the first write (one LDS round trip per
// column).  All 64 lanes take part with the window's row stride; lanes past
// the span add exactly 0 (p = 0) and so write back the value they read in
// this batch.  Such a write can only land on a pixel of this column by
// wrapping past the end of a window row, i.e. on the pixel of a later row,
// whose real update comes later in program order (the DS operations of a
// wave execute in order): every pixel ends with its real update.  The
// highest index touched is RW * RH + 127 <= wcap + 127.
template <int N, typename T>
__device__ __forceinline__ void outer_rows2(T *dst, int RW, const T *cv, T p0, T p1) {
  T t0[N], t1[N];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    t0[m] = dst[m * RW];
    t1[m] = dst[m * RW + 64];
  }
#pragma unroll
  for (int m = 0; m < N; ++m) {
    dst[m * RW] = t0[m] + cv[m] * p0;
    dst[m * RW + 64] = t1[m] + cv[m] * p1;
  }
}

template <int W, typename T>
__device__ __forceinline__ void outer_n2(int n, T *dst, int RW, const T (&cv)[W], T p0, T p1) {
  switch (n) {
#define TRTT_OUTER2(NN) \
  case NN:              \
    if (NN <= W) outer_rows2<(NN <= W ? NN : 1), T>(dst, RW, cv, p0, p1); \
    break;
    TRTT_OUTER2(1) TRTT_OUTER2(2) TRTT_OUTER2(3) TRTT_OUTER2(4) TRTT_OUTER2(5) TRTT_OUTER2(6)
    TRTT_OUTER2(7) TRTT_OUTER2(8) TRTT_OUTER2(9) TRTT_OUTER2(10) TRTT_OUTER2(11)
    TRTT_OUTER2(12) TRTT_OUTER2(13)
#undef TRTT_OUTER2
    default: break;
  }
}

// outer_rows2 with a compile-time row pitch: the row offsets become
// immediate offsets of the LDS reads and writes.
template <int N, int PITCH, typename T>
__device__ __forceinline__ void outer_rows2c(T *dst, const T *cv, T p0, T p1) {
  T t0[N], t1[N];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    t0[m] = dst[m * PITCH];
    t1[m] = dst[m * PITCH + 64];
  }
#pragma unroll
  for (int m = 0; m < N; ++m) {
    dst[m * PITCH] = t0[m] + cv[m] * p0;
    dst[m * PITCH + 64] = t1[m] + cv[m] * p1;
  }
}

template <int W, int PITCH, typename T>
__device__ __forceinline__ void outer_n2c(int n, T *dst, const T (&cv)[W], T p0, T p1) {
  switch (n) {
#define TRTT_OUTER2C(NN) \
  case NN:               \
    if (NN <= W) outer_rows2c<(NN <= W ? NN : 1), PITCH, T>(dst, cv, p0, p1); \
    break;
    TRTT_OUTER2C(1) TRTT_OUTER2C(2) TRTT_OUTER2C(3) TRTT_OUTER2C(4) TRTT_OUTER2C(5)
    TRTT_OUTER2C(6) TRTT_OUTER2C(7) TRTT_OUTER2C(8) TRTT_OUTER2C(9) TRTT_OUTER2C(10)
    TRTT_OUTER2C(11) TRTT_OUTER2C(12) TRTT_OUTER2C(13)
#undef TRTT_OUTER2C
    default: break;
  }
}

template <int W, typename T>
__device__ __forceinline__ void outer_n(int n, T *dst, int RW, const T (&cv)[W], T pu) {
  switch (n) {
#define TRTT_OUTER(NN) \
  case NN:             \
    if (NN <= W) outer_rows<(NN <= W ? NN : 1), T>(dst, RW, cv, pu); \
    break;
    TRTT_OUTER(1) TRTT_OUTER(2) TRTT_OUTER(3) TRTT_OUTER(4) TRTT_OUTER(5) TRTT_OUTER(6)
    TRTT_OUTER(7) TRTT_OUTER(8) TRTT_OUTER(9) TRTT_OUTER(10) TRTT_OUTER(11) TRTT_OUTER(12)
    TRTT_OUTER(13)
#undef TRTT_OUTER
    default: break;
  }
}

template <int W, typename T>
__device__ __forceinline__ T q_n(int n, const T *src, int RW, const T (&cv)[W]) {
  switch (n) {
#define TRTT_Q(NN) \
  case NN:         \
    return NN <= W ? q_rows<(NN <= W ? NN : 1), T>(src, RW, cv) : T(0);
    TRTT_Q(1) TRTT_Q(2) TRTT_Q(3) TRTT_Q(4) TRTT_Q(5) TRTT_Q(6) TRTT_Q(7) TRTT_Q(8) TRTT_Q(9)
    TRTT_Q(10) TRTT_Q(11) TRTT_Q(12) TRTT_Q(13)
#undef TRTT_Q
    default: return T(0);
  }
}

// Forward profile steps of one voxel whose window has at most NP pixels
// (pp: the profile at the window's first pixel, unclamped: W margin entries
// precede the profile, spare entries follow it, and neither is read).  The
// pixels past the footprint get zero weights.  passes (wave-uniform) comes
// from the host: 1 where it has shown that consecutive voxels start their
// windows on distinct pixels (plain read-modify-write steps, the lanes hit
// distinct entries at each step and LDS serves one wave's steps in order);
// m > 1 where voxels m slices apart do, so the lanes of one residue of k
// mod m run the plain steps together, one residue after the other; 0:
// LDS atomics.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int NP, typename T>
__device__ __forceinline__ void accumulate_steps(const T (&w)[NP], T *pp, T val, int passes) {
  // plain steps: one FMA per pixel into the profile entry (the product is
  // not formed apart, so it is not shared with the atomic branch)
  if (passes == 1) {
#pragma unroll
    for (int l = 0; l < NP; ++l) {
      pp[l] = fma_t(w[l], val, pp[l]);
      __builtin_amdgcn_wave_barrier();
    }
  } else if (passes > 1) {
    const int r = int(threadIdx.x & 63) % passes;
    for (int ps = 0; ps < passes; ++ps) {
      if (r == ps) {
#pragma unroll
        for (int l = 0; l < NP; ++l) {
          pp[l] = fma_t(w[l], val, pp[l]);
          __builtin_amdgcn_wave_barrier();
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  } else {
#pragma unroll
    for (int l = 0; l < NP; ++l) atomicAdd(pp + l, w[l] * val);
  }
}

template <int NP, int TAB, typename T>
__device__ __forceinline__ void profile_steps(const DevPlan &P, const Vox<T> &v, T *pp, T val,
                                              int passes) {
  T w[NP];
  uweights<NP, TAB>(P, v, w);
  accumulate_steps<NP>(w, pp, val, passes);
}

template <int NP, typename T>
__device__ __forceinline__ T dot_steps(const T (&w)[NP], const T *qq) {
  T sum = T(0);
#pragma unroll
  for (int l = 0; l < NP; ++l) sum += w[l] * qq[l];
  return sum;
}

// Backward gather of one voxel whose window has at most NP pixels against
// the column's Q row (qq: Q at the window's first pixel; Q is zero past the
// column span).
template <int NP, int TAB, typename T>
__device__ __forceinline__ T gather_steps(const DevPlan &P, const Vox<T> &v, const T *qq) {
  T w[NP];
  uweights<NP, TAB>(P, v, w);
  return dot_steps<NP>(w, qq);
}

// Symmetric footprints with a column-uniform J (2J + 3 <= W): profile steps
// (pp: the profile at pixel fl - J - 1) and gathers (qq: Q at that pixel).
template <int W, typename T>
__device__ __forceinline__ void sym_profile_n(int J, const Sym<T> &s, T *pp, T val, int passes) {
  switch (J) {
#define TRTT_SP(JJ)                                 \
  case JJ:                                          \
    if constexpr (2 * JJ + 3 <= W) {                \
      T w[2 * JJ + 3];                              \
      sym_weights<JJ>(s, w);                        \
      accumulate_steps<2 * JJ + 3>(w, pp, val, passes); \
    }                                               \
    break;
    TRTT_SP(0) TRTT_SP(1) TRTT_SP(2) TRTT_SP(3) TRTT_SP(4) TRTT_SP(5)
#undef TRTT_SP
    default: break;
  }
}


template <int W, typename T>
__device__ __forceinline__ T sym_gather_n(int J, const Sym<T> &s, const T *qq) {
  switch (J) {
#define TRTT_SG(JJ)                        \
  case JJ:                                 \
    if constexpr (2 * JJ + 3 <= W) {       \
      T w[2 * JJ + 3];                     \
      sym_weights<JJ>(s, w);               \
      return dot_steps<2 * JJ + 3>(w, qq); \
    }                                      \
    break;
    TRTT_SG(0) TRTT_SG(1) TRTT_SG(2) TRTT_SG(3) TRTT_SG(4) TRTT_SG(5)
#undef TRTT_SG
    default: break;
  }
  return T(0);
}

// Dispatch on the column's wave-uniform pixel bound np (1..W).
template <int W, int TAB, typename T>
__device__ __forceinline__ void profile_n(const DevPlan &P, int np, const Vox<T> &v, T *pp, T val,
                                          int pl) {
  switch (np) {
#define TRTT_PS(NN) \
  case NN:          \
    if (NN <= W) profile_steps<(NN <= W ? NN : 1), TAB, T>(P, v, pp, val, pl); \
    break;
    TRTT_PS(1) TRTT_PS(2) TRTT_PS(3) TRTT_PS(4) TRTT_PS(5) TRTT_PS(6) TRTT_PS(7) TRTT_PS(8)
    TRTT_PS(9) TRTT_PS(10) TRTT_PS(11) TRTT_PS(12) TRTT_PS(13)
#undef TRTT_PS
    default: break;
  }
}

template <int W, int TAB, typename T>
__device__ __forceinline__ T gather_n(const DevPlan &P, int np, const Vox<T> &v, const T *qq) {
  switch (np) {
#define TRTT_GS(NN) \
  case NN:          \
    return NN <= W ? gather_steps<(NN <= W ? NN : 1), TAB, T>(P, v, qq) : T(0);
    TRTT_GS(1) TRTT_GS(2) TRTT_GS(3) TRTT_GS(4) TRTT_GS(5) TRTT_GS(6) TRTT_GS(7) TRTT_GS(8)
    TRTT_GS(9) TRTT_GS(10) TRTT_GS(11) TRTT_GS(12) TRTT_GS(13)
#undef TRTT_GS
    default: return T(0);
  }
}

// Slow path for one voxel: arbitrary window sizes, direct global memory.
template <bool FWD, int TAB, typename T>
__device__ void slow_voxel(const DevPlan &P, const Col<T> &c, const Vox<T> &v, T val, T *out,
                           const T *in, T &acc) {
  const int ul = max(v.ul, 0);  // the part of the window on the detector
  const int npu = v.ur - ul, npv = c.vr - c.vl;
  if (npu <= 0 || npv <= 0) return;
  const T xu0 = T(double(ul) - v.uc) * v.su;
  const T xv0 = c.xvs;
  // tabulated footprint: u coordinates from the window start ul (>= v.ul)
  auto fu = [&](int l) {
    return TAB ? table_cdf<TAB, T>(P, fma(double(ul - v.ul + l), v.sdu, v.xu0))
               : cdf3c<T>(xu0 + T(l) * v.su);
  };
  auto fv = [&](int m) {
    return TAB ? table_cdf<TAB, T>(P, fma(double(m), c.sdv, c.xv0)) : cdf3c<T>(xv0 + T(m) * c.sv);
  };
  T fv0 = fv(0);
  T sum = T(0);
  for (int m = 0; m < npv; ++m) {
    const T fv1 = fv(m + 1);
    const T wv = fv1 - fv0;
    fv0 = fv1;
    T fu0 = fu(0);
    for (int l = 0; l < npu; ++l) {
      const T fu1 = fu(l + 1);
      const T w = wv * (fu1 - fu0);
      fu0 = fu1;
      const long pix = long(c.vl + m) * P.nu + (ul + l);
      if (FWD)
        unsafeAtomicAdd(out + pix, w * val);
      else
        sum += w * in[pix];
    }
  }
  if (!FWD) acc += sum * v.rho;
}

struct TileMap {
  int bx, by, bz;
  __device__ __forceinline__ void init(const DevPlan &P, int cx, int cy, int b) {
    const int ntx = (P.nx + cx - 1) / cx, nty = (P.ny + cy - 1) / cy;
    bx = b % ntx;
    b /= ntx;
    by = b % nty;
    bz = b / nty;
  }
};

__device__ __forceinline__ void block_union(int &a, int &b, int &c, int &d, int *red) {
  // a, c: min; b, d: max; values are wave-uniform on entry
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[4 * w + 0] = a;
    red[4 * w + 1] = b;
    red[4 * w + 2] = c;
    red[4 * w + 3] = d;
  }
  __syncthreads();
  a = red[0];
  b = red[1];
  c = red[2];
  d = red[3];
#pragma unroll
  for (int q = 1; q < kWaves; ++q) {
    a = min(a, red[4 * q + 0]);
    b = max(b, red[4 * q + 1]);
    c = min(c, red[4 * q + 2]);
    d = max(d, red[4 * q + 3]);
  }
}

// Column hull of the u windows over the k range [kf, kl] (uc is affine in k
// and the window bounds are monotone in uc): windows at the two ends, padded
// by one pixel.  Voxels are still checked against it.  np: pixel bound of
// the column's voxel windows.
template <typename T>
__device__ __forceinline__ void column_hull(const DevPlan &P, const ViewConst &g, const Col<T> &c,
                                            int kf, int kl, int &lo, int &hi, int &np) {
  Vox<T> a, b;
  voxel(P, g.a12, c, kf, a);
  voxel(P, g.a12, c, kl, b);
  lo = max(0, min(a.ul, b.ul) - 1);
  hi = min(P.nu, max(a.ur, b.ur) + 1);
  np = pixel_bound(a.hs, b.hs);
}

// ---------------------------------------------------------------------------
// Column table of a wave: lane L evaluates column L of the wave (geometry,
// v window, u hull over the tile's k range); the column loop then broadcasts
// column cc from lane cc with readlane.
// ---------------------------------------------------------------------------
template <typename T>
struct ColLane {
  Col<T> c;
  int lo, hi;   // u hull of the column over [kf, kl] on the detector
  int np;       // pixel bound of its voxel windows (symmetric windows: J)
  int i, j;     // voxel column
  int base;     // (j nx + i) nz: the column's offset in the k-fastest volume copy
  // symmetric windows: first pixel of the column profile (not clamped to
  // the detector) and its length | J << 16
  int ulo, span_j;
  bool live;    // inside the volume, wk != 0, non-empty windows
  bool bad;     // wk == 0 (voxel in the source plane)
};

// Column hull of the symmetric windows over [kf, kl]: J = floor(hs_max),
// hs_max the larger half support of the two ends (hs is convex in k), and
// the profile spans fl_min - J - 1 .. fl_max + J + 1 with fl the floors of
// the centres at the two ends (uc is affine in k and col_uc monotone in k,
// so every centre of the column lies between them).
template <typename T>
__device__ __forceinline__ void column_hull_sym(const DevPlan &P, const ViewConst &g,
                                                const Col<T> &c, int kf, int kl, ColLane<T> &cl) {
  Vox<T> a, b;
  voxel(P, g.a12, c, kf, a);
  voxel(P, g.a12, c, kl, b);
  const float hs = fmaxf(a.hs, b.hs);
  const int J = hs < 64.0f ? int(floorf(hs)) : 64;
  const double fa = floor(col_uc(c, double(kf))), fb = floor(col_uc(c, double(kl)));
  const double lo = fmin(fa, fb) - double(J + 1), hi = fmax(fa, fb) + double(J + 2);
  // far-off columns saturate: their span fails every profile bound
  cl.ulo = int(fmin(fmax(lo, -1.0e9), 1.0e9));
  cl.span_j = int(fmin(fmax(hi - lo, 0.0), 65535.0)) | (min(J, 255) << 16);
  cl.lo = int(fmin(fmax(lo, 0.0), double(P.nu)));
  cl.hi = int(fmin(fmax(hi, 0.0), double(P.nu)));
  cl.np = J;
}

// Column `col` of the tile on view g; `active` false leaves the lane empty.
template <bool CONE, int TAB, int CX, int CY, typename T>
__device__ __forceinline__ void column_entry(const DevPlan &P, const ViewConst &g, const TileMap &tm,
                                             int col, bool active, int kf, int kl, ColLane<T> &cl) {
  cl.i = tm.bx * CX + col % CX;
  cl.j = tm.by * CY + col / CX;
  cl.base = (cl.j * P.nx + cl.i) * P.nz;
  cl.live = false;
  cl.bad = false;
  cl.lo = cl.hi = 0;
  cl.np = 0;
  cl.ulo = 0;
  cl.span_j = 0;
  if (active && cl.i < P.nx && cl.j < P.ny) {
    column<CONE, TAB>(P, g, cl.i, cl.j, cl.c);
    cl.bad = !cl.c.ok;
    if (cl.c.ok && cl.c.vl < cl.c.vr) {
      if (TAB == 0)
        column_hull_sym(P, g, cl.c, kf, kl, cl);
      else
        column_hull(P, g, cl.c, kf, kl, cl.lo, cl.hi, cl.np);
      cl.live = cl.lo < cl.hi;
    }
  }
}

template <bool CONE, int TAB, int CX, int CY, int NCOL, typename T>
__device__ __forceinline__ void column_table(const DevPlan &P, const ViewConst &g, const TileMap &tm,
                                             int wave, int lane, int kf, int kl, ColLane<T> &cl) {
  column_entry<CONE, TAB, CX, CY>(P, g, tm, wave * NCOL + lane, lane < NCOL, kf, kl, cl);
}

__device__ __forceinline__ int bcast_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

template <int TAB = 0, typename T>
__device__ __forceinline__ void bcast_col(const ColLane<T> &src, int cc, Col<T> &c, int &lo, int &hi,
                                          int &np, int &i, int &j) {
  if (TAB) {
    c.xv0 = bcast(src.c.xv0, cc);
    c.sdv = bcast(src.c.sdv, cc);
    c.hu0d = bcast(src.c.hu0d, cc);
    c.inv_ad = bcast(src.c.inv_ad, cc);
  }
  c.xvs = bcast(src.c.xvs, cc);
  c.uc0 = bcast(src.c.uc0, cc);
  c.duc = bcast(src.c.duc, cc);
  c.rho = bcast(src.c.rho, cc);
  c.hu0 = bcast(src.c.hu0, cc);
  c.inv_a = bcast(src.c.inv_a, cc);
  c.sv = bcast(src.c.sv, cc);
  // index pairs packed in 16 bits each (the host limits the sizes)
  const int v2 = bcast_i(src.c.vl | (src.c.vr << 16), cc);
  c.vl = v2 & 0xffff;
  c.vr = int(unsigned(v2) >> 16);
  c.ok = true;
  const int h2 = bcast_i(src.lo | (src.hi << 16), cc);
  lo = h2 & 0xffff;
  hi = int(unsigned(h2) >> 16);
  np = bcast_i(src.np, cc);
  const int ij = bcast_i(src.i | (src.j << 16), cc);
  i = ij & 0xffff;
  j = int(unsigned(ij) >> 16);
}

// Column cc for the symmetric windows: lo, span and J of its profile.
template <typename T>
__device__ __forceinline__ void bcast_sym(const ColLane<T> &src, int cc, Col<T> &c, int &lo,
                                          int &span, int &J) {
  c.xvs = bcast(src.c.xvs, cc);
  c.uc0 = bcast(src.c.uc0, cc);
  c.duc = bcast(src.c.duc, cc);
  c.rho = bcast(src.c.rho, cc);
  c.ca = bcast(src.c.ca, cc);
  c.inv_a = bcast(src.c.inv_a, cc);
  c.sv = bcast(src.c.sv, cc);
  const int v2 = bcast_i(src.c.vl | (src.c.vr << 16), cc);
  c.vl = v2 & 0xffff;
  c.vr = int(unsigned(v2) >> 16);
  c.ok = true;
  lo = bcast_i(src.ulo, cc);
  const int sj = bcast_i(src.span_j, cc);
  span = sj & 0xffff;
  J = int(unsigned(sj) >> 16);
}

}  // namespace axial
}  // namespace trtt

// Wave reductions of the ROCm device library (DPP lane shifts and a
// readlane, no LDS round trips), over the active lanes: every caller runs
// them with the whole wave active.
extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);

namespace trtt {
namespace axial {

// (the earlier butterfly of __shfl_xor steps, six ds_bpermute round trips
// per reduction, cost the 256^3 cone forward 0.8 ms and the backward 0.8 ms)
__device__ __forceinline__ int wred_min(int x) { return __ockl_wfred_min_i32(x); }
__device__ __forceinline__ int wred_max(int x) { return __ockl_wfred_max_i32(x); }

// Block window: union over the live columns of the four waves.
template <typename T>
__device__ __forceinline__ void block_window(const ColLane<T> &cl, bool live, int *red, int &ul,
                                             int &ur, int &vl, int &vr) {
  ul = wred_min(live ? cl.lo : INT_MAX);
  ur = wred_max(live ? cl.hi : INT_MIN);
  vl = wred_min(live ? cl.c.vl : INT_MAX);
  vr = wred_max(live ? cl.c.vr : INT_MIN);
  block_union(ul, ur, vl, vr, red);
}

// Detector window of the columns held by the lanes where m is true (wave-
// uniform results): it starts on a multiple of 4 pixels and its rows are a
// multiple of 4 long; true when it is not empty and fits cap elements.
template <typename T>
__device__ __forceinline__ bool window_fits(const ColLane<T> &cl, bool m, int cap, int &ul, int &ur,
                                            int &vl, int &vr) {
  ul = bcast_i(wred_min(m ? cl.lo : INT_MAX), 0) & ~3;
  ur = bcast_i(wred_max(m ? cl.hi : INT_MIN), 0);
  vl = bcast_i(wred_min(m ? cl.c.vl : INT_MAX), 0);
  vr = bcast_i(wred_max(m ? cl.c.vr : INT_MIN), 0);
  return ur > ul && vr > vl && ((ur - ul + 3) & ~3) * (vr - vl) <= cap;
}

// ---------------------------------------------------------------------------
// Forward.  One wave per workgroup: the wave owns a tile of CX x CY columns
// x 64 k and a private detector window in LDS, so no barrier or atomic is
// needed between waves.  Per column the lanes accumulate the axial profile
// P[u] = sum_k x_k rho_k w_k(u) in LDS (plain read-modify-write steps in
// `plain` interleaved passes when the host has shown that the lanes of a
// pass hit distinct pixels at each step, LDS atomics when plain is 0), then add the outer product P x cv to the window with the lanes
// on distinct window columns.  The window goes to a scratch slot; fwd_reduce
// sums the slots covering each detector pixel.
// ---------------------------------------------------------------------------
// Occupancy targets (waves per SIMD) for the register allocator, 0 leaves
// the allocation to the compiler.  Measured on the 256^3 cone benchmark:
// forward 6 (80 VGPRs, a few spilled; its LDS then allows 21 waves per CU)
// beats 5 (96 VGPRs, 20 waves: 28.2 against 27.9 ms) and 4 (128 VGPRs:
// 32 ms); backward 6 (78) beats 4 (127) and 5 (90).
#ifndef TRTT_AXIAL_FWD_WAVES
#define TRTT_AXIAL_FWD_WAVES 6
#endif
#ifndef TRTT_AXIAL_BWD_WAVES
#define TRTT_AXIAL_BWD_WAVES 6
#endif
#if TRTT_AXIAL_FWD_WAVES > 0
#define TRTT_FWD_OCC __attribute__((amdgpu_waves_per_eu(TRTT_AXIAL_FWD_WAVES)))
#else
#define TRTT_FWD_OCC
#endif
#if TRTT_AXIAL_BWD_WAVES > 0
#define TRTT_BWD_OCC __attribute__((amdgpu_waves_per_eu(TRTT_AXIAL_BWD_WAVES)))
#else
#define TRTT_BWD_OCC
#endif

// TWO: some view interpolates between two phases; without, every view
// reads one phase (t0 = t1) with the weight w0 + w1, and the second phase's
// pointer, weight and branches leave the column loop.
template <bool CONE, int W, int CX, int CY, int TAB, bool TWO, typename T>
__global__ __launch_bounds__(64) TRTT_FWD_OCC void fwd(DevPlan P, const T *__restrict__ vol,
                                          T *__restrict__ sino, int views_per_block, int wcap,
                                          int pmax, int view_base, int view_end,
                                          T *__restrict__ scratch, int4 *__restrict__ hdr,
                                          int plain, int flush, int ntiles, int xcds) {
  constexpr int NCOL = CX * CY;  // columns of the wave
  static_assert(NCOL <= 64, "one column per lane");
  // With scratch windows (flush 0) the previous chunk's reduction runs on
  // a second stream beside this kernel: this kernel is bound by instruction
  // issue and the reduction by memory latency, so these waves issue first
  // (forward 43.6 -> 42.6 ms on the 256^3 cone benchmark; priorities 1, 2
  // and 3 measure the same).
  __builtin_amdgcn_s_setprio(1);
  const int lane = threadIdx.x;
  T *win = reinterpret_cast<T *>(g_smem);
  T *prof = win + wcap + W;  // after W margin entries
  const T uscl = T(P.u_scl);
  // block -> (tile, view group).  Grid (tiles, groups) by default.  With
  // xcds > 1 the grid is 1-D in an XCD-aware order: the hardware deals
  // consecutive workgroups to the XCDs in turn, so block L runs on XCD
  // L mod xcds.  A unit is one view group on one k block of tiles; XCD x
  // runs the units x, x + xcds, ..., every tile of a unit on the same XCD,
  // so the detector lines that a unit's atomic adds touch stay in one L2.
  int tile = int(blockIdx.x), group = int(blockIdx.y);
  if (xcds > 1) {
    const int nbz = (P.nz + kKT - 1) / kKT;
    const int txy = ntiles / nbz;
    const int L = int(blockIdx.x);
    const int sq = L / xcds;
    const int u = (L - sq * xcds) + xcds * (sq / txy);
    group = u / nbz;
    tile = sq % txy + txy * (u - group * nbz);
  }
  const int first_view = view_base + group * views_per_block;
  if (first_view >= view_end) return;  // padding of the XCD-aware grid
  TileMap tm;
  tm.init(P, CX, CY, tile);
  const int k = tm.bz * kKT + lane;
  const bool kin = k < P.nz;
  const int kf = tm.bz * kKT, kl = min(kf + kKT - 1, P.nz - 1);
  const long npix = long(P.nu) * P.nv;
  prof[lane] = T(0);
  prof[lane + 64] = T(0);

  // When the column tables of all views of the block fit in the wave, they
  // are evaluated at once: lane vv NCOL + c holds column c on view vv.
  const bool multi = views_per_block * NCOL <= 64;
  ColLane<T> cl;
  if (multi) {
    const int slot = lane / NCOL;
    const bool active = slot < views_per_block && first_view + slot < view_end;
    column_entry<CONE, TAB, CX, CY>(P, P.views[active ? first_view + slot : first_view], tm,
                                    lane % NCOL, active, kf, kl, cl);
    if (cl.bad) atomicOr(P.err, 1);
  }

  for (int vv = 0; vv < views_per_block; ++vv) {
    const int view = first_view + vv;
    if (view >= view_end) break;
    const ViewConst &g = P.views[view];
    T *out = sino + long(view) * npix;
    const long slot = long(view - view_base) * ntiles + tile;
    // per-view constants of the column loop, read once (the loop's global
    // stores would otherwise force a reload per column)
    const double a12 = g.a12;
    const T a12s = T(g.a12 * P.u_scl);
    const T w0 = TWO ? T(g.w0) : T(g.w0 + g.w1), w1 = T(g.w1);
    const bool two = TWO && g.w1 != 0.0;
    const T *vol0 = vol + long(g.t0) * P.vox_per_phase;
    const T *vol1 = TWO ? vol + long(g.t1) * P.vox_per_phase : vol0;

    if (!multi) {
      column_table<CONE, TAB, CX, CY, NCOL>(P, g, tm, 0, lane, kf, kl, cl);
      if (cl.bad) atomicOr(P.err, 1);
    }
    // this view's columns sit in lanes loff .. loff + NCOL - 1
    const int loff = multi ? vv * NCOL : 0;
    const bool inview = cl.live && lane >= loff && lane < loff + NCOL;
    // The columns of the view run in one pass with one window, or, when
    // that window does not fit the LDS capacity (atomic flush only: a
    // scratch window has one slot per view and tile), in two passes over
    // two halves of the tile (its rows of columns split, else its columns
    // of columns), else in four passes over its quadrants, each pass with
    // its own window and flush.  A view whose quadrants do not fit either
    // takes the direct path.
    const int cidx = lane - loff;
    const bool half_rows = cidx / CX < CY / 2, half_cols = cidx % CX < CX / 2;
    int split = 0;
    // window of the whole view (wave-uniform: kept in scalar registers)
    int ul0, ur0, vl0, vr0;
    const bool fits0 = window_fits(cl, inview, wcap, ul0, ur0, vl0, vr0);
    if (flush && !fits0 && ur0 > ul0 && vr0 > vl0) {
      int a, b, c, d, e, f, g, h;
      if (window_fits(cl, inview && half_rows, wcap, a, b, c, d) &&
          window_fits(cl, inview && !half_rows, wcap, e, f, g, h))
        split = 1;
      else if (window_fits(cl, inview && half_cols, wcap, a, b, c, d) &&
               window_fits(cl, inview && !half_cols, wcap, e, f, g, h))
        split = 2;
      else if (window_fits(cl, inview && half_rows && half_cols, wcap, a, b, c, d) &&
               window_fits(cl, inview && !half_rows && half_cols, wcap, e, f, g, h) &&
               window_fits(cl, inview && half_rows && !half_cols, wcap, a, b, c, d) &&
               window_fits(cl, inview && !half_rows && !half_cols, wcap, e, f, g, h))
        split = 3;
    }
    const int npass = split == 0 ? 1 : (split == 3 ? 4 : 2);
    for (int pass = 0; pass < npass; ++pass) {
    // pass p of the quadrants: rows half p & 1, columns half p >> 1
    const bool in_rows = (pass & 1) == 0 ? half_rows : !half_rows;
    const bool in_cols = (split == 3 ? pass >> 1 : pass) == 0 ? half_cols : !half_cols;
    const bool live = inview && (split == 0 || (split == 1 && in_rows) ||
                                 (split == 2 && in_cols) || (split == 3 && in_rows && in_cols));
    int ul = ul0, ur = ur0, vl = vl0, vr = vr0;
    bool in_lds = fits0;
    if (split) in_lds = window_fits(cl, live, wcap, ul, ur, vl, vr);
    const bool any = ur > ul && vr > vl;
    const int RW = any ? (ur - ul + 3) & ~3 : 0, RH = any ? vr - vl : 0;
    // Window rows at the fixed pitch kFwdPitch where they fit (atomic flush
    // only: a scratch window keeps the RW layout its header describes); the
    // outer product then reaches its rows through immediate LDS offsets.
    const bool fixed = in_lds && flush && RW <= kFwdPitch && kFwdPitch * RH <= wcap;
    const int pitch = fixed ? kFwdPitch : RW;
    // window zeroing and flush in 16-byte vectors: the capacity wcap is a
    // multiple of 64 elements, so the rounded-up count stays inside it
    using V16 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int kE = 16 / int(sizeof(T));
    const int nvec = (pitch * RH + kE - 1) / kE;
    if (in_lds) {
      V16 *w16 = reinterpret_cast<V16 *>(win);
      for (int q = lane; q < nvec; q += 64) w16[q] = V16{};
    }
    __builtin_amdgcn_wave_barrier();

    if (any) {
      // lanes past the last slice work on a clamped k with a zero value
      const int kc = kin ? k : P.nz - 1;
      const double kd = double(kc);
      // the voxels of the next live column are loaded while the current one
      // is projected (volume copy with k fastest: the 64 lanes read 64
      // consecutive words)
      unsigned long long pend = __ballot(live);
      T xa = T(0), xb = T(0);
      auto load_col = [&](int col) {
        const unsigned idx = unsigned(bcast_i(cl.base, col) + kc);
        xa = vol0[idx];
        xb = two ? vol1[idx] : T(0);
      };
      if (pend) load_col(__builtin_ctzll(pend));
      // The column loop.  JJ >= 0 (closed-form footprint, plain profile
      // steps): every column of the view takes the window width 2 JJ + 3,
      // JJ the largest J of its columns (a narrower column gets extra bounds
      // of weight 0 on each side), so the loop dispatches on J and on the
      // pass mode once per view instead of once per column.  JJ < 0: the
      // dispatch per column.
      auto columns = [&](auto jc) {
      constexpr int JJ = decltype(jc)::value;
      while (pend) {
        const int cc = __builtin_ctzll(pend);
        pend &= pend - 1;
        T x = w0 * xa;
        if (two) x += w1 * xb;
        if (pend) load_col(__builtin_ctzll(pend));
        if constexpr (TAB == 0) {
          // symmetric windows: every voxel of the column takes the profile
          // (its window lies inside the column profile by construction;
          // entries off the detector are dropped at the outer product)
          Col<T> c;
          int lo, span, J;
          bcast_sym(cl, cc, c, lo, span, J);
          if constexpr (JJ >= 0) {
            lo -= JJ - J;
            span += 2 * (JJ - J);
            J = JJ;
          }
          const int npv = c.vr - c.vl;
          const bool use_prof = in_lds && span <= pmax && npv <= W && 2 * J + 3 <= W;
          T cv[W];
          vweights<W, 0>(P, c, lane, cv);
          if (use_prof) {
            // prof[0, 128) is zero here (cleared at the start and after each
            // outer product); longer spans clear the rest
            for (int u = 128 + lane; u < span; u += 64) prof[u] = T(0);
            __builtin_amdgcn_wave_barrier();
            Sym<T> sv;
            sym_voxel(uscl, a12s, c, kd, lo, sv);
            // lanes past the last slice share the clamped slice's window:
            // they add their zeros into the W margin entries in front of the
            // profile (never read), so no two lanes of a plain step meet
            T *const pp = kin ? prof + (sv.p - J - 1) : prof - W;
            const T val = kin ? x * sv.rho : T(0);
            if constexpr (JJ >= 0) {
              T w[2 * JJ + 3];
              sym_weights<JJ>(sv, w);
              accumulate_steps<2 * JJ + 3>(w, pp, val, 1);
            } else {
              sym_profile_n<W>(J, sv, pp, val, plain);
            }
            __builtin_amdgcn_wave_barrier();
            // outer product P x cv over the column's pixels in the window
            // [ul, ur) (which lies on the detector): pixel a + t is profile
            // entry t0 + t
            const int a = max(lo, ul), cnt = min(lo + span, ur) - a, t0 = a - lo;
            T *dst0 = win + (c.vl - vl) * pitch + (a - ul);
            if (cnt <= 128) {
              const T p0 = lane < cnt ? prof[t0 + lane] : T(0);
              const T p1 = lane + 64 < cnt ? prof[t0 + lane + 64] : T(0);
              if (fixed)
                outer_n2c<W, kFwdPitch>(npv, dst0 + lane, cv, p0, p1);
              else
                outer_n2<W>(npv, dst0 + lane, RW, cv, p0, p1);
            } else {
              for (int u = lane; u < cnt; u += 64)
                outer_n<W>(npv, dst0 + u, pitch, cv, prof[t0 + u]);
            }
            __builtin_amdgcn_wave_barrier();
            prof[lane] = T(0);
            prof[lane + 64] = T(0);
            __builtin_amdgcn_wave_barrier();
          } else {
            // direct path for the whole column
            c.hu0 = bcast(cl.c.hu0, cc);
            if (kin) {
              Vox<T> v;
              voxel(P, a12, c, kc, v);
              if (on_detector(v)) {
                T dummy = T(0);
                slow_voxel<true, 0, T>(P, c, v, x * v.rho, out, nullptr, dummy);
              }
            }
          }
        } else {
          Col<T> c;
          int lo, hi, np, i, j;
          bcast_col<TAB>(cl, cc, c, lo, hi, np, i, j);
          np = min(np, W);
          const int span = hi - lo;
          const int npv = c.vr - c.vl;
          const bool use_prof = in_lds && span <= pmax && npv <= W;
          T cv[W];
          vweights<W, TAB>(P, c, lane, cv);
          T val = T(0);
          Vox<T> v;
          bool act = false;
          voxel<TAB>(P, a12, c, kc, v);
          act = kin && on_detector(v);
          val = kin ? x * v.rho : T(0);
          if (use_prof) {
            // prof[0, 128) is zero here (cleared at the start and after each
            // outer product); longer spans clear the rest, spare entries included
            for (int u = 128 + lane; u < span + W; u += 64) prof[u] = T(0);
            __builtin_amdgcn_wave_barrier();
            // Windows starting left of the detector use the W margin entries
            // in front of the profile (never read).  Every lane runs the steps:
            // lanes off the fast path add 0 into the margin, and the rare
            // voxels that need the direct path take it in a separate block.
            const bool fast = act && v.ul >= lo - W && v.ur <= hi && fits(v, np, P.nu);
            profile_n<W, TAB>(P, np, v, fast ? prof + (v.ul - lo) : prof - W, fast ? val : T(0),
                              plain);
            if (__ballot(act && !fast)) {
              T dummy = T(0);
              if (act && !fast) slow_voxel<true, TAB, T>(P, c, v, val, out, nullptr, dummy);
            }
            __builtin_amdgcn_wave_barrier();
            // outer product P x cv; the lanes sit on distinct window columns and
            // the rows of one column are distinct: one batch of reads, FMAs and
            // writes per lane
            T *dst0 = win + (c.vl - vl) * pitch + (lo - ul);
            if (span <= 128) {
              // profile entries past the span may hold weights of pixels past
              // the detector edge: masked to 0
              const T p0 = lane < span ? prof[lane] : T(0);
              const T p1 = lane + 64 < span ? prof[lane + 64] : T(0);
              if (fixed)
                outer_n2c<W, kFwdPitch>(npv, dst0 + lane, cv, p0, p1);
              else
                outer_n2<W>(npv, dst0 + lane, RW, cv, p0, p1);
            } else {
              for (int u = lane; u < span; u += 64) outer_n<W>(npv, dst0 + u, pitch, cv, prof[u]);
            }
            __builtin_amdgcn_wave_barrier();
            prof[lane] = T(0);
            prof[lane + 64] = T(0);
            __builtin_amdgcn_wave_barrier();
          } else if (act) {
            T dummy = T(0);
            slow_voxel<true, TAB, T>(P, c, v, val, out, nullptr, dummy);
          }
        }
      }
      };
      if constexpr (TAB == 0) {
        // J of the view: the largest over its live columns (wave-uniform)
        const int jv = bcast_i(wred_max(live ? int(unsigned(cl.span_j) >> 16) : 0), 0);
        if (plain == 1 && jv == 3 && 9 <= W)
          columns(std::integral_constant<int, 3>{});
        else if (plain == 1 && jv == 2)
          columns(std::integral_constant<int, 2>{});
        else
          columns(std::integral_constant<int, -1>{});
      } else {
        columns(std::integral_constant<int, -1>{});
      }
    }
    if (in_lds && flush) {
      // add the window to the view's projection: global atomic adds (in
      // L2), the lanes on consecutive pixels of a row, zeros skipped; the
      // columns past the detector's last pixel (the rounded-up row end)
      // are not added
      const int cw = min(RW, P.nu - ul);
      T *orow = out + long(vl) * P.nu + ul;
      if (cw <= 128) {
        const bool c0 = lane < cw, c1 = lane + 64 < cw;
        for (int r = 0; r < RH; ++r, orow += P.nu) {
          const T a = win[r * pitch + lane], b = win[r * pitch + lane + 64];
          if (c0 && a != T(0)) unsafeAtomicAdd(orow + lane, a);
          if (c1 && b != T(0)) unsafeAtomicAdd(orow + lane + 64, b);
        }
      } else {
        for (int r = 0; r < RH; ++r, orow += P.nu)
          for (int cI = lane; cI < cw; cI += 64) {
            const T a = win[r * pitch + cI];
            if (a != T(0)) unsafeAtomicAdd(orow + cI, a);
          }
      }
    } else if (in_lds) {
      V16 *dst = reinterpret_cast<V16 *>(scratch + slot * wcap);
      const V16 *w16 = reinterpret_cast<const V16 *>(win);
      for (int q = lane; q < nvec; q += 64) dst[q] = w16[q];
    }
    if (lane == 0 && !flush) hdr[slot] = in_lds ? make_int4(ul, vl, RW, RH) : make_int4(0, 0, 0, 0);
    __builtin_amdgcn_wave_barrier();
    }  // passes
  }
}

// vol[t][k][j][i] -> volk[t][j][i][k] through a 32 (i) x 64 (k) LDS tile,
// both sides coalesced.  grid: (ceil(nx/32) * ny, ceil(nz/64), nt).
template <typename T>
__global__ __launch_bounds__(kThreads) void to_k_fastest(const T *__restrict__ vol,
                                                         T *__restrict__ volk, int nx, int ny,
                                                         int nz) {
  __shared__ T tile[64][33];
  const int nix = (nx + 31) / 32;
  const int i0 = (blockIdx.x % nix) * 32, j = blockIdx.x / nix, k0 = blockIdx.y * 64;
  const long phase = long(blockIdx.z) * nx * ny * nz;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int kk = ty; kk < 64; kk += 8) {
    const int k = k0 + kk, i = i0 + tx;
    tile[kk][tx] = (k < nz && i < nx) ? vol[phase + (long(k) * ny + j) * nx + i] : T(0);
  }
  __syncthreads();
  const int lk = threadIdx.x & 63, li = threadIdx.x >> 6;  // 64 x 4
  for (int ii = li; ii < 32; ii += 4) {
    const int k = k0 + lk, i = i0 + ii;
    if (k < nz && i < nx) volk[phase + (long(j) * nx + i) * nz + k] = tile[lk][ii];
  }
}

// Sum of the scratch windows over the tiles, per detector pixel:
// sino[view][v][u] += sum_t window_t(u, v).  Pixel block 64 (u) x 16 (v).
// bin_windows first lists, per (view, pixel block), the tiles whose window
// overlaps the block; a block whose list overflowed scans every header.
constexpr int kRedU = 64, kRedV = 16;
constexpr int kBinCap = 2048;  // tiles listed per (view, pixel block)

__global__ __launch_bounds__(kThreads) void zero_ints(int *__restrict__ p, long n) {
  for (long i = long(blockIdx.x) * kThreads + threadIdx.x; i < n; i += long(gridDim.x) * kThreads)
    p[i] = 0;
}

__global__ __launch_bounds__(kThreads) void bin_windows(const int4 *__restrict__ hdr, int ntiles,
                                                        int nviews, int npu, int npv,
                                                        int *__restrict__ count,
                                                        int *__restrict__ list) {
  const long i = long(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= long(ntiles) * nviews) return;
  const int vloc = int(i / ntiles), t = int(i - long(vloc) * ntiles);
  const int4 r = hdr[i];
  if (r.z <= 0 || r.w <= 0) return;
  const int nb = npu * npv;
  const int pu0 = max(r.x / kRedU, 0), pu1 = min((r.x + r.z - 1) / kRedU, npu - 1);
  const int pv0 = max(r.y / kRedV, 0), pv1 = min((r.y + r.w - 1) / kRedV, npv - 1);
  for (int pv = pv0; pv <= pv1; ++pv)
    for (int pu = pu0; pu <= pu1; ++pu) {
      const long b = long(vloc) * nb + pv * npu + pu;
      const int idx = atomicAdd(count + b, 1);
      if (idx < kBinCap) list[b * kBinCap + idx] = t;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void fwd_reduce(DevPlan P, const T *__restrict__ scratch,
                                                       const int4 *__restrict__ hdr,
                                                       const int *__restrict__ count,
                                                       const int *__restrict__ list,
                                                       T *__restrict__ sino, int view_base,
                                                       int ntiles, int cap) {
  // 16 x 16 threads, each on 4 consecutive pixels of one row: windows start
  // on a multiple of 4 pixels and their rows are multiples of 4 long, so the
  // four pixels are inside or outside a window together and one 16-byte
  // load reads them.  The block walks its list of overlapping windows (or
  // every header when the list overflowed) with wave-uniform indices, read
  // through scalar loads: no LDS, which the concurrent forward kernel needs.
  static_assert(sizeof(T) == 4, "four pixels per 16-byte load");
  const int npu = (P.nu + kRedU - 1) / kRedU, npv = (P.nv + kRedV - 1) / kRedV;
  const int pu = blockIdx.x % npu, pv = blockIdx.x / npu;
  const int vloc = blockIdx.y;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int u = pu * kRedU + 4 * tx, v = pv * kRedV + ty;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  const int4 *h = hdr + long(vloc) * ntiles;
  const long b = long(vloc) * npu * npv + blockIdx.x;
  const int nlist = count[b];
  const bool binned = nlist <= kBinCap;
  const int total = binned ? nlist : ntiles;
  const int *lst = list + b * kBinCap;
  const T *base = scratch + long(vloc) * ntiles * cap;
  // unconditional loads (threads outside a window read its first entries
  // and add 0) so that the loads of several windows are in flight together
#pragma unroll 4
  for (int e = 0; e < total; ++e) {
    const int t = binned ? lst[e] : e;
    const int4 r = h[t];
    const T *src = base + long(t) * cap;
    const int du = u - r.x, dv = v - r.y;
    const bool in = r.z != 0 && du >= 0 && du < r.z && dv >= 0 && dv < r.w;
    const float4 val = *reinterpret_cast<const float4 *>(src + (in ? dv * r.z + du : 0));
    if (in) {
      acc.x += val.x;
      acc.y += val.y;
      acc.z += val.z;
      acc.w += val.w;
    }
  }
  if (v < P.nv) {
    T *row = sino + long(view_base + vloc) * P.nu * P.nv + long(v) * P.nu;
    const float a[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (u + e < P.nu) row[u + e] += a[e];
  }
}

// ---------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------
// The accumulators of a lane are one vector register tuple: the column loop
// (not unrolled) indexes them with its wave-uniform counter, which indexed
// VGPR moves serve (an array indexed that way is placed in scratch memory,
// a memory round trip per column).
template <typename T, int N>
using AccVec = T __attribute__((ext_vector_type(N)));

// VPL voxels per lane: the tile spans 64 VPL slices and lane L holds the
// voxels k = kf + 64 q + L, q < VPL, so that the per-column set-up (column
// broadcast, v weights, Q row) serves 64 VPL voxels.
template <bool CONE, int W, int CX, int CY, bool TWO, int VPL, int TAB, typename T>
__global__ __launch_bounds__(kThreads) TRTT_BWD_OCC void bwd(DevPlan P, const T *__restrict__ sino,
                                                T *__restrict__ vol,
                                                const int *__restrict__ seg_begin, int atomic_out,
                                                int lds_cap, int pmax) {
  constexpr int NCOL = CX * CY / kWaves;
  constexpr int NA = NCOL * VPL;  // accumulators per lane
  static_assert(NCOL <= 64, "one column per lane");
  T *region = reinterpret_cast<T *>(g_smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Q row of the wave: W zero entries before and after it
  T *qrow = region + lds_cap + wave * (pmax + 2 * W) + W;
  const T uscl = T(P.u_scl);
  if (lane < W) qrow[lane - W] = T(0);
  __shared__ int red[4 * kWaves];
  TileMap tm;
  tm.init(P, CX, CY, blockIdx.x);
  const int kf = tm.bz * (kKT * VPL), kl = min(kf + kKT * VPL - 1, P.nz - 1);
  const long npix = long(P.nu) * P.nv;
  const int s_begin = seg_begin[blockIdx.y], s_end = seg_begin[blockIdx.y + 1];

  for (int s = s_begin; s < s_end; ++s) {
    const Segment seg = P.segs[s];
    // TWO: some segment interpolates between two phases (second accumulator)
    AccVec<T, NA> acc0 = T(0);
    AccVec<T, TWO ? NA : 1> acc1 = T(0);

    // The column tables of NVT consecutive views of the segment are
    // evaluated at once: lane v NCOL + c holds column c on view n0 + v.
    constexpr int NVT = 64 / NCOL;
    static_assert(64 % NCOL == 0, "views per column table");
    ColLane<T> cl;
    for (int n = 0; n < seg.count; ++n) {
      const int vslot = n % NVT;
      if (vslot == 0) {
        const int s2 = lane / NCOL;
        const bool active = n + s2 < seg.count;
        const int vw = P.order[seg.start + (active ? n + s2 : n)];
        column_entry<CONE, TAB, CX, CY>(P, P.views[vw], tm, wave * NCOL + lane % NCOL, active, kf,
                                        kl, cl);
        if (cl.bad) atomicOr(P.err, 1);
      }
      const int view = P.order[seg.start + n];
      const ViewConst &g = P.views[view];
      const T *in = sino + long(view) * npix;
      const double a12 = g.a12;
      const T a12s = T(g.a12 * P.u_scl);
      const T w0 = T(g.w0), w1 = T(g.w1), wsum = T(g.w0 + g.w1);
      // this view's columns sit in lanes loff .. loff + NCOL - 1
      const int loff = vslot * NCOL;
      const bool live = cl.live && lane >= loff && lane < loff + NCOL;
      int ul, ur, vl, vr;
      block_window(cl, live, red, ul, ur, vl, vr);
      // wave-uniform (readfirstlane: kept in scalar registers)
      ul = bcast_i(ul, 0);
      ur = bcast_i(ur, 0);
      vl = bcast_i(vl, 0);
      vr = bcast_i(vr, 0);
      const bool any = ur > ul && vr > vl;
      const int RW = any ? ur - ul : 0, RH = any ? vr - vl : 0;
      const bool in_lds = any && RW * RH <= lds_cap;
      // Window rows at a fixed pitch of kPitch where they fit (the Q rows
      // then reach the window rows through immediate offsets); thread c
      // stages column c of every row.
      const bool fixed = in_lds && RW <= kPitch && kPitch * RH <= lds_cap;
      const int pitch = fixed ? kPitch : RW;
      if (fixed && sizeof(T) == 4) {
        // loads straight into LDS (no registers, no LDS write instructions,
        // every row in flight at once): lane L of wave w writes column
        // 64 w + L of the row, the wave's destination being uniform
        const int cI = int(threadIdx.x);
        if (cI < RW) {
          const T *src = in + (long(vl) * P.nu + (ul + cI));
          T *dst = region + (cI & ~63);
          for (int r = 0; r < RH; ++r, src += P.nu, dst += kPitch)
            __builtin_amdgcn_global_load_lds(src, dst, 4, 0, 0);
        }
        // the loads' LDS writes land before this wave reaches the barrier
        __builtin_amdgcn_s_waitcnt(0);
      } else if (fixed) {
        // eight rows' loads in flight at a time (one load and its wait per
        // row left the view's staging latency exposed)
        const int cI = int(threadIdx.x);
        if (cI < RW) {
          long o = long(vl) * P.nu + (ul + cI);
          for (int r = 0; r < RH; r += 8, o += 8 * long(P.nu)) {
            T t[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) t[q] = r + q < RH ? in[o + q * long(P.nu)] : T(0);
#pragma unroll
            for (int q = 0; q < 8; ++q)
              if (r + q < RH) region[(r + q) * kPitch + cI] = t[q];
          }
        }
      } else if (in_lds) {
        // flat over the window with the source offset advanced
        // incrementally (one division per view, not per element)
        const int sr = kThreads / RW, sc = kThreads - sr * RW;
        int r = int(threadIdx.x) / RW, cI = int(threadIdx.x) - r * RW;
        long off = long(vl + r) * P.nu + (ul + cI);
        const long step = long(sr) * P.nu + sc, wrap = long(P.nu) - RW;
        for (int idx = threadIdx.x; idx < RW * RH; idx += kThreads) {
          region[idx] = in[off];
          off += step;
          cI += sc;
          if (cI >= RW) {
            cI -= RW;
            off += wrap;
          }
        }
      }
      __syncthreads();
      if (any) {
        const unsigned long long live_mask = __ballot(live) >> loff;
        // (J is dispatched per column here: a dispatch once per view, with
        // the loop instantiated per J as in the forward, spilled more
        // registers and ran slower, 22.8 -> 26.0 ms on the 256^3 cone.)
#pragma unroll 1
        for (int cc = 0; cc < NCOL; ++cc) {
          if (!((live_mask >> cc) & 1ull)) continue;
          T contrib[VPL];
#pragma unroll
          for (int q = 0; q < VPL; ++q) contrib[q] = T(0);
          if constexpr (TAB == 0) {
            // symmetric windows: every voxel of the column gathers from the
            // column's Q row (its window lies inside the row by construction)
            Col<T> c;
            int lo, span, J;
            bcast_sym(cl, loff + cc, c, lo, span, J);
            const int npv = c.vr - c.vl;
            const bool use_q = in_lds && span <= pmax && npv <= W && 2 * J + 3 <= W;
            T cv[W];
            vweights<W, 0>(P, c, lane, cv);
            if (use_q) {
              // Q over the column profile: entry t is pixel lo + t, zero off
              // the window [ul, ur) (which lies on the detector); pixel a + t
              // is entry t0 + t
              const int a = max(lo, ul), cnt = min(lo + span, ur) - a, t0 = a - lo;
              const T *src0 = region + (c.vl - vl) * pitch + (a - ul);
              T *qa = qrow + t0;
              for (int u = lane; u < t0; u += 64) qrow[u] = T(0);
              for (int u = t0 + cnt + lane; u < span; u += 64) qrow[u] = T(0);
              if (cnt <= 128 * VPL) {
                // 128 pixels per batch of reads; lanes past cnt read other
                // entries of the LDS region (index < lds_cap + 128 VPL) and
                // write nothing
#pragma unroll
                for (int h = 0; h < VPL; ++h) {
                  const int b = 128 * h;
                  if (h == 0 || cnt > b) {
                    T q0, q1;
                    if (fixed)
                      q_n2c<W, kPitch>(npv, src0 + b + lane, cv, q0, q1);
                    else
                      q_n2<W>(npv, src0 + b + lane, RW, cv, q0, q1);
                    if (b + lane < cnt) qa[b + lane] = q0;
                    if (b + lane + 64 < cnt) qa[b + lane + 64] = q1;
                  }
                }
              } else {
                for (int u = lane; u < cnt; u += 64) qa[u] = q_n<W>(npv, src0 + u, pitch, cv);
              }
              __builtin_amdgcn_wave_barrier();
              // (one dispatch on J for the VPL voxels together, which lets
              // their LDS reads overlap, spilled more registers: slower)
#pragma unroll
              for (int q = 0; q < VPL; ++q) {
                const int k = kf + kKT * q + lane;
                const bool kin = k < P.nz;
                Sym<T> sv;
                sym_voxel(uscl, a12s, c, double(kin ? k : P.nz - 1), lo, sv);
                const T gq = sym_gather_n<W>(J, sv, qrow + (sv.p - J - 1));
                contrib[q] = kin ? gq * sv.rho : T(0);
              }
              __builtin_amdgcn_wave_barrier();
            } else {
              // direct path for the whole column
              c.hu0 = bcast(cl.c.hu0, loff + cc);
#pragma unroll
              for (int q = 0; q < VPL; ++q) {
                const int k = kf + kKT * q + lane;
                if (k < P.nz) {
                  Vox<T> v;
                  voxel(P, a12, c, k, v);
                  if (on_detector(v))
                    slow_voxel<false, 0, T>(P, c, v, T(0), nullptr, in, contrib[q]);
                }
              }
            }
          } else {
            Col<T> c;
            int lo, hi, np, i, j;
            bcast_col<TAB>(cl, loff + cc, c, lo, hi, np, i, j);
            np = min(np, W);
            const int span = hi - lo;
            const int npv = c.vr - c.vl;
            const bool use_q = in_lds && span <= pmax && npv <= W;
            T cv[W];
            vweights<W, TAB>(P, c, lane, cv);
            if (use_q) {
              const T *src0 = region + (c.vl - vl) * pitch + (lo - ul);
              if (span <= 128 * VPL) {
                // 128 pixels per batch of reads; lanes past the span read other
                // entries of the LDS region (index < lds_cap + 128 VPL) and
                // store zeros into the slack
  #pragma unroll
                for (int h = 0; h < VPL; ++h) {
                  const int b = 128 * h;
                  if (h == 0 || span > b) {
                    T q0, q1;
                    if (fixed)
                      q_n2c<W, kPitch>(npv, src0 + b + lane, cv, q0, q1);
                    else
                      q_n2<W>(npv, src0 + b + lane, RW, cv, q0, q1);
                    if (b + lane < span + W) qrow[b + lane] = b + lane < span ? q0 : T(0);
                    if (b + lane + 64 < span + W)
                      qrow[b + lane + 64] = b + lane + 64 < span ? q1 : T(0);
                  }
                }
              } else {
                for (int u = lane; u < span; u += 64) qrow[u] = q_n<W>(npv, src0 + u, pitch, cv);
              }
              if (lane < W) qrow[span + lane] = T(0);
              __builtin_amdgcn_wave_barrier();
  #pragma unroll
              for (int q = 0; q < VPL; ++q) {
                const int k = kf + kKT * q + lane;
                const bool kin = k < P.nz;
                Vox<T> v;
                voxel<TAB>(P, a12, c, kin ? k : P.nz - 1, v);
                // All np pixels: zero weights past the footprint, zero Q past
                // the column span (right detector edge) and in the W entries
                // in front of the row (windows starting left of the detector).
                // Every lane gathers: lanes off the fast path read the zero
                // entries, and the rare voxels that need the direct path take it
                // in a separate block.
                const bool act = kin && on_detector(v);
                const bool fast = act && v.ul >= lo - W && v.ur <= hi && fits(v, np, P.nu);
                const T gq = gather_n<W, TAB>(P, np, v, fast ? qrow + (v.ul - lo) : qrow - W);
                contrib[q] = fast ? gq * v.rho : T(0);
                if (__ballot(act && !fast)) {
                  if (act && !fast)
                    slow_voxel<false, TAB, T>(P, c, v, T(0), nullptr, in, contrib[q]);
                }
              }
              __builtin_amdgcn_wave_barrier();
            } else {
  #pragma unroll
              for (int q = 0; q < VPL; ++q) {
                const int k = kf + kKT * q + lane;
                if (k < P.nz) {
                  Vox<T> v;
                  voxel<TAB>(P, a12, c, k, v);
                  if (on_detector(v))
                    slow_voxel<false, TAB, T>(P, c, v, T(0), nullptr, in, contrib[q]);
                }
              }
            }
          }
#pragma unroll
          for (int q = 0; q < VPL; ++q) {
            if (!TWO || seg.t0 == seg.t1) {
              acc0[cc * VPL + q] += wsum * contrib[q];
            } else {
              acc0[cc * VPL + q] += w0 * contrib[q];
              if constexpr (TWO) acc1[cc * VPL + q] += w1 * contrib[q];
            }
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < VPL; ++q) {
      const int k = kf + kKT * q + lane;
      if (k >= P.nz) continue;
#pragma unroll
      for (int cc = 0; cc < NCOL; ++cc) {
        const int col = wave * NCOL + cc;
        const int i = tm.bx * CX + col % CX, j = tm.by * CY + col / CX;
        if (i >= P.nx || j >= P.ny) continue;
        const int a = cc * VPL + q;
        const long idx = (long(k) * P.ny + j) * P.nx + i;
        T *d0 = vol + long(seg.t0) * P.vox_per_phase + idx;
        if (atomic_out) {
          if (acc0[a] != T(0)) unsafeAtomicAdd(d0, acc0[a]);
        } else {
          *d0 += acc0[a];
        }
        if (TWO && seg.t0 != seg.t1) {
          T *d1 = vol + long(seg.t1) * P.vox_per_phase + idx;
          if (atomic_out) {
            if (acc1[TWO ? a : 0] != T(0)) unsafeAtomicAdd(d1, acc1[TWO ? a : 0]);
          } else {
            *d1 += acc1[TWO ? a : 0];
          }
        }
      }
    }
  }
}

}  // namespace axial

// Tile shapes (columns along x, y; forward tiles span 64 k, backward tiles
// 64 k times bwd_tile_vpl) selectable per plan: forward one wave per
// workgroup, backward four waves.
constexpr int kFwdTiles[kNumFwdTiles][2] = {{8, 4}, {4, 4}, {8, 8}};
constexpr int kBwdTiles[kNumBwdTiles][2] = {{8, 8}, {8, 4}, {4, 4}, {8, 4}, {4, 4}, {8, 8}};

int axial_tiles(const DevPlan &P, int cx, int cy, int kt) {
  return ((P.nx + cx - 1) / cx) * ((P.ny + cy - 1) / cy) * ((P.nz + kt - 1) / kt);
}

int axial_tiles(const DevPlan &P, int job, const AxialTuning &tu) {
  if (job == 0)
    return axial_tiles(P, kFwdTiles[tu.fwd_tile][0], kFwdTiles[tu.fwd_tile][1], axial::kKT);
  return axial_tiles(P, kBwdTiles[tu.bwd_tile][0], kBwdTiles[tu.bwd_tile][1],
                     axial::kKT * bwd_tile_vpl(tu.bwd_tile));
}

// Scratch of the forward reduction for `chunk` views: one detector window
// per (view, tile), and the header area: one header per slot, then one
// tile count and list per (view, pixel block).
int axial_reduce_blocks(const DevPlan &P) {
  return ((P.nu + axial::kRedU - 1) / axial::kRedU) * ((P.nv + axial::kRedV - 1) / axial::kRedV);
}

size_t axial_hdr_bytes(const DevPlan &P, int chunk, const AxialTuning &tu) {
  const size_t slots = size_t(chunk) * size_t(axial_tiles(P, 0, tu));
  const size_t lists = size_t(chunk) * size_t(axial_reduce_blocks(P));
  return slots * sizeof(int4) + lists * sizeof(int) * (1 + size_t(axial::kBinCap));
}

size_t axial_scratch_bytes(const DevPlan &P, int chunk, const AxialTuning &tu) {
  const size_t slots = size_t(chunk) * size_t(axial_tiles(P, 0, tu));
  return slots * size_t(tu.fwd_window_bytes) + axial_hdr_bytes(P, chunk, tu);
}

template <bool CONE, int W, int CX, int CY, typename T>
void launch_fwd(const DevPlan &P, dim3 grid, int lds_bytes, hipStream_t st, const T *volk, T *out,
                int vpb, int wcap, int pmax, int vb, int ve, T *scratch, int4 *hdr, int plain,
                int flush, int ntiles, int xcds, bool two) {
#define TRTT_FWD(TB, TW)                                                                   \
  hipLaunchKernelGGL((axial::fwd<CONE, W, CX, CY, TB, TW, T>), grid, dim3(64), lds_bytes, st, P, \
                     volk, out, vpb, wcap, pmax, vb, ve, scratch, hdr, plain, flush, ntiles, xcds)
  if (!P.table) {
    if (two) TRTT_FWD(0, true); else TRTT_FWD(0, false);
  } else if (P.table_deg == 2) {
    if (two) TRTT_FWD(2, true); else TRTT_FWD(2, false);
  } else {
    if (two) TRTT_FWD(3, true); else TRTT_FWD(3, false);
  }
#undef TRTT_FWD
}

template <bool CONE, int W, int CX, int CY, int VPL, typename T>
void launch_bwd(const DevPlan &P, dim3 grid, int lds_bytes, hipStream_t st, const T *in, T *out,
                const int *seg_begin, int atomic_out, int lds_cap, int pmax, bool two) {
#define TRTT_BWD(TW, TB)                                                                        \
  hipLaunchKernelGGL((axial::bwd<CONE, W, CX, CY, TW, VPL, TB, T>), grid, dim3(axial::kThreads), \
                     lds_bytes, st, P, in, out, seg_begin, atomic_out, lds_cap, pmax)
  if (!P.table) {
    if (two) TRTT_BWD(true, 0); else TRTT_BWD(false, 0);
  } else if (P.table_deg == 2) {
    if (two) TRTT_BWD(true, 2); else TRTT_BWD(false, 2);
  } else {
    if (two) TRTT_BWD(true, 3); else TRTT_BWD(false, 3);
  }
#undef TRTT_BWD
}

template <bool CONE, int W, typename T>
hipError_t launch_axial_t(const DevPlan &P, int job, const void *in, void *out,
                          const int *seg_begin, int nrows, int views_per_block, int atomic_out,
                          int pmax, int pmax2, void *scratch, void *hdr, int chunk,
                          void *volk_buf, int plain, const AxialTuning &tu, hipEvent_t *ev,
                          const AxialOverlap *ov, int nbuf, int two, hipStream_t st) {
  // ev (optional): pairs of events around each main kernel launch
  int npair = 0;
  if (job == 0) {
    const int wcap = tu.fwd_window_bytes / int(sizeof(T));
    // window, W margin entries, profile with W spare entries (at least 128
    // entries past the window for the batched outer product)
    const int lds_bytes = (wcap + W + max(pmax + W, 128)) * int(sizeof(T));
    const unsigned ntiles = unsigned(axial_tiles(P, 0, tu));
    const int npt = axial_reduce_blocks(P);
    T *volk = static_cast<T *>(volk_buf);
    hipLaunchKernelGGL(axial::to_k_fastest<T>, dim3(((P.nx + 31) / 32) * P.ny, (P.nz + 63) / 64, P.nt),
                       dim3(axial::kThreads), 0, st, static_cast<const T *>(in), volk, P.nx, P.ny,
                       P.nz);
    const int nchunks = (P.nviews + chunk - 1) / chunk;
    const int flush = tu.fwd_flush;
    const bool overlap = !flush && ov && ov->aux && nbuf == 2 && nchunks > 1;
    const size_t win_elems = size_t(chunk) * ntiles * size_t(wcap);
    const size_t hdr_bytes = axial_hdr_bytes(P, chunk, tu);
    for (int c = 0, vb = 0; vb < P.nviews; vb += chunk, ++c) {
      const int ve = vb + chunk < P.nviews ? vb + chunk : P.nviews;
      const int ngroups = (ve - vb + views_per_block - 1) / views_per_block;
      dim3 grid(ntiles, ngroups);
      // XCD-aware order (atomic flush): units of one view group and one k
      // block of tiles, dealt to the XCDs in turn; blocks past the last unit
      // return at once
      const int xcds = flush && tu.fwd_xcds > 1 ? tu.fwd_xcds : 0;
      if (xcds) {
        const long nbz = (P.nz + axial::kKT - 1) / axial::kKT;
        const long units = long(ngroups) * nbz;
        grid = dim3(unsigned(long(xcds) * (long(ntiles) / nbz) * ((units + xcds - 1) / xcds)));
      }
      const int b = overlap ? (c & 1) : 0;
      auto *o = static_cast<T *>(out);
      auto *sc = static_cast<T *>(scratch) + b * win_elems;
      auto *hd = reinterpret_cast<int4 *>(static_cast<char *>(hdr) + b * hdr_bytes);
      // buffer b is free once the reduction of chunk c - 2 has read it
      if (overlap && c >= 2) hipStreamWaitEvent(st, ov->red_done[b], 0);
      if (ev) hipEventRecord(ev[2 * npair], st);
      switch (tu.fwd_tile) {
        case 1:
          launch_fwd<CONE, W, 4, 4, T>(P, grid, lds_bytes, st, volk, o, views_per_block, wcap, pmax,
                                       vb, ve, sc, hd, plain, flush, int(ntiles), xcds, two != 0);
          break;
        case 2:
          launch_fwd<CONE, W, 8, 8, T>(P, grid, lds_bytes, st, volk, o, views_per_block, wcap, pmax,
                                       vb, ve, sc, hd, plain, flush, int(ntiles), xcds, two != 0);
          break;
        default:
          launch_fwd<CONE, W, 8, 4, T>(P, grid, lds_bytes, st, volk, o, views_per_block, wcap, pmax,
                                       vb, ve, sc, hd, plain, flush, int(ntiles), xcds, two != 0);
      }
      if (ev) hipEventRecord(ev[2 * npair + 1], st);
      ++npair;
      if (flush) continue;  // the windows are in the projections already
      hipStream_t rs = st;
      if (overlap) {
        hipEventRecord(ov->fwd_done[b], st);
        hipStreamWaitEvent(ov->aux, ov->fwd_done[b], 0);
        rs = ov->aux;
      }
      // bins after the chunk's headers (the header area holds `chunk` views)
      const int nb = axial_reduce_blocks(P);
      int *count = reinterpret_cast<int *>(hd + long(chunk) * ntiles);
      int *list = count + long(chunk) * nb;
      const long nc = long(ve - vb) * nb;
      hipLaunchKernelGGL(axial::zero_ints, dim3(unsigned((nc + axial::kThreads - 1) / axial::kThreads)),
                         dim3(axial::kThreads), 0, rs, count, nc);
      const long nh = long(ve - vb) * ntiles;
      hipLaunchKernelGGL(axial::bin_windows,
                         dim3(unsigned((nh + axial::kThreads - 1) / axial::kThreads)),
                         dim3(axial::kThreads), 0, rs, static_cast<const int4 *>(hd), int(ntiles),
                         ve - vb, (P.nu + axial::kRedU - 1) / axial::kRedU,
                         (P.nv + axial::kRedV - 1) / axial::kRedV, count, list);
      hipLaunchKernelGGL(axial::fwd_reduce<T>, dim3(npt, ve - vb), dim3(axial::kThreads), 0, rs,
                         P, static_cast<const T *>(sc), static_cast<const int4 *>(hd), count, list,
                         o, vb, int(ntiles), wcap);
      if (overlap) hipEventRecord(ov->red_done[b], ov->aux);
    }
    if (overlap) {
      // join: the launch stream resumes after both buffers' last reductions
      hipStreamWaitEvent(st, ov->red_done[0], 0);
      hipStreamWaitEvent(st, ov->red_done[1], 0);
    }
  } else {
    const int vpl = bwd_tile_vpl(tu.bwd_tile);
    if (vpl == 2 && pmax2 <= 0) return hipErrorInvalidValue;
    const int pm = vpl == 2 ? pmax2 : pmax;
    const int lds_cap = tu.bwd_window_bytes / int(sizeof(T));
    // window, one Q row per wave with W zero entries on each side; at least
    // 128 VPL entries past the window for the batched Q rows
    const int lds_bytes =
        tu.bwd_window_bytes + max(axial::kWaves * (pm + 2 * W), 128 * vpl) * int(sizeof(T));
    dim3 grid(unsigned(axial_tiles(P, 1, tu)), nrows);
    const auto *i = static_cast<const T *>(in);
    auto *o = static_cast<T *>(out);
    if (ev) hipEventRecord(ev[0], st);
    switch (tu.bwd_tile) {
      case 1:
        launch_bwd<CONE, W, 8, 4, 1, T>(P, grid, lds_bytes, st, i, o, seg_begin, atomic_out,
                                        lds_cap, pm, two != 0);
        break;
      case 2:
        launch_bwd<CONE, W, 4, 4, 1, T>(P, grid, lds_bytes, st, i, o, seg_begin, atomic_out,
                                        lds_cap, pm, two != 0);
        break;
      case 3:
        launch_bwd<CONE, W, 8, 4, 2, T>(P, grid, lds_bytes, st, i, o, seg_begin, atomic_out,
                                        lds_cap, pm, two != 0);
        break;
      case 4:
        launch_bwd<CONE, W, 4, 4, 2, T>(P, grid, lds_bytes, st, i, o, seg_begin, atomic_out,
                                        lds_cap, pm, two != 0);
        break;
      case 5:
        launch_bwd<CONE, W, 8, 8, 2, T>(P, grid, lds_bytes, st, i, o, seg_begin, atomic_out,
                                        lds_cap, pm, two != 0);
        break;
      default:
        launch_bwd<CONE, W, 8, 8, 1, T>(P, grid, lds_bytes, st, i, o, seg_begin, atomic_out,
                                        lds_cap, pm, two != 0);
    }
    if (ev) hipEventRecord(ev[1], st);
  }
  return hipGetLastError();
}

hipError_t launch_axial(bool cone, int w, bool f64, const DevPlan &P, int job, const void *in,
                        void *out, const int *seg_begin, int nrows, int views_per_block,
                        int atomic_out, int pmax, int pmax2, void *scratch, void *hdr, int chunk,
                        void *volk, int plain, const AxialTuning &tu, hipEvent_t *ev,
                        const AxialOverlap *ov, int nbuf, int two, hipStream_t st) {
  if (tu.fwd_tile < 0 || tu.fwd_tile >= kNumFwdTiles || tu.bwd_tile < 0 ||
      tu.bwd_tile >= kNumBwdTiles)
    return hipErrorInvalidValue;
#define AX_CASE(C, WW, TT)                                                                   \
  if (cone == C && w == WW && f64 == (sizeof(TT) == 8))                                     \
    return launch_axial_t<C, WW, TT>(P, job, in, out, seg_begin, nrows, views_per_block,   \
                                     atomic_out, pmax, pmax2, scratch, hdr, chunk, volk, plain, \
                                     tu, ev, ov, nbuf, two, st);
  AX_CASE(true, 9, float)
  AX_CASE(true, 13, float)
  AX_CASE(false, 9, float)
  AX_CASE(false, 13, float)
#undef AX_CASE
  return hipErrorInvalidValue;
}

}  // namespace trtt
