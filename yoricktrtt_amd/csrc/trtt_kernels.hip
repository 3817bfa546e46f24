// gfx950 kernels of the B-spline footprint projectors and of the Long &
// Fessler separable-footprint projector.
//
// Work decomposition.  A workgroup of 256 threads owns a tile of voxels and
// walks a list of views.  For each view it
//   1. evaluates the footprint of each of its voxels (FP64 detector
//      geometry in the reference's operation order, so windows and
//      tabulated-CDF indices are those of the C plugin),
//   2. reduces the union of the voxel windows over the tile (wave shuffles
//      plus one LDS exchange),
//   3. forward: scatters separable weights into that detector window held in
//      LDS, then flushes the window to HBM with one atomic add per non-zero
//      pixel; backward: stages the window from HBM into LDS with coalesced
//      row loads and gathers from LDS.
// A window larger than the LDS budget falls back to direct global atomics /
// loads for that (tile, view).
#include <algorithm>
#include <climits>

#include "trtt_core.h"

namespace trtt {

extern __shared__ __align__(16) unsigned char g_smem[];

constexpr int kThreads = 256;
constexpr int kMaxW = 12;          // weights held in registers per chunk

// ---------------------------------------------------------------------------
// Closed-form cumulative integral of the centred cubic B-spline
// (trtt_3D_projectors_cubic.c:117-138): F(x) = 1/2 + sign(x) G(|x|),
// G(t) = t(2/3 + t^2(t/8 - 1/3)) on [0,1], 1/2 - (2-t)^4/24 on [1,2].
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ __forceinline__ T cdf3(T x) {
  T t = fabs(x);
  t = t < T(2) ? t : T(2);
  const T t2 = t * t;
  const T g1 = t * (T(2.0 / 3.0) + t2 * (t * T(0.125) - T(1.0 / 3.0)));
  const T s = T(2) - t;
  const T s2 = s * s;
  const T g2 = T(0.5) - s2 * s2 * T(1.0 / 24.0);
  const T g = t > T(1) ? g2 : g1;
  return x < T(0) ? T(0.5) - g : T(0.5) + g;
}

// Footprint of one voxel on one view.
struct Foot {
  double uc, vc;       // projected centre (pixel-bound coordinates)
  double hu, hv;       // footprint scales (pixels)
  double su, sv;       // normalisations (1/h, or 1/(h*step) for 2-D tables)
  double rho0;         // amplitude
  int ul, ur, vl, vr;  // clamped window [ul,ur) x [vl,vr)
};

// Window clamp of trtt_3D_projectors.c:1293-1307; the 2-D closed-form
// projectors clamp the right bound to n-1 (trtt_2D_projectors_cubic.c:256-258).
template <bool QUIRK>
__host__ __device__ __forceinline__ void clamp_window(double c, double hs, int n, int &lo, int &hi) {
#pragma clang fp contract(off)
  const double l = c - hs, r = c + hs;
  const double nm1 = double(n - 1);
  lo = l < 0.0 ? 0 : (l <= nm1 ? int(floor(l)) : n - 1);
  if (!QUIRK) {
    hi = r < 0.0 ? 0 : (r <= nm1 ? int(floor(r)) + 1 : n);
  } else {
    hi = r < 0.0 ? 0 : (r <= double(n - 2) ? int(floor(r)) + 1 : n - 1);
  }
}

// Per-voxel footprint (FP64).  Returns false when the voxel lies in the
// source plane (the reference's TRTT_FAILURE case).
template <int KIND, bool TABLE, bool QUIRK>
__host__ __device__ __forceinline__ bool voxel_footprint(const DevPlan &P, const ViewConst &g, int i, int j,
                                                int k, Foot &f) {
#pragma clang fp contract(off)
  const double di = double(i), dj = double(j), dk = double(k);
  constexpr bool CONE = (KIND == CB3D || KIND == FB2D);
  constexpr bool IS2D = (KIND == PB2D || KIND == FB2D);
  if (CONE) {
    const double wk = ((g.o3 * dk + g.o4) + g.o2 * dj) + g.o1 * di;
    const double num_v = ((g.nvz * dk + g.nvc) + g.nvy * dj) + g.nvx * di;
    const double den = ((g.dz * dk + g.dc) + g.dy * dj) + g.dx * di;
    const double q = g.a4m / den;
    f.vc = q * num_v + g.b3;
    if (!(wk != 0.0)) return false;
    const double magn = g.a4m / wk;
    const double vca8 = (f.vc - g.a8) * P.v_scl;
    const double vca8sqr = vca8 * vca8;
    const double dv = sqrt(1.0 + (vca8sqr / g.a4sqr));
    if (!IS2D) {
      const double num_u = ((g.nuz * dk + g.nuc) + g.nuy * dj) + g.nux * di;
      f.uc = q * num_u + g.c3;
      const double uca12 = (f.uc - g.a12) * P.u_scl;
      const double du = sqrt(1.0 + ((uca12 * uca12) / (g.a4sqr + vca8sqr)));
      f.rho0 = ((P.rho_init * (magn * magn)) * du) * dv;
      f.hu = (P.hu_init * magn) * du;
      f.su = 1.0 / f.hu;
      f.hv = (P.hv_init * magn) * dv;
      f.sv = 1.0 / f.hv;
      clamp_window<false>(f.uc, P.hs0 * f.hu, P.nu, f.ul, f.ur);
      clamp_window<false>(f.vc, P.hs0 * f.hv, P.nv, f.vl, f.vr);
    } else {
      f.uc = 0.5;
      f.rho0 = (P.rho_init * magn) * dv;
      f.hv = (P.hv_init * magn) * dv;
      f.sv = TABLE ? 1 / (f.hv * P.step) : 1.0 / f.hv;
      f.hu = 1.0;
      f.su = 1.0;
      f.ul = 0;
      f.ur = 1;
      clamp_window<QUIRK>(f.vc, P.hs0 * f.hv, P.nv, f.vl, f.vr);
    }
  } else {
    f.vc = ((g.nvz * dk + g.nvc) + g.nvy * dj) + g.nvx * di;
    f.rho0 = g.pb_rho0;
    f.hv = g.pb_hv;
    f.sv = g.pb_sv;
    if (!IS2D) {
      f.uc = ((g.nuz * dk + g.nuc) + g.nuy * dj) + g.nux * di;
      f.hu = g.pb_hu;
      f.su = g.pb_su;
      clamp_window<false>(f.uc, g.pb_hsu, P.nu, f.ul, f.ur);
      clamp_window<false>(f.vc, g.pb_hsv, P.nv, f.vl, f.vr);
    } else {
      f.uc = 0.5;
      f.hu = 1.0;
      f.su = 1.0;
      f.ul = 0;
      f.ur = 1;
      clamp_window<QUIRK>(f.vc, g.pb_hsv, P.nv, f.vl, f.vr);
    }
  }
  return true;
}

__host__ __device__ __forceinline__ long table_index(double x, long m) {
  double c = x >= 0.0 ? x : 0.0;
  c = c <= double(m - 1) ? c : double(m - 1);
  return long(floor(c));
}

// Streaming evaluator of one footprint axis: value(l) = F(x_{l+1}) - F(x_l)
// over pixels [lo, lo+n).  Closed form: x_l = (lo-c)*s + l*s in T.  Table:
// the reference's FP64 start expressions (trtt_3D_projectors.c:275-302 for
// 3-D, trtt_2D_projectors.c:92-108 for 2-D) and floor() lookups.
template <typename T, bool TABLE, bool TAB2D>
struct AxisFoot {
  double x0, s;  // table coordinate start and step (FP64)
  T xf0, sf;     // closed-form start and step
  T prev;
  const T *cum;
  long m;

  __host__ __device__ __forceinline__ void init(const DevPlan &P, int lo, double c, double h,
                                                double scl) {
#pragma clang fp contract(off)
    if (TABLE) {
      cum = static_cast<const T *>(P.table);
      m = P.m;
      if (TAB2D) {
        s = scl;
        x0 = (double(lo) - (c + P.coord0 * h)) * s;
      } else {
        s = scl / P.step;
        x0 = ((double(lo) - c) - h * P.coord0) * s;
      }
      prev = cum[table_index(x0, m)];
    } else {
      xf0 = T((double(lo) - c) * scl);
      sf = T(scl);
      prev = cdf3<T>(xf0);
    }
  }
  // weight of pixel lo+l (must be called with l = 0, 1, 2, ... in order)
  __host__ __device__ __forceinline__ T next(int l) {
#pragma clang fp contract(off)
    T nxt;
    if (TABLE) {
      nxt = cum[table_index(x0 + double(l + 1) * s, m)];
    } else {
      nxt = cdf3<T>(xf0 + T(l + 1) * sf);
    }
    const T w = nxt - prev;
    prev = nxt;
    return w;
  }
};

__device__ __forceinline__ int wave_min(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
  return x;
}

// Union of the voxel windows of the workgroup.  `red` holds 4 ints per wave.
__device__ __forceinline__ void block_window(int &ul, int &ur, int &vl, int &vr, int *red) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  ul = wave_min(ul);
  vl = wave_min(vl);
  ur = wave_max(ur);
  vr = wave_max(vr);
  if (lane == 0) {
    red[4 * w + 0] = ul;
    red[4 * w + 1] = vl;
    red[4 * w + 2] = ur;
    red[4 * w + 3] = vr;
  }
  __syncthreads();
  ul = red[0];
  vl = red[1];
  ur = red[2];
  vr = red[3];
#pragma unroll
  for (int q = 1; q < kThreads / 64; ++q) {
    ul = min(ul, red[4 * q + 0]);
    vl = min(vl, red[4 * q + 1]);
    ur = max(ur, red[4 * q + 2]);
    vr = max(vr, red[4 * q + 3]);
  }
}

template <typename T>
__device__ __forceinline__ void atomic_add(T *p, T v) {
  unsafeAtomicAdd(p, v);
}

// Tile of TX x TY x TZ voxels; each of the 256 threads owns NPT voxels,
// stacked along k, or along j for single-slice tiles larger than the
// workgroup (the per-view window set-up then serves several pixels).
template <int TX, int TY, int TZ>
struct Tile {
  static constexpr int NPT = (TX * TY * TZ) / kThreads;
  static constexpr bool ALONG_J = TZ == 1 && TX * TY > kThreads;
  static constexpr int KSTEP = ALONG_J ? 0 : kThreads / (TX * TY);
  static constexpr int JSTEP = ALONG_J ? kThreads / TX : 0;
  int i, j, k0;
  int i0, j0, kb;  // tile origin
  __device__ __forceinline__ void init(const DevPlan &P) {
    const int ntx = (P.nx + TX - 1) / TX, nty = (P.ny + TY - 1) / TY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty;
    const int bz = b / nty;
    const int lid = threadIdx.x;
    i0 = bx * TX;
    j0 = by * TY;
    kb = bz * TZ;
    i = bx * TX + lid % TX;
    if (ALONG_J) {
      j = by * TY + lid / TX;
      k0 = 0;
    } else {
      j = by * TY + (lid / TX) % TY;
      k0 = bz * TZ + lid / (TX * TY);
    }
  }
  // coordinates of the thread's voxel q
  __device__ __forceinline__ int vj(int q) const { return ALONG_J ? j + q * JSTEP : j; }
  __device__ __forceinline__ int vk(int q) const { return ALONG_J ? k0 : k0 + q * KSTEP; }
  // Square single-slice tiles, laid out per view: pixel q of the thread
  // sits at offset lid % TX across the rays and (lid / TX) NPT + q along
  // them; the across axis is i, or j when flip is set.
  __device__ __forceinline__ int ai(int q, bool flip) const {
    return flip ? i0 + int(threadIdx.x / TX) * NPT + q : i;
  }
  __device__ __forceinline__ int aj(int q, bool flip) const {
    return flip ? j0 + int(threadIdx.x % TX) : j0 + int(threadIdx.x / TX) * NPT + q;
  }
  // Cubic 3-D tiles, laid out per view: a wave's lanes cover TX offsets
  // across the rays times TZ slices, and the thread's NPT voxels follow one
  // another along the rays; the across axis is i, or j when flip is set.
  static constexpr bool CUBE = !ALONG_J && TX == TY && TX * TZ == 64 && TX * TY * TZ > kThreads;
  __device__ __forceinline__ void cube(int q, bool flip, int &vi, int &vjj, int &vkk) const {
    const int a = int(threadIdx.x % TX);
    const int b = int(threadIdx.x / (TX * TZ)) * NPT + q;
    vkk = kb + int(threadIdx.x / TX) % TZ;
    vi = i0 + (flip ? b : a);
    vjj = j0 + (flip ? a : b);
  }
};

// True when the detector coordinate across the rays (v of a 2-D view, u of
// a 3-D one) varies faster along the image's j axis than along i near
// voxel (ic, jc, kc): the derivatives of the footprint centre of
// voxel_footprint, up to a common factor.
template <int KIND>
__device__ __forceinline__ bool detector_along_j(const ViewConst &g, double ic, double jc,
                                                 double kc) {
  constexpr bool IS2D = (KIND == PB2D || KIND == FB2D);
  const double nx = IS2D ? g.nvx : g.nux, ny = IS2D ? g.nvy : g.nuy;
  const double nz = IS2D ? g.nvz : g.nuz, nc = IS2D ? g.nvc : g.nuc;
  double gi = nx, gj = ny;
  if (KIND == FB2D || KIND == CB3D) {
    const double num = ((nz * kc + nc) + ny * jc) + nx * ic;
    const double den = ((g.dz * kc + g.dc) + g.dy * jc) + g.dx * ic;
    gi = nx * den - num * g.dx;
    gj = ny * den - num * g.dy;
  }
  return fabs(gj) > fabs(gi);
}

// ---------------------------------------------------------------------------
// Forward: sino[view] += A_view (w0 x[t0] + w1 x[t1])
// ---------------------------------------------------------------------------
template <typename T, int KIND, bool TABLE, bool QUIRK, int TX, int TY, int TZ>
__global__ __launch_bounds__(kThreads) void fp_forward(DevPlan P, const T *__restrict__ vol,
                                                       T *__restrict__ sino, int views_per_block,
                                                       int lds_cap) {
  constexpr bool IS2D = (KIND == PB2D || KIND == FB2D);
  using TL = Tile<TX, TY, TZ>;
  // The LDS atomic adds into the window bound this kernel.  The window
  // holds FP64 sums whatever T is: on CDNA4 an LDS atomic add of FP64 runs
  // several times faster than one of FP32 (configuration-2 sizes, FP32
  // plan: 5.2 ms with an FP32 window, 1.9 ms with an FP64 one).  Square 2-D
  // tiles and cubic 3-D tiles choose per view whether a wave's lanes run
  // along i or along j (times k in 3-D): across the rays, so that the lanes'
  // footprints start on different detector cells (adds to one address
  // serialise).
  constexpr bool FLIP = IS2D && TL::ALONG_J && TX == TY;
  constexpr bool CUBE = TL::CUBE && !IS2D;
  using W = double;
  W *reg = reinterpret_cast<W *>(g_smem);
  __shared__ int red[4 * (kThreads / 64)];
  TL tile;
  tile.init(P);
  const long npix = long(P.nu) * P.nv;

  for (int vv = 0; vv < views_per_block; ++vv) {
    const int view = blockIdx.y * views_per_block + vv;
    if (view >= P.nviews) break;
    const ViewConst &g = P.views[view];
    T *out = sino + long(view) * npix;
    const bool flip = (FLIP || CUBE) &&
                      detector_along_j<KIND>(g, tile.i0 + 0.5 * (TX - 1), tile.j0 + 0.5 * (TY - 1),
                                             tile.kb + 0.5 * (TZ - 1));

    Foot f[TL::NPT];
    T val[TL::NPT];
    bool act[TL::NPT];
    int ul = INT_MAX, vl = INT_MAX, ur = INT_MIN, vr = INT_MIN;
#pragma unroll
    for (int q = 0; q < TL::NPT; ++q) {
      int iq, jq, k;
      if (CUBE) {
        tile.cube(q, flip, iq, jq, k);
      } else {
        k = tile.vk(q);
        iq = FLIP ? tile.ai(q, flip) : tile.i;
        jq = FLIP ? tile.aj(q, flip) : tile.vj(q);
      }
      act[q] = iq < P.nx && jq < P.ny && k < P.nz;
      val[q] = T(0);
      if (act[q]) {
        if (!voxel_footprint<KIND, TABLE, QUIRK>(P, g, iq, jq, k, f[q])) {
          atomicOr(P.err, 1);
          act[q] = false;
          continue;
        }
        const long idx = (long(k) * P.ny + jq) * P.nx + iq;
        T x = T(g.w0) * vol[long(g.t0) * P.vox_per_phase + idx];
        if (g.w1 != 0.0) x += T(g.w1) * vol[long(g.t1) * P.vox_per_phase + idx];
        val[q] = x * T(f[q].rho0);
        act[q] = f[q].ul < f[q].ur && f[q].vl < f[q].vr;
        if (act[q]) {
          ul = min(ul, f[q].ul);
          ur = max(ur, f[q].ur);
          vl = min(vl, f[q].vl);
          vr = max(vr, f[q].vr);
        }
      }
    }
    block_window(ul, ur, vl, vr, red);
    const bool any = ur > ul && vr > vl;
    const int RW = any ? ur - ul : 0, RH = any ? vr - vl : 0;
    const bool in_lds = any && long(RW) * RH <= lds_cap;
    if (in_lds) {
      for (int idx = threadIdx.x; idx < RW * RH; idx += kThreads) reg[idx] = W(0);
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < TL::NPT; ++q) {
      if (!act[q]) continue;
      const Foot &F = f[q];
      const int npu = F.ur - F.ul, npv = F.vr - F.vl;
      for (int l0 = 0; l0 < npu; l0 += kMaxW) {
        T wu[kMaxW];
        AxisFoot<T, TABLE, IS2D> au;
        if (!IS2D) {
          au.init(P, F.ul + l0, F.uc, F.hu, F.su);
#pragma unroll
          for (int l = 0; l < kMaxW; ++l) wu[l] = (l0 + l < npu) ? au.next(l) : T(0);
        } else {
#pragma unroll
          for (int l = 0; l < kMaxW; ++l) wu[l] = l == 0 ? T(1) : T(0);
        }
        AxisFoot<T, TABLE, IS2D> av;
        av.init(P, F.vl, F.vc, F.hv, F.sv);
        for (int mm = 0; mm < npv; ++mm) {
          const T c = av.next(mm) * val[q];
          if (in_lds) {
            W *row = reg + (F.vl + mm - vl) * RW + (F.ul + l0 - ul);
#pragma unroll
            for (int l = 0; l < kMaxW; ++l)
              if (l0 + l < npu) atomicAdd(row + l, W(c * wu[l]));
          } else {
            T *row = out + long(F.vl + mm) * P.nu + (F.ul + l0);
#pragma unroll
            for (int l = 0; l < kMaxW; ++l)
              if (l0 + l < npu) atomic_add(row + l, c * wu[l]);
          }
        }
      }
    }
    if (in_lds) {
      __syncthreads();
      for (int idx = threadIdx.x; idx < RW * RH; idx += kThreads) {
        const W v = reg[idx];
        if (v != W(0)) {
          const int r = idx / RW, c = idx - r * RW;
          atomic_add(out + long(vl + r) * P.nu + (ul + c), T(v));
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Backward: x[t0] += w0 A' y, x[t1] += w1 A' y over the views of the
// segments [seg_begin[blockIdx.y], seg_begin[blockIdx.y+1]).
// ---------------------------------------------------------------------------
template <typename T, int KIND, bool TABLE, bool QUIRK, int TX, int TY, int TZ>
__global__ __launch_bounds__(kThreads) void fp_backward(DevPlan P, const T *__restrict__ sino,
                                                        T *__restrict__ vol,
                                                        const int *__restrict__ seg_begin,
                                                        int atomic_out, int lds_cap) {
  constexpr bool IS2D = (KIND == PB2D || KIND == FB2D);
  using TL = Tile<TX, TY, TZ>;
  T *reg = reinterpret_cast<T *>(g_smem);
  __shared__ int red[4 * (kThreads / 64)];
  TL tile;
  tile.init(P);
  const long npix = long(P.nu) * P.nv;
  const int s_begin = seg_begin[blockIdx.y], s_end = seg_begin[blockIdx.y + 1];

  for (int s = s_begin; s < s_end; ++s) {
    const Segment seg = P.segs[s];
    T acc0[TL::NPT], acc1[TL::NPT];
#pragma unroll
    for (int q = 0; q < TL::NPT; ++q) acc0[q] = acc1[q] = T(0);

    for (int n = 0; n < seg.count; ++n) {
      const int view = P.order[seg.start + n];
      const ViewConst &g = P.views[view];
      const T *in = sino + long(view) * npix;
      Foot f[TL::NPT];
      bool act[TL::NPT];
      int ul = INT_MAX, vl = INT_MAX, ur = INT_MIN, vr = INT_MIN;
#pragma unroll
      for (int q = 0; q < TL::NPT; ++q) {
        const int k = tile.vk(q), jq = tile.vj(q);
        act[q] = tile.i < P.nx && jq < P.ny && k < P.nz;
        if (act[q]) {
          if (!voxel_footprint<KIND, TABLE, QUIRK>(P, g, tile.i, jq, k, f[q])) {
            atomicOr(P.err, 1);
            act[q] = false;
            continue;
          }
          act[q] = f[q].ul < f[q].ur && f[q].vl < f[q].vr;
          if (act[q]) {
            ul = min(ul, f[q].ul);
            ur = max(ur, f[q].ur);
            vl = min(vl, f[q].vl);
            vr = max(vr, f[q].vr);
          }
        }
      }
      block_window(ul, ur, vl, vr, red);
      const bool any = ur > ul && vr > vl;
      const int RW = any ? ur - ul : 0, RH = any ? vr - vl : 0;
      const bool in_lds = any && long(RW) * RH <= lds_cap;
      if (in_lds) {
        for (int idx = threadIdx.x; idx < RW * RH; idx += kThreads) {
          const int r = idx / RW, c = idx - r * RW;
          reg[idx] = in[long(vl + r) * P.nu + (ul + c)];
        }
        __syncthreads();
      }
#pragma unroll
      for (int q = 0; q < TL::NPT; ++q) {
        if (!act[q]) continue;
        const Foot &F = f[q];
        const int npu = F.ur - F.ul, npv = F.vr - F.vl;
        T sum = T(0);
        for (int l0 = 0; l0 < npu; l0 += kMaxW) {
          T wu[kMaxW];
          AxisFoot<T, TABLE, IS2D> au;
          if (!IS2D) {
            au.init(P, F.ul + l0, F.uc, F.hu, F.su);
#pragma unroll
            for (int l = 0; l < kMaxW; ++l) wu[l] = (l0 + l < npu) ? au.next(l) : T(0);
          } else {
#pragma unroll
            for (int l = 0; l < kMaxW; ++l) wu[l] = l == 0 ? T(1) : T(0);
          }
          AxisFoot<T, TABLE, IS2D> av;
          av.init(P, F.vl, F.vc, F.hv, F.sv);
          for (int mm = 0; mm < npv; ++mm) {
            const T c = av.next(mm);
            T rs = T(0);
            if (in_lds) {
              const T *row = reg + (F.vl + mm - vl) * RW + (F.ul + l0 - ul);
#pragma unroll
              for (int l = 0; l < kMaxW; ++l)
                if (l0 + l < npu) rs += wu[l] * row[l];
            } else {
              const T *row = in + long(F.vl + mm) * P.nu + (F.ul + l0);
#pragma unroll
              for (int l = 0; l < kMaxW; ++l)
                if (l0 + l < npu) rs += wu[l] * row[l];
            }
            sum += c * rs;
          }
        }
        const T contrib = sum * T(F.rho0);
        if (seg.t0 == seg.t1) {
          acc0[q] += T(g.w0 + g.w1) * contrib;
        } else {
          acc0[q] += T(g.w0) * contrib;
          acc1[q] += T(g.w1) * contrib;
        }
      }
      __syncthreads();
    }
    // flush this segment's accumulators (each voxel is owned by one thread)
#pragma unroll
    for (int q = 0; q < TL::NPT; ++q) {
      const int k = tile.vk(q), jq = tile.vj(q);
      if (!(tile.i < P.nx && jq < P.ny && k < P.nz)) continue;
      const long idx = (long(k) * P.ny + jq) * P.nx + tile.i;
      T *d0 = vol + long(seg.t0) * P.vox_per_phase + idx;
      if (atomic_out) {
        if (acc0[q] != T(0)) atomic_add(d0, acc0[q]);
      } else {
        *d0 += acc0[q];
      }
      if (seg.t0 != seg.t1) {
        T *d1 = vol + long(seg.t1) * P.vox_per_phase + idx;
        if (atomic_out) {
          if (acc1[q] != T(0)) atomic_add(d1, acc1[q]);
        } else {
          *d1 += acc1[q];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Long & Fessler separable footprint (LongFessler2D.c:41-141, :254-355)
// ---------------------------------------------------------------------------
struct LFFoot {
  double t[4];
  double amp;       // sqrt(1 + q^2) voxelSize pixelSize: the trapezoid's integral model
  double amp_line;  // sqrt(1 + q^2) voxelSize: the point-sampled (line-integral) model
  int kf, kl;       // inclusive bin range
};

__host__ __device__ __forceinline__ double lf_ramp(double a, double b, double c) {
  return a < b ? (b + a) * (b - a) / (2.0 * c) : 0.0;
}
__host__ __device__ __forceinline__ double lf_box(double a, double b) { return a < b ? b - a : 0.0; }

__host__ __device__ __forceinline__ void lf_footprint(const DevPlan &P, const ViewConst &g, int i1,
                                                      int i2, LFFoot &f) {
#pragma clang fp contract(off)
  const double vs = P.lf_vs;
  const double y = g.lf_y0 + double(i2) * vs, ya = y - 0.5 * vs, yb = y + 0.5 * vs;
  const double x = g.lf_x0 + double(i1) * vs, xa = x - 0.5 * vs, xb = x + 0.5 * vs;
  const double cb = g.lf_cb, sb = g.lf_sb, R = P.lf_R;
  const double cx[4] = {xa, xb, xa, xb}, cy[4] = {ya, ya, yb, yb};
#pragma unroll
  for (int c = 0; c < 4; ++c)
    f.t[c] = P.lf_scale * (cy[c] * sb + cx[c] * cb) / (R - (cy[c] * cb - cx[c] * sb)) + g.lf_off;
  // the 5-comparator network of LongFessler2D.c:103-109
#define LF_SORT(a, b)                 \
  if (f.t[a] > f.t[b]) {              \
    const double tmp = f.t[a];        \
    f.t[a] = f.t[b];                  \
    f.t[b] = tmp;                     \
  }
  LF_SORT(0, 1) LF_SORT(2, 3) LF_SORT(0, 2) LF_SORT(1, 3) LF_SORT(1, 2)
#undef LF_SORT
  const double qx = x + R * sb, qy = y - R * cb;
  const double q = fabs(qy) > fabs(qx) ? qx / qy : qy / qx;
  f.amp = sqrt(1.0 + q * q) * (vs * P.lf_ps);
  f.amp_line = sqrt(1.0 + q * q) * vs;
  const double lo = floor(f.t[0] - 0.5 * P.lf_size);
  const double hi = ceil(f.t[3] + 0.5 * P.lf_size);
  f.kf = lo < 0.0 ? 0 : (lo > double(P.nv) ? P.nv : int(lo));
  f.kl = hi >= double(P.nv) ? P.nv - 1 : (hi < -1.0 ? -1 : int(hi));
}

__host__ __device__ __forceinline__ double lf_gamma(const LFFoot &f, int k, double size) {
#pragma clang fp contract(off)
  const double s1 = double(k) - 0.5 * size, s2 = s1 + size;
  const double *t = f.t;
  return lf_ramp((s1 >= t[0] ? s1 : t[0]) - t[0], (s2 <= t[1] ? s2 : t[1]) - t[0], t[1] - t[0]) +
         lf_box(s1 >= t[1] ? s1 : t[1], s2 <= t[2] ? s2 : t[2]) +
         lf_ramp(t[3] - (s2 <= t[3] ? s2 : t[3]), t[3] - (s1 >= t[2] ? s1 : t[2]), t[3] - t[2]);
}

// Line-integral model (lf2d_line_integ_direct / _adjoint / _sparser,
// LongFessler2D.c:143-252, 357-470, 573-681): the trapezoid is sampled at
// the bin centre k instead of integrated over the bin; a times its height
// there, with the reference's expression order.
__host__ __device__ __forceinline__ double lf_point(const LFFoot &f, int k, double a) {
#pragma clang fp contract(off)
  const double *t = f.t;
  const double x = double(k);
  if (x >= t[0] && x < t[1]) return a * (x - t[0]) / (t[1] - t[0]);
  if (x >= t[1] && x < t[2]) return a;
  if (x >= t[2] && x < t[3]) return a * (x - t[3]) / (t[2] - t[3]);
  return 0.0;
}

// True when the detector coordinate of a Long & Fessler view varies faster
// along the image's y axis (i2) than along x (i1) at (x, y): the
// derivatives of lf_footprint's corner projection, up to a common factor.
__device__ __forceinline__ bool lf_detector_along_y(const ViewConst &g, double R, double x,
                                                    double y) {
  const double cb = g.lf_cb, sb = g.lf_sb;
  const double num = y * sb + x * cb, den = R - (y * cb - x * sb);
  return fabs(sb * den + num * cb) > fabs(cb * den - num * sb);
}

// Square tiles laid out per view and an FP64 LDS window, as in
// fp_forward's 2-D path.
template <typename T, bool LINE>
__global__ __launch_bounds__(kThreads) void lf_forward(DevPlan P, const T *__restrict__ img,
                                                       T *__restrict__ sino, int views_per_block,
                                                       int lds_cap) {
  constexpr int S = kTile2DSquare;
  using TL = Tile<S, S, 1>;
  double *reg = reinterpret_cast<double *>(g_smem);
  __shared__ int red[4 * (kThreads / 64)];
  TL tile;
  tile.init(P);
  for (int vv = 0; vv < views_per_block; ++vv) {
    const int view = blockIdx.y * views_per_block + vv;
    if (view >= P.nviews) break;
    const ViewConst &g = P.views[view];
    T *out = sino + long(view) * P.nv;
    const double xc = g.lf_x0 + (tile.i0 + 0.5 * (S - 1)) * P.lf_vs;
    const double yc = g.lf_y0 + (tile.j0 + 0.5 * (S - 1)) * P.lf_vs;
    const bool flip = lf_detector_along_y(g, P.lf_R, xc, yc);
    LFFoot f[TL::NPT];
    T val[TL::NPT];
    bool act[TL::NPT];
    int lo = INT_MAX, hi = INT_MIN, d0 = 0, d1 = 0;
#pragma unroll
    for (int q = 0; q < TL::NPT; ++q) {
      const int iq = tile.ai(q, flip), jq = tile.aj(q, flip);
      act[q] = iq < P.nx && jq < P.ny;
      val[q] = T(0);
      if (act[q]) {
        lf_footprint(P, g, iq, jq, f[q]);
        val[q] = img[long(jq) * P.nx + iq];
        act[q] = f[q].kf <= f[q].kl;
        if (act[q]) {
          lo = min(lo, f[q].kf);
          hi = max(hi, f[q].kl + 1);
        }
      }
    }
    block_window(lo, hi, d0, d1, red);
    const int RW = hi > lo ? hi - lo : 0;
    const bool in_lds = RW > 0 && RW <= lds_cap;
    if (in_lds) {
      for (int idx = threadIdx.x; idx < RW; idx += kThreads) reg[idx] = 0.0;
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < TL::NPT; ++q) {
      if (!act[q]) continue;
      const double a = (LINE ? f[q].amp_line : f[q].amp) * double(val[q]);
      for (int k = f[q].kf; k <= f[q].kl; ++k) {
        const T c = T(LINE ? lf_point(f[q], k, a) : lf_gamma(f[q], k, P.lf_size) * a);
        if (in_lds)
          atomicAdd(reg + (k - lo), double(c));
        else
          atomic_add(out + k, c);
      }
    }
    if (in_lds) {
      __syncthreads();
      for (int idx = threadIdx.x; idx < RW; idx += kThreads) {
        const double v = reg[idx];
        if (v != 0.0) atomic_add(out + lo + idx, T(v));
      }
    }
    __syncthreads();
  }
}

template <typename T, bool LINE>
__global__ __launch_bounds__(kThreads) void lf_backward(DevPlan P, const T *__restrict__ sino,
                                                        T *__restrict__ img, int views_per_block,
                                                        int atomic_out) {
  Tile<16, 16, 1> tile;
  tile.init(P);
  if (!(tile.i < P.nx && tile.j < P.ny)) return;
  T acc = T(0);
  for (int vv = 0; vv < views_per_block; ++vv) {
    const int view = blockIdx.y * views_per_block + vv;
    if (view >= P.nviews) break;
    const ViewConst &g = P.views[view];
    const T *in = sino + long(view) * P.nv;
    LFFoot f;
    lf_footprint(P, g, tile.i, tile.j, f);
    double s = 0.0;
    if (LINE) {
      // sum += (alpha gamma) pixel[k] with gamma the trapezoid height at k
      for (int k = f.kf; k <= f.kl; ++k) s += lf_point(f, k, f.amp_line) * double(in[k]);
    } else {
      for (int k = f.kf; k <= f.kl; ++k) s += f.amp * lf_gamma(f, k, P.lf_size) * double(in[k]);
    }
    acc += T(s);
  }
  T *d = img + long(tile.j) * P.nx + tile.i;
  if (atomic_out)
    atomic_add(d, acc);
  else
    *d += acc;
}

// ---------------------------------------------------------------------------
// Weighted residual and FP64 data-fidelity sum (Ytrtt4D_optim.i:418-446)
// ---------------------------------------------------------------------------
// SPLIT: s <- W s (the reference's split gradient A'W A x - WCyT,
// Ytrtt4D_optim.i:437-446, 764-782) instead of s <- W (s - y); fx is the
// same.
template <typename T, bool SPLIT>
__global__ __launch_bounds__(kThreads) void weighted_residual(T *__restrict__ s,
                                                              const T *__restrict__ y,
                                                              const T *__restrict__ w, long n,
                                                              double *__restrict__ fx) {
  double local = 0.0;
  for (long idx = long(blockIdx.x) * kThreads + threadIdx.x; idx < n;
       idx += long(gridDim.x) * kThreads) {
    const T ax = s[idx];
    const T d = ax - y[idx];
    const T r = w ? w[idx] * d : d;
    local += double(r) * double(d);
    s[idx] = SPLIT ? (w ? w[idx] * ax : ax) : r;
  }
  __shared__ double red[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < kThreads / 64; ++q) t += red[q];
    atomicAdd(fx, 0.5 * t);
  }
}

// ---------------------------------------------------------------------------
// Sparse system matrix of a 2-D plan.  The coefficients of pixel p = j nx + i
// on view v are what its footprint burns into the view's bins for a unit
// value (trtt_PB2D_sparser / trtt_FB2D_sparser, trtt_2D_projectors.c:393-536;
// lf2d_sparser, LongFessler2D.c:472-571), evaluated with the expressions of
// the plan's own forward kernel.  Only non-zero coefficients are entries
// (Ytrtt_2D_projectors.i:268-279), in the order view, pixel, bin.  One thread
// per (view, pixel): FILL = false writes the entry count, FILL = true writes
// the entries from the exclusive scan of the counts.
// ---------------------------------------------------------------------------

// x rounded to 6 decimals, half to even, as the reference's "%f" text round
// trip of the coefficients (Ytrtt_2D_projectors.i:268-290) leaves it: the
// exact product x 1e6 = t + e (FMA residual) decides the ties.
__device__ __forceinline__ double round_decimals6(double x) {
  const double t = x * 1.0e6;
  const double e = fma(x, 1.0e6, -t);
  double k = rint(t);
  const double d = t - k;  // exact
  if (d == 0.5 && e > 0.0) k += 1.0;
  if (d == -0.5 && e < 0.0) k -= 1.0;
  return k / 1.0e6;
}

template <typename T>
__device__ __forceinline__ void sparse_entry(bool fill, T c, int round6, long long &o, int &n,
                                             T *val, int *row, int *col, int r, int p) {
  if (c == T(0)) return;
  if (fill) {
    val[o] = round6 ? T(round_decimals6(double(c))) : c;
    row[o] = r;
    col[o] = p;
    ++o;
  } else {
    ++n;
  }
}

template <typename T, int KIND, bool TABLE, bool QUIRK, bool FILL>
__global__ __launch_bounds__(kThreads) void sparse_columns(DevPlan P, int *__restrict__ count,
                                                           const long long *__restrict__ offset,
                                                           T *__restrict__ val, int *__restrict__ row,
                                                           int *__restrict__ col, int round6) {
  const long npix = long(P.nx) * P.ny;
  const long p = long(blockIdx.x) * kThreads + threadIdx.x;
  const int view = blockIdx.y;
  if (p >= npix) return;
  const int i = int(p % P.nx), j = int(p / P.nx);
  const long q = long(view) * npix + p;
  const ViewConst &g = P.views[view];
  long long o = FILL ? offset[q] : 0;
  int n = 0;
  if constexpr (KIND == LF2D) {
    LFFoot f;
    lf_footprint(P, g, i, j, f);
    for (int k = f.kf; k <= f.kl; ++k) {
      // lf_forward's coefficient for a unit value
      const T c = T(P.lf_line ? lf_point(f, k, f.amp_line) : lf_gamma(f, k, P.lf_size) * f.amp);
      sparse_entry(FILL, c, round6, o, n, val, row, col, view * P.nv + k, int(p));
    }
  } else {
    Foot f;
    if (!voxel_footprint<KIND, TABLE, QUIRK>(P, g, i, j, 0, f)) {
      if (!FILL) atomicOr(P.err, 1);
    } else if (f.vl < f.vr) {
      // fp_forward's coefficient for a unit value: weight x rho0
      AxisFoot<T, TABLE, true> av;
      av.init(P, f.vl, f.vc, f.hv, f.sv);
      const T amp = T(f.rho0);
      for (int mm = 0; mm < f.vr - f.vl; ++mm) {
        const T c = av.next(mm) * amp;
        sparse_entry(FILL, c, round6, o, n, val, row, col, view * P.nv + f.vl + mm, int(p));
      }
    }
  }
  if (!FILL) count[q] = n;
}

// Host evaluation of one pixel's coefficient column on one view, with the
// kernels' own footprint code: the per-pixel sparser symbols
// (trtt_PB2D_sparser / trtt_FB2D_sparser, trtt_2D_projectors.c:393-536, on
// the tabulated footprint; lf2d_sparser / lf2d_line_integ_sparser,
// LongFessler2D.c:472-681).  out[bin] += coefficient for a unit value.
// Returns false when the pixel lies in the source plane.
bool host_column_2d(bool fan, const DevPlan &P, const ViewConst &g, int i, int j, double *out) {
  Foot f;
  const bool ok = fan ? voxel_footprint<FB2D, true, false>(P, g, i, j, 0, f)
                      : voxel_footprint<PB2D, true, false>(P, g, i, j, 0, f);
  if (!ok) return false;
  if (f.vl < f.vr) {
    AxisFoot<double, true, true> av;
    av.init(P, f.vl, f.vc, f.hv, f.sv);
    for (int m = 0; m < f.vr - f.vl; ++m) out[f.vl + m] += f.rho0 * av.next(m);
  }
  return true;
}

// The cubic B-spline CDF on the host (the function cdf3 of the kernels).
double host_cdf3(double x) { return cdf3<double>(x); }

void host_lf_column(const DevPlan &P, const ViewConst &g, int i1, int i2, double *out) {
  LFFoot f;
  lf_footprint(P, g, i1, i2, f);
  for (int k = f.kf; k <= f.kl; ++k)
    out[k] += P.lf_line ? lf_point(f, k, f.amp_line) : lf_gamma(f, k, P.lf_size) * f.amp;
}

template <typename T, bool FILL>
hipError_t launch_sparse_fill(int kind, bool table, bool quirk, const DevPlan &P, int *count,
                              const long long *offset, void *val, int *row, int *col, int round6,
                              hipStream_t st) {
  const long npix = long(P.nx) * P.ny;
  dim3 grid(unsigned((npix + kThreads - 1) / kThreads), unsigned(P.nviews));
  T *v = static_cast<T *>(val);
#define TRTT_SPARSE_CASE(K, TB, Q)                                                              \
  if (kind == K && (K == LF2D || (table == TB && quirk == Q))) {                               \
    hipLaunchKernelGGL((sparse_columns<T, K, TB, Q, FILL>), grid, dim3(kThreads), 0, st, P,     \
                       count, offset, v, row, col, round6);                                     \
    return hipGetLastError();                                                                   \
  }
  TRTT_SPARSE_CASE(FB2D, false, true)
  TRTT_SPARSE_CASE(FB2D, true, false)
  TRTT_SPARSE_CASE(PB2D, false, true)
  TRTT_SPARSE_CASE(PB2D, true, false)
  TRTT_SPARSE_CASE(LF2D, false, false)
#undef TRTT_SPARSE_CASE
  return hipErrorInvalidValue;
}

// fill = false: count[view * npix + pixel] = entries; fill = true: write them
// at offset[view * npix + pixel].  Grid (pixels / 256, views).
hipError_t launch_sparse_columns(int kind, bool table, bool quirk, bool f64, const DevPlan &P,
                                 bool fill, int *count, const long long *offset, void *val,
                                 int *row, int *col, int round6, hipStream_t st) {
  if (f64)
    return fill ? launch_sparse_fill<double, true>(kind, table, quirk, P, count, offset, val, row,
                                                   col, round6, st)
                : launch_sparse_fill<double, false>(kind, table, quirk, P, count, offset, val,
                                                    row, col, round6, st);
  return fill ? launch_sparse_fill<float, true>(kind, table, quirk, P, count, offset, val, row, col,
                                                round6, st)
              : launch_sparse_fill<float, false>(kind, table, quirk, P, count, offset, val, row,
                                                 col, round6, st);
}

// ---------------------------------------------------------------------------
// Distance-driven parallel-beam projector (trtt_DD_PB3D,
// trtt_3D_projectors.c:1077-1173; burns :406-525 and :817-940).  A voxel's
// footprint is the rectangle [uc - Du, uc + Du] x [vc - Dv, vc + Dv] with its
// corners clamped to [0, n-1]; a pixel's weight is h times its overlap with
// the rectangle along u (times u_scl) and along v (times v_scl).  One thread
// per voxel: the forward scatters with atomics, the transpose gathers.
// ---------------------------------------------------------------------------

// Overlap of pixel floor(b1) + p with [b1, b2], scaled.  `stuck` gives the
// reference transpose's v weights (:909-916): its loop over the inner pixels
// advances the u counter, so all inner pixels write one slot, the upper
// partial pixel overwrites it, and the slots after it stay zero.
__device__ __forceinline__ double dd_overlap(double b1, double b2, int p, double scl, bool stuck) {
#pragma clang fp contract(off)
  const double up = ceil(b1), down = floor(b2);
  const double inner = down - up;
  const bool low = up > b1, high = down < b2;
  if (inner < 0.0) return p == 0 ? scl * (b2 - b1) : 0.0;
  if (!stuck) {
    const int n = int(ceil(b2) - floor(b1));
    if (low && p == 0) return scl * (up - b1);
    if (high && p == n - 1) return scl * (b2 - down);
    return scl;
  }
  if (low && p == 0) return scl * (up - b1);
  if (p == (low ? 1 : 0)) return high ? scl * (b2 - down) : (inner > 0.0 ? scl : 0.0);
  return 0.0;
}

struct DDRect {
  double ub1, ub2, vb1, vb2;
  int nu_e, nv_e;
  long corner;
};

__device__ __forceinline__ void dd_rect(const DDConst &c, long i, long j, long k, DDRect &r) {
#pragma clang fp contract(off)
  const double vc = ((c.v_z * double(k) + c.v_c) + c.v_y * double(j)) + c.v_x * double(i);
  const double uc = ((c.u_z * double(k) + c.u_c) + c.u_y * double(j)) + c.u_x * double(i);
  const double nu1 = double(c.nu - 1), nv1 = double(c.nv - 1);
  r.ub1 = fmin(fmax(uc - c.Du, 0.0), nu1);
  r.ub2 = fmin(fmax(uc + c.Du, 0.0), nu1);
  r.vb1 = fmin(fmax(vc - c.Dv, 0.0), nv1);
  r.vb2 = fmin(fmax(vc + c.Dv, 0.0), nv1);
  r.nu_e = int(ceil(r.ub2) - floor(r.ub1));
  r.nv_e = int(ceil(r.vb2) - floor(r.vb1));
  r.corner = long(floor(r.vb1)) * c.nu + long(floor(r.ub1));
}

// Views of a batch: blockIdx.y selects view `view0 + blockIdx.y` of `cs` in
// the forward; the transpose loops over the views in each thread.
template <typename T>
__global__ __launch_bounds__(kThreads) void dd_forward(const DDConst *__restrict__ cs,
                                                       const T *__restrict__ vol,
                                                       T *__restrict__ sino) {
#pragma clang fp contract(off)
  const DDConst &c = cs[blockIdx.y];
  const long vox = long(blockIdx.x) * kThreads + threadIdx.x;
  if (vox >= c.nx * c.ny * c.nz) return;
  const long i = vox % c.nx, j = (vox / c.nx) % c.ny, k = vox / (c.nx * c.ny);
  DDRect r;
  dd_rect(c, i, j, k, r);
  if (r.nu_e <= 0 || r.nv_e <= 0) return;
  T *out = sino + long(blockIdx.y) * c.nu * c.nv;
  const double x = double(vol[vox]);
  for (int m = 0; m < r.nv_e; ++m) {
    const double hv = c.h * dd_overlap(r.vb1, r.vb2, m, c.vs, false);
    for (int l = 0; l < r.nu_e; ++l)
      atomic_add(out + r.corner + long(m) * c.nu + l,
                 T(hv * dd_overlap(r.ub1, r.ub2, l, c.us, false) * x));
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void dd_transpose(const DDConst *__restrict__ cs,
                                                         int nviews, const T *__restrict__ sino,
                                                         T *__restrict__ vol, int stuck) {
#pragma clang fp contract(off)
  const long vox = long(blockIdx.x) * kThreads + threadIdx.x;
  if (vox >= cs[0].nx * cs[0].ny * cs[0].nz) return;
  const long i = vox % cs[0].nx, j = (vox / cs[0].nx) % cs[0].ny, k = vox / (cs[0].nx * cs[0].ny);
  double s = 0.0;
  for (int view = 0; view < nviews; ++view) {
    const DDConst &c = cs[view];
    DDRect r;
    dd_rect(c, i, j, k, r);
    if (r.nu_e <= 0 || r.nv_e <= 0) continue;
    const T *in = sino + long(view) * c.nu * c.nv;
    for (int m = 0; m < r.nv_e; ++m) {
      const double hv = c.h * dd_overlap(r.vb1, r.vb2, m, c.vs, stuck != 0);
      for (int l = 0; l < r.nu_e; ++l)
        s += hv * dd_overlap(r.ub1, r.ub2, l, c.us, false) *
             double(in[r.corner + long(m) * c.nu + l]);
    }
  }
  vol[vox] += T(s);
}

// Every view of `cs` (device array, nviews entries, same sizes): job 0:
// sino[view] += A_view vol; job 1: vol += sum over the views of the
// reference's transpose (stuck = 1) or of the exact adjoint (stuck = 0).
hipError_t launch_dd(bool f64, const DDConst *cs, const DDConst &c0, int nviews, int job,
                     int stuck, const void *in, void *out, hipStream_t st) {
  const long n = c0.nx * c0.ny * c0.nz;
  const unsigned blocks = unsigned((n + kThreads - 1) / kThreads);
  if (job == 0) {
    const dim3 grid(blocks, unsigned(nviews));
    if (f64)
      hipLaunchKernelGGL(dd_forward<double>, grid, dim3(kThreads), 0, st, cs,
                         static_cast<const double *>(in), static_cast<double *>(out));
    else
      hipLaunchKernelGGL(dd_forward<float>, grid, dim3(kThreads), 0, st, cs,
                         static_cast<const float *>(in), static_cast<float *>(out));
  } else {
    if (f64)
      hipLaunchKernelGGL(dd_transpose<double>, dim3(blocks), dim3(kThreads), 0, st, cs, nviews,
                         static_cast<const double *>(in), static_cast<double *>(out), stuck);
    else
      hipLaunchKernelGGL(dd_transpose<float>, dim3(blocks), dim3(kThreads), 0, st, cs, nviews,
                         static_cast<const float *>(in), static_cast<float *>(out), stuck);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host-side launchers (explicit instantiation table)
// ---------------------------------------------------------------------------
#ifndef TRTT_LF_BWD_BLOCKS
#define TRTT_LF_BWD_BLOCKS 8192
#endif
#ifndef TRTT_FWD3D_LDS
#define TRTT_FWD3D_LDS 32768
#endif

template <typename T, int KIND, bool TABLE, bool QUIRK>
hipError_t launch_fp(const DevPlan &P, int job, const void *in, void *out, const int *seg_begin,
                     int nrows, int views_per_block, int atomic_out, hipStream_t st) {
  constexpr bool IS2D = (KIND == PB2D || KIND == FB2D);
  constexpr int TX = IS2D ? 16 : 8, TY = IS2D ? kTile2DRows : 8, TZ = IS2D ? 1 : 8;
  const int lds_bytes = 32768;
  const int lds_cap = lds_bytes / int(sizeof(T));
  const int ntx = (P.nx + TX - 1) / TX, nty = (P.ny + TY - 1) / TY, ntz = (P.nz + TZ - 1) / TZ;
  const unsigned ntiles = unsigned(ntx) * nty * ntz;
  if (job == 0) {
    // the 2-D forward takes square tiles (lanes laid out per view)
    // (FP64 window of up to 2048 cells in 2-D, TRTT_FWD3D_LDS bytes in 3-D)
    constexpr int FX = IS2D ? kTile2DSquare : TX, FY = IS2D ? kTile2DSquare : TY;
    const unsigned nft = unsigned((P.nx + FX - 1) / FX) * ((P.ny + FY - 1) / FY) * ntz;
    const int lds_f = IS2D ? 8 * 2048 : TRTT_FWD3D_LDS;
    const int cap_f = lds_f / 8;
    dim3 grid(nft, (P.nviews + views_per_block - 1) / views_per_block);
    hipLaunchKernelGGL((fp_forward<T, KIND, TABLE, QUIRK, FX, FY, TZ>), grid, dim3(kThreads),
                       lds_f, st, P, static_cast<const T *>(in), static_cast<T *>(out),
                       views_per_block, cap_f);
  } else {
    dim3 grid(ntiles, nrows);
    hipLaunchKernelGGL((fp_backward<T, KIND, TABLE, QUIRK, TX, TY, TZ>), grid, dim3(kThreads),
                       lds_bytes, st, P, static_cast<const T *>(in), static_cast<T *>(out),
                       seg_begin, atomic_out, lds_cap);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fp_any(int kind, bool table, bool quirk, const DevPlan &P, int job,
                         const void *in, void *out, const int *seg_begin, int nrows,
                         int views_per_block, int atomic_out, hipStream_t st) {
#define TRTT_CASE(K, TB, Q)                                                                \
  if (kind == K && table == TB && quirk == Q)                                             \
    return launch_fp<T, K, TB, Q>(P, job, in, out, seg_begin, nrows, views_per_block,      \
                                  atomic_out, st);
  TRTT_CASE(CB3D, false, false)
  TRTT_CASE(CB3D, true, false)
  TRTT_CASE(PB3D, false, false)
  TRTT_CASE(PB3D, true, false)
  TRTT_CASE(FB2D, false, true)
  TRTT_CASE(FB2D, true, false)
  TRTT_CASE(PB2D, false, true)
  TRTT_CASE(PB2D, true, false)
#undef TRTT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_projector(int kind, bool table, bool quirk, bool f64, const DevPlan &P, int job,
                            const void *in, void *out, const int *seg_begin, int nrows,
                            int views_per_block, int atomic_out, hipStream_t st) {
  if (kind == LF2D) {
    const int ntx = (P.nx + 15) / 16, nty = (P.ny + 15) / 16;
    dim3 grid(unsigned(ntx) * nty, (P.nviews + views_per_block - 1) / views_per_block);
    if (job == 0) {
      // square tiles, an FP64 window of up to 2048 cells
      constexpr int S = kTile2DSquare, cap = 2048;
      const dim3 fgrid(unsigned((P.nx + S - 1) / S) * ((P.ny + S - 1) / S), grid.y);
#define TRTT_LF_FWD(T, L)                                                                   \
  hipLaunchKernelGGL((lf_forward<T, L>), fgrid, dim3(kThreads), cap * 8, st, P,              \
                     static_cast<const T *>(in), static_cast<T *>(out), views_per_block, cap)
      if (f64) {
        if (P.lf_line) TRTT_LF_FWD(double, true); else TRTT_LF_FWD(double, false);
      } else {
        if (P.lf_line) TRTT_LF_FWD(float, true); else TRTT_LF_FWD(float, false);
      }
#undef TRTT_LF_FWD
    } else {
      // the views split over grid rows (atomic adds into the image) until
      // the grid offers TRTT_LF_BWD_BLOCKS workgroups: one row of 16x16
      // tiles alone leaves one wave per SIMD on a 512^2 image
      const long tiles = long(ntx) * nty;
      const long span = std::max(1, std::min(views_per_block, P.nviews));
      const long want = (long(TRTT_LF_BWD_BLOCKS) + tiles - 1) / tiles;
      const int vpb = int(std::max(1L, std::min(span, (P.nviews + want - 1) / want)));
      const dim3 bgrid(unsigned(tiles), unsigned((P.nviews + vpb - 1) / vpb));
      const int atomic = (atomic_out || bgrid.y > 1) ? 1 : 0;
#define TRTT_LF_BWD(T, L)                                                                   \
  hipLaunchKernelGGL((lf_backward<T, L>), bgrid, dim3(kThreads), 0, st, P,                    \
                     static_cast<const T *>(in), static_cast<T *>(out), vpb, atomic)
      if (f64) {
        if (P.lf_line) TRTT_LF_BWD(double, true); else TRTT_LF_BWD(double, false);
      } else {
        if (P.lf_line) TRTT_LF_BWD(float, true); else TRTT_LF_BWD(float, false);
      }
#undef TRTT_LF_BWD
    }
    return hipGetLastError();
  }
  if (f64)
    return launch_fp_any<double>(kind, table, quirk, P, job, in, out, seg_begin, nrows,
                                 views_per_block, atomic_out, st);
  return launch_fp_any<float>(kind, table, quirk, P, job, in, out, seg_begin, nrows,
                              views_per_block, atomic_out, st);
}

hipError_t launch_residual(bool f64, void *s, const void *y, const void *w, long n, double *fx,
                           hipStream_t st, bool split) {
  const long blocks_needed = (n + kThreads - 1) / kThreads;
  const unsigned blocks = unsigned(blocks_needed < 4096 ? (blocks_needed > 0 ? blocks_needed : 1)
                                                        : 4096);
#define TRTT_RES(T, S)                                                                     \
  hipLaunchKernelGGL((weighted_residual<T, S>), dim3(blocks), dim3(kThreads), 0, st,       \
                     static_cast<T *>(s), static_cast<const T *>(y), static_cast<const T *>(w), \
                     n, fx)
  if (f64) {
    if (split) TRTT_RES(double, true); else TRTT_RES(double, false);
  } else {
    if (split) TRTT_RES(float, true); else TRTT_RES(float, false);
  }
#undef TRTT_RES
  return hipGetLastError();
}

}  // namespace trtt
