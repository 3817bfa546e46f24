// Shared definitions of the gfx950 footprint projectors: per-view constants,
// the kernel-side plan and the host-side plan object.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/trtt_hip.h"

namespace trtt {

enum Kind : int { PB3D = 0, CB3D = 1, PB2D = 2, FB2D = 3, LF2D = 4 };

// Per-view constants, FP64, derived on the host from XOS/XSD with the
// reference expressions (trtt_3D_projectors.c:1190-1232 for cone,
// :967-1003 for parallel, trtt_2D_projectors.c:195-221 / :279-315 for 2-D).
// The device evaluates, per voxel (i,j,k), in the reference's incremental
// order ((z*k + c) + y*j) + x*i:
//   cone/fan:  depth w = o.(i,j,k,1), numerators nu/nv, denominator d
//   parallel:  affine uc/vc and view-constant footprint scales.
struct ViewConst {
  double nux, nuy, nuz, nuc;
  double nvx, nvy, nvz, nvc;
  double dx, dy, dz, dc;
  double o1, o2, o3, o4;
  double a4m, a4sqr, a8, a12, b3, c3;
  // parallel beam: view-constant footprint
  double pb_rho0, pb_hu, pb_hv, pb_su, pb_sv, pb_hsu, pb_hsv;
  // time interpolation: the view sees w0*x[t0] + w1*x[t1]
  double w0, w1;
  int t0, t1;
  // Long & Fessler: cos/sin of beta, detector offset in pixels, object
  // origin (x0, y0) in the rotated frame
  double lf_cb, lf_sb, lf_off, lf_x0, lf_y0;
};

// A contiguous run of views (in plan order) sharing the same (t0, t1).
struct Segment {
  int start;   // index into the view order array
  int count;
  int t0, t1;
};

// Everything a kernel needs, passed by value.
struct DevPlan {
  const ViewConst *views;
  const int *order;        // plan order -> view index (sinogram slot)
  const Segment *segs;
  const void *table;       // cumulative footprint table (T), or nullptr
  int *err;                // set to 1 when a voxel lies in the source plane
  long m;                  // table size
  double step, coord0;
  int nx, ny, nz, nu, nv, nviews, nt;
  long vox_per_phase;      // nx*ny*nz
  double u_scl, v_scl;
  double hu_init, hv_init, rho_init, hs0;
  int table_deg;           // axial kernels: degree of the B-spline CDF the table holds
  // Long & Fessler scalars
  double lf_vs, lf_ps, lf_R, lf_scale, lf_size;
  int lf_line;             // 1: point-sampled (line-integral) trapezoid model
};

// Constants of one view of the distance-driven parallel-beam projector
// (trtt_DD_PB3D, trtt_3D_projectors.c:1077-1120), FP64, reference order.
struct DDConst {
  double v_x, v_y, v_z, v_c;  // detector v of voxel (i,j,k): ((v_z k + v_c) + v_y j) + v_x i
  double u_x, u_y, u_z, u_c;
  double Du, Dv;              // half sizes of the rectangular footprint (pixels)
  double h;                   // s^3 / (u_scl v_scl |u_z| max(|v_x|, |v_y|))
  double us, vs;
  long nx, ny, nz, nu, nv;
};

// Rows of the 16-wide pixel tiles of the general 2-D footprint kernels
// (16 x kTile2DRows pixels per 256-thread workgroup, kTile2DRows / 16 per
// thread).  Measured on a 512^2 fan plan of 720 views (FP32 cubic, forward /
// backward): 16 rows 6.51 / 2.24 ms, 32 rows 5.88 / 1.97, 64 rows 5.58 /
// 1.85, 128 rows 5.60 / 2.49, 256 rows 5.94 / 4.56.
#ifndef TRTT_TILE2D_Y
#define TRTT_TILE2D_Y 64
#endif
constexpr int kTile2DRows = TRTT_TILE2D_Y;
// Side of the square pixel tiles of the general 2-D forward (kTile2DSquare^2
// / 256 pixels per thread, laid out per view across the rays).
#ifndef TRTT_TILE2D_SQUARE
#define TRTT_TILE2D_SQUARE 32
#endif
constexpr int kTile2DSquare = TRTT_TILE2D_SQUARE;

// Launch shape of the axial kernels of a plan (trtt_hip_plan_set_tuning).
constexpr int kNumFwdTiles = 3;  // 8x4, 4x4, 8x8 columns x 64 k
constexpr int kNumBwdTiles = 6;  // 8x8, 8x4, 4x4 columns x 64 k; 8x4, 4x4, 8x8 x 128 k
// voxels per lane of a backward tile (its k extent is 64 times this)
constexpr int bwd_tile_vpl(int t) { return t >= 3 ? 2 : 1; }
struct AxialTuning {
  int fwd_tile = 0;
  int bwd_tile = 0;
  int fwd_window_bytes = 16384;   // private LDS window of a forward wave
  int fwd_views_per_block = 0;    // 0: from the grid size
  int bwd_window_bytes = 32768;   // shared LDS window of a backward workgroup
  int fwd_chunk_views = 0;        // views per forward scratch round, 0: fill the budget
  int fwd_flush = 1;              // 1: windows added to the projections with global atomics;
                                  // 0: written to scratch slots, summed by fwd_reduce
  int fwd_xcds = 0;               // > 1: XCD-aware forward block order for that many XCDs
};

// Second stream of the axial forward: the reduction of view chunk c runs on
// `aux` while the main stream projects chunk c + 1 (two scratch buffers).
struct AxialOverlap {
  hipStream_t aux = nullptr;  // nullptr: reduce on the launch stream
  hipEvent_t fwd_done[2] = {nullptr, nullptr};
  hipEvent_t red_done[2] = {nullptr, nullptr};
};

}  // namespace trtt

// Host-side plan (opaque to C callers).
struct trtt_hip_plan {
  int kind = 0;
  int precision = TRTT_HIP_F32;
  int device = 0;
  bool table = false;
  int table_deg = 0;  // the table holds the CDF of the B-spline of this degree (2 or 3) at
                      // its nodes (axial path allowed), 0 otherwise
  bool quirk_last_bin = false;
  long nx = 0, ny = 0, nz = 0, nt = 1, nu = 1, nv = 1, nviews = 0;
  double s_scl = 1.0, u_scl = 1.0, v_scl = 1.0;
  int s_deg = 3;
  std::vector<trtt::ViewConst> h_views;
  std::vector<int> h_order;
  std::vector<trtt::Segment> h_segs;
  std::vector<int> h_seg_begin;        // per back-projection block row
  trtt::ViewConst *d_views = nullptr;
  int *d_order = nullptr;
  trtt::Segment *d_segs = nullptr;
  int *d_seg_begin = nullptr;
  void *d_table = nullptr;
  int *d_err = nullptr;
  long m = 0;
  double step = 0.0, coord0 = 0.0;
  trtt::DevPlan dev{};
  // fast axial path (closed-form cubic, FP32, rotation axis = z = u):
  // footprint width bound W (0: not eligible) and column profile length
  int axial_w = 0;
  int axial_pmax = 0;
  int axial_pmax2 = 0;       // column profile length over 128 k (0: over 256)
  int axial_plain = 0;       // profile passes without LDS atomics (0: atomics)
  int axial_chunk = 1;      // views per forward scratch round
  void *d_scratch = nullptr;  // forward windows, one per (view, tile)
  void *d_hdr = nullptr;      // int4 (ul, vl, width, height) per window
  void *d_volk = nullptr;     // volume series with k fastest (forward input)
  trtt::AxialTuning tune;
  // optional HIP events around the main kernels of the last forward (one
  // pair per view chunk) and backward (one pair) of the axial path
  bool timing = false;
  std::vector<hipEvent_t> ev_fwd, ev_bwd;
  int fwd_launches = 0, bwd_launches = 0;
  int axial_nbuf = 1;           // forward scratch buffers (2: reduction overlapped)
  bool two_phase_segs = false;  // some back-projection segment has t0 != t1
  trtt::AxialOverlap overlap;
  // host-array evaluation of the data term (trtt_hip_eval*): device copies
  // of y, W, WCyT and the work buffers, made by trtt_hip_eval_set_data
  struct EvalData *eval = nullptr;
};
