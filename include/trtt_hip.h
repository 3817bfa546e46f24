/*
 * trtt_hip.h -- C ABI of libtrtt_hip.so, the MI355X (gfx950) implementation
 * of TRTT's B-spline footprint projectors, the Long & Fessler separable
 * footprint projector and the per-phase operator loop.
 *
 * Part 1 exports the reference plugin symbols with their exact prototypes so
 * that the unchanged Yorick glue (plugin/ywrap.c, tomography-LongFessler-
 * model/ywrap.c) links against this library instead of the C plugin.  Each
 * call is synchronous, takes caller-owned host FP64 arrays valid only during
 * the call, ACCUMULATES into `out` (lf2d_*: zeroes it first when clr != 0) and
 * returns TRTT_SUCCESS (0) or TRTT_FAILURE (-1) with a message on stderr.
 *
 * Part 2 is the device-resident batched interface: a plan uploads the
 * per-view geometry once; executing it projects every view of every phase in
 * one launch sequence on a caller stream, with device pointers only.
 *
 * Layouts (row-major, as the reference C sees Yorick's column-major arrays):
 *   volume   [nz][ny][nx]     4-D series [nt][nz][ny][nx]
 *   image    [ny][nx]
 *   detector [nv][nu] per view (u axial, v transaxial); sinogram [view][nv][nu]
 *   xos/xsd  12 row-major coefficients per view.
 */
#ifndef TRTT_HIP_H
#define TRTT_HIP_H 1

#ifdef __cplusplus
extern "C" {
#endif

#define TRTT_SUCCESS (0)
#define TRTT_FAILURE (-1)

/* ===================================================================== */
/* Part 1 -- reference plugin symbols (drop-in)                           */
/* ===================================================================== */

/* 2-D parallel / fan beam, tabulated footprint CDF.
 * Replaces plugin/trtt_2D_projectors.h:72-74 (trtt_2D_projectors.c:186, :273). */
int trtt_PB2D(double *out, double *in, long nx, long ny, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nv, double v_scl,
              double *xos, double *xsd, int job);
int trtt_FB2D(double *out, double *in, long nx, long ny, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nv, double v_scl,
              double *xos, double *xsd, int job);

/* 2-D closed-form cubic footprint (the last detector bin never receives a
 * contribution, as in the reference).
 * Replaces plugin/trtt_2D_projectors.h:76-78 (trtt_2D_projectors_cubic.c:209, :292). */
int trtt_PB2D_cubic(double *out, double *in, long nx, long ny, double s_scl, int s_deg,
                    long nv, double v_scl, double *xos, double *xsd, int job);
int trtt_FB2D_cubic(double *out, double *in, long nx, long ny, double s_scl, int s_deg,
                    long nv, double v_scl, double *xos, double *xsd, int job);

/* 3-D parallel / cone beam, tabulated footprint CDF.
 * Replaces plugin/trtt_3D_projectors.h:299, :423 (trtt_3D_projectors.c:959, :1179). */
int trtt_PB3D(double *out, double *in, long nx, long ny, long nz, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nu, long nv,
              double u_scl, double v_scl, double *xos, double *xsd, int job);
int trtt_CB3D(double *out, double *in, long nx, long ny, long nz, double s_scl, int s_deg,
              double *footprint, double *cum_footprint, long size_footprint,
              double step_footprint, double *coord_footprint, long nu, long nv,
              double u_scl, double v_scl, double *xos, double *xsd, int job);

/* 3-D closed-form cubic footprint.
 * Replaces plugin/trtt_3D_projectors_cubic.h:299, :363 (trtt_3D_projectors_cubic.c:506, :623). */
int trtt_PB3D_cubic(double *out, double *in, long nx, long ny, long nz, double s_scl,
                    int s_deg, long nu, long nv, double u_scl, double v_scl, double *xos,
                    double *xsd, int job);
int trtt_CB3D_cubic(double *out, double *in, long nx, long ny, long nz, double s_scl,
                    int s_deg, long nu, long nv, double u_scl, double v_scl, double *xos,
                    double *xsd, int job);

/* Long & Fessler parameters, field order of LongFessler2D.h:16-28. */
typedef struct _lf2d_parameters lf2d_parameters_t;
struct _lf2d_parameters {
  double pixelSize, pixelStep, voxelSize, sourceDistance, detectorDistance;
  double objectOffset1, objectOffset2;
  long objectDimension1, objectDimension2, detectorDimension;
};

/* Replaces tomography-LongFessler-model/LongFessler2D.h:30-59
 * (LongFessler2D.c:41 lf2d_direct, :254 lf2d_adjoint). */
void lf2d_direct(const lf2d_parameters_t *p, double beta, double detectorOffset,
                 const double voxel[], double pixel[], int clr);
void lf2d_adjoint(const lf2d_parameters_t *p, double beta, double detectorOffset,
                  const double pixel[], double voxel[], int clr);

/* Line-integral (point-sampled trapezoid) Long & Fessler model: the
 * footprint of a pixel is sampled at each bin centre instead of integrated
 * over the bin, with amplitude sqrt(1+q^2) voxelSize.  GPU, like
 * lf2d_direct / lf2d_adjoint.
 * Replaces tomography-LongFessler-model/LongFessler2D.h:40-49, 61-71
 * (LongFessler2D.c:143 lf2d_line_integ_direct, :357 lf2d_line_integ_adjoint). */
void lf2d_line_integ_direct(const lf2d_parameters_t *p, double beta, double detectorOffset,
                            const double voxel[], double pixel[], int clr);
void lf2d_line_integ_adjoint(const lf2d_parameters_t *p, double beta, double detectorOffset,
                             const double pixel[], double voxel[], int clr);

/* Coefficient column of one object pixel (i1, i2) on one view:
 * pixel[k] += coefficient of bin k for a unit pixel value (clr: pixel zeroed
 * first).  Host code with the kernels' own footprint functions (one pixel is
 * a few dozen operations); trtt_hip_sparse_create builds the whole matrix on
 * the device.  Glue: tomography-LongFessler-model/ywrap.c:55-77.
 * Replaces LongFessler2D.c:472 (lf2d_sparser) and :573
 * (lf2d_line_integ_sparser). */
void lf2d_sparser(const lf2d_parameters_t *p, double beta, double detectorOffset, const long i1,
                  const long i2, double pixel[], int clr);
void lf2d_line_integ_sparser(const lf2d_parameters_t *p, double beta, double detectorOffset,
                             const long i1, const long i2, double pixel[], int clr);

/* Coefficient column of image pixel (i, j) on one view with the tabulated
 * footprint: out[bin] += coefficient for a unit pixel value (nx, ny and
 * footprint are not read, as in the reference).  Host code, as for
 * lf2d_sparser.  Glue: plugin/ywrap.c:295-321.
 * Replaces plugin/trtt_2D_projectors.h:80-82 (trtt_2D_projectors.c:393
 * trtt_PB2D_sparser, :454 trtt_FB2D_sparser). */
int trtt_PB2D_sparser(double *out, long i, long j, long nx, long ny, double s_scl, int s_deg,
                      double *footprint, double *cum_footprint, long size_footprint,
                      double step_footprint, double *coord_footprint, long nv, double v_scl,
                      double *xos, double *xsd);
int trtt_FB2D_sparser(double *out, long i, long j, long nx, long ny, double s_scl, int s_deg,
                      double *footprint, double *cum_footprint, long size_footprint,
                      double step_footprint, double *coord_footprint, long nv, double v_scl,
                      double *xos, double *xsd);

/* Host-side affine transforms used by the Yorick XForm3D object
 * (plugin/trtt_xform3d.h; plugin/Ytrtt.c:9-116).  Row-major 3x4 storage. */
typedef struct _trtt_xform3d trtt_xform3d_t;
trtt_xform3d_t *trtt_xform3d_new(void);
void trtt_xform3d_delete(trtt_xform3d_t *xform);
int trtt_xform3d_init(trtt_xform3d_t *xform);
int trtt_xform3d_reset(trtt_xform3d_t *xform);
int trtt_xform3d_copy(trtt_xform3d_t *dst, const trtt_xform3d_t *src);
int trtt_xform3d_apply(double dst[3], const trtt_xform3d_t *xform, const double src[3]);
int trtt_xform3d_combine(trtt_xform3d_t *dst, const trtt_xform3d_t *left,
                         const trtt_xform3d_t *right);
int trtt_xform3d_move(trtt_xform3d_t *dst, double tx, double ty, double tz,
                      const trtt_xform3d_t *src);
int trtt_xform3d_rotate(trtt_xform3d_t *dst, double qw, double qx, double qy, double qz,
                        const trtt_xform3d_t *src);
int trtt_xform3d_scale(trtt_xform3d_t *dst, double x_scl, double y_scl, double z_scl,
                       const trtt_xform3d_t *src);
int trtt_xform3d_invert(trtt_xform3d_t *dst, const trtt_xform3d_t *src);
int trtt_xform3d_get_coefficients(double dst[12], const trtt_xform3d_t *src);
int trtt_xform3d_set_coefficients(trtt_xform3d_t *dst, const double src[12]);

/* Distance-driven parallel-beam projector (plugin/trtt_3D_projectors.c:
 * 1077-1173, prototype trtt_3D_projectors.h): job 0 accumulates the forward
 * projection, any other job the reference's transpose.  That transpose is
 * not the adjoint when a footprint spans three or more pixels along v: its
 * loop over the inner pixels (trtt_3D_projectors.c:909-916) advances the u
 * counter, and it is reproduced as is. */
int trtt_DD_PB3D(double *out, double *in, long nx, long ny, long nz, double s_scl, long nu,
                 long nv, double u_scl, double v_scl, double *xos, double *xsd, int job);
/* The exact adjoint of trtt_DD_PB3D's forward projection (no reference
 * counterpart; same arguments without job): out (volume) += A' in. */
int trtt_hip_DD_PB3D_adjoint(double *out, double *in, long nx, long ny, long nz, double s_scl,
                             long nu, long nv, double u_scl, double v_scl, double *xos,
                             double *xsd);

/* Allocation-free variant of the eight footprint drop-ins above (no
 * counterpart in the reference).  It is for callers that cannot allocate
 * device memory: code run while a HIP graph is being set up, serving loops,
 * harnesses that forbid hipMalloc.  kind is TRTT_HIP_PB3D, _CB3D, _PB2D or
 * _FB2D; cum = NULL selects the closed-form cubic footprint (the *_cubic
 * symbols), else the tabulated one (cum_footprint, size_footprint,
 * step_footprint and coord_footprint[0] = coord0); the other arguments and
 * the semantics (out += A in for job 0, out += A' in otherwise, precision of
 * trtt_hip_set_precision) are those of the drop-ins.  It always runs the
 * general footprint kernels.  Every device buffer lives in one static 64 MiB
 * device array per GPU (held until the process ends): calls whose tables,
 * staging and output need more fail with TRTT_FAILURE and a message naming
 * the limit.  Calls on one GPU share that array; they are serialised within
 * the process, and concurrent calls from several processes on one GPU each
 * use their own copy. */
int trtt_hip_dropin_noalloc(int kind, double *out, const double *in, long nx, long ny, long nz,
                            double s_scl, int s_deg, const double *cum, long size_footprint,
                            double step_footprint, double coord0, long nu, long nv, double u_scl,
                            double v_scl, const double *xos, const double *xsd, int job);

/* Sparse matrix-vector product out[i[k]-1] += a[k]*in[j[k]-1] (1-based indices).
 * Replaces plugin/trtt_common.c:62-69. */
int trtt_mvmult(double *out, double *in, long number, double *a, long *i, long *j);

/* ===================================================================== */
/* Part 2 -- device-resident batched interface (no reference counterpart:  */
/* it replaces the per-view Yorick loop of Tomography_4D/Ytrtt4D_optim.i: */
/* 418-446 and its MPY form :87-223 with whole-iteration launches)        */
/* ===================================================================== */

/* Data precision of a plan: the projectors compute and accumulate in this
 * type; per-voxel detector geometry is always FP64. */
#define TRTT_HIP_F32 0
#define TRTT_HIP_F64 1

/* Geometry kinds. */
#define TRTT_HIP_PB3D 0  /* trtt_PB3D[_cubic] */
#define TRTT_HIP_CB3D 1  /* trtt_CB3D[_cubic] */
#define TRTT_HIP_PB2D 2  /* trtt_PB2D[_cubic] */
#define TRTT_HIP_FB2D 3  /* trtt_FB2D[_cubic] */
#define TRTT_HIP_LF2D 4  /* lf2d_direct / lf2d_adjoint */

typedef struct trtt_hip_plan trtt_hip_plan_t;

/* Select the device used by the drop-in symbols and new plans (default 0). */
int trtt_hip_set_device(int device);
/* Precision of the drop-in symbols of Part 1 (default TRTT_HIP_F32). */
int trtt_hip_set_precision(int precision);
/* Last error message of the calling thread ("" when none). */
const char *trtt_hip_last_error(void);

/* Create a plan for `nviews` views of one footprint projector.
 *   kind       TRTT_HIP_PB3D, _CB3D, _PB2D or _FB2D (2-D: nz = nu = 1)
 *   cum, size_footprint, step_footprint, coord0: tabulated CDF, or cum=NULL
 *              for the closed-form cubic footprint
 *   xos, xsd   nviews*12 FP64 coefficients
 *   t0, t1, w0, w1  per-view time interpolation (NULL: every view reads phase
 *              0 with weight 1): view k sees volume w0[k]*x[t0[k]] + w1[k]*x[t1[k]]
 *   nt         number of phases of the volume series
 * Views may come in any order; the plan groups them by phase internally. */
trtt_hip_plan_t *trtt_hip_plan_create(int kind, int precision, long nx, long ny, long nz,
                                      long nt, double s_scl, int s_deg, long nu, long nv,
                                      double u_scl, double v_scl, const double *xos,
                                      const double *xsd, long nviews, const double *cum,
                                      long size_footprint, double step_footprint,
                                      double coord0, const int *t0, const int *t1,
                                      const double *w0, const double *w1);

/* Long & Fessler plan: beta[k], offset[k] per view (Ytrtt_2D_projectors_Long_
 * Fessler.i:216-235 builds them as theta-pi/2 and voff). */
trtt_hip_plan_t *trtt_hip_plan_create_lf2d(int precision, const lf2d_parameters_t *p,
                                           const double *beta, const double *offset,
                                           long nviews);

/* Long & Fessler plans: line = 1 selects the line-integral model of
 * lf2d_line_integ_direct / _adjoint (and its sparse matrix), 0 the
 * bin-integrated one (default). */
int trtt_hip_plan_set_lf_line(trtt_hip_plan_t *plan, int line);

/* sino[view] += A_view vol for every view (device pointers, plan precision).
 * `stream` is a hipStream_t (NULL: the default stream).  Asynchronous. */
int trtt_hip_forward(trtt_hip_plan_t *plan, const void *vol, void *sino, void *stream);
/* vol += sum_view A_view' sino[view] (device pointers).  Asynchronous. */
int trtt_hip_backward(trtt_hip_plan_t *plan, const void *sino, void *vol, void *stream);
/* Weighted residual of the data term (Tomography_4D/Ytrtt4D_optim.i:418-446):
 *   r = w .* (sino - y) in place of sino;  *fx_dev += 0.5 * sum r .* (sino - y)
 * in FP64 (fx_dev is a device double).  w may be NULL (unit weights). */
int trtt_hip_weighted_residual(trtt_hip_plan_t *plan, void *sino, const void *y,
                               const void *w, double *fx_dev, void *stream);

/* Host-array evaluation of the weighted least-squares data term over every
 * view of a plan, for callers that hold host arrays (the Yorick loop): the
 * batched replacement of trtt_4D_mpy_eval (Tomography_4D/Ytrtt4D_optim.i:
 * 87-223) and of its serial loop (:418-446).  Host FP64 arrays throughout.
 * eval_set_data uploads the data y[view][nv][nu] and the weights w (NULL:
 * unit weights; the reference's W_k = diag(dweights_k weight_k)) in the
 * plan's precision and forms WCyT = sum_k A_k' W_k y_k once on the device
 * (:764-782); call it again when the data change.  Synchronous.
 * eval uploads x (volume series [nt][nz][ny][nx]) and returns
 *   *fx = 1/2 sum_k (A_k x - y_k)' W_k (A_k x - y_k)   (FP64 sum)
 *   gx  = sum_k A_k' W_k (A_k x - y_k)                  (split = 0), or
 *       = sum_k A_k' W_k A_k x - WCyT                   (split != 0: the
 *         reference's own split, :437-446; FP64 difference of the two terms).
 * Synchronous; the device copies stay with the plan until it is destroyed. */
int trtt_hip_eval_set_data(trtt_hip_plan_t *plan, const double *y, const double *w);
int trtt_hip_eval(trtt_hip_plan_t *plan, double *fx, double *gx, const double *x, int split);

/* 1 when a launch of this plan met a voxel in the source plane (the
 * reference's TRTT_FAILURE case), 0 otherwise, -1 on error.  Synchronous. */
int trtt_hip_plan_error_flag(trtt_hip_plan_t *plan);

/* Kernel family a plan uses: 0 = general footprint kernels, W > 0 = the
 * axial closed-form kernels with a footprint width bound of W pixels. */
int trtt_hip_plan_path(const trtt_hip_plan_t *plan);

/* Launch shape of the closed-form axial kernels of a plan (no effect on the
 * general kernels).  Results do not depend on it beyond FP32 summation order;
 * the autotuner (yoricktrtt_amd/autotune.py) times the candidates and keeps
 * the fastest per geometry.  set_tuning reallocates the forward scratch when
 * the tile or window changes: call it while no launch of the plan is in
 * flight. */
typedef struct trtt_hip_tuning {
  int fwd_tile;            /* 0: 8x4, 1: 4x4, 2: 8x8 voxel columns x 64 k */
  int bwd_tile;            /* 0: 8x8, 1: 8x4, 2: 4x4 voxel columns x 64 k;
                              3: 8x4, 4: 4x4, 5: 8x8 x 128 k (plans whose column
                              profiles over 128 slices exceed 256 pixels refuse them) */
  int fwd_window_bytes;    /* LDS detector window of a forward wave, 4096..65536 */
  int fwd_views_per_block; /* views per forward workgroup, 0 = from the grid size */
  int bwd_window_bytes;    /* LDS detector window of a backward workgroup, 4096..131072 */
  int fwd_chunk_views;     /* views per forward scratch round, 0 = as many as fit 4 GiB */
  int fwd_flush;           /* 1 (default): each tile's LDS window is added to the
                              projections with global atomic adds; 0: windows go to
                              scratch slots and a reduction kernel sums them (on a
                              second stream, overlapped with the next chunk) */
  int fwd_xcds;            /* 0 or 1: forward grid of tiles x view groups; n > 1 (atomic
                              flush): a 1-D grid in an XCD-aware order for a device that
                              deals consecutive workgroups to n XCDs in turn, so that all
                              tiles of one view group and k block run on one XCD and
                              their atomic adds stay in its L2 (MI355X: 8) */
} trtt_hip_tuning_t;
int trtt_hip_plan_get_tuning(const trtt_hip_plan_t *plan, trtt_hip_tuning_t *tuning);
int trtt_hip_plan_set_tuning(trtt_hip_plan_t *plan, const trtt_hip_tuning_t *tuning);

/* Kernel timing of the axial path: when on, the plan records HIP events on
 * the launch stream around its main forward kernel (one pair per chunk of
 * views) and its backward kernel (leave it off under graph capture).
 * kernel_time waits for the events of the last forward (job 0) or backward
 * (job 1) call and returns the summed duration in ms and the launch count. */
int trtt_hip_plan_set_timing(trtt_hip_plan_t *plan, int on);
int trtt_hip_plan_kernel_time(trtt_hip_plan_t *plan, int job, double *ms, int *launches);

/* Sparse system matrix of a 2-D plan (TRTT_HIP_PB2D, _FB2D or _LF2D; one
 * phase, no time weights), built and kept on the device.  Replaces the
 * matrix assembly of Ytrtt_2D_projectors.i:212-369 (_trtt_2D_sparse_coefs,
 * _trtt_2D_sparse_matrix), whose per-pixel loop calls trtt_PB2D_sparser /
 * trtt_FB2D_sparser (plugin/trtt_2D_projectors.c:393-536) or lf2d_sparser
 * (tomography-LongFessler-model/LongFessler2D.c:472-571), and the products
 * with the matrix (trtt_mvmult, plugin/trtt_common.c:62-69).
 *   row (view, bin) = view * nv + bin, column (pixel) = j * nx + i;
 *   entries: the non-zero coefficients of the plan's own footprint model,
 *   in the order view, pixel, bin (the reference's order);
 *   decimals  6: each coefficient rounded to 6 decimals, half to even, as
 *             the reference's "%f" text round trip leaves it;
 *            -1: exact coefficients.
 * Limits: views <= 65535, views x pixels < 2^31, entries < 2^31 - 512.
 * Synchronous; the matrix keeps its own device memory (COO, CSR and CSC:
 * entries x (3 x value size + 16) bytes). */
typedef struct trtt_hip_sparse trtt_hip_sparse_t;
trtt_hip_sparse_t *trtt_hip_sparse_create(trtt_hip_plan_t *plan, int decimals);
/* Number of entries, -1 for a null matrix. */
long trtt_hip_sparse_nnz(const trtt_hip_sparse_t *m);
/* Entries in the reference's order with 1-based indices, as the
 * MCops_coefs / MCops_row_indices / MCops_col_indices arrays of
 * Ytrtt_2D_projectors.i:326-350 (host arrays of nnz entries).  Synchronous. */
int trtt_hip_sparse_coo(const trtt_hip_sparse_t *m, double *coefs, long *rows, long *cols);
/* sino += M img; img += M' sino (device pointers, plan precision; img is
 * [ny][nx], sino [nviews][nv]).  Asynchronous on `stream`. */
int trtt_hip_sparse_forward(trtt_hip_sparse_t *m, const void *img, void *sino, void *stream);
int trtt_hip_sparse_backward(trtt_hip_sparse_t *m, const void *sino, void *img, void *stream);
int trtt_hip_sparse_destroy(trtt_hip_sparse_t *m);

/* Distance-driven parallel-beam projector over many views on the device
 * (the batched form of trtt_DD_PB3D; no reference counterpart): views k <
 * nviews of xos/xsd (nviews*12 each), volume [nz][ny][nx], sinogram
 * [nviews][nv][nu], plan precision.  backward: exact = 1 gives the adjoint of
 * forward, exact = 0 the reference's transpose (trtt_3D_projectors.c:
 * 909-916).  Asynchronous on `stream`; outputs accumulate. */
typedef struct trtt_hip_dd trtt_hip_dd_t;
trtt_hip_dd_t *trtt_hip_dd_create(int precision, long nx, long ny, long nz, double s_scl,
                                  long nu, long nv, double u_scl, double v_scl, const double *xos,
                                  const double *xsd, long nviews);
int trtt_hip_dd_forward(trtt_hip_dd_t *p, const void *vol, void *sino, void *stream);
int trtt_hip_dd_backward(trtt_hip_dd_t *p, const void *sino, void *vol, int exact, void *stream);
int trtt_hip_dd_destroy(trtt_hip_dd_t *p);

/* Temporal interpolation of a series of nt frames (images or volumes of npix
 * values, plan precision): the B_k operators of the dynamic systems,
 * _trtt_2D_time_interpolator (Ytrtt_2D_projectors.i:761-814) and the
 * interpolator of Ytrtt_3D_projectors.i:254-375, whose taps come from
 * spl_interp_coefs_new_spline.  View k has `taps` taps (idx[k*taps+q] in
 * [0, nt), weight w[k*taps+q]; host arrays, copied):
 *   job 0: out[k] += sum_q w[k*taps+q] in[idx[k*taps+q]]   (nviews output frames)
 *   job != 0: the transpose, out[t] += sum over the taps naming t of w in[k]
 *             (nt output frames; repeated phases of one view add up).
 * Frames <= 65535, phases <= 65535, 1..64 taps.  apply is asynchronous on
 * `stream` (device pointers, in != out). */
typedef struct trtt_hip_tinterp trtt_hip_tinterp_t;
trtt_hip_tinterp_t *trtt_hip_tinterp_create(int precision, long npix, long nt, long nviews,
                                            int taps, const int *idx, const double *w);
int trtt_hip_tinterp_apply(trtt_hip_tinterp_t *h, int job, const void *in, void *out,
                           void *stream);
int trtt_hip_tinterp_destroy(trtt_hip_tinterp_t *h);

long trtt_hip_plan_nviews(const trtt_hip_plan_t *plan);
long trtt_hip_plan_volume_size(const trtt_hip_plan_t *plan);   /* nt*nz*ny*nx */
long trtt_hip_plan_sino_size(const trtt_hip_plan_t *plan);     /* nviews*nv*nu */
int trtt_hip_plan_destroy(trtt_hip_plan_t *plan);

/* ===================================================================== */
/* Part 3 -- regularisation on device (stand-ins for YTotVar rgl_totvar and */
/* SPL spl_interp_apply, Ytrtt4D_optim.i:448-482; no source in the        */
/* reference tree)                                                         */
/* ===================================================================== */

/* Isotropic TV with threshold of a series of nt volumes [nt][nz][ny][nx]:
 *   *f_dev += weight * sum_i (sqrt(eps^2 + |D x|_i^2) - eps),  g += its gradient,
 * D = forward differences in x, y, z (zero at the last sample) and, for
 * nt > 1, wt times the periodic difference in t.  `work` holds
 * (nt > 1 ? 4 : 3) * nt*nz*ny*nx elements of the plan precision. */
int trtt_hip_tv(const void *x, void *g, void *work, double *f_dev, long nx, long ny, long nz,
                long nt, double weight, double wt, double eps, int precision, void *stream);

/* Samples at the integer nodes of cubic B-spline coefficients, per volume:
 * dst (+)= S src with S = [1/6, 2/3, 1/6] along x, y and z (x and y only
 * when nz == 1: a 2-D image), zero padding
 * (S is symmetric: it is also its own transpose).  src != dst. */
int trtt_hip_bspline_filter(const void *src, void *dst, long nx, long ny, long nz, long nt,
                            int accumulate, int precision, void *stream);

/* ===================================================================== */
/* Part 4 -- learned warm start: BF16 3-D U-Net layers on the matrix      */
/* cores (no counterpart in the reference; used by the warm start of the  */
/* VMLM-B reconstruction)                                                  */
/* ===================================================================== */
/* Tensors are NDHWC: [n][d][h][w][c].  BF16 tensors are raw 16-bit words
 * (`void *`); FP32 master weights are w[cout][kp] with k = tap*cin + c,
 * tap = 9 (dz+1) + 3 (dy+1) + (dx+1), zero past taps*cin; kp % 16 == 0.
 * Every function returns 0, or -1 on a bad argument or launch error. */

/* y = act(conv(x, wk) + bias) with zero padding, ks = 3 or 1 (taps = ks^3),
 * act = ReLU when relu != 0; y (BF16) and optional y32 (FP32) get
 * [n][d][h][w][cout].  cout <= 128, n*d <= 65535.  Run on dY with the
 * transposed weights wt (trtt_unet_prep_weights) it is the data gradient. */
int trtt_unet_conv_fwd(const void *x, const void *wk, const float *bias, void *y, float *y32,
                       int n, int d, int h, int w, int cin, int cout, int ks, int kp, int relu,
                       void *stream);

/* dw[cout][kp] += sum_p dy[p][co] xcol[p][k] and db[cout] += sum_p dy[p][co]
 * (db may be NULL); FP32 atomics. */
int trtt_unet_conv_wgrad(const void *x, const void *dy, float *dw, float *db, int n, int d, int h,
                         int w, int cin, int cout, int ks, int kp, void *stream);

/* BF16 operands of a layer from its FP32 master w[cout][kp]: wk[cout][kp]
 * and (if wt != NULL) the mirrored transpose wt[cin][kpt],
 * wt[c][tap'*cout + co] = w[co][(taps-1-tap')*cin + c]. */
int trtt_unet_prep_weights(const float *w, void *wk, void *wt, int taps, int cin, int cout, int kp,
                           int kpt, void *stream);

/* 2x2x2 max pooling of x [n][d][h][w][c] into y with the argmax map arg
 * (forward), or dx (FP32, fine grid, zeroed by the caller) from dy and arg
 * (backward).  d, h, w even. */
int trtt_unet_pool(const void *x, void *y, void *arg, float *dx, const void *dy, int n, int d,
                   int h, int w, int c, int backward, void *stream);

/* x2 nearest upsampling of src [n][d/2][h/2][w/2][c] into channels
 * [off, off + c) of dst [n][d][h][w][ctot] (forward), or its gradient: dst
 * (coarse, c channels) = sum over the 8 children of src's channel slice. */
int trtt_unet_upsample(const void *src, void *dst, int n, int d, int h, int w, int c, int ctot,
                       int off, int backward, void *stream);

/* dir 0: dst[p][off + ch] = src[p][ch]; dir 1: dst[p][ch] = src[p][off + ch]
 * (src/dst with ctot channels on the sliced side). */
int trtt_unet_channel_copy(const void *src, void *dst, long npos, int c, int ctot, int off,
                           int dir, void *stream);

/* dx = (dy32 + dy16) * (act > 0); either gradient may be NULL. */
int trtt_unet_relu_bwd(const float *dy32, const void *dy16, const void *act, void *dx, long n,
                       void *stream);

/* b = BF16(a), round to nearest even. */
int trtt_unet_cast_bf16(const float *a, void *b, long n, void *stream);

/* out = xin + head; with target != NULL: *loss += mean (out - target)^2
 * (FP64, loss may be NULL) and dhead (BF16, may be NULL) = 2 (out - target)/n. */
int trtt_unet_residual_mse(const float *xin, const float *head, const float *target, float *out,
                           void *dhead, long n, double *loss, void *stream);

/* One Adam step on n FP32 parameters; lr_t = lr sqrt(1-b2^t) / (1-b1^t). */
int trtt_unet_adam(float *p, const float *g, float *m, float *v, long n, float lr_t, float b1,
                   float b2, float eps, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TRTT_HIP_H */
