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
 * Replaces plugin/trtt_3D_projectors.h:396, :520 (trtt_3D_projectors.c:959, :1179). */
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

/* 1 when a launch of this plan met a voxel in the source plane (the
 * reference's TRTT_FAILURE case), 0 otherwise, -1 on error.  Synchronous. */
int trtt_hip_plan_error_flag(trtt_hip_plan_t *plan);

/* Kernel family a plan uses: 0 = general footprint kernels, W > 0 = the
 * axial closed-form kernels with a footprint width bound of W pixels. */
int trtt_hip_plan_path(const trtt_hip_plan_t *plan);

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
 * dst (+)= S src with S = [1/6, 2/3, 1/6] along x, y and z, zero padding
 * (S is symmetric: it is also its own transpose).  src != dst. */
int trtt_hip_bspline_filter(const void *src, void *dst, long nx, long ny, long nz, long nt,
                            int accumulate, int precision, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TRTT_HIP_H */
