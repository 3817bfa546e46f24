/*
 * trtt_oracle.c -- FP64 CPU restatement of the TRTT B-spline footprint
 * projectors and of the Long & Fessler separable-footprint projector.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never
 * by the product library.  Every routine cites the reference function whose
 * arithmetic it restates (paths relative to the YorickTRTT source tree).
 *
 * Layout conventions (identical to the reference C plugin):
 *   volume   [nz][ny][nx]  index (k*ny + j)*nx + i
 *   image    [ny][nx]      index j*nx + i
 *   detector [nv][nu]      index v*nu + u   (u axial, v transaxial)
 *   xos, xsd: 12 row-major coefficients of a 3x4 affine transform.
 * All projectors ACCUMULATE into `out` (out += A x or out += A' y).
 *
 * Where a table-mode footprint is used the detector-coordinate arithmetic is
 * evaluated in the same FP64 operation order as the reference, because the
 * table index is a floor() of that coordinate and a different rounding would
 * move the index.  Build with -ffp-contract=off for the same reason.
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - golden vectors written by tests/golden/make_golden.py from the reference
 *     C plugin compiled by oracle/Makefile into oracle/_ref/;
 *   - closed forms (B-spline CDF, single-voxel footprint mass), adjoint
 *     identities, linearity, brute-force integration for the Long & Fessler
 *     trapezoid (its reference source needs yapi.h and is not buildable here).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_FAIL (-1)

/* ------------------------------------------------------------------------ */
/* Cubic B-spline and its cumulative integral                                */
/* ------------------------------------------------------------------------ */

/* Centred cubic B-spline beta3 on [-2,2].  Used to build the footprint table
 * (Ytrtt_tomobj.i:140-181 draws it through SPL; the values are the textbook
 * piecewise cubic). */
double orc_bspline3(double x)
{
    double t = fabs(x);
    if (t >= 2.0) return 0.0;
    if (t >= 1.0) {
        double s = 2.0 - t;
        return s * s * s / 6.0;
    }
    return 2.0 / 3.0 - t * t + 0.5 * t * t * t;
}

/* B-spline of degree 1 and 2 (centred), for footprint tables of other degrees.
 * SPL's convention for even degrees is not published: parity unpinned. */
double orc_bspline(double x, int deg)
{
    double t = fabs(x);
    switch (deg) {
    case 0:
        return t < 0.5 ? 1.0 : (t == 0.5 ? 0.5 : 0.0);
    case 1:
        return t < 1.0 ? 1.0 - t : 0.0;
    case 2:
        if (t < 0.5) return 0.75 - t * t;
        if (t < 1.5) return 0.5 * (1.5 - t) * (1.5 - t);
        return 0.0;
    default:
        return orc_bspline3(x);
    }
}

/* Integral of beta3 from -inf to x (trtt_3D_projectors_cubic.c:117-138,
 * trtt_2D_projectors_cubic.c:73-94).  Written with the symmetry
 * F(x) = 1/2 + sign(x) G(|x|); algebraically identical to the reference's
 * expanded quartics. */
double orc_bspline3_cdf(double x)
{
    double t = fabs(x), g;
    if (t >= 2.0) {
        g = 0.5;
    } else if (t > 1.0) {
        double s = 2.0 - t, s2 = s * s;
        g = 0.5 - s2 * s2 / 24.0;
    } else {
        double t2 = t * t;
        g = t * (2.0 / 3.0 + t2 * (t / 8.0 - 1.0 / 3.0));
    }
    return x < 0.0 ? 0.5 - g : 0.5 + g;
}

/* Footprint table (Ytrtt_tomobj.i:140-181 `_monodim_bspline`, :296-337
 * `_make_footprint`; trapezoid cumulative sum trtt_common.c:134-142).
 * n = 2*deg+3 nodes, m samples on [-(n-1)/2, (n-1)/2]. */
void orc_footprint_table(long m, int deg, double *footprint, double *cum,
                         double *coord, double *step_out)
{
    long n = 2 * (long)deg + 3, i;
    double step = (double)(n - 1) / (double)(m - 1);
    double half = 0.5 * (double)(n - 1);
    for (i = 0; i < m; ++i) {
        coord[i] = (double)i * step - half;
        footprint[i] = orc_bspline(coord[i], deg);
    }
    cum[0] = 0.0;
    for (i = 1; i < m; ++i)
        cum[i] = cum[i - 1] + 0.5 * step * (footprint[i - 1] + footprint[i]);
    *step_out = step;
}

/* ------------------------------------------------------------------------ */
/* One axis of the separable footprint                                       */
/* ------------------------------------------------------------------------ */

enum { FP_CUBIC = 0, FP_TABLE3D = 1, FP_TABLE2D = 2 };

typedef struct {
    int kind;            /* FP_CUBIC, FP_TABLE3D or FP_TABLE2D */
    const double *cum;   /* cumulative table (table kinds) */
    long m;              /* table size */
    double step;         /* table step */
    double coord0;       /* first table coordinate */
} orc_fp_t;

static long table_index(double x, long m)
{
    double c = x;
    if (!(c >= 0.0)) c = 0.0;              /* TRTT_MAX(x,0) */
    if (!(c <= (double)(m - 1))) c = (double)(m - 1); /* TRTT_MIN(.,m-1) */
    return (long)floor(c);
}

/* Weights of detector pixels [lo, lo+n) for a footprint centred at c with
 * scale h; scl is the caller's normalisation (1/h for cubic and 3-D table
 * kinds, 1/(h*step) for the 2-D table kind).
 *   cubic:   x0 = (lo-c)*scl                     (trtt_3D_projectors_cubic.c:262-275)
 *   table3d: x0 = ((lo-c) - h*coord0)*(scl/step) (trtt_3D_projectors.c:275-302)
 *   table2d: x0 = (lo - (c + coord0*h))*scl      (trtt_2D_projectors.c:92-108)
 * then x_{l+1} = x0 + (l+1)*s and w_l = F(x_{l+1}) - F(x_l). */
static void axis_weights(const orc_fp_t *fp, double lo, long n, double c,
                         double h, double scl, double *w)
{
    long l;
    if (fp->kind == FP_CUBIC) {
        double x0 = (lo - c) * scl, prev = orc_bspline3_cdf(x0);
        for (l = 0; l < n; ++l) {
            double nx = orc_bspline3_cdf(x0 + (double)(l + 1) * scl);
            w[l] = nx - prev;
            prev = nx;
        }
    } else {
        double s, x0, prev;
        if (fp->kind == FP_TABLE3D) {
            s = scl / fp->step;
            x0 = ((lo - c) - h * fp->coord0) * s;
        } else {
            s = scl;
            x0 = (lo - (c + fp->coord0 * h)) * s;
        }
        prev = fp->cum[table_index(x0, fp->m)];
        for (l = 0; l < n; ++l) {
            double nx = fp->cum[table_index(x0 + (double)(l + 1) * s, fp->m)];
            w[l] = nx - prev;
            prev = nx;
        }
    }
}

/* Footprint window clamp (trtt_3D_projectors.c:1293-1307,
 * trtt_2D_projectors.c:229-235).  `right_max` is n for every projector except
 * the 2-D closed-form ones, which clamp the right bound to n-1 so that the
 * last detector bin never receives a contribution
 * (trtt_2D_projectors_cubic.c:256-258, 367-369). */
static void window(double c, double hs, long n, int quirk_last_bin,
                   double *lo, double *hi)
{
    double l = c - hs, r = c + hs;
    if (l < 0.0) l = 0.0;
    else if (l <= (double)(n - 1)) l = floor(l);
    else l = (double)(n - 1);
    if (!quirk_last_bin) {
        if (r < 0.0) r = 0.0;
        else if (r <= (double)(n - 1)) r = floor(r) + 1.0;
        else r = (double)n;
    } else {
        if (r < 0.0) r = 0.0;
        else if (r <= (double)(n - 2)) r = floor(r) + 1.0;
        else r = (double)(n - 1);
    }
    *lo = l;
    *hi = r;
}

/* ------------------------------------------------------------------------ */
/* 3-D parallel / cone beam                                                  */
/* ------------------------------------------------------------------------ */

/* Per-view constants derived from XOS (object->source) and XSD
 * (source->detector index) as in trtt_CB3D (trtt_3D_projectors.c:1190-1232)
 * and trtt_PB3D (:967-1003). */
typedef struct {
    double a4, a8, a12, a4m, a4sqr, b3, c3;
    double nvx, nvy, nvz, nvc, nux, nuy, nuz, nuc;  /* cone numerators */
    double dx, dy, dz, dc;                            /* cone denominator */
    double o1, o2, o3, o4;                            /* source-frame depth */
    double vx, vy, vz, vcst, ux, uy, uz, ucst;        /* parallel affine */
    double pb_rho0, pb_hu, pb_hv, pb_su, pb_sv, pb_hsu, pb_hsv;
} view3d_t;

static void view3d_setup(view3d_t *g, const double *xos, const double *xsd,
                         double s, int s_deg, double u_scl, double v_scl)
{
    double a1 = xsd[0], a2 = xsd[1], a3 = xsd[2], a4 = xsd[3];
    double a5 = xsd[4], a6 = xsd[5], a7 = xsd[6], a8 = xsd[7];
    double a9 = xsd[8], a10 = xsd[9], a11 = xsd[10], a12 = xsd[11];
    double a1inv = -1.0 * (1.0 / a1);
    double b1 = a6 + a1inv * a5 * a2, c1 = a10 + a1inv * a9 * a2;
    double b2 = a7 + a1inv * a5 * a3, c2 = a11 + a1inv * a9 * a3;
    double b3 = a8 + a1inv * a5 * a4, c3 = a12 + a1inv * a9 * a4;
    double hs0 = 0.5 * (s_deg + 1);

    g->a4 = a4; g->a8 = a8; g->a12 = a12;
    g->a4m = -1.0 * a4; g->a4sqr = a4 * a4; g->b3 = b3; g->c3 = c3;
    /* cone: projective numerators / denominator (trtt_3D_projectors.c:1212-1222) */
    g->nvx = b1 * xos[4] + b2 * xos[8];  g->nux = c1 * xos[4] + c2 * xos[8];
    g->nvy = b1 * xos[5] + b2 * xos[9];  g->nuy = c1 * xos[5] + c2 * xos[9];
    g->nvz = b1 * xos[6] + b2 * xos[10]; g->nuz = c1 * xos[6] + c2 * xos[10];
    g->nvc = b1 * xos[7] + b2 * xos[11]; g->nuc = c1 * xos[7] + c2 * xos[11];
    g->dx = a1 * xos[0] + a2 * xos[4] + a3 * xos[8];
    g->dy = a1 * xos[1] + a2 * xos[5] + a3 * xos[9];
    g->dz = a1 * xos[2] + a2 * xos[6] + a3 * xos[10];
    g->dc = a1 * xos[3] + a2 * xos[7] + a3 * xos[11];
    g->o1 = xos[0]; g->o2 = xos[1]; g->o3 = xos[2]; g->o4 = xos[3];
    /* parallel: affine detector coordinates (trtt_3D_projectors.c:983-986) */
    g->vx = g->nvx; g->ux = g->nux;
    g->vy = g->nvy; g->uy = g->nuy;
    g->vz = g->nvz; g->uz = g->nuz;
    g->vcst = b1 * xos[7] + b2 * xos[11] + b3;
    g->ucst = c1 * xos[7] + c2 * xos[11] + c3;
    {
        double a1sqr = a1 * a1;
        double a5sqr = v_scl * v_scl * a5 * a5;
        double a9sqr = u_scl * u_scl * a9 * a9;
        double du = sqrt(1.0 + (a9sqr / (a1sqr + a5sqr)));
        double dv = sqrt(1.0 + (a5sqr / a1sqr));
        g->pb_rho0 = s * s * s * du * dv;
        g->pb_hu = s / u_scl;
        g->pb_hv = s / v_scl;
        g->pb_su = 1.0 / g->pb_hu;
        g->pb_sv = 1.0 / g->pb_hv;
        g->pb_hsu = hs0 * g->pb_hu;
        g->pb_hsv = hs0 * g->pb_hv;
    }
}

/* Generic 3-D footprint projector.  cone=0: trtt_PB3D / trtt_PB3D_cubic;
 * cone=1: trtt_CB3D / trtt_CB3D_cubic.  fp->kind selects the footprint.
 * job==0: out[nv][nu] += A in[nz][ny][nx]; job!=0: out[vol] += A' in[det].
 * Voxel geometry is accumulated in the reference's incremental order
 * (k, then j, then i; trtt_3D_projectors.c:1234-1261). */
static int sb3d_slab(double *out, const double *in, long nx, long ny, long nz,
                     double s, int s_deg, const orc_fp_t *fp, long nu, long nv,
                     double u_scl, double v_scl, const double *xos,
                     const double *xsd, int cone, int job, long k0, long k1);

int orc_sb3d(double *out, const double *in, long nx, long ny, long nz,
             double s, int s_deg, const orc_fp_t *fp, long nu, long nv,
             double u_scl, double v_scl, const double *xos, const double *xsd,
             int cone, int job)
{
    return sb3d_slab(out, in, nx, ny, nz, s, s_deg, fp, nu, nv, u_scl, v_scl,
                     xos, xsd, cone, job, 0, nz);
}

/* The projector restricted to the slices k0 <= k < k1.  The per-voxel
 * arithmetic does not depend on the restriction (each slice starts its
 * incremental sums from k).  job != 0 writes the slab only:
 * out[((k-k0)*ny + j)*nx + i]; job == 0 projects the slab's voxels of the
 * full volume `in`. */
static int sb3d_slab(double *out, const double *in, long nx, long ny, long nz,
                     double s, int s_deg, const orc_fp_t *fp, long nu, long nv,
                     double u_scl, double v_scl, const double *xos,
                     const double *xsd, int cone, int job, long k0, long k1)
{
    view3d_t g;
    long i, j, k, l, m;
    double hs0 = 0.5 * (s_deg + 1);
    double hu_init = s / u_scl, hv_init = s / v_scl, rho_init = s * s * s;
    double *wu = (double *)malloc(sizeof(double) * (size_t)(nu + 1));
    double *wv = (double *)malloc(sizeof(double) * (size_t)(nv + 1));
    if (wu == NULL || wv == NULL) { free(wu); free(wv); return ORC_FAIL; }
    view3d_setup(&g, xos, xsd, s, s_deg, u_scl, v_scl);
    (void)nz;

    for (k = k0; k < k1; ++k) {
        double wk1 = g.o3 * k + g.o4, nu1 = g.nuz * k + g.nuc;
        double nv1 = g.nvz * k + g.nvc, de1 = g.dz * k + g.dc;
        double pv1 = g.vz * k + g.vcst, pu1 = g.uz * k + g.ucst;
        for (j = 0; j < ny; ++j) {
            double wk2 = wk1 + g.o2 * j, nu2 = nu1 + g.nuy * j;
            double nv2 = nv1 + g.nvy * j, de2 = de1 + g.dy * j;
            double pv2 = pv1 + g.vy * j, pu2 = pu1 + g.uy * j;
            for (i = 0; i < nx; ++i) {
                double uc, vc, rho0, hu, hv, su, sv, hsu, hsv;
                double ul, ur, vl, vr;
                long npu, npv, base;
                if (cone) {
                    double wk = wk2 + g.o1 * i;
                    double num_u = nu2 + g.nux * i, num_v = nv2 + g.nvx * i;
                    double den = de2 + g.dx * i, q = g.a4m / den;
                    double uca12, vca8, vca8sqr, magn, du, dv;
                    uc = q * num_u + g.c3;
                    vc = q * num_v + g.b3;
                    if (wk == 0) {
                        fprintf(stderr, "orc_sb3d: voxel in the source plane\n");
                        free(wu); free(wv);
                        return ORC_FAIL;
                    }
                    uca12 = (uc - g.a12) * u_scl;
                    vca8 = (vc - g.a8) * v_scl;
                    vca8sqr = vca8 * vca8;
                    magn = g.a4m / wk;
                    du = sqrt(1.0 + ((uca12 * uca12) / (g.a4sqr + vca8sqr)));
                    dv = sqrt(1.0 + (vca8sqr / g.a4sqr));
                    rho0 = rho_init * (magn * magn) * du * dv;
                    hu = hu_init * magn * du;
                    hv = hv_init * magn * dv;
                    su = 1.0 / hu;
                    sv = 1.0 / hv;
                    hsu = hs0 * hu;
                    hsv = hs0 * hv;
                } else {
                    vc = pv2 + g.vx * i;
                    uc = pu2 + g.ux * i;
                    rho0 = g.pb_rho0;
                    hu = g.pb_hu; hv = g.pb_hv;
                    su = g.pb_su; sv = g.pb_sv;
                    hsu = g.pb_hsu; hsv = g.pb_hsv;
                }
                window(uc, hsu, nu, 0, &ul, &ur);
                window(vc, hsv, nv, 0, &vl, &vr);
                if (!(ul < ur && vl < vr)) continue;
                npu = (long)(ur - ul);
                npv = (long)(vr - vl);
                axis_weights(fp, ul, npu, uc, hu, su, wu);
                axis_weights(fp, vl, npv, vc, hv, sv, wv);
                base = (long)vl * nu + (long)ul;
                if (job == 0) {
                    double x = in[(k * ny + j) * nx + i];
                    for (m = 0; m < npv; ++m)
                        for (l = 0; l < npu; ++l)
                            out[base + m * nu + l] += wv[m] * wu[l] * x * rho0;
                } else {
                    double acc = 0.0;
                    for (m = 0; m < npv; ++m)
                        for (l = 0; l < npu; ++l)
                            acc += rho0 * wu[l] * wv[m] * in[base + m * nu + l];
                    out[((k - k0) * ny + j) * nx + i] += acc;
                }
            }
        }
    }
    free(wu);
    free(wv);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* 2-D parallel / fan beam                                                   */
/* ------------------------------------------------------------------------ */

/* Generic 2-D footprint projector.  fan=0: trtt_PB2D[_cubic]
 * (trtt_2D_projectors.c:186-268, trtt_2D_projectors_cubic.c:209-287);
 * fan=1: trtt_FB2D[_cubic] (trtt_2D_projectors.c:273-383,
 * trtt_2D_projectors_cubic.c:292-398).  Image [ny][nx], detector [nv]. */
int orc_sb2d(double *out, const double *in, long nx, long ny, double s,
             int s_deg, const orc_fp_t *fp, long nv, double v_scl,
             const double *xos, const double *xsd, int fan, int job)
{
    long i, j, l;
    double a1 = xsd[0], a2 = xsd[1], a4 = xsd[3], a5 = xsd[4];
    double a6 = xsd[5], a8 = xsd[7];
    double a1inv = -1.0 * (1.0 / a1);
    double a5a1 = a5 * a1inv;
    double b1 = a6 + a1inv * a5 * a2, b3 = a8 + a1inv * a5 * a4;
    double hs0 = 0.5 * (s_deg + 1);
    int cubic = fp->kind == FP_CUBIC;
    double *w = (double *)malloc(sizeof(double) * (size_t)(nv + 1));
    /* parallel-beam constants (trtt_2D_projectors.c:195-221) */
    double a5a1scl = a5a1 * v_scl;
    double pb_dv = sqrt(1 + a5a1scl * a5a1scl);
    double pb_rho0 = s * s * pb_dv;
    double pb_h = (pb_dv * s) / v_scl;
    double pb_scl = cubic ? 1 / pb_h : 1 / (pb_h * fp->step);
    double pb_hs = 0.5 * (s_deg + 1) * pb_h;
    double pb_b1o5 = b1 * xos[4], pb_b1o6 = b1 * xos[5];
    double pb_b1o8b3 = b1 * xos[7] + b3;
    /* fan-beam constants (trtt_2D_projectors.c:279-315) */
    double a4m = -1.0 * a4, a4sqr = a4 * a4;
    double b1o5 = b1 * xos[4], b1o6 = b1 * xos[5], b1o8 = b1 * xos[7];
    double d_i = a1 * xos[0] + a2 * xos[4];
    double d_j = a1 * xos[1] + a2 * xos[5];
    double d_c = a1 * xos[3] + a2 * xos[7];
    double rho_init = s * s, h_init = s / v_scl;
    if (w == NULL) return ORC_FAIL;
    /* The table kind for 2-D is FP_TABLE2D; the closed form keeps FP_CUBIC. */
    for (j = 0; j < ny; ++j) {
        double vct = pb_b1o6 * j + pb_b1o8b3;
        double numt = b1o6 * j + b1o8, dent = d_j * j + d_c;
        double wkt = xos[1] * j + xos[3];
        for (i = 0; i < nx; ++i) {
            double vc, rho0, h, scl, hs, vl, vr;
            long npix;
            if (fan) {
                double wk = wkt + xos[0] * i;
                double num = numt + b1o5 * i, den = dent + d_i * i;
                double magn, vca8, dv;
                vc = (a4m / den) * num + b3;
                if (wk == 0) {
                    fprintf(stderr, "orc_sb2d: pixel in the source plane\n");
                    free(w);
                    return ORC_FAIL;
                }
                magn = a4m / wk;
                vca8 = (vc - a8) * v_scl;
                dv = sqrt(1.0 + ((vca8 * vca8) / (a4sqr)));
                rho0 = rho_init * magn * dv;
                h = h_init * magn * dv;
                scl = cubic ? 1.0 / h : 1 / (h * fp->step);
                hs = hs0 * h;
            } else {
                vc = vct + pb_b1o5 * i;
                rho0 = pb_rho0; h = pb_h; scl = pb_scl; hs = pb_hs;
            }
            window(vc, hs, nv, cubic, &vl, &vr);
            if (!(vl < vr)) continue;
            npix = (long)(vr - vl);
            axis_weights(fp, vl, npix, vc, h, scl, w);
            if (job == 0) {
                double x = in[j * nx + i];
                for (l = 0; l < npix; ++l)
                    out[(long)vl + l] += rho0 * x * w[l];
            } else {
                double acc = 0.0;
                for (l = 0; l < npix; ++l)
                    acc += rho0 * in[(long)vl + l] * w[l];
                out[j * nx + i] += acc;
            }
        }
    }
    free(w);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Convenience entry points with the reference argument lists                */
/* ------------------------------------------------------------------------ */

static orc_fp_t make_fp(int kind, const double *cum, long m, double step,
                        const double *coord)
{
    orc_fp_t fp;
    fp.kind = kind;
    fp.cum = cum;
    fp.m = m;
    fp.step = step;
    fp.coord0 = coord ? coord[0] : 0.0;
    return fp;
}

int orc_PB3D(double *out, const double *in, long nx, long ny, long nz,
             double s, int deg, const double *cum, long m, double step,
             const double *coord, long nu, long nv, double us, double vs,
             const double *xos, const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_TABLE3D, cum, m, step, coord);
    return orc_sb3d(out, in, nx, ny, nz, s, deg, &fp, nu, nv, us, vs, xos, xsd, 0, job);
}

int orc_CB3D(double *out, const double *in, long nx, long ny, long nz,
             double s, int deg, const double *cum, long m, double step,
             const double *coord, long nu, long nv, double us, double vs,
             const double *xos, const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_TABLE3D, cum, m, step, coord);
    return orc_sb3d(out, in, nx, ny, nz, s, deg, &fp, nu, nv, us, vs, xos, xsd, 1, job);
}

int orc_PB3D_cubic(double *out, const double *in, long nx, long ny, long nz,
                   double s, int deg, long nu, long nv, double us, double vs,
                   const double *xos, const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_CUBIC, NULL, 0, 1.0, NULL);
    return orc_sb3d(out, in, nx, ny, nz, s, deg, &fp, nu, nv, us, vs, xos, xsd, 0, job);
}

int orc_CB3D_cubic(double *out, const double *in, long nx, long ny, long nz,
                   double s, int deg, long nu, long nv, double us, double vs,
                   const double *xos, const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_CUBIC, NULL, 0, 1.0, NULL);
    return orc_sb3d(out, in, nx, ny, nz, s, deg, &fp, nu, nv, us, vs, xos, xsd, 1, job);
}

int orc_PB2D(double *out, const double *in, long nx, long ny, double s, int deg,
             const double *cum, long m, double step, const double *coord,
             long nv, double vs, const double *xos, const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_TABLE2D, cum, m, step, coord);
    return orc_sb2d(out, in, nx, ny, s, deg, &fp, nv, vs, xos, xsd, 0, job);
}

int orc_FB2D(double *out, const double *in, long nx, long ny, double s, int deg,
             const double *cum, long m, double step, const double *coord,
             long nv, double vs, const double *xos, const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_TABLE2D, cum, m, step, coord);
    return orc_sb2d(out, in, nx, ny, s, deg, &fp, nv, vs, xos, xsd, 1, job);
}

int orc_PB2D_cubic(double *out, const double *in, long nx, long ny, double s,
                   int deg, long nv, double vs, const double *xos,
                   const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_CUBIC, NULL, 0, 1.0, NULL);
    return orc_sb2d(out, in, nx, ny, s, deg, &fp, nv, vs, xos, xsd, 0, job);
}

int orc_FB2D_cubic(double *out, const double *in, long nx, long ny, double s,
                   int deg, long nv, double vs, const double *xos,
                   const double *xsd, int job)
{
    orc_fp_t fp = make_fp(FP_CUBIC, NULL, 0, 1.0, NULL);
    return orc_sb2d(out, in, nx, ny, s, deg, &fp, nv, vs, xos, xsd, 1, job);
}

/* ------------------------------------------------------------------------ */
/* Long & Fessler separable footprint, 2-D flat-detector fan beam            */
/* ------------------------------------------------------------------------ */

/* Parameter block with the field order of lf2d_parameters_t
 * (LongFessler2D.h:16-28). */
typedef struct {
    double pixelSize, pixelStep, voxelSize, sourceDistance, detectorDistance;
    double objectOffset1, objectOffset2;
    long objectDimension1, objectDimension2, detectorDimension;
} orc_lf2d_t;

/* Integral over [a,b] of the ramp 0..1 rising over length c, i.e.
 * (b^2-a^2)/(2c) for a<b (LongFessler2D.c:26-30). */
static double ramp_integral(double a, double b, double c)
{
    return a < b ? (b + a) * (b - a) / (2.0 * c) : 0.0;
}

static double box_integral(double a, double b)
{
    return a < b ? b - a : 0.0;
}

static void sort4(double t[4])
{
    /* the same comparator network as LongFessler2D.c:103-109 */
    static const int pr[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
    int q;
    for (q = 0; q < 5; ++q) {
        double *a = &t[pr[q][0]], *b = &t[pr[q][1]];
        if (*a > *b) { double x = *a; *a = *b; *b = x; }
    }
}

/* lf2d_direct (job 0, LongFessler2D.c:41-141) and lf2d_adjoint (job 1,
 * :254-355).  Image [objectDimension2][objectDimension1], detector
 * [detectorDimension]; `clr` zeroes the destination first. */
/* Point-sampled trapezoid (the line-integral model of
 * lf2d_line_integ_direct / _adjoint / _sparser, LongFessler2D.c:143-252,
 * 357-470, 573-681): the height at bin centre k of the trapezoid rising over
 * [t0,t1], flat to t2, falling to t3, times a. */
static double lf_point(const double t[4], long k, double a)
{
    double x = (double)k;
    if (x >= t[0] && x < t[1]) return a * (x - t[0]) / (t[1] - t[0]);
    if (x >= t[1] && x < t[2]) return a;
    if (x >= t[2] && x < t[3]) return a * (x - t[3]) / (t[2] - t[3]);
    return 0.0;
}

static void lf2d_model(const orc_lf2d_t *p, double beta, double det_offset,
                       const double *src, double *dst, int clr, int job,
                       int line);

void orc_lf2d(const orc_lf2d_t *p, double beta, double det_offset,
              const double *src, double *dst, int clr, int job)
{
    lf2d_model(p, beta, det_offset, src, dst, clr, job, 0);
}

/* line = 0: lf2d_direct / lf2d_adjoint (bin-integrated trapezoid);
 * line = 1: lf2d_line_integ_direct / _adjoint (trapezoid sampled at the bin
 * centres, amplitude sqrt(1+q^2) voxelSize). */
void orc_lf2d_ex(const orc_lf2d_t *p, double beta, double det_offset,
                 const double *src, double *dst, int clr, int job, int line)
{
    lf2d_model(p, beta, det_offset, src, dst, clr, job, line);
}

static void lf2d_model(const orc_lf2d_t *p, double beta, double det_offset,
                       const double *src, double *dst, int clr, int job,
                       int line)
{
    long n1 = p->objectDimension1, n2 = p->objectDimension2;
    long nd = p->detectorDimension, i1, i2, k;
    double vs = p->voxelSize, ps = p->pixelSize, pst = p->pixelStep;
    double R = p->sourceDistance, D = p->detectorDistance;
    double cb = cos(beta), sb = sin(beta);
    double oy = cb * p->objectOffset1 + sb * p->objectOffset2;
    double ox = -sb * p->objectOffset1 + cb * p->objectOffset2;
    double x0 = ox - 0.5 * (n1 - 1) * vs, y0 = oy - 0.5 * (n2 - 1) * vs;
    double off = det_offset / pst + 0.5 * (nd - 1);
    double scale = (R + D) / pst, size = ps / pst;

    if (clr) memset(dst, 0, sizeof(double) * (size_t)(job == 0 ? nd : n1 * n2));
    for (i2 = 0; i2 < n2; ++i2) {
        double y = y0 + i2 * vs, ya = y - 0.5 * vs, yb = y + 0.5 * vs;
        for (i1 = 0; i1 < n1; ++i1) {
            double x = x0 + i1 * vs, xa = x - 0.5 * vs, xb = x + 0.5 * vs;
            double t[4], qx, qy, q, amp, acc = 0.0;
            long kf, kl;
            /* projected corner abscissae, in pixel units (:94-100) */
            double cx[4] = {xa, xb, xa, xb}, cy[4] = {ya, ya, yb, yb};
            int c;
            for (c = 0; c < 4; ++c)
                t[c] = scale * (cy[c] * sb + cx[c] * cb) /
                       (R - (cy[c] * cb - cx[c] * sb)) + off;
            sort4(t);
            qx = x + R * sb;
            qy = y - R * cb;
            q = fabs(qy) > fabs(qx) ? qx / qy : qy / qx;
            amp = line ? sqrt(1.0 + q * q) * vs : sqrt(1.0 + q * q) * (vs * ps);
            if (job == 0) amp *= src[i1 + i2 * n1];
            kf = (long)floor(t[0] - 0.5 * size);
            if (kf < 0) kf = 0;
            kl = (long)ceil(t[3] + 0.5 * size);
            if (kl >= nd) kl = nd - 1;
            for (k = kf; k <= kl && line; ++k) {
                if (job == 0) dst[k] += lf_point(t, k, amp);
                else acc += lf_point(t, k, amp) * src[k];
            }
            for (k = kf; k <= kl && !line; ++k) {
                double s1 = k - 0.5 * size, s2 = s1 + size;
                double lo1 = s1 > t[0] ? s1 : t[0], hi1 = s2 < t[1] ? s2 : t[1];
                double lo2 = s1 > t[1] ? s1 : t[1], hi2 = s2 < t[2] ? s2 : t[2];
                double lo3 = s2 < t[3] ? s2 : t[3], hi3 = s1 > t[2] ? s1 : t[2];
                double gam = ramp_integral(lo1 - t[0], hi1 - t[0], t[1] - t[0]) +
                             box_integral(lo2, hi2) +
                             ramp_integral(t[3] - lo3, t[3] - hi3, t[3] - t[2]);
                if (job == 0) dst[k] += gam * amp;
                else acc += amp * gam * src[k];
            }
            if (job != 0) dst[i1 + i2 * n1] += acc;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Multi-view drivers for the CPU baseline (OpenMP over views)               */
/* ------------------------------------------------------------------------ */

/* Same projector with the footprint given field by field (for ctypes). */
/* ------------------------------------------------------------------------ */
/* Distance-driven parallel-beam projector                                    */
/* (trtt_DD_PB3D, trtt_3D_projectors.c:1077-1173; footprint burns :406-525,   */
/* transpose :817-940).  Each voxel projects to the rectangle                 */
/* [uc - Du, uc + Du] x [vc - Dv, vc + Dv] with its corners clamped to        */
/* [0, n-1]; the weight of a pixel is the length of its overlap with the     */
/* rectangle along each axis, times u_scl / v_scl and h_norm.                 */
/* ------------------------------------------------------------------------ */

/* Overlap of pixel `left + p` with [b1, b2] (clamped bounds), scaled.
 * `stuck` restates the reference transpose along v (:909-916), whose loop
 * over the pixels inside the rectangle advances the u counter instead of the
 * v one: every inner pixel writes the same slot, the last slot write wins and
 * the slots after it stay zero. */
static double dd_overlap(double b1, double b2, long p, double scl, int stuck)
{
    const double up = ceil(b1), down = floor(b2);
    const long inner = (long)(down - up); /* pixels fully inside, or < 0 */
    const int low = up > b1, high = down < b2;
    if (inner < 0) return p == 0 ? scl * (b2 - b1) : 0.0;
    if (!stuck) {
        const long n = (long)(ceil(b2) - floor(b1));
        if (low && p == 0) return scl * (up - b1);
        if (high && p == n - 1) return scl * (b2 - down);
        return scl;
    }
    /* slots: [0] the lower partial pixel if any, then one shared slot */
    if (low && p == 0) return scl * (up - b1);
    if (p == (low ? 1 : 0)) return high ? scl * (b2 - down) : (inner > 0 ? scl : 0.0);
    return 0.0;
}

int orc_dd_pb3d(double *out, const double *in, long nx, long ny, long nz, double s,
                long nu, long nv, double us, double vs, const double *xos,
                const double *xsd, int job, int adjoint)
{
#pragma STDC FP_CONTRACT OFF
    const double a1inv = -1.0 * (1.0 / xsd[0]);
    const double b1 = xsd[5] + a1inv * xsd[4] * xsd[1], c1 = xsd[9] + a1inv * xsd[8] * xsd[1];
    const double b2 = xsd[6] + a1inv * xsd[4] * xsd[2], c2 = xsd[10] + a1inv * xsd[8] * xsd[2];
    const double b3 = xsd[7] + a1inv * xsd[4] * xsd[3], c3 = xsd[11] + a1inv * xsd[8] * xsd[3];
    const double v_x = b1 * xos[4] + b2 * xos[8], u_x = c1 * xos[4] + c2 * xos[8];
    const double v_y = b1 * xos[5] + b2 * xos[9], u_y = c1 * xos[5] + c2 * xos[9];
    const double v_z = b1 * xos[6] + b2 * xos[10], u_z = c1 * xos[6] + c2 * xos[10];
    const double v_c = b1 * xos[7] + b2 * xos[11] + b3;
    const double u_c = c1 * xos[7] + c2 * xos[11] + c3;
    double Du = fabs(u_z);
    double Dv = fabs(v_y) > fabs(v_x) ? fabs(v_y) : fabs(v_x);
    const double h = (s * s * s) / (us * vs * Du * Dv);
    Du = 0.5 * Du;
    Dv = 0.5 * Dv;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                const double vc = ((v_z * k + v_c) + v_y * j) + v_x * i;
                const double uc = ((u_z * k + u_c) + u_y * j) + u_x * i;
                const double ub1 = fmin(fmax(uc - Du, 0.0), (double)(nu - 1));
                const double ub2 = fmin(fmax(uc + Du, 0.0), (double)(nu - 1));
                const double vb1 = fmin(fmax(vc - Dv, 0.0), (double)(nv - 1));
                const double vb2 = fmin(fmax(vc + Dv, 0.0), (double)(nv - 1));
                const long nu_e = (long)(ceil(ub2) - floor(ub1));
                const long nv_e = (long)(ceil(vb2) - floor(vb1));
                if (nu_e <= 0 || nv_e <= 0) continue;
                const long corner = (long)floor(vb1) * nu + (long)floor(ub1);
                const long vox = (k * ny + j) * nx + i;
                const int stuck = job != 0 && !adjoint;
                for (long m = 0; m < nv_e; ++m) {
                    const double cv = dd_overlap(vb1, vb2, m, vs, stuck);
                    for (long l = 0; l < nu_e; ++l) {
                        const double cu = dd_overlap(ub1, ub2, l, us, 0);
                        const long pix = corner + m * nu + l;
                        if (job == 0)
                            out[pix] += h * cv * cu * in[vox];
                        else
                            out[vox] += h * cv * cu * in[pix];
                    }
                }
            }
    return 0;
}

int orc_sb3d_ex(double *out, const double *in, long nx, long ny, long nz,
                double s, int deg, int kind, const double *cum, long m,
                double step, double coord0, long nu, long nv, double us,
                double vs, const double *xos, const double *xsd, int cone,
                int job)
{
    orc_fp_t fp;
    fp.kind = kind; fp.cum = cum; fp.m = m; fp.step = step; fp.coord0 = coord0;
    return orc_sb3d(out, in, nx, ny, nz, s, deg, &fp, nu, nv, us, vs, xos, xsd,
                    cone, job);
}

int orc_sb2d_ex(double *out, const double *in, long nx, long ny, double s,
                int deg, int kind, const double *cum, long m, double step,
                double coord0, long nv, double vs, const double *xos,
                const double *xsd, int fan, int job)
{
    orc_fp_t fp;
    fp.kind = kind; fp.cum = cum; fp.m = m; fp.step = step; fp.coord0 = coord0;
    return orc_sb2d(out, in, nx, ny, s, deg, &fp, nv, vs, xos, xsd, fan, job);
}

/* Forward of `nview` views: sino[view][nv][nu] += A_view vol.  Views are
 * independent, so they are distributed over OpenMP threads. */
int orc_sb3d_views_fwd(double *sino, const double *vol, long nx, long ny,
                       long nz, double s, int deg, int kind, const double *cum,
                       long m, double step, double coord0, long nu, long nv,
                       double us, double vs, const double *xos,
                       const double *xsd, long nview, int cone)
{
    long v;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (v = 0; v < nview; ++v)
        bad |= orc_sb3d_ex(sino + v * nu * nv, vol, nx, ny, nz, s, deg, kind,
                           cum, m, step, coord0, nu, nv, us, vs, xos + 12 * v,
                           xsd + 12 * v, cone, 0) != ORC_OK;
    return bad ? ORC_FAIL : ORC_OK;
}

/* Number of footprint weights summed over the voxels of one view (the
 * npix_u*npix_v products of trtt_3D_projectors.c:1293-1314); used to count
 * the algorithmic work of a projection. */
double orc_sb3d_count_pairs(long nx, long ny, long nz, double s, int deg,
                            long nu, long nv, double us, double vs,
                            const double *xos, const double *xsd, int cone)
{
    view3d_t g;
    long i, j, k;
    double hs0 = 0.5 * (deg + 1), total = 0.0;
    view3d_setup(&g, xos, xsd, s, deg, us, vs);
    for (k = 0; k < nz; ++k)
        for (j = 0; j < ny; ++j)
            for (i = 0; i < nx; ++i) {
                double uc, vc, hsu, hsv, ul, ur, vl, vr;
                if (cone) {
                    double wk = g.o4 + g.o3 * k + g.o2 * j + g.o1 * i;
                    double den = g.dc + g.dz * k + g.dy * j + g.dx * i;
                    double q = g.a4m / den, magn = g.a4m / wk;
                    double vca8, uca12, du, dv;
                    uc = q * (g.nuc + g.nuz * k + g.nuy * j + g.nux * i) + g.c3;
                    vc = q * (g.nvc + g.nvz * k + g.nvy * j + g.nvx * i) + g.b3;
                    uca12 = (uc - g.a12) * us;
                    vca8 = (vc - g.a8) * vs;
                    du = sqrt(1.0 + uca12 * uca12 / (g.a4sqr + vca8 * vca8));
                    dv = sqrt(1.0 + vca8 * vca8 / g.a4sqr);
                    hsu = hs0 * (s / us) * magn * du;
                    hsv = hs0 * (s / vs) * magn * dv;
                } else {
                    uc = g.ucst + g.uz * k + g.uy * j + g.ux * i;
                    vc = g.vcst + g.vz * k + g.vy * j + g.vx * i;
                    hsu = g.pb_hsu;
                    hsv = g.pb_hsv;
                }
                window(uc, hsu, nu, 0, &ul, &ur);
                window(vc, hsv, nv, 0, &vl, &vr);
                if (ul < ur && vl < vr) total += (ur - ul) * (vr - vl);
            }
    return total;
}

/* Back projection of `nview` views restricted to the slices k0 <= k < k1:
 * vol[(k-k0)][ny][nx] += sum_view A_view' sino[view].  OpenMP over views,
 * each thread summing into a private slab that is added at the end (the
 * summation order of a view's contributions is that of orc_sb3d; the order
 * over views depends on the thread schedule, within FP64 rounding). */
int orc_sb3d_views_bwd(double *vol, const double *sino, long nx, long ny,
                       long nz, double s, int deg, int kind, const double *cum,
                       long m, double step, double coord0, long nu, long nv,
                       double us, double vs, const double *xos,
                       const double *xsd, long nview, int cone, long k0,
                       long k1)
{
    orc_fp_t fp;
    long nslab = (k1 - k0) * ny * nx;
    int bad = 0;
    if (k0 < 0 || k1 > nz || k0 >= k1) return ORC_FAIL;
    fp.kind = kind; fp.cum = cum; fp.m = m; fp.step = step; fp.coord0 = coord0;
#pragma omp parallel reduction(| : bad)
    {
        long v, q;
        double *acc = (double *)calloc((size_t)nslab, sizeof(double));
        if (acc == NULL) bad = 1;
#pragma omp for schedule(dynamic, 1)
        for (v = 0; v < nview; ++v)
            if (acc != NULL)
                bad |= sb3d_slab(acc, sino + v * nu * nv, nx, ny, nz, s, deg, &fp,
                                 nu, nv, us, vs, xos + 12 * v, xsd + 12 * v,
                                 cone, 1, k0, k1) != ORC_OK;
        if (acc != NULL) {
#pragma omp critical
            for (q = 0; q < nslab; ++q) vol[q] += acc[q];
            free(acc);
        }
    }
    return bad ? ORC_FAIL : ORC_OK;
}

/* 2-D projector over `nview` views (OpenMP over views).  job 0:
 * sino[view][nv] += A_view img; job != 0: img += sum_view A_view' sino[view]
 * (private per-thread images added at the end). */
int orc_sb2d_views(double *out, const double *in, long nx, long ny, double s,
                   int deg, int kind, const double *cum, long m, double step,
                   double coord0, long nv, double vs, const double *xos,
                   const double *xsd, long nview, int fan, int job)
{
    orc_fp_t fp;
    int bad = 0;
    fp.kind = kind; fp.cum = cum; fp.m = m; fp.step = step; fp.coord0 = coord0;
    if (job == 0) {
        long v;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
        for (v = 0; v < nview; ++v)
            bad |= orc_sb2d(out + v * nv, in, nx, ny, s, deg, &fp, nv, vs,
                            xos + 12 * v, xsd + 12 * v, fan, 0) != ORC_OK;
        return bad ? ORC_FAIL : ORC_OK;
    }
#pragma omp parallel reduction(| : bad)
    {
        long v, q;
        double *acc = (double *)calloc((size_t)(nx * ny), sizeof(double));
        if (acc == NULL) bad = 1;
#pragma omp for schedule(dynamic, 4)
        for (v = 0; v < nview; ++v)
            if (acc != NULL)
                bad |= orc_sb2d(acc, in + v * nv, nx, ny, s, deg, &fp, nv, vs,
                                xos + 12 * v, xsd + 12 * v, fan, 1) != ORC_OK;
        if (acc != NULL) {
#pragma omp critical
            for (q = 0; q < nx * ny; ++q) out[q] += acc[q];
            free(acc);
        }
    }
    return bad ? ORC_FAIL : ORC_OK;
}

/* Long & Fessler projector over `nview` views (beta[view], offset[view]).
 * job 0: sino[view][nd] = A_view img (each view's row cleared first);
 * job != 0: img += sum_view A_view' sino[view]. */
int orc_lf2d_views(const orc_lf2d_t *p, const double *beta,
                   const double *offset, const double *src, double *dst,
                   long nview, int job, int line)
{
    long nd = p->detectorDimension, npix = p->objectDimension1 * p->objectDimension2;
    int bad = 0;
    if (job == 0) {
        long v;
#pragma omp parallel for schedule(dynamic, 4)
        for (v = 0; v < nview; ++v)
            lf2d_model(p, beta[v], offset[v], src, dst + v * nd, 1, 0, line);
        return ORC_OK;
    }
#pragma omp parallel reduction(| : bad)
    {
        long v, q;
        double *acc = (double *)calloc((size_t)npix, sizeof(double));
        if (acc == NULL) bad = 1;
#pragma omp for schedule(dynamic, 4)
        for (v = 0; v < nview; ++v)
            if (acc != NULL)
                lf2d_model(p, beta[v], offset[v], src + v * nd, acc, 0, 1, line);
        if (acc != NULL) {
#pragma omp critical
            for (q = 0; q < npix; ++q) dst[q] += acc[q];
            free(acc);
        }
    }
    return bad ? ORC_FAIL : ORC_OK;
}
