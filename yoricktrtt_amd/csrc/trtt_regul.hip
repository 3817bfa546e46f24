// Regularisation kernels of the reconstruction: isotropic total variation
// with a threshold, and the cubic B-spline coefficient -> sample filter.
//
// TV (this build's definition; the reference calls YTotVar's rgl_totvar,
// which is not available: Ytrtt4D_optim.i:448-482):
//   TV(x) = w * sum_i ( sqrt(eps^2 + |D x|_i^2) - eps )
// with forward differences along x, y, z (zero at the last sample) and, for
// a series of nt > 1 volumes, a periodic difference along t scaled by wt
// (the reference extends t periodically, Ytrtt4D_optim.i:225-255).
//
// B-spline samples of cubic coefficients at the integer nodes: separable
// [1/6, 2/3, 1/6] along x, y, z (x, y only for a single slice) with zero
// padding (symmetric, so the same filter is its own transpose).
#include "trtt_core.h"

namespace trtt {
namespace regul {

constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ T at(const T *x, long nx, long ny, long nz, long i, long j, long k) {
  return x[(k * ny + j) * nx + i];
}

// Pass 1: normalised differences n_d = D_d x / s and the FP64 value.
template <typename T>
__global__ __launch_bounds__(kThreads) void tv_pass1(const T *__restrict__ x, T *__restrict__ nrm,
                                                     long nx, long ny, long nz, long nt, double wt,
                                                     double eps, double weight,
                                                     double *__restrict__ f) {
  const long nvox = nx * ny * nz, ntot = nvox * nt;
  const int ncomp = nt > 1 ? 4 : 3;
  double local = 0.0;
  for (long idx = long(blockIdx.x) * kThreads + threadIdx.x; idx < ntot;
       idx += long(gridDim.x) * kThreads) {
    const long t = idx / nvox, r = idx - t * nvox;
    const long k = r / (nx * ny), r2 = r - k * nx * ny, j = r2 / nx, i = r2 - j * nx;
    const T *xt = x + t * nvox;
    const T c = xt[r];
    const T dx = i + 1 < nx ? xt[r + 1] - c : T(0);
    const T dy = j + 1 < ny ? xt[r + nx] - c : T(0);
    const T dz = k + 1 < nz ? xt[r + nx * ny] - c : T(0);
    T dt = T(0);
    if (nt > 1) {
      const long tn = t + 1 < nt ? t + 1 : 0;
      dt = T(wt) * (x[tn * nvox + r] - c);
    }
    const T s = sqrt(T(eps * eps) + dx * dx + dy * dy + dz * dz + dt * dt);
    local += double(s) - eps;
    const T inv = s > T(0) ? T(1) / s : T(0);
    nrm[0 * ntot + idx] = dx * inv;
    nrm[1 * ntot + idx] = dy * inv;
    nrm[2 * ntot + idx] = dz * inv;
    if (ncomp == 4) nrm[3 * ntot + idx] = T(wt) * dt * inv;
  }
  __shared__ double red[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kThreads / 64; ++q) s += red[q];
    atomicAdd(f, weight * s);
  }
}

// Pass 2: g += w * (D^T n) (the adjoint of the forward differences).
template <typename T>
__global__ __launch_bounds__(kThreads) void tv_pass2(const T *__restrict__ nrm, T *__restrict__ g,
                                                     long nx, long ny, long nz, long nt,
                                                     double weight) {
  const long nvox = nx * ny * nz, ntot = nvox * nt;
  for (long idx = long(blockIdx.x) * kThreads + threadIdx.x; idx < ntot;
       idx += long(gridDim.x) * kThreads) {
    const long t = idx / nvox, r = idx - t * nvox;
    const long k = r / (nx * ny), r2 = r - k * nx * ny, j = r2 / nx, i = r2 - j * nx;
    T acc = T(0);
    // d/dx_c of sum_d s_d: -n_d[c] (as the base of its difference) + n_d[c - e_d]
    if (i + 1 < nx) acc -= nrm[0 * ntot + idx];
    if (i > 0) acc += nrm[0 * ntot + idx - 1];
    if (j + 1 < ny) acc -= nrm[1 * ntot + idx];
    if (j > 0) acc += nrm[1 * ntot + idx - nx];
    if (k + 1 < nz) acc -= nrm[2 * ntot + idx];
    if (k > 0) acc += nrm[2 * ntot + idx - nx * ny];
    if (nt > 1) {
      const long tp = t > 0 ? t - 1 : nt - 1;
      acc += nrm[3 * ntot + tp * nvox + r] - nrm[3 * ntot + idx];
    }
    g[idx] += T(weight) * acc;
  }
}

// dst = S src, S the separable [1/6, 2/3, 1/6] filter with zero padding,
// per volume of a series.
template <typename T>
__global__ __launch_bounds__(kThreads) void bspline_filter(const T *__restrict__ src,
                                                           T *__restrict__ dst, long nx, long ny,
                                                           long nz, long nt, int accumulate) {
  const long nvox = nx * ny * nz, ntot = nvox * nt;
  const T wv[3] = {T(1.0 / 6.0), T(2.0 / 3.0), T(1.0 / 6.0)};
  for (long idx = long(blockIdx.x) * kThreads + threadIdx.x; idx < ntot;
       idx += long(gridDim.x) * kThreads) {
    const long t = idx / nvox, r = idx - t * nvox;
    const long k = r / (nx * ny), r2 = r - k * nx * ny, j = r2 / nx, i = r2 - j * nx;
    const T *s = src + t * nvox;
    T acc = T(0);
    // a single slice (nz == 1) is a 2-D image: no filtering along z
    const int cz = nz > 1 ? 1 : 0;
    for (int c = -cz; c <= cz; ++c) {
      const long kk = k + c;
      if (kk < 0 || kk >= nz) continue;
      for (int b = -1; b <= 1; ++b) {
        const long jj = j + b;
        if (jj < 0 || jj >= ny) continue;
        T row = T(0);
        for (int a = -1; a <= 1; ++a) {
          const long ii = i + a;
          if (ii < 0 || ii >= nx) continue;
          row += wv[a + 1] * s[(kk * ny + jj) * nx + ii];
        }
        acc += (cz ? wv[c + 1] : T(1)) * wv[b + 1] * row;
      }
    }
    if (accumulate)
      dst[idx] += acc;
    else
      dst[idx] = acc;
  }
}

unsigned grid_for(long n) {
  const long b = (n + kThreads - 1) / kThreads;
  return unsigned(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace regul
}  // namespace trtt

namespace trtt {
namespace regul {

// Temporal interpolation of a series of frames (images or volumes of npix
// values): B_k of the dynamic systems (_trtt_2D_time_interpolator,
// Ytrtt_2D_projectors.i:761-814; Ytrtt_3D_projectors.i:254-375), with taps
// (phase index, weight) per view.  Forward: frame k = sum over its taps of
// w x[phase]; transpose: phase t += sum over the taps that name t of w y[k],
// gathered from a per-phase list (no atomics, fixed order).
template <typename T>
__global__ __launch_bounds__(kThreads) void tinterp_fwd(const T *__restrict__ x, T *__restrict__ out,
                                                        long npix, int taps,
                                                        const int *__restrict__ idx,
                                                        const T *__restrict__ w) {
  const long p = long(blockIdx.x) * kThreads + threadIdx.x;
  const int k = blockIdx.y;
  if (p >= npix) return;
  T s = T(0);
  for (int q = 0; q < taps; ++q) s += w[k * taps + q] * x[long(idx[k * taps + q]) * npix + p];
  out[long(k) * npix + p] += s;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void tinterp_bwd(const T *__restrict__ y, T *__restrict__ out,
                                                        long npix, const int *__restrict__ ptr,
                                                        const int *__restrict__ src,
                                                        const T *__restrict__ w) {
  const long p = long(blockIdx.x) * kThreads + threadIdx.x;
  const int t = blockIdx.y;
  if (p >= npix) return;
  T s = T(0);
  for (int e = ptr[t]; e < ptr[t + 1]; ++e) s += w[e] * y[long(src[e]) * npix + p];
  out[long(t) * npix + p] += s;
}

}  // namespace regul

// job 0: out[k] += sum_q w[k taps + q] in[idx[k taps + q]] for the nframes
// views; job 1: out[t] += sum over the list of phase t of w in[src].
hipError_t launch_tinterp(bool f64, int job, const void *in, void *out, long npix, int nframes,
                          int taps, const int *idx, const void *w, const int *ptr,
                          const int *src, const void *wb, hipStream_t st) {
  using namespace regul;
  dim3 grid(unsigned((npix + kThreads - 1) / kThreads), unsigned(nframes));
  if (job == 0) {
    if (f64)
      hipLaunchKernelGGL(tinterp_fwd<double>, grid, dim3(kThreads), 0, st,
                         static_cast<const double *>(in), static_cast<double *>(out), npix, taps,
                         idx, static_cast<const double *>(w));
    else
      hipLaunchKernelGGL(tinterp_fwd<float>, grid, dim3(kThreads), 0, st,
                         static_cast<const float *>(in), static_cast<float *>(out), npix, taps,
                         idx, static_cast<const float *>(w));
  } else {
    if (f64)
      hipLaunchKernelGGL(tinterp_bwd<double>, grid, dim3(kThreads), 0, st,
                         static_cast<const double *>(in), static_cast<double *>(out), npix, ptr,
                         src, static_cast<const double *>(wb));
    else
      hipLaunchKernelGGL(tinterp_bwd<float>, grid, dim3(kThreads), 0, st,
                         static_cast<const float *>(in), static_cast<float *>(out), npix, ptr, src,
                         static_cast<const float *>(wb));
  }
  return hipGetLastError();
}

}  // namespace trtt

extern "C" {

int trtt_hip_tv(const void *x, void *g, void *work, double *f_dev, long nx, long ny, long nz,
                long nt, double weight, double wt, double eps, int precision, void *stream) {
  using namespace trtt::regul;
  if (!x || !g || !work || !f_dev || nx <= 0 || ny <= 0 || nz <= 0 || nt <= 0) return -1;
  const long n = nx * ny * nz * nt;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == TRTT_HIP_F64) {
    hipLaunchKernelGGL(tv_pass1<double>, dim3(grid_for(n)), dim3(kThreads), 0, st,
                       static_cast<const double *>(x), static_cast<double *>(work), nx, ny, nz,
                       nt, wt, eps, weight, f_dev);
    hipLaunchKernelGGL(tv_pass2<double>, dim3(grid_for(n)), dim3(kThreads), 0, st,
                       static_cast<const double *>(work), static_cast<double *>(g), nx, ny, nz,
                       nt, weight);
  } else {
    hipLaunchKernelGGL(tv_pass1<float>, dim3(grid_for(n)), dim3(kThreads), 0, st,
                       static_cast<const float *>(x), static_cast<float *>(work), nx, ny, nz, nt,
                       wt, eps, weight, f_dev);
    hipLaunchKernelGGL(tv_pass2<float>, dim3(grid_for(n)), dim3(kThreads), 0, st,
                       static_cast<const float *>(work), static_cast<float *>(g), nx, ny, nz, nt,
                       weight);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_hip_bspline_filter(const void *src, void *dst, long nx, long ny, long nz, long nt,
                            int accumulate, int precision, void *stream) {
  using namespace trtt::regul;
  if (!src || !dst || src == dst || nx <= 0 || ny <= 0 || nz <= 0 || nt <= 0) return -1;
  const long n = nx * ny * nz * nt;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == TRTT_HIP_F64)
    hipLaunchKernelGGL(bspline_filter<double>, dim3(grid_for(n)), dim3(kThreads), 0, st,
                       static_cast<const double *>(src), static_cast<double *>(dst), nx, ny, nz,
                       nt, accumulate);
  else
    hipLaunchKernelGGL(bspline_filter<float>, dim3(grid_for(n)), dim3(kThreads), 0, st,
                       static_cast<const float *>(src), static_cast<float *>(dst), nx, ny, nz, nt,
                       accumulate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
