// Sparse system matrix of a 2-D plan on the device: assembly into CSR (rows =
// (view, bin)) and CSC (columns = pixels) by stable radix sorts of the
// coordinate entries, and the two sparse matrix-vector products.  The
// products read each stored entry once (value + 32-bit index), so they are
// bound by HBM bandwidth; the dense vectors they gather from (an image of
// nx ny values, a sinogram of nviews nv values) stay in L2 / MALL.
#include <hipcub/hipcub.hpp>

#include "trtt_core.h"

namespace trtt {

namespace {

// Waves (rows) per workgroup of the products: neighbouring rows gather
// from the same lines of the vector, which a workgroup's CU keeps in L1.
// Measured on a 944 M-entry matrix (FP32, forward / transposed): 1 wave
// 2.03 / 2.35 ms, 4 waves 1.86 / 2.20, 8 waves 1.70 / 2.22, 16 waves
// 1.68 / 2.21.
#ifndef TRTT_SPMV_WAVES
#define TRTT_SPMV_WAVES 16
#endif
constexpr int kSpThreads = 256;
constexpr int kRowsPerBlock = TRTT_SPMV_WAVES;  // one wave per row
constexpr int kSpmvThreads = 64 * kRowsPerBlock;

struct ToLong {
  __host__ __device__ long long operator()(int x) const { return x; }
};

__global__ void iota_u32(unsigned *p, long n) {
  const long i = long(blockIdx.x) * kSpThreads + threadIdx.x;
  if (i < n) p[i] = unsigned(i);
}

// out_val[e] = val[perm[e]], out_idx[e] = idx[perm[e]]
template <typename T>
__global__ void gather_entries(const unsigned *__restrict__ perm, const T *__restrict__ val,
                               const int *__restrict__ idx, T *__restrict__ out_val,
                               int *__restrict__ out_idx, long n) {
  const long e = long(blockIdx.x) * kSpThreads + threadIdx.x;
  if (e >= n) return;
  const unsigned s = perm[e];
  out_val[e] = val[s];
  out_idx[e] = idx[s];
}

// ptr[r] = first entry of row r in the sorted keys, ptr[nrows] = n; every
// entry fills the pointers of the empty rows before its own.
__global__ void row_pointers(const unsigned *__restrict__ key, int *__restrict__ ptr, long n,
                             int nrows) {
  const long e = long(blockIdx.x) * kSpThreads + threadIdx.x;
  if (e > n) return;
  const int hi = e < n ? int(key[e]) : nrows;
  const int lo = e > 0 ? int(key[e - 1]) + 1 : 0;
  for (int r = lo; r <= hi; ++r) ptr[r] = int(e);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// y[r] += sum over the entries e of row r of val[e] x[idx[e]].  One wave per
// row; the 64 lanes take consecutive entries, so the gathers of x of one load
// touch few cache lines (the entries of a row are sorted by their index), and
// four independent partial sums per lane keep four loads of each stream in
// flight.  (Four entries per lane from 16-byte loads was slower: each gather
// then spans four times as many lines.)
template <typename T>
__global__ __launch_bounds__(kSpmvThreads) void spmv(const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y,
                                                   int nrows) {
  const int r = int(blockIdx.x) * kRowsPerBlock + int(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  const int b = ptr[r], e = ptr[r + 1];
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
  int k = b + lane;
  for (; k + 192 < e; k += 256) {
    const T v0 = val[k], v1 = val[k + 64], v2 = val[k + 128], v3 = val[k + 192];
    const int c0 = idx[k], c1 = idx[k + 64], c2 = idx[k + 128], c3 = idx[k + 192];
    s0 += v0 * x[c0];
    s1 += v1 * x[c1];
    s2 += v2 * x[c2];
    s3 += v3 * x[c3];
  }
  for (; k < e; k += 64) s0 += val[k] * x[idx[k]];
  const T s = wave_sum((s0 + s1) + (s2 + s3));
  if (lane == 0) y[r] += s;
}

}  // namespace

// Exclusive scan of n entry counts into 64-bit offsets.  temp / temp_bytes:
// hipcub's size query when temp is null.
hipError_t sparse_scan(const int *count, long long *offset, long n, void *temp,
                       size_t &temp_bytes, hipStream_t st) {
  hipcub::TransformInputIterator<long long, ToLong, const int *> it(count, ToLong());
  return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, it, offset, int(n), st);
}

// Stable sort of the entry ids 0..n-1 by key (keys < 2^bits); the sorted
// keys land in key_out and the ids in perm.  temp as for sparse_scan.
hipError_t sparse_sort(const unsigned *key, unsigned *key_out, unsigned *ids, unsigned *perm,
                       long n, int bits, void *temp, size_t &temp_bytes, hipStream_t st) {
  if (temp && n > 0) {
    hipLaunchKernelGGL(iota_u32, dim3(unsigned((n + kSpThreads - 1) / kSpThreads)),
                       dim3(kSpThreads), 0, st, ids, n);
  }
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, key, key_out, ids, perm, int(n), 0,
                                            bits, st);
}

// Compressed form from the sorted keys and the permutation: values and the
// other index of each entry in sorted order, and the n_major + 1 pointers.
hipError_t sparse_compress(bool f64, const unsigned *key_sorted, const unsigned *perm,
                           const void *val, const int *other, void *out_val, int *out_idx,
                           int *ptr, long n, int n_major, hipStream_t st) {
  const unsigned blocks = unsigned((n + kSpThreads - 1) / kSpThreads);
  if (n > 0) {
    if (f64)
      hipLaunchKernelGGL(gather_entries<double>, dim3(blocks), dim3(kSpThreads), 0, st, perm,
                         static_cast<const double *>(val), other, static_cast<double *>(out_val),
                         out_idx, n);
    else
      hipLaunchKernelGGL(gather_entries<float>, dim3(blocks), dim3(kSpThreads), 0, st, perm,
                         static_cast<const float *>(val), other, static_cast<float *>(out_val),
                         out_idx, n);
  }
  hipLaunchKernelGGL(row_pointers, dim3(unsigned((n + 1 + kSpThreads - 1) / kSpThreads)),
                     dim3(kSpThreads), 0, st, key_sorted, ptr, n, n_major);
  return hipGetLastError();
}

// y += M x for a compressed matrix with n_major rows (ptr, idx, val).
hipError_t sparse_spmv(bool f64, const int *ptr, const int *idx, const void *val, const void *x,
                       void *y, int n_major, hipStream_t st) {
  if (n_major <= 0) return hipSuccess;
  const unsigned blocks = unsigned((n_major + kRowsPerBlock - 1) / kRowsPerBlock);
  if (f64)
    hipLaunchKernelGGL(spmv<double>, dim3(blocks), dim3(kSpmvThreads), 0, st, ptr, idx,
                       static_cast<const double *>(val), static_cast<const double *>(x),
                       static_cast<double *>(y), n_major);
  else
    hipLaunchKernelGGL(spmv<float>, dim3(blocks), dim3(kSpmvThreads), 0, st, ptr, idx,
                       static_cast<const float *>(val), static_cast<const float *>(x),
                       static_cast<float *>(y), n_major);
  return hipGetLastError();
}

}  // namespace trtt
