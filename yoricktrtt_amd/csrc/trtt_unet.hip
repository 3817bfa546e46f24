// Kernels of the learned warm start: a small 3-D U-Net in BF16 with FP32
// accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x16_bf16).
//
// Tensors are NDHWC (channels fastest); activations and their gradients are
// BF16, master weights, weight gradients and the Adam state FP32.  A 3x3x3
// (or 1x1x1) convolution with zero padding is an implicit GEMM
//   Y[p][co] = sum_k Xcol[p][k] W[co][k],  k = tap*cin + c,
// with K = taps*cin padded to kp (a multiple of 16).  Tap t is the offset
// (t/9 - 1, (t/3)%3 - 1, t%3 - 1) along (z, y, x).  The data gradient is the
// same kernel run on dY with the mirrored, transposed weights
//   Wt[c][tap'*cout + co] = W[co][(taps-1-tap')*cin + c];
// the weight gradient is a GEMM over positions, split over workgroups and
// reduced with FP32 atomics.
//
// MFMA 16x16x16 BF16 fragments (wave64): lane l holds
//   A[l % 16][4*(l/16) + j], B[4*(l/16) + j][l % 16], D[4*(l/16) + j][l % 16]
// for j = 0..3.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace trtt {
namespace unet {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return uint16_t((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return uint16_t(u >> 16);
}

struct Shape {
  int n, d, h, w;  // batch and volume
};

__device__ __forceinline__ void tap_offsets(int KS, int t, int &dz, int &dy, int &dx) {
  if (KS == 3) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
  } else {
    dz = dy = dx = 0;
  }
}

__device__ __forceinline__ bool inside(const Shape &s, int z, int y, int x) {
  return z >= 0 && z < s.d && y >= 0 && y < s.h && x >= 0 && x < s.w;
}

__device__ __forceinline__ long pos_index(const Shape &s, int n, int z, int y, int x) {
  return ((long(n) * s.d + z) * s.h + y) * s.w + x;
}

__device__ __forceinline__ bf16x4 unpack4(uint2 raw) {
  bf16x4 v;
  v[0] = short(raw.x & 0xffff);
  v[1] = short(raw.x >> 16);
  v[2] = short(raw.y & 0xffff);
  v[3] = short(raw.y >> 16);
  return v;
}

// Y = act(conv(X, W) + b), written as BF16 (and FP32 when y32 is given).
// grid: (ceil(w/16), h, n*d); one wave computes 16 positions along x times
// NT*16 output channels.  VEC: cin is a multiple of 16, so the 4 values of a
// lane's A fragment are 4 consecutive channels of one tap (one 8-byte load)
// and the tap advances at most once per 16-wide K step.
template <int KS, int NT, bool VEC>
__global__ __launch_bounds__(64) void conv_fwd(const uint16_t *__restrict__ x,
                                               const uint16_t *__restrict__ wk,
                                               const float *__restrict__ bias,
                                               uint16_t *__restrict__ y, float *__restrict__ y32,
                                               Shape s, int cin, int cout, int kp, int relu) {
  const int lane = threadIdx.x;
  const int x0 = blockIdx.x * 16, yrow = blockIdx.y;
  const int n = blockIdx.z / s.d, z = blockIdx.z - n * s.d;
  const int K = KS * KS * KS * cin;
  const int m = lane & 15, kq = 4 * (lane >> 4);
  const int xm = x0 + m;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  int tap = 0, c = kq;  // VEC: k0 + kq = tap*cin + c
  for (int k0 = 0; k0 < kp; k0 += 16) {
    bf16x4 a = bf16x4{0, 0, 0, 0};
    if (VEC) {
      int dz, dy, dx;
      tap_offsets(KS, tap, dz, dy, dx);
      if (k0 + kq < K && xm < s.w && inside(s, z + dz, yrow + dy, xm + dx))
        a = unpack4(*reinterpret_cast<const uint2 *>(
            x + pos_index(s, n, z + dz, yrow + dy, xm + dx) * cin + c));
      c += 16;
      if (c >= cin) {
        c -= cin;
        ++tap;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + kq + j;
        if (k >= K || xm >= s.w) continue;
        const int t = k / cin, ch = k - t * cin;
        int dz, dy, dx;
        tap_offsets(KS, t, dz, dy, dx);
        if (inside(s, z + dz, yrow + dy, xm + dx))
          a[j] = short(x[pos_index(s, n, z + dz, yrow + dy, xm + dx) * cin + ch]);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = t * 16 + m;
      bf16x4 b = bf16x4{0, 0, 0, 0};
      if (co < cout) b = unpack4(*reinterpret_cast<const uint2 *>(wk + long(co) * kp + k0 + kq));
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = t * 16 + m;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int xp = x0 + kq + j;
      if (xp >= s.w || co >= cout) continue;
      float v = acc[t][j] + (bias ? bias[co] : 0.f);
      if (relu) v = v > 0.f ? v : 0.f;
      const long p = pos_index(s, n, z, yrow, xp);
      y[p * cout + co] = f2bf(v);
      if (y32) y32[p * cout + co] = v;
    }
  }
}

// dW[co][k] += sum_p dY[p][co] Xcol[p][k] over a range of 16-position groups
// (16 consecutive x of one row).  grid: (group chunks, kp/16, ceil(cout/16)).
// A = dY^T (rows co), B = Xcol (columns k), K of the MFMA = positions.
template <int KS>
__global__ __launch_bounds__(64) void conv_wgrad(const uint16_t *__restrict__ x,
                                                 const uint16_t *__restrict__ dy,
                                                 float *__restrict__ dw, Shape s, int cin, int cout,
                                                 int kp, int chunk) {
  const int lane = threadIdx.x;
  const int gw = (s.w + 15) / 16;
  const int ngroups = s.n * s.d * s.h * gw;
  const int g0 = blockIdx.x * chunk;
  const int g1 = g0 + chunk < ngroups ? g0 + chunk : ngroups;
  const int k0 = blockIdx.y * 16, c0 = blockIdx.z * 16;
  const int K = KS * KS * KS * cin;
  const int r = lane & 15, kq = 4 * (lane >> 4);
  // this lane's B column k = k0 + r: its tap and channel are loop invariant
  const int kk = k0 + r;
  const bool kvalid = kk < K;
  const int tap = kvalid ? kk / cin : 0, ch = kvalid ? kk - tap * cin : 0;
  int ddz, ddy, ddx;
  tap_offsets(KS, tap, ddz, ddy, ddx);
  const int co = c0 + r;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int g = g0; g < g1; ++g) {
    const int row = g / gw, xg = (g - row * gw) * 16;
    const int yy = row % s.h, rz = row / s.h;
    const int zz = rz % s.d, nn = rz / s.d;
    const bool rowin = kvalid && inside(s, zz + ddz, yy + ddy, 0);
    bf16x4 a = bf16x4{0, 0, 0, 0}, b = bf16x4{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int xp = xg + kq + j;
      if (xp >= s.w) continue;
      if (co < cout) a[j] = short(dy[pos_index(s, nn, zz, yy, xp) * cout + co]);
      const int xs = xp + ddx;
      if (rowin && xs >= 0 && xs < s.w)
        b[j] = short(x[pos_index(s, nn, zz + ddz, yy + ddy, xs) * cin + ch]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cr = c0 + kq + j;
    if (cr < cout && kk < kp && acc[j] != 0.f) unsafeAtomicAdd(dw + long(cr) * kp + kk, acc[j]);
  }
}

// db[co] += sum_p dY[p][co]
__global__ __launch_bounds__(256) void bias_grad(const uint16_t *__restrict__ dy,
                                                 float *__restrict__ db, long npos, int cout) {
  const int lanes = 256 / cout;
  const int sub = threadIdx.x / cout, co = threadIdx.x - sub * cout;
  if (sub >= lanes) return;
  float s = 0.f;
  for (long p = long(blockIdx.x) * lanes + sub; p < npos; p += long(gridDim.x) * lanes)
    s += bf2f(dy[p * cout + co]);
  unsafeAtomicAdd(db + co, s);
}

// BF16 operands from the FP32 master w[co][kp] (k = tap*cin + c, zero
// padding past taps*cin): wk[co][kp] and the mirrored transpose wt[c][kpt].
__global__ void prep_weights(const float *__restrict__ w, uint16_t *__restrict__ wk,
                             uint16_t *__restrict__ wt, int taps, int cin, int cout, int kp,
                             int kpt) {
  const long n1 = long(cout) * kp, n2 = wt ? long(cin) * kpt : 0;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n1 + n2;
       i += long(gridDim.x) * blockDim.x) {
    if (i < n1) {
      const int k = int(i % kp);
      wk[i] = k < taps * cin ? f2bf(w[i]) : uint16_t(0);
    } else {
      const long j = i - n1;
      const int ci = int(j / kpt), k = int(j - long(ci) * kpt);
      uint16_t v = 0;
      if (k < taps * cout) {
        const int tp = k / cout, co = k - tp * cout;
        v = f2bf(w[long(co) * kp + (taps - 1 - tp) * cin + ci]);
      }
      wt[j] = v;
    }
  }
}

__device__ __forceinline__ void split_index(long i, int c, int w, int h, int d, int &ch, int &xx,
                                            int &yy, int &zz, int &nn) {
  long q = i;
  ch = int(q % c);
  q /= c;
  xx = int(q % w);
  q /= w;
  yy = int(q % h);
  q /= h;
  zz = int(q % d);
  nn = int(q / d);
}

// 2x2x2 max pooling; arg keeps the first maximal child (z-major order).
__global__ void pool_fwd(const uint16_t *__restrict__ x, uint16_t *__restrict__ y,
                         uint8_t *__restrict__ arg, Shape s, int c) {
  const int d2 = s.d / 2, h2 = s.h / 2, w2 = s.w / 2;
  const long total = long(s.n) * d2 * h2 * w2 * c;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += long(gridDim.x) * blockDim.x) {
    int ch, xx, yy, zz, nn;
    split_index(i, c, w2, h2, d2, ch, xx, yy, zz, nn);
    float best = 0.f;
    int bi = 0;
    for (int o = 0; o < 8; ++o) {
      const int z1 = 2 * zz + (o >> 2), y1 = 2 * yy + ((o >> 1) & 1), x1 = 2 * xx + (o & 1);
      const float v = bf2f(x[pos_index(s, nn, z1, y1, x1) * c + ch]);
      if (o == 0 || v > best) {
        best = v;
        bi = o;
      }
    }
    y[i] = f2bf(best);
    arg[i] = uint8_t(bi);
  }
}

// dX[fine] (FP32, zeroed by the caller) = dY at the winning child.
__global__ void pool_bwd(const uint16_t *__restrict__ dy, const uint8_t *__restrict__ arg,
                         float *__restrict__ dx, Shape s, int c) {
  const int d2 = s.d / 2, h2 = s.h / 2, w2 = s.w / 2;
  const long total = long(s.n) * d2 * h2 * w2 * c;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += long(gridDim.x) * blockDim.x) {
    int ch, xx, yy, zz, nn;
    split_index(i, c, w2, h2, d2, ch, xx, yy, zz, nn);
    const int o = arg[i];
    const int z1 = 2 * zz + (o >> 2), y1 = 2 * yy + ((o >> 1) & 1), x1 = 2 * xx + (o & 1);
    dx[pos_index(s, nn, z1, y1, x1) * c + ch] = bf2f(dy[i]);
  }
}

// Nearest-neighbour x2 upsampling of a coarse tensor (c channels) into
// channels [off, off + c) of a fine tensor with ctot channels.
__global__ void upsample_into(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, Shape fine,
                              int c, int ctot, int off) {
  const long total = long(fine.n) * fine.d * fine.h * fine.w * c;
  const Shape coarse{fine.n, fine.d / 2, fine.h / 2, fine.w / 2};
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += long(gridDim.x) * blockDim.x) {
    int ch, xx, yy, zz, nn;
    split_index(i, c, fine.w, fine.h, fine.d, ch, xx, yy, zz, nn);
    y[(i / c) * ctot + off + ch] = x[pos_index(coarse, nn, zz / 2, yy / 2, xx / 2) * c + ch];
  }
}

// Gradient of upsample_into: dX[coarse] = sum over its 8 children.
__global__ void upsample_bwd(const uint16_t *__restrict__ dy, uint16_t *__restrict__ dx,
                             Shape fine, int c, int ctot, int off) {
  const int d2 = fine.d / 2, h2 = fine.h / 2, w2 = fine.w / 2;
  const long total = long(fine.n) * d2 * h2 * w2 * c;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += long(gridDim.x) * blockDim.x) {
    int ch, xx, yy, zz, nn;
    split_index(i, c, w2, h2, d2, ch, xx, yy, zz, nn);
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
      const int z1 = 2 * zz + (o >> 2), y1 = 2 * yy + ((o >> 1) & 1), x1 = 2 * xx + (o & 1);
      s += bf2f(dy[pos_index(fine, nn, z1, y1, x1) * ctot + off + ch]);
    }
    dx[i] = f2bf(s);
  }
}

// dir 0: channels [off, off + c) of dst (ctot channels) = src (c channels);
// dir 1: dst (c channels) = channels [off, off + c) of src (ctot channels).
__global__ void channel_copy(const uint16_t *__restrict__ src, uint16_t *__restrict__ dst,
                             long npos, int c, int ctot, int off, int dir) {
  const long total = npos * c;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += long(gridDim.x) * blockDim.x) {
    const long p = i / c;
    const int ch = int(i - p * c);
    if (dir == 0)
      dst[p * ctot + off + ch] = src[i];
    else
      dst[i] = src[p * ctot + off + ch];
  }
}

// dZ = (dY32 + dY16) * (act > 0): ReLU backward from the post-activation,
// summing the gradients that reach a tensor along two paths.
__global__ void relu_bwd(const float *__restrict__ dy32, const uint16_t *__restrict__ dy16,
                         const uint16_t *__restrict__ act, uint16_t *__restrict__ dx, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += long(gridDim.x) * blockDim.x) {
    float g = 0.f;
    if (dy32) g += dy32[i];
    if (dy16) g += bf2f(dy16[i]);
    dx[i] = bf2f(act[i]) > 0.f ? f2bf(g) : uint16_t(0);
  }
}

__global__ void cast_f2bf(const float *__restrict__ a, uint16_t *__restrict__ b, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += long(gridDim.x) * blockDim.x)
    b[i] = f2bf(a[i]);
}

// out = xin + head; with a target: loss += mean (out - target)^2 (FP64) and
// dhead = 2 (out - target) / n in BF16.
__global__ __launch_bounds__(256) void residual_mse(const float *__restrict__ xin,
                                                    const float *__restrict__ head,
                                                    const float *__restrict__ target,
                                                    float *__restrict__ out,
                                                    uint16_t *__restrict__ dhead, long n,
                                                    double *__restrict__ loss) {
  double local = 0.0;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += long(gridDim.x) * blockDim.x) {
    const float o = xin[i] + head[i];
    out[i] = o;
    if (target) {
      const float d = o - target[i];
      local += double(d) * double(d);
      if (dhead) dhead[i] = f2bf(2.f * d / float(n));
    }
  }
  if (!loss) return;
  __shared__ double red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / double(n));
}

// Adam on FP32 parameters; lr_t carries the bias correction.
__global__ void adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                     float *__restrict__ v, long n, float lr_t, float b1, float b2, float eps) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += long(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
  }
}

unsigned grid1(long n) {
  const long b = (n + 255) / 256;
  return unsigned(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

template <int KS, int NT>
void launch_conv(dim3 grid, hipStream_t st, bool vec, const uint16_t *x, const uint16_t *wk,
                 const float *bias, uint16_t *y, float *y32, Shape s, int cin, int cout, int kp,
                 int relu) {
  if (vec)
    hipLaunchKernelGGL((conv_fwd<KS, NT, true>), grid, dim3(64), 0, st, x, wk, bias, y, y32, s,
                       cin, cout, kp, relu);
  else
    hipLaunchKernelGGL((conv_fwd<KS, NT, false>), grid, dim3(64), 0, st, x, wk, bias, y, y32, s,
                       cin, cout, kp, relu);
}

template <int KS>
bool dispatch_conv(int nt, dim3 grid, hipStream_t st, bool vec, const uint16_t *x,
                   const uint16_t *wk, const float *bias, uint16_t *y, float *y32, Shape s, int cin,
                   int cout, int kp, int relu) {
  switch (nt) {
    case 1: launch_conv<KS, 1>(grid, st, vec, x, wk, bias, y, y32, s, cin, cout, kp, relu); break;
    case 2: launch_conv<KS, 2>(grid, st, vec, x, wk, bias, y, y32, s, cin, cout, kp, relu); break;
    case 3: launch_conv<KS, 3>(grid, st, vec, x, wk, bias, y, y32, s, cin, cout, kp, relu); break;
    case 4: launch_conv<KS, 4>(grid, st, vec, x, wk, bias, y, y32, s, cin, cout, kp, relu); break;
    case 6: launch_conv<KS, 6>(grid, st, vec, x, wk, bias, y, y32, s, cin, cout, kp, relu); break;
    case 8: launch_conv<KS, 8>(grid, st, vec, x, wk, bias, y, y32, s, cin, cout, kp, relu); break;
    default: return false;
  }
  return true;
}

}  // namespace unet
}  // namespace trtt

using namespace trtt::unet;

extern "C" {

int trtt_unet_conv_fwd(const void *x, const void *wk, const float *bias, void *y, float *y32,
                       int n, int d, int h, int w, int cin, int cout, int ks, int kp, int relu,
                       void *stream) {
  const int K = ks * ks * ks * cin;
  if (!x || !wk || !y || (ks != 1 && ks != 3) || cin <= 0 || cout <= 0 || cout > 128 ||
      kp % 16 || kp < K || n <= 0 || d <= 0 || h <= 0 || w <= 0 || long(n) * d > 65535 ||
      h > 65535)
    return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Shape s{n, d, h, w};
  const dim3 grid((w + 15) / 16, h, n * d);
  int nt = (cout + 15) / 16;
  if (nt == 5) nt = 6;
  if (nt == 7) nt = 8;
  const bool vec = cin % 16 == 0;
  const auto *xi = static_cast<const uint16_t *>(x);
  const auto *wi = static_cast<const uint16_t *>(wk);
  auto *yo = static_cast<uint16_t *>(y);
  const bool ok = ks == 3 ? dispatch_conv<3>(nt, grid, st, vec, xi, wi, bias, yo, y32, s, cin,
                                             cout, kp, relu)
                          : dispatch_conv<1>(nt, grid, st, vec, xi, wi, bias, yo, y32, s, cin,
                                             cout, kp, relu);
  if (!ok) return -1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_conv_wgrad(const void *x, const void *dy, float *dw, float *db, int n, int d, int h,
                         int w, int cin, int cout, int ks, int kp, void *stream) {
  if (!x || !dy || !dw || (ks != 1 && ks != 3) || kp % 16 || kp < ks * ks * ks * cin ||
      cout <= 0 || cout > 256 || cin <= 0 || n <= 0 || d <= 0 || h <= 0 || w <= 0)
    return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Shape s{n, d, h, w};
  const long npos = long(n) * d * h * w;
  const long ngroups = long(n) * d * h * ((w + 15) / 16);
  if (ngroups > 0x7fffffffL) return -1;
  const int chunk = 128;
  const dim3 grid(unsigned((ngroups + chunk - 1) / chunk), kp / 16, (cout + 15) / 16);
  if (ks == 3)
    hipLaunchKernelGGL(conv_wgrad<3>, grid, dim3(64), 0, st, static_cast<const uint16_t *>(x),
                       static_cast<const uint16_t *>(dy), dw, s, cin, cout, kp, chunk);
  else
    hipLaunchKernelGGL(conv_wgrad<1>, grid, dim3(64), 0, st, static_cast<const uint16_t *>(x),
                       static_cast<const uint16_t *>(dy), dw, s, cin, cout, kp, chunk);
  if (db) {
    const long lanes = 256 / cout;
    long blocks = (npos + lanes * 64 - 1) / (lanes * 64);
    blocks = blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096;
    hipLaunchKernelGGL(bias_grad, dim3(unsigned(blocks)), dim3(256), 0, st,
                       static_cast<const uint16_t *>(dy), db, npos, cout);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_prep_weights(const float *w, void *wk, void *wt, int taps, int cin, int cout, int kp,
                           int kpt, void *stream) {
  if (!w || !wk || kp < taps * cin || (wt && kpt < taps * cout)) return -1;
  const long n = long(cout) * kp + (wt ? long(cin) * kpt : 0);
  hipLaunchKernelGGL(prep_weights, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     w, static_cast<uint16_t *>(wk), static_cast<uint16_t *>(wt), taps, cin, cout,
                     kp, kpt);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_pool(const void *x, void *y, void *arg, float *dx, const void *dy, int n, int d,
                   int h, int w, int c, int backward, void *stream) {
  if (d % 2 || h % 2 || w % 2 || !arg) return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Shape s{n, d, h, w};
  const long m = long(n) * (d / 2) * (h / 2) * (w / 2) * c;
  if (!backward)
    hipLaunchKernelGGL(pool_fwd, dim3(grid1(m)), dim3(256), 0, st,
                       static_cast<const uint16_t *>(x), static_cast<uint16_t *>(y),
                       static_cast<uint8_t *>(arg), s, c);
  else
    hipLaunchKernelGGL(pool_bwd, dim3(grid1(m)), dim3(256), 0, st,
                       static_cast<const uint16_t *>(dy), static_cast<const uint8_t *>(arg), dx, s,
                       c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_upsample(const void *src, void *dst, int n, int d, int h, int w, int c, int ctot,
                       int off, int backward, void *stream) {
  if (d % 2 || h % 2 || w % 2 || off + c > ctot) return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Shape fine{n, d, h, w};
  const long m = long(n) * d * h * w * c;
  if (!backward)
    hipLaunchKernelGGL(upsample_into, dim3(grid1(m)), dim3(256), 0, st,
                       static_cast<const uint16_t *>(src), static_cast<uint16_t *>(dst), fine, c,
                       ctot, off);
  else
    hipLaunchKernelGGL(upsample_bwd, dim3(grid1(m / 8)), dim3(256), 0, st,
                       static_cast<const uint16_t *>(src), static_cast<uint16_t *>(dst), fine, c,
                       ctot, off);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_channel_copy(const void *src, void *dst, long npos, int c, int ctot, int off,
                           int dir, void *stream) {
  if (off + c > ctot) return -1;
  hipLaunchKernelGGL(channel_copy, dim3(grid1(npos * c)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const uint16_t *>(src),
                     static_cast<uint16_t *>(dst), npos, c, ctot, off, dir);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_relu_bwd(const float *dy32, const void *dy16, const void *act, void *dx, long n,
                       void *stream) {
  hipLaunchKernelGGL(relu_bwd, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     dy32, static_cast<const uint16_t *>(dy16), static_cast<const uint16_t *>(act),
                     static_cast<uint16_t *>(dx), n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_cast_bf16(const float *a, void *b, long n, void *stream) {
  hipLaunchKernelGGL(cast_f2bf, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream), a,
                     static_cast<uint16_t *>(b), n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_residual_mse(const float *xin, const float *head, const float *target, float *out,
                           void *dhead, long n, double *loss, void *stream) {
  hipLaunchKernelGGL(residual_mse, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     xin, head, target, out, static_cast<uint16_t *>(dhead), n, loss);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trtt_unet_adam(float *p, const float *g, float *m, float *v, long n, float lr_t, float b1,
                   float b2, float eps, void *stream) {
  hipLaunchKernelGGL(adam, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream), p, g, m,
                     v, n, lr_t, b1, b2, eps);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
