// Learned warm start: a small 3D CNN that predicts the patch-proper (8^3) part of the converged potential
// from the density and the interpolated coarse potential, trained on solutions of the solver itself.
// No reference counterpart (the reference always starts from the interpolated cube).
//
// Network (activations BF16 channel-last [n][z][y][x][c], weights BF16, accumulation and master weights
// FP32):
//   x0: 18^3 x 8 channels: c0 = interp/s, c1 = Const*rho/s on the solve interior, c2..7 = 0, where s is
//       the largest |interp| of the patch (scale invariance of the discrete Poisson problem)
//   y1 = relu(conv3(x0) + b1)                 18^3 -> 16^3, 16 channels
//   yl = crop(y_{l-1}) + relu(conv3(y_{l-1}) + bl),  l = 2..5   16^3 -> 8^3 (residual blocks)
//   d  = conv3_same(y5) + b6                  8^3, 1 channel
//   guess(core) = interp(core) + s*d          (the rest of the interior keeps the interpolated values)
//
// Convolutions are implicit GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16): rows = output
// voxels (16 per tile), columns = 16 output channels, K = 27 taps x input channels, with the tap-major
// K order k = tap*CIN + ci so that each lane's 8 consecutive K values are 8 channels of one input voxel
// (one 16-byte LDS read).  MFMA lane layout (16x16x32): A[i][k] with i = lane%16, k = 8*(lane/16)+0..7;
// B[k][j] with j = lane%16, k = 8*(lane/16)+0..7; D[i][j] with j = lane%16, i = 4*(lane/16)+0..3.

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstdint>

#include "pg_cnn.h"
#include "pg_geometry.h"
#include "pg_interp.h"

namespace pgamd {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

__device__ __forceinline__ float bf2f(u16 v) { return __uint_as_float(((uint32_t)v) << 16); }
// round to nearest even
__device__ __forceinline__ u16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

union Pack8 {
  bf16x8 v;
  uint4 u;
  u16 h[8];
};

// ---- forward convolution (layers 1-5) -----------------------------------------------------------------

constexpr int kThreads = 256;
constexpr int kTZ = 2;  // output z-planes per workgroup

// One workgroup: kTZ output planes of one sample.  X: [n][S][S][S][CIN] bf16, W: [16][27*CIN] bf16,
// bias: [16] f32.  Y = crop(X) + relu(h) (RESID) or relu(h); H (optional) keeps h for the backward pass.
template <int CIN, bool RESID>
__global__ void __launch_bounds__(kThreads) conv_fwd_kernel(const u16* __restrict__ X, const u16* __restrict__ W,
                                                            const float* __restrict__ bias, u16* __restrict__ Y,
                                                            u16* __restrict__ H, int S) {
  extern __shared__ uint4 smem[];
  u16* xs = reinterpret_cast<u16*>(smem);
  const int So = S - 2;
  const int nslab = (So + kTZ - 1) / kTZ;
  const int n = blockIdx.x / nslab, slab = blockIdx.x % nslab;
  const int z0 = slab * kTZ;
  const int tz = min(kTZ, So - z0);
  const int planes = tz + 2;
  // stage the input planes z0 .. z0+tz+1 (16-byte vectors)
  const size_t plane_elems = (size_t)S * S * CIN;
  const uint4* src = reinterpret_cast<const uint4*>(X + ((size_t)n * S + z0) * plane_elems);
  const int nvec = (int)(planes * plane_elems / 8);
  for (int i = threadIdx.x; i < nvec; i += kThreads) smem[i] = src[i];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int K = 27 * CIN;
  constexpr int KC = (K + 31) / 32;
  // B operand (weights) for every K chunk, kept in registers
  bf16x8 wreg[KC];
  const int co = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
  for (int c = 0; c < KC; c++) {
    Pack8 p;
    const int kb = c * 32 + kq;
    if (kb < K)
      p.u = *reinterpret_cast<const uint4*>(W + co * K + kb);
    else
      p.u = make_uint4(0, 0, 0, 0);
    wreg[c] = p.v;
  }
  const float b = bias[co];
  __syncthreads();
  const int M = tz * So * So;
  const int ntiles = (M + 15) / 16;
  for (int t = wave; t < ntiles; t += kThreads / 64) {
    const int v = t * 16 + (lane & 15);
    const bool vin = v < M;
    const int vv = vin ? v : 0;
    const int lz = vv / (So * So), oy = (vv / So) % So, ox = vv % So;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < KC; c++) {
      const int kb = c * 32 + kq;
      const int tap = kb / CIN, ci0 = kb % CIN;
      Pack8 a;
      if (vin && tap < 27) {
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        a.u = *reinterpret_cast<const uint4*>(xs + (((lz + dz) * S + oy + dy) * S + ox + dx) * CIN + ci0);
      } else {
        a.u = make_uint4(0, 0, 0, 0);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, wreg[c], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int vo = t * 16 + 4 * (lane >> 4) + r;
      if (vo >= M) continue;
      const int z = vo / (So * So), y = (vo / So) % So, x = vo % So;
      const float h = acc[r] + b;
      float out = fmaxf(h, 0.0f);
      if (RESID) out += bf2f(xs[(((z + 1) * S + y + 1) * S + x + 1) * CIN + co]);
      const size_t o = ((((size_t)n * So + z0 + z) * So + y) * So + x) * 16 + co;
      Y[o] = f2bf(out);
      if (H) H[o] = f2bf(h);
    }
  }
}

// Layer 6: 16 -> 1 channel, 'same' 3^3 convolution at 8^3 (zero padding).  One workgroup per sample.
__global__ void __launch_bounds__(kThreads) conv6_fwd_kernel(const u16* __restrict__ Y5, const float* __restrict__ w6,
                                                             const float* __restrict__ b6, float* __restrict__ D) {
  __shared__ u16 ys[10 * 10 * 10 * 16];  // zero-padded y5 (BF16 values, exact)
  const int n = blockIdx.x;
  for (int i = threadIdx.x; i < 10 * 10 * 10 * 16; i += kThreads) {
    const int c = i & 15, p = i >> 4;
    const int x = p % 10 - 1, y = (p / 10) % 10 - 1, z = p / 100 - 1;
    u16 v = 0;
    if (x >= 0 && x < 8 && y >= 0 && y < 8 && z >= 0 && z < 8) v = Y5[((((size_t)n * 8 + z) * 8 + y) * 8 + x) * 16 + c];
    ys[i] = v;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < 512; v += kThreads) {
    const int x = v % 8, y = (v / 8) % 8, z = v / 64;
    float s = b6[0];
    for (int tap = 0; tap < 27; tap++) {
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      const u16* yp = ys + (((z + dz) * 10 + y + dy) * 10 + x + dx) * 16;
#pragma unroll
      for (int c = 0; c < 16; c++) s += w6[tap * 16 + c] * bf2f(yp[c]);
    }
    D[(size_t)n * 512 + v] = s;
  }
}

// ---- loss and backward ---------------------------------------------------------------------------------

// MSE over all predicted values: dD = 2 (D - T) / count; per-block partial sums of (D - T)^2.
__global__ void __launch_bounds__(kThreads) mse_kernel(const float* __restrict__ D, const float* __restrict__ T,
                                                       float* __restrict__ dD, float* __restrict__ partial, int count) {
  __shared__ float red[kThreads];
  float s = 0.0f;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < count; i += gridDim.x * kThreads) {
    const float e = D[i] - T[i];
    s += e * e;
    dD[i] = 2.0f * e / (float)count;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Layer-6 backward: dY5 (fp32 [n][8^3][16]) and per-sample partial gradients of w6 (432) and b6 (1).
__global__ void __launch_bounds__(kThreads) conv6_bwd_kernel(const u16* __restrict__ Y5, const float* __restrict__ w6,
                                                             const float* __restrict__ dD, float* __restrict__ dY5,
                                                             float* __restrict__ gpart) {
  __shared__ u16 ys[10 * 10 * 10 * 16];  // zero-padded y5 (BF16 values, exact)
  __shared__ float ds[10 * 10 * 10];
  const int n = blockIdx.x;
  for (int i = threadIdx.x; i < 10 * 10 * 10 * 16; i += kThreads) {
    const int c = i & 15, p = i >> 4;
    const int x = p % 10 - 1, y = (p / 10) % 10 - 1, z = p / 100 - 1;
    u16 v = 0;
    if (x >= 0 && x < 8 && y >= 0 && y < 8 && z >= 0 && z < 8) v = Y5[((((size_t)n * 8 + z) * 8 + y) * 8 + x) * 16 + c];
    ys[i] = v;
  }
  for (int p = threadIdx.x; p < 1000; p += kThreads) {
    const int x = p % 10 - 1, y = (p / 10) % 10 - 1, z = p / 100 - 1;
    ds[p] = (x >= 0 && x < 8 && y >= 0 && y < 8 && z >= 0 && z < 8) ? dD[(size_t)n * 512 + (z * 8 + y) * 8 + x] : 0.0f;
  }
  __syncthreads();
  // dY5[q][c] = sum_tap w6[tap][c] * dD[q - tap + 1]
  for (int i = threadIdx.x; i < 512 * 16; i += kThreads) {
    const int c = i & 15, q = i >> 4;
    const int x = q % 8, y = (q / 8) % 8, z = q / 64;
    float s = 0.0f;
    for (int tap = 0; tap < 27; tap++) {
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      s += w6[tap * 16 + c] * ds[((z + 2 - dz) * 10 + y + 2 - dy) * 10 + x + 2 - dx];
    }
    dY5[(size_t)n * 8192 + i] = s;
  }
  // dw6[tap][c] = sum_v dD[v] * Y5pad[v + tap - 1][c];  db6 = sum_v dD[v]
  for (int k = threadIdx.x; k < 433; k += kThreads) {
    float s = 0.0f;
    if (k < 432) {
      const int tap = k / 16, c = k % 16;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      for (int v = 0; v < 512; v++) {
        const int x = v % 8, y = (v / 8) % 8, z = v / 64;
        s += ds[((z + 1) * 10 + y + 1) * 10 + x + 1] * bf2f(ys[(((z + dz) * 10 + y + dy) * 10 + x + dx) * 16 + c]);
      }
    } else {
      for (int v = 0; v < 512; v++) {
        const int x = v % 8, y = (v / 8) % 8, z = v / 64;
        s += ds[((z + 1) * 10 + y + 1) * 10 + x + 1];
      }
    }
    gpart[(size_t)n * 433 + k] = s;
  }
}

// dH = dY * (H > 0), stored BF16 for the matrix cores; per-block partial bias gradients.
__global__ void __launch_bounds__(kThreads) relu_bwd_kernel(const float* __restrict__ dY, const u16* __restrict__ H,
                                                            u16* __restrict__ dH, float* __restrict__ bpart,
                                                            size_t count) {
  __shared__ float red[kThreads];
  // thread t handles channel t%16 of voxels strided by (grid*threads/16); count is a multiple of 16
  float s = 0.0f;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads) {
    const float g = bf2f(H[i]) > 0.0f ? dY[i] : 0.0f;
    const u16 gb = f2bf(g);
    dH[i] = gb;
    s += bf2f(gb);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o >= 16; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 16) bpart[(size_t)blockIdx.x * 16 + threadIdx.x] = red[threadIdx.x];
}

// Input gradient of a 16->16 layer: dX[p][ci] = sum_{tap,co} W[co][tap][ci] dH[p - tap][co]  (+ dY of
// the residual path, centred).  Implicit GEMM with WT[ci][tap*16+co] (the weights transposed per tap).
__global__ void __launch_bounds__(kThreads) conv_bwd_data_kernel(const u16* __restrict__ dH, const u16* __restrict__ WT,
                                                                 const float* __restrict__ dYres, float* __restrict__ dX,
                                                                 int S) {
  extern __shared__ uint4 smem[];
  u16* hs = reinterpret_cast<u16*>(smem);
  const int So = S - 2;
  const int nslab = (S + kTZ - 1) / kTZ;
  const int n = blockIdx.x / nslab, slab = blockIdx.x % nslab;
  const int z0 = slab * kTZ;
  const int tz = min(kTZ, S - z0);
  // dH planes q_z in [z0-2, z0+tz-1], zero outside [0, So)
  const int qz0 = z0 - 2, planes = tz + 2;
  const int pv = So * So * 2;  // uint4 per plane (16 channels = 2 x uint4)
  for (int i = threadIdx.x; i < planes * pv; i += kThreads) {
    const int pl = i / pv, r = i % pv;
    const int qz = qz0 + pl;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (qz >= 0 && qz < So) v = reinterpret_cast<const uint4*>(dH + (((size_t)n * So + qz) * So * So) * 16)[r];
    smem[i] = v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int K = 432, KC = 14;
  bf16x8 wreg[KC];
  const int ci = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
  for (int c = 0; c < KC; c++) {
    Pack8 p;
    const int kb = c * 32 + kq;
    if (kb < K)
      p.u = *reinterpret_cast<const uint4*>(WT + ci * K + kb);
    else
      p.u = make_uint4(0, 0, 0, 0);
    wreg[c] = p.v;
  }
  __syncthreads();
  const int M = tz * S * S;
  const int ntiles = (M + 15) / 16;
  for (int t = wave; t < ntiles; t += kThreads / 64) {
    const int v = t * 16 + (lane & 15);
    const bool vin = v < M;
    const int vv = vin ? v : 0;
    const int lz = vv / (S * S), py = (vv / S) % S, px = vv % S;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < KC; c++) {
      const int kb = c * 32 + kq;
      const int tap = kb / 16, co0 = kb % 16;
      Pack8 a;
      a.u = make_uint4(0, 0, 0, 0);
      if (vin && tap < 27) {
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        const int qy = py - dy, qx = px - dx;
        // staged plane index: (z0 + lz - dz) - qz0 = lz - dz + 2
        if (qy >= 0 && qy < So && qx >= 0 && qx < So)
          a.u = *reinterpret_cast<const uint4*>(hs + (((lz - dz + 2) * So + qy) * So + qx) * 16 + co0);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, wreg[c], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int vo = t * 16 + 4 * (lane >> 4) + r;
      if (vo >= M) continue;
      const int z = z0 + vo / (S * S), y = (vo / S) % S, x = vo % S;
      float g = acc[r];
      if (dYres && z >= 1 && z <= So && y >= 1 && y <= So && x >= 1 && x <= So)
        g += dYres[((((size_t)n * So + z - 1) * So + y - 1) * So + x - 1) * 16 + ci];
      dX[((((size_t)n * S + z) * S + y) * S + x) * 16 + ci] = g;
    }
  }
}

// Weight gradient: dW[co][tap*CIN+ci] = sum_{n,v} dH[n][v][co] X[n][v+tap][ci], partial per workgroup.
// GEMM with M = 16 output channels, N = 27*CIN, K = voxels (32 per MFMA); operands gathered from LDS.
template <int CIN>
__global__ void __launch_bounds__(kThreads) conv_bwd_weight_kernel(const u16* __restrict__ X, const u16* __restrict__ dH,
                                                                   float* __restrict__ wpart, int S) {
  extern __shared__ uint4 smem[];
  const int So = S - 2;
  const int nslab = (So + kTZ - 1) / kTZ;
  const int n = blockIdx.x / nslab, slab = blockIdx.x % nslab;
  const int z0 = slab * kTZ;
  const int tz = min(kTZ, So - z0);
  const size_t xplane = (size_t)S * S * CIN;
  const int xvec = (int)((tz + 2) * xplane / 8);
  const int hvec = tz * So * So * 2;
  u16* xs = reinterpret_cast<u16*>(smem);
  u16* hs = xs + (size_t)(kTZ + 2) * xplane;
  const uint4* xsrc = reinterpret_cast<const uint4*>(X + ((size_t)n * S + z0) * xplane);
  for (int i = threadIdx.x; i < xvec; i += kThreads) smem[i] = xsrc[i];
  const uint4* hsrc = reinterpret_cast<const uint4*>(dH + ((((size_t)n * So + z0) * So) * So) * 16);
  uint4* hdst = reinterpret_cast<uint4*>(hs);
  for (int i = threadIdx.x; i < hvec; i += kThreads) hdst[i] = hsrc[i];
  __syncthreads();
  constexpr int NC = 27 * CIN;
  constexpr int NT = (NC + 15) / 16;          // column tiles
  constexpr int TPW = (NT + 3) / 4;           // tiles per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  floatx4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; i++) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int M = tz * So * So;
  const int kq = 8 * (lane >> 4);
  for (int k0 = 0; k0 < M; k0 += 32) {
    // A[i = co][k = voxel]: dH[v][co] for the lane's 8 voxels
    Pack8 a;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int v = k0 + kq + e;
      a.h[e] = v < M ? hs[v * 16 + (lane & 15)] : (u16)0;
    }
#pragma unroll
    for (int i = 0; i < TPW; i++) {
      const int tile = wave * TPW + i;
      if (tile >= NT) break;
      const int col = tile * 16 + (lane & 15);
      const int tap = col / CIN, ci = col % CIN;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      Pack8 b;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int v = k0 + kq + e;
        u16 val = 0;
        if (v < M && col < NC) {
          const int lz = v / (So * So), oy = (v / So) % So, ox = v % So;
          val = xs[(((lz + dz) * S + oy + dy) * S + ox + dx) * CIN + ci];
        }
        b.h[e] = val;
      }
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[i], 0, 0, 0);
    }
  }
  // D[i = co][j = col]: lane holds col = tile*16 + lane%16, co = 4*(lane/16) + r
  float* out = wpart + (size_t)blockIdx.x * 16 * NC;
#pragma unroll
  for (int i = 0; i < TPW; i++) {
    const int tile = wave * TPW + i;
    if (tile >= NT) break;
    const int col = tile * 16 + (lane & 15);
    if (col >= NC) continue;
#pragma unroll
    for (int r = 0; r < 4; r++) out[(4 * (lane >> 4) + r) * NC + col] = acc[i][r];
  }
}

// Sum `parts` partial vectors of length len (fixed order: deterministic) and add into dst.
__global__ void reduce_parts_kernel(const float* __restrict__ part, int parts, int len, float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int p = 0; p < parts; p++) s += part[(size_t)p * len + i];
  dst[i] += s;
}

// Adam (FP32 master weights), then refresh the BF16 forward copy and the per-tap transposed copy.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void to_bf16_kernel(const float* __restrict__ src, u16* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

// dst = src rounded to the nearest BF16 value, kept as FP32 (the VALU output layer's weights)
__global__ void round_bf16_kernel(const float* __restrict__ src, float* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = bf2f(f2bf(src[i]));
}

// WT[ci][tap*16+co] = W[co][tap*16+ci] for the 16->16 layers
__global__ void transpose_w_kernel(const float* __restrict__ W, u16* __restrict__ WT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 16 * 432) return;
  const int ci = i / 432, k = i % 432, tap = k / 16, co = k % 16;
  WT[i] = f2bf(W[co * 432 + tap * 16 + ci]);
}

// ---- input preparation and guess assembly -------------------------------------------------------------

// One workgroup per patch of ghost size gs (<= 5): interpolate the coarse potential into the solve cube,
// embed it at the centre of the 18^3 input (zeros around smaller cubes), scale by s = max|interp|.
template <int GS>
__device__ void prep_patch(const float* __restrict__ rho, const float* __restrict__ cpot, float cst, int int_scheme,
                           u16* __restrict__ x, float* __restrict__ scale, float* cube /* LDS N^3 */) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, C = g.C, U = g.U, NC = C - 2, R = g.R, OFF = 5 - GS;
  __shared__ float red[kThreads];
  for (int q = threadIdx.x; q < NC * NC * NC; q += kThreads) {
    const int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
    float v[3][3][3];
    for (int dk = 0; dk < 3; dk++)
      for (int dj = 0; dj < 3; dj++)
        for (int di = 0; di < 3; di++) v[dk][dj][di] = cpot[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
    float ch[2][2][2];
    interp_children(v, int_scheme, ch);
    for (int cz = 0; cz < 2; cz++)
      for (int cy = 0; cy < 2; cy++)
        for (int cx = 0; cx < 2; cx++) {
          const int X = 2 * (i - 1) + cx - U, Y = 2 * (j - 1) + cy - U, Z = 2 * (k - 1) + cz - U;
          if (X >= 0 && X < N && Y >= 0 && Y < N && Z >= 0 && Z < N) cube[(Z * N + Y) * N + X] = ch[cz][cy][cx];
        }
  }
  __syncthreads();
  float mx = 0.0f;
  for (int t = threadIdx.x; t < N * N * N; t += kThreads) mx = fmaxf(mx, fabsf(cube[t]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const float s = red[0] > 0.0f ? red[0] : 1.0f;
  const float inv = 1.0f / s;
  if (threadIdx.x == 0) *scale = s;
  for (int t = threadIdx.x; t < 18 * 18 * 18; t += kThreads) {
    const int X = t % 18 - OFF, Y = (t / 18) % 18 - OFF, Z = t / 324 - OFF;
    float c0 = 0.0f, c1 = 0.0f;
    if (X >= 0 && X < N && Y >= 0 && Y < N && Z >= 0 && Z < N) {
      c0 = cube[(Z * N + Y) * N + X] * inv;
      if (X >= 1 && X <= R && Y >= 1 && Y <= R && Z >= 1 && Z <= R)
        c1 = cst * rho[((Z - 1) * R + (Y - 1)) * R + (X - 1)] * inv;
    }
    uint4 w;
    w.x = (uint32_t)f2bf(c0) | ((uint32_t)f2bf(c1) << 16);
    w.y = w.z = w.w = 0;
    reinterpret_cast<uint4*>(x)[t] = w;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kThreads) prep_kernel(const float* __restrict__ rho, const float* __restrict__ cpot,
                                                        int gs, float cst, int int_scheme, u16* __restrict__ x,
                                                        float* __restrict__ scale) {
  extern __shared__ float cube[];
  const int p = blockIdx.x;
  const Geo g = geo_of(gs);
  const float* r = rho + (size_t)p * g.R * g.R * g.R;
  const float* c = cpot + (size_t)p * g.C * g.C * g.C;
  u16* xp = x + (size_t)p * 18 * 18 * 18 * 8;
  switch (gs) {
    case 1: prep_patch<1>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    case 2: prep_patch<2>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    case 3: prep_patch<3>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    case 4: prep_patch<4>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    default: prep_patch<5>(r, c, cst, int_scheme, xp, scale + p, cube); break;
  }
}

// Guess of the interior: interpolated values, plus s*d on the patch-proper core.
template <int GS>
__device__ void guess_patch(const float* __restrict__ cpot, const float* __restrict__ d, float s, int int_scheme,
                            float* __restrict__ guess, float* cube) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, C = g.C, U = g.U, NC = C - 2, R = g.R;
  for (int q = threadIdx.x; q < NC * NC * NC; q += kThreads) {
    const int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
    float v[3][3][3];
    for (int dk = 0; dk < 3; dk++)
      for (int dj = 0; dj < 3; dj++)
        for (int di = 0; di < 3; di++) v[dk][dj][di] = cpot[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
    float ch[2][2][2];
    interp_children(v, int_scheme, ch);
    for (int cz = 0; cz < 2; cz++)
      for (int cy = 0; cy < 2; cy++)
        for (int cx = 0; cx < 2; cx++) {
          const int X = 2 * (i - 1) + cx - U, Y = 2 * (j - 1) + cy - U, Z = 2 * (k - 1) + cz - U;
          if (X >= 0 && X < N && Y >= 0 && Y < N && Z >= 0 && Z < N) cube[(Z * N + Y) * N + X] = ch[cz][cy][cx];
        }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < R * R * R; t += kThreads) {
    const int x = t % R, y = (t / R) % R, z = t / (R * R);  // interior-local, cube coords +1
    float v = cube[((z + 1) * N + y + 1) * N + x + 1];
    const int cx = x + 1 - GS, cy = y + 1 - GS, cz = z + 1 - GS;  // patch-proper coordinates
    if (cx >= 0 && cx < 8 && cy >= 0 && cy < 8 && cz >= 0 && cz < 8) v = v + s * d[(cz * 8 + cy) * 8 + cx];
    guess[t] = v;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kThreads) guess_kernel(const float* __restrict__ cpot, const float* __restrict__ D,
                                                         const float* __restrict__ scale, int gs, int int_scheme,
                                                         float* __restrict__ guess) {
  extern __shared__ float cube[];
  const int p = blockIdx.x;
  const Geo g = geo_of(gs);
  const float* c = cpot + (size_t)p * g.C * g.C * g.C;
  float* out = guess + (size_t)p * g.R * g.R * g.R;
  switch (gs) {
    case 1: guess_patch<1>(c, D + (size_t)p * 512, scale[p], int_scheme, out, cube); break;
    case 2: guess_patch<2>(c, D + (size_t)p * 512, scale[p], int_scheme, out, cube); break;
    case 3: guess_patch<3>(c, D + (size_t)p * 512, scale[p], int_scheme, out, cube); break;
    case 4: guess_patch<4>(c, D + (size_t)p * 512, scale[p], int_scheme, out, cube); break;
    default: guess_patch<5>(c, D + (size_t)p * 512, scale[p], int_scheme, out, cube); break;
  }
}

}  // namespace cnn

// ---- host launchers ----------------------------------------------------------------------------------------

using namespace cnn;

static size_t fwd_lds(int S, int cin) { return (size_t)(kTZ + 2) * S * S * cin * 2; }

hipError_t cnn_conv_fwd(const uint16_t* X, const uint16_t* W, const float* b, uint16_t* Y, uint16_t* H, int n, int S,
                        int cin, bool resid, hipStream_t s) {
  const int So = S - 2, nslab = (So + kTZ - 1) / kTZ;
  const size_t lds = fwd_lds(S, cin);
  if (cin == 8 && !resid)
    hipLaunchKernelGGL((conv_fwd_kernel<8, false>), dim3(n * nslab), dim3(kThreads), lds, s, X, W, b, Y, H, S);
  else if (cin == 16 && resid)
    hipLaunchKernelGGL((conv_fwd_kernel<16, true>), dim3(n * nslab), dim3(kThreads), lds, s, X, W, b, Y, H, S);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t cnn_conv6_fwd(const uint16_t* Y5, const float* w6, const float* b6, float* D, int n, hipStream_t s) {
  hipLaunchKernelGGL(conv6_fwd_kernel, dim3(n), dim3(kThreads), 0, s, Y5, w6, b6, D);
  return hipGetLastError();
}

hipError_t cnn_mse(const float* D, const float* T, float* dD, float* partial, int count, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mse_kernel, dim3(nblocks), dim3(kThreads), 0, s, D, T, dD, partial, count);
  return hipGetLastError();
}

hipError_t cnn_conv6_bwd(const uint16_t* Y5, const float* w6, const float* dD, float* dY5, float* gpart, int n,
                         hipStream_t s) {
  hipLaunchKernelGGL(conv6_bwd_kernel, dim3(n), dim3(kThreads), 0, s, Y5, w6, dD, dY5, gpart);
  return hipGetLastError();
}

hipError_t cnn_relu_bwd(const float* dY, const uint16_t* H, uint16_t* dH, float* bpart, size_t count, int nblocks,
                        hipStream_t s) {
  hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblocks), dim3(kThreads), 0, s, dY, H, dH, bpart, count);
  return hipGetLastError();
}

hipError_t cnn_conv_bwd_data(const uint16_t* dH, const uint16_t* WT, const float* dYres, float* dX, int n, int S,
                             hipStream_t s) {
  const int nslab = (S + kTZ - 1) / kTZ;
  const size_t lds = (size_t)(kTZ + 2) * (S - 2) * (S - 2) * 16 * 2;
  hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(n * nslab), dim3(kThreads), lds, s, dH, WT, dYres, dX, S);
  return hipGetLastError();
}

int cnn_bwd_weight_blocks(int n, int S) { return n * ((S - 2 + kTZ - 1) / kTZ); }

hipError_t cnn_conv_bwd_weight(const uint16_t* X, const uint16_t* dH, float* wpart, int n, int S, int cin,
                               hipStream_t s) {
  const int So = S - 2, nslab = (So + kTZ - 1) / kTZ;
  const size_t lds = (size_t)(kTZ + 2) * S * S * cin * 2 + (size_t)kTZ * So * So * 16 * 2;
  if (cin == 8)
    hipLaunchKernelGGL((conv_bwd_weight_kernel<8>), dim3(n * nslab), dim3(kThreads), lds, s, X, dH, wpart, S);
  else if (cin == 16)
    hipLaunchKernelGGL((conv_bwd_weight_kernel<16>), dim3(n * nslab), dim3(kThreads), lds, s, X, dH, wpart, S);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t cnn_reduce_parts(const float* part, int parts, int len, float* dst, hipStream_t s) {
  hipLaunchKernelGGL(reduce_parts_kernel, dim3((len + 255) / 256), dim3(256), 0, s, part, parts, len, dst);
  return hipGetLastError();
}

hipError_t cnn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t cnn_to_bf16(const float* src, uint16_t* dst, int n, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t cnn_round_bf16(const float* src, float* dst, int n, hipStream_t s) {
  hipLaunchKernelGGL(round_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t cnn_transpose_w(const float* W, uint16_t* WT, hipStream_t s) {
  hipLaunchKernelGGL(transpose_w_kernel, dim3((16 * 432 + 255) / 256), dim3(256), 0, s, W, WT);
  return hipGetLastError();
}

hipError_t cnn_prep(const float* rho, const float* cpot, int n, int gs, float cst, int int_scheme, uint16_t* x,
                    float* scale, hipStream_t s) {
  const Geo g = geo_of(gs);
  hipLaunchKernelGGL(prep_kernel, dim3(n), dim3(kThreads), sizeof(float) * g.N * g.N * g.N, s, rho, cpot, gs, cst,
                     int_scheme, x, scale);
  return hipGetLastError();
}

hipError_t cnn_guess(const float* cpot, const float* D, const float* scale, int n, int gs, int int_scheme,
                     float* guess, hipStream_t s) {
  const Geo g = geo_of(gs);
  hipLaunchKernelGGL(guess_kernel, dim3(n), dim3(kThreads), sizeof(float) * g.N * g.N * g.N, s, cpot, D, scale, gs,
                     int_scheme, guess);
  return hipGetLastError();
}

}  // namespace pgamd
