// Learned warm start: a small 3D CNN that predicts the correction from the interpolated coarse potential
// to the converged potential on the whole solve interior, trained on solutions of the solver itself.
// No reference counterpart (the reference always starts from the interpolated cube).
//
// Network (activations BF16 channel-last, weights BF16, accumulation and master weights FP32):
//   frame   the 18^3 solve cube of ghost size 5 (smaller cubes centred, zeros around), with the fields
//           q0 = interp/s and q1 = 64 r/s (solve interior only), s = max|interp|, r the SOR residual of
//           the interpolated cube (reference operand order), i.e. the right-hand side of A e = r that the
//           correction e to the interpolated start satisfies
//   x0      [n][10^3][16]: space-to-depth of the zero-extended 20^3 frame (fine index f = 2c - 1 + d per
//           axis, channel ((dz*2+dy)*2+dx)*2 + q), so that coarse voxels 1..8 hold the 2^3 blocks of the
//           16^3 interior and voxels 0 and 9 hold the Dirichlet layer
//   y1      relu(conv3(x0) + b1)                        10^3 -> 8^3, 16 channels
//   yl      y_{l-1} + relu(conv3(pad(y_{l-1})) + bl)    l = 2..5, 8^3 (residual blocks, zero padding)
//   D       conv3(pad(y5)) + b6                         8^3, 8 channels = the 2^3 sub-cells (linear)
//   guess   interp + s * D[o][p] at interior frame index 2o + p (depth-to-space)
// Activations are stored padded (10^3 with a zero border) so that every layer is the same valid 3^3
// convolution: 10^3 x 16 channels in, 8^3 x 16 channels out, K = 27 taps x 16 channels = 432.
//
// Convolutions are implicit GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16): the weights are
// the A operand (rows = 16 output channels), the activations the B operand (columns = 16 output voxels of a
// tile), K made of (tap, channel) pairs so that each lane's 8 consecutive K values are 8 channels of one input
// voxel (one 16-byte LDS read).  MFMA lane layout (16x16x32): A[i][k] with i = lane%16, k = 8*(lane/16)+0..7;
// B[k][j] with j = lane%16, k = 8*(lane/16)+0..7; D[i][j] with j = lane%16, i = 4*(lane/16)+0..3.  So a lane
// loads the channels of voxel lane%16 and receives output channels 4*(lane/16)..+3 of that same voxel: the
// epilogue reads and writes 4 consecutive channels (8 bytes) at once.
// Forward K order: 15 steps c = dz*5 + s; step c covers the in-plane positions 2s and 2s+1 (position
// dy*3 + dx, 16 channels each) of tap plane dz; position 9 is a pad with zero weights whose A operand
// repeats position 8.  Grouping the steps by dz lets the fused kernel read an input plane once and feed it
// to the three output planes that use it.  The backward passes keep the tap-major order k = tap*16 + ci.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pg_cnn.h"
#include "pg_geometry.h"
#include "pg_interp.h"
#include "pg_kernels.h"

#pragma clang fp contract(off)

namespace pgamd {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef uint16_t u16;

__device__ __forceinline__ float bf2f(u16 v) { return __uint_as_float(((uint32_t)v) << 16); }
// round to nearest even (one v_cvt_pk_bf16_f32 on gfx950); every kernel here uses this one conversion
__device__ __forceinline__ u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }

// Two floats as two BF16 values in one word (the first one low): one two-source v_cvt_pk_bf16_f32, the same
// rounding as f2bf.
__device__ __forceinline__ uint32_t pack2_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((floatx2){a, b}, bf16x2));
}
// Four floats as four BF16 values (8 bytes, the first one lowest).
__device__ __forceinline__ uint2 pack4_bf16(float a, float b, float c, float d) {
  return make_uint2(pack2_bf16(a, b), pack2_bf16(c, d));
}
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

union Pack8 {
  bf16x8 v;
  uint4 u;
  u16 h[8];
};

constexpr int kThreads = 256;
constexpr int S = kCnnPadded;   // 10: padded input edge
constexpr int So = kCnnEdge;    // 8: output edge
constexpr int CH = 16;          // channels
constexpr int K = 27 * CH;      // 432
constexpr int KC = (K + 31) / 32;  // 14 MFMA K-steps of the tap-major order (the last one half used)
constexpr int KZ = 15;          // forward K steps: 3 dz x 5 position pairs
constexpr int KF = KZ * 32;     // 480 = kCnnKF
static_assert(KF == kCnnKF, "forward weight layout");
constexpr int kTZ = 2;          // output z-planes per workgroup
constexpr int kSlabs = So / kTZ;
constexpr int kPlane = S * S * CH;  // bf16 values per padded plane

// B operand (16 columns of a [16][432] weight matrix, row j = lane%16) for every K step.
__device__ __forceinline__ void load_weights(const u16* __restrict__ W, bf16x8 (&wreg)[KC]) {
  const int j = threadIdx.x & 15, kq = 8 * ((threadIdx.x & 63) >> 4);
#pragma unroll
  for (int c = 0; c < KC; c++) {
    Pack8 p;
    const int kb = c * 32 + kq;
    if (kb < K)
      p.u = *reinterpret_cast<const uint4*>(W + j * K + kb);
    else
      p.u = make_uint4(0, 0, 0, 0);
    wreg[c] = p.v;
  }
}

// B operand of the forward layout (row j = lane%16 of [16][480]) for the 15 steps.
__device__ __forceinline__ void load_weights_fwd(const u16* __restrict__ W, bf16x8 (&wreg)[KZ]) {
  const int j = threadIdx.x & 15, kq = 8 * ((threadIdx.x & 63) >> 4);
#pragma unroll
  for (int c = 0; c < KZ; c++) {
    Pack8 p;
    p.u = *reinterpret_cast<const uint4*>(W + j * KF + c * 32 + kq);
    wreg[c] = p.v;
  }
}

// In-plane position (dy*3 + dx) and channel half read by lane group g = lane/16 at position pair s.
__device__ __forceinline__ int fwd_pos(int s, int g) { return min(2 * s + (g >> 1), 8); }

// ---- forward convolution ------------------------------------------------------------------------------
// MODE 0: y = relu(h) (layer 1);  MODE 1: y = x_centre + relu(h) (layers 2-5), both written into the
// interior of a padded [n][10^3][16] BF16 buffer whose border stays zero, with h kept for the backward
// pass when H != null;  MODE 2: D = h for the first 8 channels, FP32 [n][8^3][8] (output layer).
// One workgroup: kTZ output planes of one sample.
template <int MODE>
__global__ void __launch_bounds__(kThreads) conv_fwd_kernel(const u16* __restrict__ X, const u16* __restrict__ W,
                                                            const float* __restrict__ bias, u16* __restrict__ Y,
                                                            u16* __restrict__ H, float* __restrict__ D) {
  __shared__ uint4 smem[(kTZ + 2) * kPlane / 8];
  const u16* xs = reinterpret_cast<const u16*>(smem);
  const int n = blockIdx.x / kSlabs, z0 = (blockIdx.x % kSlabs) * kTZ;
  const uint4* src = reinterpret_cast<const uint4*>(X + ((size_t)n * S + z0) * kPlane);
  for (int i = threadIdx.x; i < (kTZ + 2) * kPlane / 8; i += kThreads) smem[i] = src[i];
  bf16x8 wreg[KZ];
  load_weights_fwd(W, wreg);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, co0 = 4 * g;   // this lane's output channels co0..co0+3
  const float4 b4 = *reinterpret_cast<const float4*>(bias + co0);
  __syncthreads();
  constexpr int M = kTZ * So * So;   // 128 output voxels: 8 tiles, 2 per wave
  for (int t = wave; t < M / 16; t += kThreads / 64) {
    const int v = t * 16 + (lane & 15);   // this lane's voxel: its B-operand column and its D column
    const int lz = v / (So * So), oy = (v / So) % So, ox = v % So;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < KZ; c++) {
      const int dz = c / 5, pos = fwd_pos(c % 5, g), dy = pos / 3, dx = pos % 3;
      Pack8 a;
      a.u = *reinterpret_cast<const uint4*>(xs + (((lz + dz) * S + oy + dy) * S + ox + dx) * CH + 8 * (g & 1));
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[c], a.v, acc, 0, 0, 0);
    }
    const float h0 = acc[0] + b4.x, h1 = acc[1] + b4.y, h2 = acc[2] + b4.z, h3 = acc[3] + b4.w;
    const size_t o = (((size_t)n * So + z0 + lz) * So + oy) * So + ox;  // voxel in the 8^3 output
    if (MODE == 2) {
      if (co0 < kCnnOut) *reinterpret_cast<float4*>(D + o * kCnnOut + co0) = make_float4(h0, h1, h2, h3);
    } else {
      float o0 = fmaxf(h0, 0.0f), o1 = fmaxf(h1, 0.0f), o2 = fmaxf(h2, 0.0f), o3 = fmaxf(h3, 0.0f);
      if (MODE == 1) {
        const uint2 rx = *reinterpret_cast<const uint2*>(xs + (((lz + 1) * S + oy + 1) * S + ox + 1) * CH + co0);
        o0 += bf_lo(rx.x);
        o1 += bf_hi(rx.x);
        o2 += bf_lo(rx.y);
        o3 += bf_hi(rx.y);
      }
      *reinterpret_cast<uint2*>(Y + ((((size_t)n * S + z0 + lz + 1) * S + oy + 1) * S + ox + 1) * CH + co0) =
          pack4_bf16(o0, o1, o2, o3);
      if (H) *reinterpret_cast<uint2*>(H + o * CH + co0) = pack4_bf16(h0, h1, h2, h3);
    }
  }
}

// ---- loss and backward ---------------------------------------------------------------------------------

// MSE over the n*512*8 outputs: dY [n][512][16] = 2 (D - T) / count on the 8 live channels, 0 on the
// others; per-block partial sums of (D - T)^2.
__global__ void __launch_bounds__(kThreads) mse_kernel(const float* __restrict__ D, const float* __restrict__ T,
                                                       float* __restrict__ dY, float* __restrict__ partial,
                                                       int nvox) {
  __shared__ float red[kThreads];
  const float inv = 2.0f / ((float)nvox * kCnnOut);
  float s = 0.0f;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nvox * CH; i += gridDim.x * kThreads) {
    const int v = i / CH, c = i % CH;
    float g = 0.0f;
    if (c < kCnnOut) {
      const float e = D[(size_t)v * kCnnOut + c] - T[(size_t)v * kCnnOut + c];
      s += e * e;
      g = e * inv;
    }
    dY[i] = g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// 7-point stencil (six neighbours - 6 a) at cell f of a 16^3 frame held in LDS, zero outside the frame.
__device__ __forceinline__ float stencil16(const float* a, int f) {
  constexpr int F = 16;
  const int x = f % F, y = (f / F) % F, z = f / (F * F);
  const float zp = z + 1 < F ? a[f + F * F] : 0.0f, zm = z > 0 ? a[f - F * F] : 0.0f;
  const float yp = y + 1 < F ? a[f + F] : 0.0f, ym = y > 0 ? a[f - F] : 0.0f;
  const float xp = x + 1 < F ? a[f + 1] : 0.0f, xm = x > 0 ? a[f - 1] : 0.0f;
  return (((((zp + zm) + yp) + ym) + xp) + xm) - 6.0f * a[f];
}

// Residual term of the loss, run after mse_kernel on the same stream: per sample, e = D - T on the 16^3
// frame (fine cell 2o + p of output voxel o, sub-cell p) and r = stencil(e), the residual the error leaves in
// the guess.  Adds w * 2 / (n * 4096) * stencil(r) to dY (the stencil is symmetric with the zero exterior)
// and w * sum(r^2) to this block's loss partial.  Block b handles samples b, b + gridDim.x, ...
__global__ void __launch_bounds__(kThreads) res_loss_kernel(const float* __restrict__ D, const float* __restrict__ T,
                                                            float* __restrict__ dY, float* __restrict__ partial,
                                                            int n, float w) {
  constexpr int F = 16, F3 = F * F * F;
  static_assert(F3 == kCnnOutputs, "one frame cell per output value");
  __shared__ float e[F3], r[F3];
  __shared__ float red[kThreads];
  const float k = w * (2.0f / ((float)n * kCnnOutputs));
  float acc = 0.0f;
  for (int smp = blockIdx.x; smp < n; smp += gridDim.x) {
    const float* d = D + (size_t)smp * kCnnOutputs;
    const float* t = T + (size_t)smp * kCnnOutputs;
    for (int f = threadIdx.x; f < F3; f += kThreads) {
      const int x = f % F, y = (f / F) % F, z = f / (F * F);
      const int o = ((z >> 1) * So + (y >> 1)) * So + (x >> 1), p = ((z & 1) * 2 + (y & 1)) * 2 + (x & 1);
      e[f] = d[o * kCnnOut + p] - t[o * kCnnOut + p];
    }
    __syncthreads();
    for (int f = threadIdx.x; f < F3; f += kThreads) {
      const float v = stencil16(e, f);
      r[f] = v;
      acc += v * v;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < F3; f += kThreads) {
      const int x = f % F, y = (f / F) % F, z = f / (F * F);
      const int o = ((z >> 1) * So + (y >> 1)) * So + (x >> 1), p = ((z & 1) * 2 + (y & 1)) * 2 + (x & 1);
      dY[((size_t)smp * kCnnVox + o) * CH + p] += k * stencil16(r, f);
    }
    __syncthreads();   // e and r are rewritten for the next sample
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] += w * red[0];
}

// dH = dY * (H > 0) (or dY when H is null: the linear output layer), stored BF16 for the matrix cores;
// per-block partial bias gradients (thread t always handles channel t % 16).
__global__ void __launch_bounds__(kThreads) relu_bwd_kernel(const float* __restrict__ dY, const u16* __restrict__ H,
                                                            u16* __restrict__ dH, float* __restrict__ bpart,
                                                            size_t count) {
  __shared__ float red[kThreads];
  float s = 0.0f;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kThreads) {
    const float g = (H == nullptr || bf2f(H[i]) > 0.0f) ? dY[i] : 0.0f;
    const u16 gb = f2bf(g);
    dH[i] = gb;
    s += bf2f(gb);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o >= CH; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < CH) bpart[(size_t)blockIdx.x * CH + threadIdx.x] = red[threadIdx.x];
}

// Input gradient on the interior of the padded input: dX[p][ci] = sum_{tap,co} W[co][tap][ci] dH[p+1-tap][co]
// (+ dYres[p][ci], the residual path), p in 8^3.  Implicit GEMM with WT[ci][tap*16+co].
__global__ void __launch_bounds__(kThreads) conv_bwd_data_kernel(const u16* __restrict__ dH, const u16* __restrict__ WT,
                                                                 const float* __restrict__ dYres,
                                                                 float* __restrict__ dX) {
  constexpr int pv = So * So * 2;  // uint4 per dH plane (16 channels = 2 x uint4)
  __shared__ uint4 smem[(kTZ + 2) * pv];
  const u16* hs = reinterpret_cast<const u16*>(smem);
  const int n = blockIdx.x / kSlabs, z0 = (blockIdx.x % kSlabs) * kTZ;
  // dH planes o_z in [z0-1, z0+kTZ], zero outside [0, So)
  for (int i = threadIdx.x; i < (kTZ + 2) * pv; i += kThreads) {
    const int pl = i / pv, r = i % pv;
    const int qz = z0 - 1 + pl;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (qz >= 0 && qz < So) v = reinterpret_cast<const uint4*>(dH + (((size_t)n * So + qz) * So * So) * CH)[r];
    smem[i] = v;
  }
  bf16x8 wreg[KC];
  load_weights(WT, wreg);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci = lane & 15, kq = 8 * (lane >> 4);
  __syncthreads();
  constexpr int M = kTZ * So * So;
  for (int t = wave; t < M / 16; t += kThreads / 64) {
    const int v = t * 16 + (lane & 15);
    const int lz = v / (So * So), py = (v / So) % So, px = v % So;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < KC; c++) {
      const int kb = c * 32 + kq;
      const int tap = kb / CH, co0 = kb % CH;
      Pack8 a;
      a.u = make_uint4(0, 0, 0, 0);
      if (tap < 27) {
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        const int qy = py + 1 - dy, qx = px + 1 - dx;
        // staged plane of o_z = z0 + lz + 1 - dz is lz + 2 - dz
        if (qy >= 0 && qy < So && qx >= 0 && qx < So)
          a.u = *reinterpret_cast<const uint4*>(hs + (((lz + 2 - dz) * So + qy) * So + qx) * CH + co0);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, wreg[c], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int vo = t * 16 + 4 * (lane >> 4) + r;
      const int z = vo / (So * So), y = (vo / So) % So, x = vo % So;
      const size_t o = ((((size_t)n * So + z0 + z) * So + y) * So + x) * CH + ci;
      float g = acc[r];
      if (dYres) g += dYres[o];
      dX[o] = g;
    }
  }
}

// Weight gradient: dW[co][tap*16+ci] = sum_{n,v} dH[n][v][co] X[n][v+tap][ci], one partial per workgroup.
// GEMM with M = 16 output channels, N = 432, K = voxels (32 per MFMA); operands gathered from LDS.
__global__ void __launch_bounds__(kThreads) conv_bwd_weight_kernel(const u16* __restrict__ X, const u16* __restrict__ dH,
                                                                   float* __restrict__ wpart) {
  constexpr int xvec = (kTZ + 2) * kPlane / 8;
  constexpr int hvec = kTZ * So * So * 2;
  __shared__ uint4 smem[xvec + hvec];
  const int n = blockIdx.x / kSlabs, z0 = (blockIdx.x % kSlabs) * kTZ;
  const u16* xs = reinterpret_cast<const u16*>(smem);
  const u16* hs = reinterpret_cast<const u16*>(smem + xvec);
  const uint4* xsrc = reinterpret_cast<const uint4*>(X + ((size_t)n * S + z0) * kPlane);
  for (int i = threadIdx.x; i < xvec; i += kThreads) smem[i] = xsrc[i];
  const uint4* hsrc = reinterpret_cast<const uint4*>(dH + (((size_t)n * So + z0) * So * So) * CH);
  for (int i = threadIdx.x; i < hvec; i += kThreads) smem[xvec + i] = hsrc[i];
  __syncthreads();
  constexpr int NT = K / 16;                  // 27 column tiles
  constexpr int TPW = (NT + 3) / 4;           // tiles per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  floatx4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; i++) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
  constexpr int M = kTZ * So * So;            // 128 voxels = 4 K steps
  const int kq = 8 * (lane >> 4);
  for (int k0 = 0; k0 < M; k0 += 32) {
    // A[i = co][k = voxel]: dH[v][co] for the lane's 8 voxels
    Pack8 a;
#pragma unroll
    for (int e = 0; e < 8; e++) a.h[e] = hs[(k0 + kq + e) * CH + (lane & 15)];
#pragma unroll
    for (int i = 0; i < TPW; i++) {
      const int tile = wave * TPW + i;
      if (tile >= NT) break;
      const int col = tile * 16 + (lane & 15);
      const int tap = col / CH, ci = col % CH;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      Pack8 b;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int v = k0 + kq + e;
        const int lz = v / (So * So), oy = (v / So) % So, ox = v % So;
        b.h[e] = xs[(((lz + dz) * S + oy + dy) * S + ox + dx) * CH + ci];
      }
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[i], 0, 0, 0);
    }
  }
  // D[i = co][j = col]: lane holds col = tile*16 + lane%16, co = 4*(lane/16) + r
  float* out = wpart + (size_t)blockIdx.x * 16 * K;
#pragma unroll
  for (int i = 0; i < TPW; i++) {
    const int tile = wave * TPW + i;
    if (tile >= NT) break;
    const int col = tile * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) out[(4 * (lane >> 4) + r) * K + col] = acc[i][r];
  }
}

// Sum `parts` partial vectors of length len (fixed order: deterministic) and store into dst.
__global__ void reduce_parts_kernel(const float* __restrict__ part, int parts, int len, float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int p = 0; p < parts; p++) s += part[(size_t)p * len + i];
  dst[i] = s;
}

// Adam on the FP32 master weights.
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

// BF16 copies of the six weight matrices: the forward layout Wb[l][co][c*32 + kk] (step c = dz*5 + s,
// position 2s + kk/16, channel kk%16; zero at the pad position 9), and per-tap transposed
// WT[l][ci][tap*16+co] = W[l][co][tap*16+ci] for the input-gradient pass.
__global__ void weight_copies_kernel(const float* __restrict__ P, u16* __restrict__ Wb, u16* __restrict__ WT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCnnLayers * 16 * KF) {
    const int l = i / (16 * KF), r = i % (16 * KF);
    const int co = r / KF, k = r % KF, c = k / 32, kk = k % 32;
    const int dz = c / 5, pos = 2 * (c % 5) + (kk >> 4);
    const float* W = P + (size_t)l * kCnnLayerParams;
    Wb[i] = pos < 9 ? f2bf(W[co * K + (dz * 9 + pos) * CH + (kk & 15)]) : (u16)0;
  }
  if (i < kCnnLayers * 16 * K) {
    const int l = i / (16 * K), r = i % (16 * K);
    const float* W = P + (size_t)l * kCnnLayerParams;
    const int ci = r / K, k = r % K, tap = k / CH, co = k % CH;
    WT[i] = f2bf(W[co * K + tap * CH + ci]);
  }
}

// ---- input preparation and guess assembly -------------------------------------------------------------

constexpr float kResScale = 64.0f;   // q1 = kResScale * r / s

// SOR residual of the interpolated cube at interior cube cell (X, Y, Z) (1..R), reference operand order.
template <int N>
__device__ __forceinline__ float interp_residual(const float* cube, int X, int Y, int Z, float crho) {
  const float* c = cube + (Z * N + Y) * N + X;
  return c[N * N] + c[-N * N] + c[N] + c[-N] + c[1] + c[-1] - 6.0f * c[0] - crho;
}

// Interpolated solve cube of one patch into LDS (N^3 floats).
template <int GS>
__device__ void interp_cube(const float* __restrict__ cpot, int int_scheme, float* cube) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, C = g.C, U = g.U, NC = C - 2;
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
}

// One workgroup per patch: interpolate, take s = max|interp|, write the packed input.
template <int GS>
__device__ void prep_patch(const float* __restrict__ rho, const float* __restrict__ cpot, float cst, int int_scheme,
                           u16* __restrict__ x, float* __restrict__ scale, float* cube) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, R = g.R, OFF = 5 - GS;
  __shared__ float red[kThreads];
  interp_cube<GS>(cpot, int_scheme, cube);
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
  // one thread per (coarse voxel, sub-cell): 1000 x 8 items, 2 channels each
  for (int t = threadIdx.x; t < S * S * S * 8; t += kThreads) {
    const int vox = t >> 3, d = t & 7;
    const int cx = vox % S, cy = (vox / S) % S, cz = vox / (S * S);
    const int dz = d >> 2, dy = (d >> 1) & 1, dx = d & 1;
    // frame index f = 2c - 1 + d, cube index f - OFF
    const int X = 2 * cx - 1 + dx - OFF, Y = 2 * cy - 1 + dy - OFF, Z = 2 * cz - 1 + dz - OFF;
    float c0 = 0.0f, c1 = 0.0f;
    if (X >= 0 && X < N && Y >= 0 && Y < N && Z >= 0 && Z < N) {
      c0 = cube[(Z * N + Y) * N + X] * inv;
      if (X >= 1 && X <= R && Y >= 1 && Y <= R && Z >= 1 && Z <= R)
        c1 = (interp_residual<N>(cube, X, Y, Z, cst * rho[((Z - 1) * R + (Y - 1)) * R + (X - 1)]) * inv) * kResScale;
    }
    reinterpret_cast<uint32_t*>(x)[t] = (uint32_t)f2bf(c0) | ((uint32_t)f2bf(c1) << 16);
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
  u16* xp = x + (size_t)p * kCnnInput;
  switch (gs) {
    case 1: prep_patch<1>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    case 2: prep_patch<2>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    case 3: prep_patch<3>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    case 4: prep_patch<4>(r, c, cst, int_scheme, xp, scale + p, cube); break;
    default: prep_patch<5>(r, c, cst, int_scheme, xp, scale + p, cube); break;
  }
}

// Guess of the interior: interp + s * D at interior frame index i = 2o + p.
template <int GS>
__device__ void guess_patch(const float* __restrict__ cpot, const float* __restrict__ d, float s, int int_scheme,
                            float* __restrict__ guess, float* cube) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, R = g.R, OFF = 5 - GS;
  interp_cube<GS>(cpot, int_scheme, cube);
  for (int t = threadIdx.x; t < R * R * R; t += kThreads) {
    const int x = t % R, y = (t / R) % R, z = t / (R * R);   // interior-local; cube index +1
    // interior frame index: cube index + OFF - 1 = interior-local + OFF
    const int ix = x + OFF, iy = y + OFF, iz = z + OFF;
    const int o = ((iz >> 1) * So + (iy >> 1)) * So + (ix >> 1);
    const int p = ((iz & 1) * 2 + (iy & 1)) * 2 + (ix & 1);
    guess[t] = cube[((z + 1) * N + y + 1) * N + x + 1] + s * d[o * kCnnOut + p];
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
  const float* d = D + (size_t)p * kCnnOutputs;
  float* out = guess + (size_t)p * g.R * g.R * g.R;
  switch (gs) {
    case 1: guess_patch<1>(c, d, scale[p], int_scheme, out, cube); break;
    case 2: guess_patch<2>(c, d, scale[p], int_scheme, out, cube); break;
    case 3: guess_patch<3>(c, d, scale[p], int_scheme, out, cube); break;
    case 4: guess_patch<4>(c, d, scale[p], int_scheme, out, cube); break;
    default: guess_patch<5>(c, d, scale[p], int_scheme, out, cube); break;
  }
}

// ---- fused inference: one workgroup per patch, activations resident in LDS ------------------------------
// interp + packing -> six convolutions ping-ponging between two padded 10^3 x 16 BF16 buffers -> guess.
// Each output value goes through exactly the MFMA sequence of conv_fwd_kernel, so the guess is bitwise
// identical to prep + forward + guess with the layer-by-layer kernels.
//
// * An activation buffer is split into two half-planes of 8 channels, [half][voxel][8] (16 KB each), so
//   that the 16-byte A-operand reads of 8 consecutive lanes (8 consecutive voxels along x) are one
//   contiguous 128-byte span.  The two rows of a tile start 160 bytes apart, so their spans share 32
//   bytes of banks: measured about 3.6 extra LDS cycles per read; a pitch-8 layout without the x border
//   removes that overlap but needs a select per read and measured slower.
// * Global traffic is issued early and coalesced: the coarse cube is staged once (float4), the density is
//   read with float4 loads and scattered into the packed input, and the interpolated interior is kept in
//   LDS (over the coarse staging, which is dead by then) for the final guess.

// Half-plane stride (bf16 values).  Unpadded: measured on MI355X, padding the stride by 32, 64, 128 or
// 256 bytes made the kernel 2-13% slower (more LDS bank-conflict cycles on the 16-byte A-operand reads).
constexpr int kHalfP = kCnnPadVox * 8;
constexpr int kBufVec = kHalfP * 2 / 8;      // uint4 per activation buffer
constexpr int kCoarseMax = 12 * 12 * 12;     // C^3 at ghost size 5
constexpr int kInteriorMax = 16 * 16 * 16;   // R^3 at ghost size 5
constexpr int kFT = 256;                     // threads of a fused workgroup (4 or 8 waves)
constexpr int kZSplit = kFT / 256;           // groups of output planes the waves split a layer into
static_assert(kFT == 256 || kFT == 512, "4 row pairs x 1 or 2 plane groups");

// 80.4 KB: two workgroups fit in the 160 KB of LDS of a CU
struct FusedLds {
  uint4 A[kBufVec];
  uint4 B[kBufVec];
  union {
    float cs[kCoarseMax];         // coarse cube, until the interpolation is done
    float in[kInteriorMax];       // then the interpolated interior (solve cells) for the guess
  } c;
  float red[kFT / 64];
};
static_assert(sizeof(FusedLds) <= 160 * 1024 / 2, "two workgroups per CU");

// Layer l's weights (A operand) and the biases of this lane's output channels 4*(lane/16)..+3.
__device__ __forceinline__ void load_layer(const u16* __restrict__ Wb, const float* __restrict__ P, int l,
                                           bf16x8 (&w)[KZ], float4& b) {
  load_weights_fwd(Wb + (size_t)l * 16 * KF, w);
  b = *reinterpret_cast<const float4*>(P + (size_t)l * kCnnLayerParams + 16 * K + 4 * ((threadIdx.x & 63) >> 4));
}

// This lane's A-operand offsets (bf16 elements, relative to its row voxel in an input plane) for the five
// position pairs: position fwd_pos(s, g) of the plane, channel half g & 1 (g = lane/16).
__device__ __forceinline__ void pair_offsets(int (&off)[5]) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const int pos = fwd_pos(s, g);
    off[s] = (g & 1) * kHalfP + ((pos / 3) * S + pos % 3) * 8;
  }
}

// Epilogue of output plane o for row pair rp: output channels co0..co0+3 (co0 = 4*(lane/16)) of voxel
// lane%16 of tile 4o + rp, 4 consecutive channels of one half-plane (8 bytes).
template <int MODE>
__device__ __forceinline__ void conv_store(const u16* in, u16* out, float* D, const floatx4& acc, const float4& b,
                                           int o, int rp) {
  const int lane = threadIdx.x & 63;
  const int co0 = 4 * (lane >> 4);
  const int vo = (o * So + rp + 4 * ((lane & 15) >> 3)) * So + (lane & 7);
  const float h0 = acc[0] + b.x, h1 = acc[1] + b.y, h2 = acc[2] + b.z, h3 = acc[3] + b.w;
  if (MODE == 2) {
    if (co0 < kCnnOut) *reinterpret_cast<float4*>(D + vo * kCnnOut + co0) = make_float4(h0, h1, h2, h3);
  } else {
    const int q = (((vo >> 6) + 1) * S + ((vo >> 3) & 7) + 1) * S + (vo & 7) + 1;
    const int c0 = (co0 >> 3) * kHalfP + q * 8 + (co0 & 7);
    float o0 = fmaxf(h0, 0.0f), o1 = fmaxf(h1, 0.0f), o2 = fmaxf(h2, 0.0f), o3 = fmaxf(h3, 0.0f);
    if (MODE == 1) {
      const uint2 rx = *reinterpret_cast<const uint2*>(in + c0);
      o0 += bf_lo(rx.x);
      o1 += bf_hi(rx.x);
      o2 += bf_lo(rx.y);
      o3 += bf_hi(rx.y);
    }
    *reinterpret_cast<uint2*>(out + c0) = pack4_bf16(o0, o1, o2, o3);
  }
}

// One convolution layer from LDS to LDS (MODE 0/1) or to the FP32 output D in LDS (MODE 2).  Wave w owns
// output rows y = 2rp, 2rp+1 (rp = w % 4) of the NO = 8 / kZSplit output planes P0.. of its plane group and
// sweeps the NO + 2 input planes they read once: input plane P0 + q feeds output planes P0 + q (dz = 0),
// P0 + q - 1 (dz = 1) and P0 + q - 2 (dz = 2), so every A operand read from LDS serves up to three MFMAs
// (with 4 waves: 50 reads for 120 MFMAs per wave instead of one read per MFMA).  Output plane o still
// accumulates its 15 steps in the order of conv_fwd_kernel: dz = 0, 1, 2, and s = 0..4 within each.
template <int MODE>
__device__ __forceinline__ void conv_lds(const u16* in, u16* out, const bf16x8 (&wreg)[KZ], const float4& b, float* D,
                                         const int (&off)[5]) {
  constexpr int NO = So / kZSplit, NQ = NO + 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15;
  const int rp = wave & 3, P0 = (wave >> 2) * NO;
  const u16* col = in + ((P0 * S + rp + 4 * (i >> 3)) * S + (i & 7)) * 8;   // plane P0, row voxel of lane i
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  floatx4 a0 = zero, a1 = zero;   // output planes P0+q-2 and P0+q-1
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    Pack8 a[5];
#pragma unroll
    for (int s = 0; s < 5; s++) a[s].u = *reinterpret_cast<const uint4*>(col + q * (S * S * 8) + off[s]);
    floatx4 a2 = zero;            // output plane P0+q
#pragma unroll
    for (int s = 0; s < 5; s++) {
      if (q < NO) a2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[s], a[s].v, a2, 0, 0, 0);
      if (q >= 1 && q <= NO) a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[5 + s], a[s].v, a1, 0, 0, 0);
      if (q >= 2) a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[10 + s], a[s].v, a0, 0, 0, 0);
    }
    if (q >= 2) conv_store<MODE>(in, out, D, a0, b, P0 + q - 2, rp);
    a0 = a1;
    a1 = a2;
  }
}

// Zero the border voxels (any coordinate 0 or 9) of both half-planes of a padded activation buffer.  The
// 488 border voxels of a half-plane are enumerated directly: the planes z = 0 and z = 9 (100 each), then per
// inner plane the rows y = 0 and y = 9 (10 each) and the x = 0 / x = 9 ends of the 8 inner rows.
__device__ __forceinline__ void zero_border(uint4* buf) {
  constexpr int SS = S * S, RING = 4 * S - 4, NB = 2 * SS + (S - 2) * RING;
  static_assert(NB == kCnnPadVox - (S - 2) * (S - 2) * (S - 2), "border count");
  for (int i = threadIdx.x; i < 2 * NB; i += kFT) {
    const int h = i / NB;
    int k = i % NB, vox;
    if (k < 2 * SS) {
      vox = k < SS ? k : k + (S - 2) * SS;
    } else {
      k -= 2 * SS;
      const int z = 1 + k / RING, r = k % RING;
      if (r < 2 * S)
        vox = z * SS + (r < S ? r : (S - 1) * S + r - S);
      else
        vox = z * SS + (1 + (r - 2 * S) / 2) * S + ((r - 2 * S) & 1) * (S - 1);
    }
    buf[h * (kHalfP / 8) + vox] = make_uint4(0, 0, 0, 0);
  }
}

// Interpolated solve cube from a coarse cube already staged in LDS (same arithmetic as interp_cube), into a
// cube stored with pitch N + 2 and one spare cell on each side (cube cell (X, Y, Z) at
// cube[(Z * (N + 2) + Y) * (N + 2) + X], -1 <= X, Y, Z <= N), so that the children that fall outside the solve
// cube land in the spare layer and every child is stored without a test.  Returns this thread's max |value|
// over the children inside the solve cube (the cube's max once reduced over the workgroup).
template <int GS>
__device__ float interp_cube_lds(const float* cs, int int_scheme, float* cube) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, NP = N + 2, C = g.C, U = g.U, NC = C - 2;
  float mx = 0.0f;
  for (int q = threadIdx.x; q < NC * NC * NC; q += kFT) {
    const int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
    float v[3][3][3];
    for (int dk = 0; dk < 3; dk++)
      for (int dj = 0; dj < 3; dj++)
        for (int di = 0; di < 3; di++) v[dk][dj][di] = cs[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
    float ch[2][2][2];
    interp_children(v, int_scheme, ch);
    const int X0 = 2 * (i - 1) - U, Y0 = 2 * (j - 1) - U, Z0 = 2 * (k - 1) - U;
    const bool vx[2] = {X0 >= 0, X0 + 1 < N}, vy[2] = {Y0 >= 0, Y0 + 1 < N}, vz[2] = {Z0 >= 0, Z0 + 1 < N};
    float* base = cube + (Z0 * NP + Y0) * NP + X0;
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
      for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++) {
          base[(cz * NP + cy) * NP + cx] = ch[cz][cy][cx];
          mx = fmaxf(mx, vx[cx] && vy[cy] && vz[cz] ? fabsf(ch[cz][cy][cx]) : 0.0f);
        }
  }
  return mx;
}

template <int GS>
__device__ void fused_patch(const float* __restrict__ rho, const float* __restrict__ cpot, const u16* __restrict__ Wb,
                            const float* __restrict__ P, float cst, int int_scheme, float* __restrict__ guess,
                            FusedLds& L) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, R = g.R, C = g.C, OFF = 5 - GS;
  constexpr int R3 = R * R * R, C3 = C * C * C;
  static_assert(N * N * N * 4 <= kBufVec * 16 && kCnnOutputs * 4 <= kBufVec * 16, "scratch must fit a buffer");
  static_assert(C3 <= kCoarseMax && R3 <= kInteriorMax && R3 % 4 == 0 && C3 % 4 == 0, "float4 staging");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NP = N + 2;   // cube pitch (one spare cell on each side)
  static_assert(NP * NP * NP * 4 <= kBufVec * 16, "padded cube must fit a buffer");
  float* cube = reinterpret_cast<float*>(L.B) + (NP * NP + NP + 1);   // cube cell (0, 0, 0)
  u16* A = reinterpret_cast<u16*>(L.A);
  u16* B = reinterpret_cast<u16*>(L.B);
  // first layer's weights, the coarse cube and this lane's share of the density are requested together
  bf16x8 wa[KZ], wb[KZ];
  float4 ba, bb;
  load_layer(Wb, P, 0, wa, ba);
  for (int t = threadIdx.x; t < C3 / 4; t += kFT)
    reinterpret_cast<float4*>(L.c.cs)[t] = reinterpret_cast<const float4*>(cpot)[t];
  // Thread t < R^2 owns the solve-interior column (x, y) = (t % R, t / R) for the packing and the guess; its
  // density column is requested now (each wavefront load reads 64 consecutive floats).
  constexpr int RR = R * R;
  static_assert(RR <= kFT, "one thread per interior column");
  const bool owns_col = threadIdx.x < RR;
  float rcol[R];
#pragma unroll
  for (int z = 0; z < R; z++) rcol[z] = owns_col ? rho[z * RR + threadIdx.x] : 0.0f;
  __syncthreads();
  float mx = interp_cube_lds<GS>(L.c.cs, int_scheme, cube);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if (lane == 0) L.red[wave] = mx;
  __syncthreads();   // the cube is complete and the coarse staging is dead
  mx = L.red[0];
#pragma unroll
  for (int w = 1; w < kFT / 64; w++) mx = fmaxf(mx, L.red[w]);
  const float s = mx > 0.0f ? mx : 1.0f;
  const float inv = 1.0f / s;
  // Packed input over the 20^3 frame (frame cell e = cube cell + 1 + OFF): channel pair (q0, q1) of sub-cell
  // d of voxel e/2 is one 32-bit word (q0 low) at half-plane d/4, in-half channels 2(d%4), 2(d%4)+1.
  // Solve-interior cells get q0 = interp/s and q1 = 64 r/s (r from the cube, still in B, and the density);
  // the other frame cells get q0 = interp/s on the cube's Dirichlet layer, 0 outside the cube, and q1 = 0.
  // The column owners sweep z with the column's cube values in registers; the frame columns outside the
  // interior are spread over all threads.  Both write disjoint words, so they share one barrier.
  uint32_t* A32 = reinterpret_cast<uint32_t*>(A);
  constexpr int HW = kHalfP / 2;                 // 32-bit words per half-plane
  constexpr int E = 2 * S, LO = 2 + OFF, HI = LO + R;   // frame edge, interior frame range [LO, HI)
  if (owns_col) {
    const int x = threadIdx.x % R, y = threadIdx.x / R;
    const int ex = x + LO, ey = y + LO;
    const int colw = (((ey >> 1) * S + (ex >> 1)) * 4) + (ey & 1) * 2 + (ex & 1);   // word of frame z = 0
    const float* cc = cube + (y + 1) * NP + (x + 1);                              // cube column, stride NP^2
#pragma unroll
    for (int ez = 0; ez < LO; ez++) {   // below the interior: Dirichlet layer, then outside the cube
      const int Z = ez - 1 - OFF;
      const float c0 = Z >= 0 ? cc[Z * NP * NP] * inv : 0.0f;
      A32[(ez & 1) * HW + (ez >> 1) * (S * S * 4) + colw] = (uint32_t)f2bf(c0);
    }
#pragma unroll
    for (int ez = HI; ez < E; ez++) {   // above the interior
      const int Z = ez - 1 - OFF;
      const float c0 = Z < N ? cc[Z * NP * NP] * inv : 0.0f;
      A32[(ez & 1) * HW + (ez >> 1) * (S * S * 4) + colw] = (uint32_t)f2bf(c0);
    }
    float cm = cc[0], c = cc[NP * NP];
#pragma unroll
    for (int z = 0; z < R; z++) {
      const float* cz = cc + (z + 1) * NP * NP;
      const float cp = cz[NP * NP];
      const float r = cp + cm + cz[NP] + cz[-NP] + cz[1] + cz[-1] - 6.0f * c - cst * rcol[z];
      const int ez = z + LO;
      A32[(ez & 1) * HW + (ez >> 1) * (S * S * 4) + colw] = pack2_bf16(c * inv, (r * inv) * kResScale);
      L.c.in[(z * R + y) * R + x] = c;
      cm = c;
      c = cp;
    }
  }
  // Frame columns outside the interior (rows below LO and above HI, then the sides of the interior rows),
  // as half-columns of E/2 cells: item i is column i/2, cells ez in [(i&1)*E/2, (i&1)*E/2 + E/2).
  constexpr int NOC = E * E - RR;
  for (int i = threadIdx.x; i < 2 * NOC; i += kFT) {
    const int j = i >> 1, z0 = (i & 1) * (E / 2);
    int ex, ey;
    if (j < LO * E) {
      ey = j / E;
      ex = j % E;
    } else if (j < E * E - HI * E + LO * E) {
      ey = HI + (j - LO * E) / E;
      ex = (j - LO * E) % E;
    } else {
      const int k = j - (E * E - HI * E + LO * E), w = E - R;
      ey = LO + k / w;
      ex = k % w < LO ? k % w : k % w - LO + HI;
    }
    const int X = ex - 1 - OFF, Y = ey - 1 - OFF;
    const bool xy_in = X >= 0 && X < N && Y >= 0 && Y < N;
    const float* cc = cube + (xy_in ? Y * NP + X : 0);
    uint32_t* w0 = A32 + ((ey >> 1) * S + (ex >> 1)) * 4 + (ey & 1) * 2 + (ex & 1);
#pragma unroll
    for (int dz = 0; dz < E / 2; dz++) {
      const int ez = z0 + dz, Z = ez - 1 - OFF;
      const float c0 = xy_in && Z >= 0 && Z < N ? cc[Z * NP * NP] * inv : 0.0f;
      w0[(ez & 1) * HW + (ez >> 1) * (S * S * 4)] = (uint32_t)f2bf(c0);
    }
  }
  __syncthreads();   // every lane is done reading the cube in B
  for (int i = threadIdx.x; i < kBufVec; i += kFT) L.B[i] = make_uint4(0, 0, 0, 0);
  int off[5];
  pair_offsets(off);
  __syncthreads();
  load_layer(Wb, P, 1, wb, bb);
  conv_lds<0>(A, B, wa, ba, nullptr, off);     // layer 1: A -> B
  __syncthreads();
  zero_border(L.A);                            // A held the input's border
  load_layer(Wb, P, 2, wa, ba);
  conv_lds<1>(B, A, wb, bb, nullptr, off);     // layer 2: B -> A
  __syncthreads();
  load_layer(Wb, P, 3, wb, bb);
  conv_lds<1>(A, B, wa, ba, nullptr, off);     // layer 3
  __syncthreads();
  load_layer(Wb, P, 4, wa, ba);
  conv_lds<1>(B, A, wb, bb, nullptr, off);     // layer 4
  __syncthreads();
  load_layer(Wb, P, 5, wb, bb);
  conv_lds<1>(A, B, wa, ba, nullptr, off);     // layer 5
  __syncthreads();
  float* D = reinterpret_cast<float*>(L.A);
  conv_lds<2>(B, nullptr, wb, bb, D, off);     // layer 6: B -> D (in A)
  __syncthreads();
  if (owns_col) {   // the guess, column by column (each wavefront store covers 64 consecutive floats)
    const int x = threadIdx.x % R, y = threadIdx.x / R;
    const int ix = x + OFF, iy = y + OFF;
    const int colo = ((iy >> 1) * So + (ix >> 1)) * kCnnOut + (iy & 1) * 2 + (ix & 1);
#pragma unroll
    for (int z = 0; z < R; z++) {
      const int iz = z + OFF, t = (z * R + y) * R + x;
      guess[t] = L.c.in[t] + s * D[colo + (iz >> 1) * (So * So * kCnnOut) + (iz & 1) * 4];
    }
  }
}

__device__ __forceinline__ void fused_dispatch(int gs, const float* r, const float* c, const u16* Wb, const float* P,
                                               float cst, int int_scheme, float* out, FusedLds& L) {
  switch (gs) {
    case 1: fused_patch<1>(r, c, Wb, P, cst, int_scheme, out, L); break;
    case 2: fused_patch<2>(r, c, Wb, P, cst, int_scheme, out, L); break;
    case 3: fused_patch<3>(r, c, Wb, P, cst, int_scheme, out, L); break;
    case 4: fused_patch<4>(r, c, Wb, P, cst, int_scheme, out, L); break;
    default: fused_patch<5>(r, c, Wb, P, cst, int_scheme, out, L); break;
  }
}

__global__ void __launch_bounds__(kFT, kFT / 128) fused_guess_kernel(const float* __restrict__ rho,
                                                               const float* __restrict__ cpot,
                                                               const u16* __restrict__ Wb, const float* __restrict__ P,
                                                               int npatch, int gs, float cst, int int_scheme,
                                                               float* __restrict__ guess) {
  __shared__ FusedLds L;
  const int p = blockIdx.x;
  if (p >= npatch) return;
  const Geo g = geo_of(gs);
  const float* r = rho + (size_t)p * g.R * g.R * g.R;
  const float* c = cpot + (size_t)p * g.C * g.C * g.C;
  float* out = guess + (size_t)p * g.R * g.R * g.R;
  fused_dispatch(gs, r, c, Wb, P, cst, int_scheme, out, L);
}

// Mixed ghost sizes: patch p reads its ghost size and packed offsets from the plan's descriptors; the
// guess is packed like the density (offset rho_off).
__global__ void __launch_bounds__(kFT, kFT / 128) fused_guess_mixed_kernel(const float* __restrict__ rho,
                                                                     const float* __restrict__ cpot,
                                                                     const PatchDesc* __restrict__ desc,
                                                                     const u16* __restrict__ Wb,
                                                                     const float* __restrict__ P, int npatch, float cst,
                                                                     int int_scheme, float* __restrict__ guess) {
  __shared__ FusedLds L;
  const int p = blockIdx.x;
  if (p >= npatch) return;
  const PatchDesc d = desc[p];
  fused_dispatch(d.gs, rho + d.rho_off, cpot + d.cpot_off, Wb, P, cst, int_scheme, guess + d.rho_off, L);
}

}  // namespace cnn

// ---- host launchers ----------------------------------------------------------------------------------------

using namespace cnn;

hipError_t cnn_conv_fwd(const uint16_t* X, const uint16_t* W, const float* b, uint16_t* Y, uint16_t* H, float* D,
                        int n, int mode, hipStream_t s) {
  const dim3 grid(n * kSlabs), block(kThreads);
  if (mode == 0)
    hipLaunchKernelGGL(conv_fwd_kernel<0>, grid, block, 0, s, X, W, b, Y, H, D);
  else if (mode == 1)
    hipLaunchKernelGGL(conv_fwd_kernel<1>, grid, block, 0, s, X, W, b, Y, H, D);
  else if (mode == 2)
    hipLaunchKernelGGL(conv_fwd_kernel<2>, grid, block, 0, s, X, W, b, Y, H, D);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t cnn_mse(const float* D, const float* T, float* dY, float* partial, int nvox, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mse_kernel, dim3(nblocks), dim3(kThreads), 0, s, D, T, dY, partial, nvox);
  return hipGetLastError();
}

hipError_t cnn_res_loss(const float* D, const float* T, float* dY, float* partial, int n, float w, int nblocks,
                        hipStream_t s) {
  hipLaunchKernelGGL(res_loss_kernel, dim3(nblocks), dim3(kThreads), 0, s, D, T, dY, partial, n, w);
  return hipGetLastError();
}

hipError_t cnn_relu_bwd(const float* dY, const uint16_t* H, uint16_t* dH, float* bpart, size_t count, int nblocks,
                        hipStream_t s) {
  hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblocks), dim3(kThreads), 0, s, dY, H, dH, bpart, count);
  return hipGetLastError();
}

hipError_t cnn_conv_bwd_data(const uint16_t* dH, const uint16_t* WT, const float* dYres, float* dX, int n,
                             hipStream_t s) {
  hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(n * kSlabs), dim3(kThreads), 0, s, dH, WT, dYres, dX);
  return hipGetLastError();
}

int cnn_bwd_weight_blocks(int n) { return n * kSlabs; }

hipError_t cnn_conv_bwd_weight(const uint16_t* X, const uint16_t* dH, float* wpart, int n, hipStream_t s) {
  hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3(n * kSlabs), dim3(kThreads), 0, s, X, dH, wpart);
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

hipError_t cnn_weight_copies(const float* P, uint16_t* Wb, uint16_t* WT, hipStream_t s) {
  const int n = kCnnLayers * 16 * KF;   // covers both copies (KF > K)
  hipLaunchKernelGGL(weight_copies_kernel, dim3((n + 255) / 256), dim3(256), 0, s, P, Wb, WT);
  return hipGetLastError();
}

hipError_t cnn_prep(const float* rho, const float* cpot, int n, int gs, float cst, int int_scheme, uint16_t* x,
                    float* scale, hipStream_t s) {
  const Geo g = geo_of(gs);
  hipLaunchKernelGGL(prep_kernel, dim3(n), dim3(kThreads), sizeof(float) * g.N * g.N * g.N, s, rho, cpot, gs, cst,
                     int_scheme, x, scale);
  return hipGetLastError();
}

hipError_t cnn_fused_guess(const float* rho, const float* cpot, const uint16_t* Wb, const float* P, int n, int gs,
                           float cst, int int_scheme, float* guess, hipStream_t s) {
  hipLaunchKernelGGL(fused_guess_kernel, dim3(n), dim3(kFT), 0, s, rho, cpot, Wb, P, n, gs, cst, int_scheme,
                     guess);
  return hipGetLastError();
}

hipError_t cnn_fused_guess_mixed(const float* rho, const float* cpot, const PatchDesc* desc, const uint16_t* Wb,
                                 const float* P, int n, float cst, int int_scheme, float* guess, hipStream_t s) {
  hipLaunchKernelGGL(fused_guess_mixed_kernel, dim3(n), dim3(kFT), 0, s, rho, cpot, desc, Wb, P, n, cst,
                     int_scheme, guess);
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
