// Learned warm start: a small 3D CNN that predicts the correction from the interpolated coarse potential
// to the converged potential on the whole solve interior, trained on solutions of the solver itself.
// No reference counterpart (the reference always starts from the interpolated cube).
//
// Network (activations BF16 channel-last, weights BF16, accumulation and master weights FP32):
//   frame   the 18^3 solve cube of ghost size 5 (smaller cubes centred, zeros around), with the fields
//           q0 = interp/s and q1 = 64 r/s on the solve interior and zero elsewhere, s = max|interp| over the
//           solve cube, r the SOR residual of the interpolated cube (reference operand order), i.e. the
//           right-hand side of A e = r that the correction e to the interpolated start satisfies (e has zero
//           Dirichlet values, so the boundary enters through r alone)
//   x0      [n][10^3][16]: space-to-depth of the zero-extended 20^3 frame (fine index f = 2c - 1 + d per
//           axis, channel ((dz*2+dy)*2+dx)*2 + q), so that coarse voxels 1..8 hold the 2^3 blocks of the
//           16^3 interior and the voxels with a coordinate 0 or 9 are zero (the first layer's padding)
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

// Step c of this lane's A operand from the lane-major copy WL of one layer ([15][64][8]).
__device__ __forceinline__ bf16x8 load_weight_step(const u16* __restrict__ WL, int c) {
  Pack8 p;
  p.u = *reinterpret_cast<const uint4*>(WL + ((size_t)c * 64 + (threadIdx.x & 63)) * 8);
  return p.v;
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
// position 2s + kk/16, channel kk%16; zero at the pad position 9), the same values lane-major
// WL[l][c][lane][8] (lane = co + 16*(kk/8) reads kk%8 = 0..7 of its row: one contiguous 1 KB load per step),
// and per-tap transposed WT[l][ci][tap*16+co] = W[l][co][tap*16+ci] for the input-gradient pass.
__global__ void weight_copies_kernel(const float* __restrict__ P, u16* __restrict__ Wb, u16* __restrict__ WT,
                                     u16* __restrict__ WL) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCnnLayers * 16 * KF) {
    const int l = i / (16 * KF), r = i % (16 * KF);
    const int co = r / KF, k = r % KF, c = k / 32, kk = k % 32;
    const int dz = c / 5, pos = 2 * (c % 5) + (kk >> 4);
    const float* W = P + (size_t)l * kCnnLayerParams;
    const u16 v = pos < 9 ? f2bf(W[co * K + (dz * 9 + pos) * CH + (kk & 15)]) : (u16)0;
    Wb[i] = v;
    WL[(((size_t)l * KZ + c) * 64 + co + 16 * (kk >> 3)) * 8 + (kk & 7)] = v;
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

// One workgroup per patch: interpolate, take s = max|interp|, write the packed input (zero outside the solve
// interior, so every voxel with a coordinate 0 or 9 is zero).
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
    if (X >= 1 && X <= R && Y >= 1 && Y <= R && Z >= 1 && Z <= R) {   // the solve interior only
      c0 = cube[(Z * N + Y) * N + X] * inv;
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
// interpolation -> packing -> six convolutions ping-ponging between two LDS buffers -> guess written by the
// output layer's epilogue.  Each output value goes through the MFMA sequence of
// conv_fwd_kernel with the products of the all-zero input planes (padded z = 0 and 9) left out (adding those
// zero products to an accumulator that started at +0 leaves it bitwise unchanged), so the guess is bitwise
// that of prep + forward + guess with the layer-by-layer kernels.
//
// * An activation buffer holds only the 8 planes of the output grid, each [half][10 x 10 voxels][8 channels]
//   BF16 with a zero ring (y or x = 0 or 9), 25.6 KB.  The 16-byte A-operand reads of 8 lanes (8 consecutive
//   voxels along x) are one 128-byte span; a tile pairs rows r and r + 4, 640 bytes apart (128 modulo the
//   256-byte bank width), so the 16 lanes of a read phase cover all 64 banks once.
// * The packed input (the solve interior only; its border voxels are zero like the activations' padding)
//   has the same layout, so the first layer is a layer like the others.
// * Wave w owns output rows w and w + 4 of every plane and sweeps the 8 input planes once; every A operand
//   read from LDS feeds up to three MFMAs, and the next plane's operands are requested before the current
//   plane's MFMAs.
// * Two buffers only: the coarse staging, the packed input and the activations share them with the FP32
//   interpolated cube.  The interpolated values the guess adds to are written to the guess array while the
//   input is packed; the output layer's epilogue reads them back (same lane, same address: L2), adds s * D and
//   stores the guess.  LDS per workgroup 51.2 KB: three workgroups per CU.
constexpr int kHalfV = S * S * 8;      // bf16 per half-plane (10 x 10 voxels x 8 channels)
constexpr int kPlaneV = 2 * kHalfV;    // bf16 per plane
constexpr int kActVec = So * kPlaneV / 8;   // uint4 per buffer (8 planes)
constexpr int kCoarseMax = 12 * 12 * 12;    // C^3 at ghost size 5
constexpr int kCubeMax = 18 * 18 * 18;      // N^3 at ghost size 5
constexpr int kInteriorMax = 16 * 16 * 16;  // R^3 at ghost size 5
constexpr int kFT = 256;                    // threads of a fused workgroup (4 waves)
constexpr int kRing = 4 * S - 4;            // ring voxels of a half-plane (36)

constexpr int kFusedBlocksPerCU = 3;

struct FusedLds {
  uint4 X[kActVec];   // coarse staging -> packed input -> activations
  uint4 C[kActVec];   // interpolated cube (FP32, unpadded N^3) -> activations
  float red[kFT / 64];
};
static_assert(kCoarseMax * 4 <= kActVec * 16 && kInteriorMax * 4 <= kActVec * 16 && kCubeMax * 4 <= kActVec * 16,
              "staging must fit a buffer");
static_assert(kFusedBlocksPerCU * sizeof(FusedLds) <= 160 * 1024, "workgroups per CU");

// The biases of layer l for this lane's output channels 4*(lane/16)..+3.
__device__ __forceinline__ void load_bias(const float* __restrict__ P, int l, float4& b) {
  b = *reinterpret_cast<const float4*>(P + (size_t)l * kCnnLayerParams + 16 * K + 4 * ((threadIdx.x & 63) >> 4));
}

// Layer l's weights (A operand, from the lane-major copy WL) and the biases of this lane's output channels
// 4*(lane/16)..+3.
__device__ __forceinline__ void load_layer(const u16* __restrict__ WL, const float* __restrict__ P, int l,
                                           bf16x8 (&w)[KZ], float4& b) {
#pragma unroll
  for (int c = 0; c < KZ; c++) w[c] = load_weight_step(WL + (size_t)l * 16 * KF, c);
  load_bias(P, l, b);
}

// This lane's A-operand offsets (bf16 elements, relative to its row voxel in an input plane) for the five
// position pairs: position fwd_pos(s, g) of the plane, channel half g & 1 (g = lane/16).
__device__ __forceinline__ void pair_offsets(int (&off)[5]) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const int pos = fwd_pos(s, g);
    off[s] = (g & 1) * kHalfV + ((pos / 3) * S + pos % 3) * 8;
  }
}

// Zero the ring voxels (y or x = 0 or 9) of both half-planes of the 8 planes of a buffer.
__device__ __forceinline__ void zero_ring(uint4* buf) {
  for (int i = threadIdx.x; i < 2 * So * kRing; i += kFT) {
    const int hp = i / kRing, k = i % kRing;
    int vox;
    if (k < S) {
      vox = k;                                   // row 0
    } else if (k < 2 * S) {
      vox = (S - 1) * S + k - S;                 // row 9
    } else {
      const int r = k - 2 * S;                   // rows 1..8, x = 0 or 9
      vox = (1 + (r >> 1)) * S + (r & 1) * (S - 1);
    }
    buf[hp * (kHalfV / 8) + vox] = make_uint4(0, 0, 0, 0);
  }
}

// What the output layer's epilogue needs: the scale, the interpolated interior (LDS) and where the guess goes.
struct GuessOut {
  float s;
  float* sd;         // s * D on the solve interior, [R^3] FP32 in the LDS buffer the output layer does not read
};

// Epilogue of output plane o for this wave's row pair: output channels co0..co0+3 (co0 = 4*(lane/16)) of voxel
// lane%16 of the tile.  MODE 0/1: 4 consecutive channels of one half-plane (8 bytes), y = relu(h) (+ the
// layer input at the same voxel, MODE 1).  MODE 2: channels 0..7 are the 2^3 sub-cells of the voxel; the
// guess interp + s * h of those fine cells is stored.
template <int GS, bool CUBE, int MODE>
__device__ __forceinline__ void conv_store(const u16* in, u16* out, const floatx4& acc, const float4& b, int o,
                                           const GuessOut& go) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int co0 = 4 * (lane >> 4);
  const int row = wave + 4 * ((lane & 15) >> 3), xv = lane & 7;
  const float h0 = acc[0] + b.x, h1 = acc[1] + b.y, h2 = acc[2] + b.z, h3 = acc[3] + b.w;
  if (MODE == 2) {
    if (co0 < kCnnOut) {
      constexpr int R = geo_of(GS).R, OFF = 5 - GS;
      const int lz = 2 * o + (co0 >> 2) - OFF, ly0 = 2 * row - OFF, lx0 = 2 * xv - OFF;
      const float h[4] = {h0, h1, h2, h3};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int ly = ly0 + (k >> 1), lx = lx0 + (k & 1);
        if (lz >= 0 && lz < R && ly >= 0 && ly < R && lx >= 0 && lx < R) go.sd[(lz * R + ly) * R + lx] = go.s * h[k];
      }
    }
  } else {
    const int c0 = o * kPlaneV + (co0 >> 3) * kHalfV + ((row + 1) * S + xv + 1) * 8 + (co0 & 7);
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

// One convolution layer (in != out).  Input plane z (0..7) feeds output plane z + 1 (dz = 0), z (dz = 1) and
// z - 1 (dz = 2); output plane o accumulates its steps in the order of conv_fwd_kernel (dz = 0, 1, 2; s = 0..4
// within each) without the steps of the zero planes z = -1 and 8.  The operands of plane z + 1 are requested
// before the MFMAs of plane z, so their LDS latency hides behind those.
// Wn (not null): the next layer's weights are requested into w as the current ones die: the dz = 0 steps before
// the last input plane (which feeds no plane z + 1), the others after its MFMAs.  The next layer then does not
// start with the whole round trip, and no extra registers are needed: 4% of the kernel, measured.  (Leaving
// the reloads out of a timing copy saves 11%: the rest is the bandwidth of 15 KB of weights per wave and layer.)
template <int GS, bool CUBE, int MODE>
__device__ __forceinline__ void conv_layer(const u16* in, u16* out, bf16x8 (&w)[KZ], const float4& b,
                                           const int (&off)[5], const GuessOut& go, const u16* __restrict__ Wn) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15;
  const u16* col = in + ((wave + 4 * (i >> 3)) * S + (i & 7)) * 8;   // padded row = output row + dy
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  floatx4 a0 = zero, a1 = zero;   // output planes z - 1 and z
  Pack8 nx[5];
#pragma unroll
  for (int s = 0; s < 5; s++) nx[s].u = *reinterpret_cast<const uint4*>(col + off[s]);
#pragma unroll
  for (int z = 0; z < So; z++) {
    Pack8 a[5];
#pragma unroll
    for (int s = 0; s < 5; s++) a[s] = nx[s];
    if (z + 1 < So) {
#pragma unroll
      for (int s = 0; s < 5; s++) nx[s].u = *reinterpret_cast<const uint4*>(col + (z + 1) * kPlaneV + off[s]);
    }
    floatx4 a2 = zero;            // output plane z + 1
    if (z + 1 == So && Wn) {
#pragma unroll
      for (int c = 0; c < 5; c++) w[c] = load_weight_step(Wn, c);
    }
#pragma unroll
    for (int s = 0; s < 5; s++) {
      if (z + 1 < So) a2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[s], a[s].v, a2, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[5 + s], a[s].v, a1, 0, 0, 0);
      if (z >= 1) a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[10 + s], a[s].v, a0, 0, 0, 0);
    }
    if (z + 1 == So && Wn) {
#pragma unroll
      for (int c = 5; c < KZ; c++) w[c] = load_weight_step(Wn, c);
    }
    if (z >= 1) conv_store<GS, CUBE, MODE>(in, out, a0, b, z - 1, go);
    a0 = a1;
    a1 = a2;
  }
  conv_store<GS, CUBE, MODE>(in, out, a0, b, So - 1, go);   // its dz = 2 input is the zero plane z = 8
}

// Interpolated solve cube from a coarse cube staged in LDS (the arithmetic of interp_cube) into an unpadded
// N^3 cube.  Returns this thread's max |value| (the cube's max once reduced over the workgroup).
template <int GS>
__device__ __forceinline__ float interp_cube_lds(const float* cs, int int_scheme, float* cube) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, C = g.C, U = g.U, NC = C - 2;
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
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
      for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++) {
          const int X = X0 + cx, Y = Y0 + cy, Z = Z0 + cz;
          if (X >= 0 && X < N && Y >= 0 && Y < N && Z >= 0 && Z < N) {
            cube[(Z * N + Y) * N + X] = ch[cz][cy][cx];
            mx = fmaxf(mx, fabsf(ch[cz][cy][cx]));
          }
        }
  }
  return mx;
}

template <int GS, bool CUBE, bool KEEP>
__device__ void fused_patch(const float* __restrict__ rho, const float* __restrict__ cpot, const u16* __restrict__ WL,
                            const float* __restrict__ P, float cst, int int_scheme, float* __restrict__ guess,
                            FusedLds& L) {
  constexpr Geo g = geo_of(GS);
  constexpr int N = g.N, R = g.R, C = g.C, OFF = 5 - GS;
  constexpr int C3 = C * C * C;
  static_assert(C3 % 4 == 0, "float4 staging");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* cs = reinterpret_cast<float*>(L.X);      // coarse staging
  float* cube = reinterpret_cast<float*>(L.C);
  u16* X0 = reinterpret_cast<u16*>(L.X);
  u16* A = reinterpret_cast<u16*>(L.C);
  // the coarse cube and this thread's density column are requested together.  Thread t owns the frame-interior
  // column (fx, fy) = (t % 16, t / 16); its solve-interior column is (fx, fy) - OFF.
  for (int t = threadIdx.x; t < C3 / 4; t += kFT)
    reinterpret_cast<float4*>(cs)[t] = reinterpret_cast<const float4*>(cpot)[t];
  const int fx = threadIdx.x & 15, fy = threadIdx.x >> 4;
  const int x = fx - OFF, y = fy - OFF;
  const bool colv = x >= 0 && x < R && y >= 0 && y < R;
  float rcol[R];
#pragma unroll
  for (int z = 0; z < R; z++) rcol[z] = colv ? rho[(z * R + y) * R + x] : 0.0f;
  __syncthreads();
  float mx = interp_cube_lds<GS>(cs, int_scheme, cube);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if (lane == 0) L.red[wave] = mx;
  __syncthreads();   // the cube is complete and the coarse staging is dead
  mx = L.red[0];
#pragma unroll
  for (int k = 1; k < kFT / 64; k++) mx = fmaxf(mx, L.red[k]);
  const float s = mx > 0.0f ? mx : 1.0f;
  const float inv = 1.0f / s;
  // The whole-cube guess takes the Dirichlet layer from the interpolation (its interior is the epilogue's).
  if (CUBE) {
    constexpr int NSHELL = 2 * N * N + 4 * (N - 1) * (N - 2);
    for (int q = threadIdx.x; q < NSHELL; q += kFT) {
      int X, Y, Z;
      shell_cell<N>(q, X, Y, Z);
      const int t = (Z * N + Y) * N + X;
      guess[t] = cube[t];
    }
  }
  // Packed input on the 16^3 frame interior (the solve interior, centred; zero elsewhere): fine cell
  // (fx, fy, fz) is sub-cell ((fz&1)*2 + (fy&1))*2 + (fx&1) of voxel (fx, fy, fz)/2, i.e. half-plane fz&1 and
  // 32-bit word (fy&1)*2 + (fx&1) of the voxel; the word is (q0, q1) = (interp/s, 64 r/s) with r the SOR
  // residual of the interpolated cube.  The column owner sweeps z with the column's cube values in registers
  // and keeps the column's interpolated values for the guess.
  uint32_t* X32 = reinterpret_cast<uint32_t*>(X0);
  const int colw = (((fy >> 1) + 1) * S + (fx >> 1) + 1) * 4 + (fy & 1) * 2 + (fx & 1);
  auto word = [&](int fz) -> uint32_t& { return X32[(fz >> 1) * (kPlaneV / 2) + (fz & 1) * (kHalfV / 2) + colw]; };
  zero_ring(L.X);
#pragma unroll
  for (int fz = 0; fz < OFF; fz++) word(fz) = 0u;
#pragma unroll
  for (int fz = OFF + R; fz < 16; fz++) word(fz) = 0u;
  // The column's interpolated values are kept for the guess: in registers in the uniform-batch instances (KEEP;
  // 164 VGPRs for the whole-cube guess, 168 for the interior guess), parked in the guess array in the
  // mixed-batch instance, whose five ghost sizes would spill.
  float icol[R];
  float* gcol = guess + (CUBE ? N * N + (y + 1) * N + (x + 1) : y * R + x);   // interior column, z = 0
  if (colv) {
    const float* cc = cube + (y + 1) * N + (x + 1);   // cube column, stride N^2
    float cm = cc[0], c = cc[N * N];
#pragma unroll
    for (int z = 0; z < R; z++) {
      const float* cz = cc + (z + 1) * N * N;
      const float cp = cz[N * N];
      const float r = cp + cm + cz[N] + cz[-N] + cz[1] + cz[-1] - 6.0f * c - cst * rcol[z];
      word(z + OFF) = pack2_bf16(c * inv, (r * inv) * kResScale);
      if (KEEP)
        icol[z] = c;
      else
        gcol[z * R * R] = c;
      cm = c;
      c = cp;
    }
  } else {
#pragma unroll
    for (int z = 0; z < R; z++) word(z + OFF) = 0u;
  }
  bf16x8 w[KZ];
  float4 b;
  load_layer(WL, P, 0, w, b);
  __syncthreads();   // the packed input is complete; the cube is dead
  zero_ring(L.C);
  int off[5];
  pair_offsets(off);
  __syncthreads();
  GuessOut go{s, reinterpret_cast<float*>(L.X)};   // X is free during the output layer (layer 5 read it)
  // each layer requests the next one's weights into w as it finishes; the bias follows it (first used after
  // the next layer's first plane of MFMAs)
  auto wnext = [&](int l) { return WL + (size_t)l * 16 * KF; };
  conv_layer<GS, CUBE, 0>(X0, A, w, b, off, go, wnext(1));   // layer 1: packed input -> A
  load_bias(P, 1, b);
  __syncthreads();   // every wave is done reading the packed input
  u16* Bb = X0;
  conv_layer<GS, CUBE, 1>(A, Bb, w, b, off, go, wnext(2));   // layers 2-5: A -> B -> A -> B -> A
  load_bias(P, 2, b);
  __syncthreads();
  conv_layer<GS, CUBE, 1>(Bb, A, w, b, off, go, wnext(3));
  load_bias(P, 3, b);
  __syncthreads();
  conv_layer<GS, CUBE, 1>(A, Bb, w, b, off, go, wnext(4));
  load_bias(P, 4, b);
  __syncthreads();
  conv_layer<GS, CUBE, 1>(Bb, A, w, b, off, go, wnext(5));
  load_bias(P, 5, b);
  __syncthreads();
  conv_layer<GS, CUBE, 2>(A, nullptr, w, b, off, go, nullptr);   // layer 6: s * D into X
  __syncthreads();
  // the guess, interp + s * D, by interior columns (whole rows per wave)
  if (colv) {
    const float* sd = reinterpret_cast<const float*>(L.X) + y * R + x;
#pragma unroll
    for (int z = 0; z < R; z++) {
      constexpr int ZS = CUBE ? N * N : R * R;
      const float base = KEEP ? icol[z] : gcol[z * ZS];
      gcol[z * ZS] = base + sd[z * R * R];
    }
  }
}

template <bool CUBE, bool KEEP>
__device__ __forceinline__ void fused_dispatch(int gs, const float* r, const float* c, const u16* WL, const float* P,
                                               float cst, int int_scheme, float* out, FusedLds& L) {
  switch (gs) {
    case 1: fused_patch<1, CUBE, KEEP>(r, c, WL, P, cst, int_scheme, out, L); break;
    case 2: fused_patch<2, CUBE, KEEP>(r, c, WL, P, cst, int_scheme, out, L); break;
    case 3: fused_patch<3, CUBE, KEEP>(r, c, WL, P, cst, int_scheme, out, L); break;
    case 4: fused_patch<4, CUBE, KEEP>(r, c, WL, P, cst, int_scheme, out, L); break;
    default: fused_patch<5, CUBE, KEEP>(r, c, WL, P, cst, int_scheme, out, L); break;
  }
}

// CUBE: the guess is the whole solve cube [n][N^3] (Dirichlet layer included), else the interior [n][R^3].
template <bool CUBE>
__global__ void __launch_bounds__(kFT, kFusedBlocksPerCU) fused_guess_kernel(const float* __restrict__ rho,
                                                             const float* __restrict__ cpot,
                                                             const u16* __restrict__ WL, const float* __restrict__ P,
                                                             int npatch, int gs, float cst, int int_scheme,
                                                             float* __restrict__ guess) {
  __shared__ FusedLds L;
  const int p = blockIdx.x;
  if (p >= npatch) return;
  const Geo g = geo_of(gs);
  const float* r = rho + (size_t)p * g.R * g.R * g.R;
  const float* c = cpot + (size_t)p * g.C * g.C * g.C;
  float* out = guess + (size_t)p * (CUBE ? g.N * g.N * g.N : g.R * g.R * g.R);
  fused_dispatch<CUBE, true>(gs, r, c, WL, P, cst, int_scheme, out, L);
}

// Mixed ghost sizes: patch p reads its ghost size and packed offsets from the plan's descriptors; the
// guess is packed like the density (offset rho_off).
__global__ void __launch_bounds__(kFT, kFusedBlocksPerCU) fused_guess_mixed_kernel(const float* __restrict__ rho,
                                                                   const float* __restrict__ cpot,
                                                                   const PatchDesc* __restrict__ desc,
                                                                   const u16* __restrict__ WL,
                                                                   const float* __restrict__ P, int npatch, float cst,
                                                                   int int_scheme, float* __restrict__ guess) {
  __shared__ FusedLds L;
  const int p = blockIdx.x;
  if (p >= npatch) return;
  const PatchDesc d = desc[p];
  fused_dispatch<false, false>(d.gs, rho + d.rho_off, cpot + d.cpot_off, WL, P, cst, int_scheme, guess + d.rho_off, L);
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

hipError_t cnn_weight_copies(const float* P, uint16_t* Wb, uint16_t* WT, uint16_t* WL, hipStream_t s) {
  const int n = kCnnLayers * 16 * KF;   // covers every copy (KF > K)
  hipLaunchKernelGGL(weight_copies_kernel, dim3((n + 255) / 256), dim3(256), 0, s, P, Wb, WT, WL);
  return hipGetLastError();
}

hipError_t cnn_prep(const float* rho, const float* cpot, int n, int gs, float cst, int int_scheme, uint16_t* x,
                    float* scale, hipStream_t s) {
  const Geo g = geo_of(gs);
  hipLaunchKernelGGL(prep_kernel, dim3(n), dim3(kThreads), sizeof(float) * g.N * g.N * g.N, s, rho, cpot, gs, cst,
                     int_scheme, x, scale);
  return hipGetLastError();
}

hipError_t cnn_fused_guess(const float* rho, const float* cpot, const uint16_t* WL, const float* P, int n, int gs,
                           float cst, int int_scheme, float* guess, int cube, hipStream_t s) {
  if (cube)
    hipLaunchKernelGGL(fused_guess_kernel<true>, dim3(n), dim3(kFT), 0, s, rho, cpot, WL, P, n, gs, cst, int_scheme,
                       guess);
  else
    hipLaunchKernelGGL(fused_guess_kernel<false>, dim3(n), dim3(kFT), 0, s, rho, cpot, WL, P, n, gs, cst, int_scheme,
                       guess);
  return hipGetLastError();
}

hipError_t cnn_fused_guess_mixed(const float* rho, const float* cpot, const PatchDesc* desc, const uint16_t* WL,
                                 const float* P, int n, float cst, int int_scheme, float* guess, hipStream_t s) {
  hipLaunchKernelGGL(fused_guess_mixed_kernel, dim3(n), dim3(kFT), 0, s, rho, cpot, desc, WL, P, n, cst,
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
