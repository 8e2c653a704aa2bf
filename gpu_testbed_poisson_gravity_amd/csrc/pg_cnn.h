// Launchers and sizes of the warm-start CNN kernels (pg_cnn.hip).  BF16 tensors are passed as uint16_t bit
// patterns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pgamd {

constexpr int kCnnEdge = 8;                                  // edge of every layer's output (coarse 2h grid)
constexpr int kCnnPadded = kCnnEdge + 2;                     // padded input edge
constexpr int kCnnChannels = 16;
constexpr int kCnnOut = 8;                                   // output channels = 2^3 sub-cells
constexpr int kCnnLayers = 6;
constexpr int kCnnK = 27 * kCnnChannels;                     // 432
constexpr int kCnnLayerParams = kCnnChannels * kCnnK + kCnnChannels;   // W [16][432] + b [16]
// Forward-pass BF16 weight layout: per output channel 15 K steps of 32 = 3 dz x 5 pairs of (dy, dx)
// positions (the 10th position is a zero pad) x 16 input channels
constexpr int kCnnKF = 3 * 5 * 2 * kCnnChannels;              // 480
constexpr int kCnnParams = kCnnLayers * kCnnLayerParams;     // 41568
constexpr int kCnnVox = kCnnEdge * kCnnEdge * kCnnEdge;      // 512
constexpr int kCnnPadVox = kCnnPadded * kCnnPadded * kCnnPadded;   // 1000
constexpr int kCnnInput = kCnnPadVox * kCnnChannels;         // bf16 values per packed input
constexpr int kCnnOutputs = kCnnVox * kCnnOut;               // fp32 values per output

// mode 0: first layer, 1: residual layer, 2: linear output layer (D)
hipError_t cnn_conv_fwd(const uint16_t* X, const uint16_t* W, const float* b, uint16_t* Y, uint16_t* H, float* D,
                        int n, int mode, hipStream_t s);
hipError_t cnn_mse(const float* D, const float* T, float* dY, float* partial, int nvox, int nblocks, hipStream_t s);
// Adds the residual term of the loss (weight w) to dY and to the nblocks loss partials cnn_mse wrote.
hipError_t cnn_res_loss(const float* D, const float* T, float* dY, float* partial, int n, float w, int nblocks,
                        hipStream_t s);
hipError_t cnn_relu_bwd(const float* dY, const uint16_t* H, uint16_t* dH, float* bpart, size_t count, int nblocks,
                        hipStream_t s);
hipError_t cnn_conv_bwd_data(const uint16_t* dH, const uint16_t* WT, const float* dYres, float* dX, int n,
                             hipStream_t s);
int cnn_bwd_weight_blocks(int n);
hipError_t cnn_conv_bwd_weight(const uint16_t* X, const uint16_t* dH, float* wpart, int n, hipStream_t s);
hipError_t cnn_reduce_parts(const float* part, int parts, int len, float* dst, hipStream_t s);
hipError_t cnn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t s);
// Wb: [6][16][kCnnKF] forward layout; WT: [6][16][432] per-tap transposed (input-gradient pass); WL: the forward
// layout lane-major, [6][15 steps][64 lanes][8], what lane l reads at step c (the fused kernel's A operand)
hipError_t cnn_weight_copies(const float* P, uint16_t* Wb, uint16_t* WT, uint16_t* WL, hipStream_t s);
hipError_t cnn_prep(const float* rho, const float* cpot, int n, int gs, float cst, int int_scheme, uint16_t* x,
                    float* scale, hipStream_t s);
// Interpolation, packing, the six layers and the guess in one launch (activations in LDS).  cube = 0: guess of
// the interior [n][R^3]; cube = 1: the whole solve cube [n][N^3], the Dirichlet layer being the interpolation.
hipError_t cnn_fused_guess(const float* rho, const float* cpot, const uint16_t* WL, const float* P, int n, int gs,
                           float cst, int int_scheme, float* guess, int cube, hipStream_t s);
struct PatchDesc;
hipError_t cnn_fused_guess_mixed(const float* rho, const float* cpot, const PatchDesc* desc, const uint16_t* WL,
                                 const float* P, int n, float cst, int int_scheme, float* guess, hipStream_t s);
hipError_t cnn_guess(const float* cpot, const float* D, const float* scale, int n, int gs, int int_scheme,
                     float* guess, hipStream_t s);

}  // namespace pgamd
