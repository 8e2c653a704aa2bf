// Launchers of the warm-start CNN kernels (pg_cnn.hip).  BF16 tensors are passed as uint16_t bit patterns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pgamd {

hipError_t cnn_conv_fwd(const uint16_t* X, const uint16_t* W, const float* b, uint16_t* Y, uint16_t* H, int n, int S,
                        int cin, bool resid, hipStream_t s);
hipError_t cnn_conv6_fwd(const uint16_t* Y5, const float* w6, const float* b6, float* D, int n, hipStream_t s);
hipError_t cnn_mse(const float* D, const float* T, float* dD, float* partial, int count, int nblocks, hipStream_t s);
hipError_t cnn_conv6_bwd(const uint16_t* Y5, const float* w6, const float* dD, float* dY5, float* gpart, int n,
                         hipStream_t s);
hipError_t cnn_relu_bwd(const float* dY, const uint16_t* H, uint16_t* dH, float* bpart, size_t count, int nblocks,
                        hipStream_t s);
hipError_t cnn_conv_bwd_data(const uint16_t* dH, const uint16_t* WT, const float* dYres, float* dX, int n, int S,
                             hipStream_t s);
int cnn_bwd_weight_blocks(int n, int S);
hipError_t cnn_conv_bwd_weight(const uint16_t* X, const uint16_t* dH, float* wpart, int n, int S, int cin,
                               hipStream_t s);
hipError_t cnn_reduce_parts(const float* part, int parts, int len, float* dst, hipStream_t s);
hipError_t cnn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t s);
hipError_t cnn_to_bf16(const float* src, uint16_t* dst, int n, hipStream_t s);
hipError_t cnn_transpose_w(const float* W, uint16_t* WT, hipStream_t s);
hipError_t cnn_round_bf16(const float* src, float* dst, int n, hipStream_t s);
hipError_t cnn_prep(const float* rho, const float* cpot, int n, int gs, float cst, int int_scheme, uint16_t* x,
                    float* scale, hipStream_t s);
hipError_t cnn_guess(const float* cpot, const float* D, const float* scale, int n, int gs, int int_scheme,
                     float* guess, hipStream_t s);

}  // namespace pgamd
