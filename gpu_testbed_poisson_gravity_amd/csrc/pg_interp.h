// Coarse-to-fine potential interpolation of one coarse cell, and the enumeration of a solve cube's Dirichlet
// layer (device code shared by the SOR, MG and network kernels).  Restates CPU_PoissonSolver_SOR.cpp:110-203 with the reference's operation order, for the FP32
// kernels and the FP64 ones (the reference's FLOAT8 build).
#pragma once

#include <hip/hip_runtime.h>

#include "pg_geometry.h"

#pragma clang fp contract(off)

namespace pgamd {

// One coarse cell's 8 children: CQUAD (13 signed slopes summed in slope order, then + parent) or QUAD
// (27-point tensor product of the weights {5,30,-3}/32 and {-3,30,5}/32).  Same operation order as the
// reference so that results are identical in either precision (every constant is exact in both).
template <typename T>
__device__ __forceinline__ void interp_children(const T (&v)[3][3][3], int int_scheme, T (&ch)[2][2][2]) {
  if (int_scheme == kIntCQuad) {
    const T c8 = T(0.125), c64 = T(0.015625), c512 = T(0.001953125);
    T S[13];
    S[0] = c8 * (v[1][1][2] - v[1][1][0]);
    S[1] = c8 * (v[1][2][1] - v[1][0][1]);
    S[2] = c8 * (v[2][1][1] - v[0][1][1]);
    S[3] = c64 * (v[0][1][2] - v[0][1][0]);
    S[4] = c64 * (v[0][2][1] - v[0][0][1]);
    S[5] = c64 * (v[1][0][2] - v[1][0][0]);
    S[6] = c64 * (v[1][2][2] - v[1][2][0]);
    S[7] = c64 * (v[2][1][2] - v[2][1][0]);
    S[8] = c64 * (v[2][2][1] - v[2][0][1]);
    S[9] = c512 * (v[0][0][2] - v[0][0][0]);
    S[10] = c512 * (v[0][2][2] - v[0][2][0]);
    S[11] = c512 * (v[2][0][2] - v[2][0][0]);
    S[12] = c512 * (v[2][2][2] - v[2][2][0]);
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
      for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++) {
          const int sx = cx ? 1 : -1, sy = cy ? 1 : -1, sz = cz ? 1 : -1;
          const int sg[13] = {sx,      sy,      sz,           -sx * sz,      -sy * sz,      -sx * sy,    sx * sy,
                              sx * sz, sy * sz, sx * sy * sz, -sx * sy * sz, -sx * sy * sz, sx * sy * sz};
          T a = sg[0] > 0 ? S[0] : -S[0];
#pragma unroll
          for (int q = 1; q < 13; q++) a = sg[q] > 0 ? a + S[q] : a - S[q];
          ch[cz][cy][cx] = a + v[1][1][1];
        }
  } else {
    const T Mp[3] = {T(-0.09375), T(0.9375), T(0.15625)};
    const T Mm[3] = {T(0.15625), T(0.9375), T(-0.09375)};
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
      for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++) {
          T a = T(0);
#pragma unroll
          for (int dk = 0; dk < 3; dk++)
#pragma unroll
            for (int dj = 0; dj < 3; dj++)
#pragma unroll
              for (int di = 0; di < 3; di++) {
                const T wz = cz ? Mp[dk] : Mm[dk];
                const T wy = cy ? Mp[dj] : Mm[dj];
                const T wx = cx ? Mp[di] : Mm[di];
                a = a + v[dk][dj][di] * wz * wy * wx;
              }
          ch[cz][cy][cx] = a;
        }
  }
}

// Cell q of the outer layer of an N^3 cube (faces z = 0 and N-1, then the ring of each layer in between).
template <int N>
__device__ __forceinline__ void shell_cell(int q, int& x, int& y, int& z) {
  constexpr int F = N * N, RING = 4 * (N - 1);
  if (q < 2 * F) {
    z = q < F ? 0 : N - 1;
    const int t = q % F;
    x = t % N;
    y = t / N;
    return;
  }
  const int r = q - 2 * F, t = r % RING;
  z = r / RING + 1;
  if (t < 2 * N) {
    y = t < N ? 0 : N - 1;
    x = t % N;
  } else {
    const int u = t - 2 * N;
    x = u < N - 2 ? 0 : N - 1;
    y = u % (N - 2) + 1;
  }
}

}  // namespace pgamd
