// Coarse-to-fine potential interpolation of one coarse cell (device code shared by the SOR and MG
// kernels).  Restates CPU_PoissonSolver_SOR.cpp:110-203 in FP32 with the reference's operation order.
#pragma once

#include <hip/hip_runtime.h>

#include "pg_geometry.h"

#pragma clang fp contract(off)

namespace pgamd {

// One coarse cell's 8 children: CQUAD (13 signed slopes summed in slope order, then + parent) or QUAD
// (27-point tensor product of the weights {5,30,-3}/32 and {-3,30,5}/32).  Same operation order as the
// reference so that FP32 results are identical.
__device__ __forceinline__ void interp_children(const float (&v)[3][3][3], int int_scheme, float (&ch)[2][2][2]) {
  if (int_scheme == kIntCQuad) {
    const float c8 = 0.125f, c64 = 0.015625f, c512 = 0.001953125f;
    float S[13];
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
          float a = sg[0] > 0 ? S[0] : -S[0];
#pragma unroll
          for (int q = 1; q < 13; q++) a = sg[q] > 0 ? a + S[q] : a - S[q];
          ch[cz][cy][cx] = a + v[1][1][1];
        }
  } else {
    const float Mp[3] = {-0.09375f, 0.9375f, 0.15625f};
    const float Mm[3] = {0.15625f, 0.9375f, -0.09375f};
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
      for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++) {
          float a = 0.0f;
#pragma unroll
          for (int dk = 0; dk < 3; dk++)
#pragma unroll
            for (int dj = 0; dj < 3; dj++)
#pragma unroll
              for (int di = 0; di < 3; di++) {
                const float wz = cz ? Mp[dk] : Mm[dk];
                const float wy = cy ? Mp[dj] : Mm[dj];
                const float wx = cx ? Mp[di] : Mm[di];
                a = a + v[dk][dj][di] * wz * wy * wx;
              }
          ch[cz][cy][cx] = a;
        }
  }
}

}  // namespace pgamd
