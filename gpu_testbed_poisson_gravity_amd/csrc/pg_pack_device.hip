// Device-resident patch-group preparation and scatter: the same gather and write-back as pg_pack.cpp, for AMR
// levels kept in HBM.  With the level resident on the device (288 GB of HBM holds levels of hundreds of millions
// of cells), a time step moves nothing over PCIe: gather kernel -> solve -> scatter kernel on one stream.
// No reference counterpart beyond the host step it mirrors (GPU_API/CUAPI_Asyn_PoissonGravitySolver.cu:37-41,
// CPU_Solver/CPU_PoissonSolver_SOR.cpp:5-7, 70-76 mark where GAMER prepares the arrays).
//
// One workgroup (256 lanes) per batch patch.  Every output element is one lane's store, consecutive lanes writing
// consecutive addresses; the gathered reads are x rows of at most three source patches each, so a wavefront reads
// a few contiguous runs.  Reuse between neighbouring patches (each fine density cell is read by its own patch and
// by up to 26 ghost regions, each coarse cell by up to 27 patches) is left to L2 and the Infinity Cache.
// HBM-bound: per patch at ghost size 5 the gather writes (16^3 + 12^3 + 5*8^3) * 4 = 33,536 bytes.

#include <hip/hip_runtime.h>

#include "../../include/pg_amd.h"
#include "pg_geometry.h"
#include "pg_internal.h"

namespace pgamd {
namespace {

constexpr int kPackThreads = 256;
constexpr int kP = kPatchSize;
constexpr int kP3 = kP * kP * kP;

__device__ __forceinline__ void split_dev(int u, int& d, int& w) {
  d = u < 0 ? -1 : (u >= kP ? 1 : 0);
  w = u - kP * d;
}

__device__ __forceinline__ int nb_dev(int dz, int dy, int dx) { return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); }

template <int GS>
__global__ void __launch_bounds__(kPackThreads) pack_groups_kernel(
    int npatch, const int* __restrict__ group_first, const int* __restrict__ group_father, int nfine,
    const float* __restrict__ fine_flu, const int* __restrict__ fine_sib, int ncoarse,
    const float* __restrict__ coarse_pot, const int* __restrict__ coarse_sib, float* __restrict__ rho,
    float* __restrict__ pot_in, float* __restrict__ flu, int* __restrict__ status) {
  constexpr Geo geo = geo_of(GS);
  constexpr int R = geo.R, C = geo.C, R3 = R * R * R, C3 = C * C * C;
  constexpr int gf = GS - 1, gc = (C - 4) / 2;
  __shared__ int sib[2][27];
  const int p = blockIdx.x, t = threadIdx.x;
  if (p >= npatch) return;
  const int g = p >> 3, l = p & 7;
  const int lx = l & 1, ly = (l >> 1) & 1, lz = l >> 2;
  const int first = group_first[g], F = group_father[g];
  const int P = first + l;
  const bool ok = first >= 0 && first <= nfine - 8 && F >= 0 && F < ncoarse;
  if (t < 54) {
    const int which = t / 27, k = t - 27 * which;
    sib[which][k] = ok ? (which == 0 ? fine_sib[(size_t)P * 27 + k] : coarse_sib[(size_t)F * 27 + k]) : -1;
  }
  __syncthreads();
  int miss_f = 0, miss_c = 0, bad = ok ? 0 : 1;
  float* r = rho + (size_t)p * R3;
  for (int i = t; i < R3; i += kPackThreads) {
    const int x = i % R, y = (i / R) % R, z = i / (R * R);
    int dz, dy, dx, wz, wy, wx;
    split_dev(z - gf, dz, wz);
    split_dev(y - gf, dy, wy);
    split_dev(x - gf, dx, wx);
    const int Q = sib[0][nb_dev(dz, dy, dx)];
    float v = 0.0f;
    if (!ok) {
    } else if (Q < 0) {
      miss_f = 1;
    } else if (Q >= nfine) {
      bad = 1;
    } else {
      v = fine_flu[(size_t)Q * kNComp * kP3 + (wz * kP + wy) * kP + wx];
    }
    r[i] = v;
  }
  float* c = pot_in + (size_t)p * C3;
  for (int i = t; i < C3; i += kPackThreads) {
    const int x = i % C, y = (i / C) % C, z = i / (C * C);
    int dz, dy, dx, wz, wy, wx;
    split_dev(4 * lz + z - gc, dz, wz);
    split_dev(4 * ly + y - gc, dy, wy);
    split_dev(4 * lx + x - gc, dx, wx);
    const int Q = sib[1][nb_dev(dz, dy, dx)];
    float v = 0.0f;
    if (!ok) {
    } else if (Q < 0) {
      miss_c = 1;
    } else if (Q >= ncoarse) {
      bad = 1;
    } else {
      v = coarse_pot[(size_t)Q * kP3 + (wz * kP + wy) * kP + wx];
    }
    c[i] = v;
  }
  if (flu) {
    float4* o = reinterpret_cast<float4*>(flu + (size_t)p * kNComp * kP3);
    if (ok) {
      const float4* s = reinterpret_cast<const float4*>(fine_flu + (size_t)P * kNComp * kP3);
      for (int i = t; i < kNComp * kP3 / 4; i += kPackThreads) o[i] = s[i];
    } else {
      for (int i = t; i < kNComp * kP3 / 4; i += kPackThreads) o[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  const int any_f = __syncthreads_or(miss_f), any_c = __syncthreads_or(miss_c), any_b = __syncthreads_or(bad);
  if (status && t == 0)
    status[p] = (any_f ? PGAMD_PACK_MISSING_FINE : 0) | (any_c ? PGAMD_PACK_MISSING_COARSE : 0) |
                (any_b ? PGAMD_PACK_BAD_INDEX : 0);
}

__global__ void __launch_bounds__(kPackThreads) scatter_groups_kernel(
    int npatch, int gg, const int* __restrict__ group_first, const float* __restrict__ pot_out,
    const float* __restrict__ flu, int nfine, float* __restrict__ fine_pot, float* __restrict__ fine_flu) {
  const int p = blockIdx.x, t = threadIdx.x;
  if (p >= npatch) return;
  const int first = group_first[p >> 3];
  if (first < 0 || first > nfine - 8) return;   // flagged by the gather; nothing is written
  const int P = first + (p & 7);
  if (fine_pot) {
    const int G = kP + 2 * gg;
    const float* w = pot_out + (size_t)p * G * G * G;
    for (int i = t; i < kP3; i += kPackThreads) {
      const int x = i & 7, y = (i >> 3) & 7, z = i >> 6;
      fine_pot[(size_t)P * kP3 + i] = w[((z + gg) * G + (y + gg)) * G + (x + gg)];
    }
  }
  if (fine_flu) {
    float4* o = reinterpret_cast<float4*>(fine_flu + (size_t)P * kNComp * kP3);
    const float4* s = reinterpret_cast<const float4*>(flu + (size_t)p * kNComp * kP3);
    for (int i = t; i < kNComp * kP3 / 4; i += kPackThreads) o[i] = s[i];
  }
}

template <int GS>
void launch_pack(int npatch, const int* gfirst, const int* gfather, int nfine, const float* fine_flu,
                 const int* fine_sib, int ncoarse, const float* coarse_pot, const int* coarse_sib, float* rho,
                 float* pot_in, float* flu, int* status, hipStream_t s) {
  hipLaunchKernelGGL(pack_groups_kernel<GS>, dim3(npatch), dim3(kPackThreads), 0, s, npatch, gfirst, gfather, nfine,
                     fine_flu, fine_sib, ncoarse, coarse_pot, coarse_sib, rho, pot_in, flu, status);
}

}  // namespace
}  // namespace pgamd

extern "C" int pgamd_prepare_patch_groups_device(int pot_ghost_size, int ngroup, const int* d_group_first,
                                                 const int* d_group_father, int nfine, const float* d_fine_flu,
                                                 const int* d_fine_sib, int ncoarse, const float* d_coarse_pot,
                                                 const int* d_coarse_sib, float* d_rho, float* d_pot_in,
                                                 float* d_flu, int* d_status, void* stream) {
  using namespace pgamd;
  if (pot_ghost_size < 1 || pot_ghost_size > kMaxGhost) return set_error(PGAMD_ERR_ARG, "POT_GHOST_SIZE outside [1,5]");
  if (ngroup < 0 || nfine < 0 || ncoarse < 0) return set_error(PGAMD_ERR_ARG, "negative count");
  if (ngroup == 0) return PGAMD_OK;
  if (ngroup > (1 << 27)) return set_error(PGAMD_ERR_ARG, "too many patch groups for one launch");
  if (!d_group_first || !d_group_father || !d_fine_flu || !d_fine_sib || !d_coarse_pot || !d_coarse_sib || !d_rho ||
      !d_pot_in)
    return set_error(PGAMD_ERR_ARG, "null array");
  if ((reinterpret_cast<uintptr_t>(d_fine_flu) | reinterpret_cast<uintptr_t>(d_flu)) & 15)
    return set_error(PGAMD_ERR_ARG, "fluid arrays must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int n = 8 * ngroup;
  switch (pot_ghost_size) {
    case 1: launch_pack<1>(n, d_group_first, d_group_father, nfine, d_fine_flu, d_fine_sib, ncoarse, d_coarse_pot,
                           d_coarse_sib, d_rho, d_pot_in, d_flu, d_status, s); break;
    case 2: launch_pack<2>(n, d_group_first, d_group_father, nfine, d_fine_flu, d_fine_sib, ncoarse, d_coarse_pot,
                           d_coarse_sib, d_rho, d_pot_in, d_flu, d_status, s); break;
    case 3: launch_pack<3>(n, d_group_first, d_group_father, nfine, d_fine_flu, d_fine_sib, ncoarse, d_coarse_pot,
                           d_coarse_sib, d_rho, d_pot_in, d_flu, d_status, s); break;
    case 4: launch_pack<4>(n, d_group_first, d_group_father, nfine, d_fine_flu, d_fine_sib, ncoarse, d_coarse_pot,
                           d_coarse_sib, d_rho, d_pot_in, d_flu, d_status, s); break;
    default: launch_pack<5>(n, d_group_first, d_group_father, nfine, d_fine_flu, d_fine_sib, ncoarse, d_coarse_pot,
                            d_coarse_sib, d_rho, d_pot_in, d_flu, d_status, s); break;
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PGAMD_OK : set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
}

extern "C" int pgamd_scatter_patch_groups_device(int gra_ghost_size, int ngroup, const int* d_group_first,
                                                 const float* d_pot_out, const float* d_flu, int nfine,
                                                 float* d_fine_pot, float* d_fine_flu, void* stream) {
  using namespace pgamd;
  if (gra_ghost_size < 1 || gra_ghost_size > kMaxGhost) return set_error(PGAMD_ERR_ARG, "GRA_GHOST_SIZE outside [1,5]");
  if (ngroup < 0 || nfine < 0) return set_error(PGAMD_ERR_ARG, "negative count");
  if (ngroup == 0) return PGAMD_OK;
  if (ngroup > (1 << 27)) return set_error(PGAMD_ERR_ARG, "too many patch groups for one launch");
  if (!d_group_first || (!d_pot_out && d_fine_pot) || (!d_flu && d_fine_flu) || (!d_fine_pot && !d_fine_flu))
    return set_error(PGAMD_ERR_ARG, "null array");
  if ((reinterpret_cast<uintptr_t>(d_flu) | reinterpret_cast<uintptr_t>(d_fine_flu)) & 15)
    return set_error(PGAMD_ERR_ARG, "fluid arrays must be 16-byte aligned");
  hipLaunchKernelGGL(scatter_groups_kernel, dim3(8 * ngroup), dim3(kPackThreads), 0, (hipStream_t)stream, 8 * ngroup,
                     gra_ghost_size, d_group_first, d_pot_out, d_flu, nfine, d_fine_pot, d_fine_flu);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PGAMD_OK : set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
}
