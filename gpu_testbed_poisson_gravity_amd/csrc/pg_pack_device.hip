// Device-resident patch-group preparation and scatter: the same gather and write-back as pg_pack.cpp, for AMR
// levels kept in HBM.  With the level resident on the device (288 GB of HBM holds levels of hundreds of millions
// of cells), a time step moves nothing over PCIe: gather kernel -> solve -> scatter kernel on one stream.
// No reference counterpart beyond the host step it mirrors (GPU_API/CUAPI_Asyn_PoissonGravitySolver.cu:37-41,
// CPU_Solver/CPU_PoissonSolver_SOR.cpp:5-7, 70-76 mark where GAMER prepares the arrays).
//
// Gather: one workgroup (1024 lanes) per patch group.  The group's density block and coarse block are staged
// in LDS from the patch slots around the group, so each source cell is fetched once per group (not once per
// batch patch that reads it), and the 8 batch patches are cut from LDS with consecutive lanes writing
// consecutive addresses.  HBM-bound: per patch at ghost size 5 the gather writes (16^3 + 12^3 + 5*8^3) * 4 =
// 33,536 bytes.  Scatter: one workgroup (256 lanes) per batch patch.

#include <hip/hip_runtime.h>

#include "../../include/pg_amd.h"
#include "pg_geometry.h"
#include "pg_internal.h"

namespace pgamd {
namespace {

constexpr int kPackThreads = 1024;    // gather: one workgroup per patch group
constexpr int kScatterThreads = 256;  // scatter: one workgroup per batch patch
constexpr int kP = kPatchSize;
constexpr int kP3 = kP * kP * kP;

__device__ __forceinline__ void split_dev(int u, int& d, int& w) {
  d = u < 0 ? -1 : (u >= kP ? 1 : 0);
  w = u - kP * d;
}

__device__ __forceinline__ int nb_dev(int dz, int dy, int dx) { return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); }

// Batch patch p = 8g + l gathered from its own neighbour table `sib` (27 fine entries) and the father's table
// `csib` (27 coarse entries): the general path, for groups whose tables are not consistent or hold an index
// outside the level.  Every thread of the workgroup takes part (the status is a workgroup reduction).
template <int GS>
__device__ void pack_patch(int p, int l, int first, bool ok, const int* sib, const int* csib, int nfine,
                           const float* __restrict__ fine_flu, int ncoarse, const float* __restrict__ coarse_pot,
                           float* __restrict__ rho, float* __restrict__ pot_in, float* __restrict__ flu,
                           int* __restrict__ status) {
  constexpr Geo geo = geo_of(GS);
  constexpr int R = geo.R, C = geo.C, R3 = R * R * R, C3 = C * C * C;
  constexpr int gf = GS - 1, gc = (C - 4) / 2;
  const int t = threadIdx.x;
  const int lx = l & 1, ly = (l >> 1) & 1, lz = l >> 2;
  const int P = first + l;
  int miss_f = 0, miss_c = 0, bad = ok ? 0 : 1;
  float* r = rho + (size_t)p * R3;
  for (int i = t; i < R3; i += kPackThreads) {
    const int x = i % R, y = (i / R) % R, z = i / (R * R);
    int dz, dy, dx, wz, wy, wx;
    split_dev(z - gf, dz, wz);
    split_dev(y - gf, dy, wy);
    split_dev(x - gf, dx, wx);
    const int Q = ok ? sib[nb_dev(dz, dy, dx)] : -1;
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
    const int Q = ok ? csib[nb_dev(dz, dy, dx)] : -1;
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

// Slot s in -1..2 of a coordinate u relative to the group's origin (-8 <= u < 24), and u's cell inside it.
__device__ __forceinline__ void slot_dev(int u, int& s, int& w) {
  s = ((u + kP) >> 3) - 1;
  w = u - kP * s;
}

template <int GS>
struct PackGeo {
  static constexpr Geo g = geo_of(GS);
  static constexpr int R = g.R, C = g.C, gf = GS - 1, gc = (C - 4) / 2;
  static constexpr int B = 2 * kP + 2 * gf;   // fine block edge: the group's 16^3 cells and the ghost shell
  static constexpr int CB = kP + 2 * gc;      // coarse block edge: the father's 8^3 cells and the shell
};

// One workgroup per patch group.  The group's density block (its 16^3 cells and the GS-1 ghost cells around
// them) and its coarse block (the father's 8^3 cells and (C-4)/2 around them) are read once into LDS, from the
// 4^3 fine and 3^3 coarse patch slots around the group, and the 8 batch patches are cut from them: every source
// cell is fetched once per group instead of once per patch that uses it.  The slots come from one
// representative patch's neighbour table; a group whose tables disagree (or hold an index outside the level)
// takes the per-patch path with each patch's own table, so the result always equals the host gather's.
template <int GS>
__global__ void __launch_bounds__(kPackThreads) pack_groups_kernel(
    int ngroup, const int* __restrict__ group_first, const int* __restrict__ group_father, int nfine,
    const float* __restrict__ fine_flu, const int* __restrict__ fine_sib, int ncoarse,
    const float* __restrict__ coarse_pot, const int* __restrict__ coarse_sib, float* __restrict__ rho,
    float* __restrict__ pot_in, float* __restrict__ flu, int* __restrict__ status) {
  using G = PackGeo<GS>;
  constexpr int R = G::R, C = G::C, gf = G::gf, gc = G::gc, B = G::B, CB = G::CB;
  constexpr int R3 = R * R * R, C3 = C * C * C;
  __shared__ float fblk[B * B * B];
  __shared__ float cblk[CB * CB * CB];
  __shared__ int sib8[8 * 27];   // each patch's fine table
  __shared__ int fslot[64];      // fine patch in group slot (sx, sy, sz), index (sz+1)*16 + (sy+1)*4 + (sx+1)
  __shared__ int cslot[27];      // the father's table
  const int g = blockIdx.x, t = threadIdx.x;
  if (g >= ngroup) return;
  const int first = group_first[g], F = group_father[g];
  const bool ok = first >= 0 && first <= nfine - 8 && F >= 0 && F < ncoarse;
  if (ok && t < 8 * 27) sib8[t] = fine_sib[(size_t)first * 27 + t];
  if (ok && t < 27) cslot[t] = coarse_sib[(size_t)F * 27 + t];
  __syncthreads();
  if (ok && t < 64) {
    const int sx = (t & 3) - 1, sy = ((t >> 2) & 3) - 1, sz = (t >> 4) - 1;
    const int lx = min(max(sx, 0), 1), ly = min(max(sy, 0), 1), lz = min(max(sz, 0), 1);
    fslot[t] = sib8[((lz * 2 + ly) * 2 + lx) * 27 + nb_dev(sz - lz, sy - ly, sx - lx)];
  }
  __syncthreads();
  // every entry a patch reads must agree with the slot table and lie inside the level (ghost size 1 reads only
  // the patch itself, entry 13); the coarse entries (all 27 are read by some patch) must lie inside the level
  int odd = ok ? 0 : 1;
  if (ok && t < 8 * 27) {
    const int l = t / 27, e = t - 27 * l;
    if (gf > 0 || e == 13) {
      const int sx = (l & 1) + e % 3 - 1, sy = ((l >> 1) & 1) + (e / 3) % 3 - 1, sz = (l >> 2) + e / 9 - 1;
      const int q = sib8[t];
      if (q != fslot[(sz + 1) * 16 + (sy + 1) * 4 + (sx + 1)] || q >= nfine) odd = 1;
    }
  }
  if (ok && t < 27 && cslot[t] >= ncoarse) odd = 1;
  if (__syncthreads_or(odd)) {
    for (int l = 0; l < 8; l++)
      pack_patch<GS>(8 * g + l, l, first, ok, sib8 + 27 * l, cslot, nfine, fine_flu, ncoarse, coarse_pot, rho,
                     pot_in, flu, status);
    return;
  }
  for (int i = t; i < B * B * B; i += kPackThreads) {
    const int bx = i % B, by = (i / B) % B, bz = i / (B * B);
    int sx, sy, sz, wx, wy, wz;
    slot_dev(bx - gf, sx, wx);
    slot_dev(by - gf, sy, wy);
    slot_dev(bz - gf, sz, wz);
    const int Q = fslot[(sz + 1) * 16 + (sy + 1) * 4 + (sx + 1)];
    fblk[i] = Q >= 0 ? fine_flu[(size_t)Q * kNComp * kP3 + (wz * kP + wy) * kP + wx] : 0.0f;
  }
  for (int i = t; i < CB * CB * CB; i += kPackThreads) {
    const int bx = i % CB, by = (i / CB) % CB, bz = i / (CB * CB);
    int sx, sy, sz, wx, wy, wz;
    slot_dev(bx - gc, sx, wx);
    slot_dev(by - gc, sy, wy);
    slot_dev(bz - gc, sz, wz);
    const int Q = cslot[nb_dev(sz, sy, sx)];
    cblk[i] = Q >= 0 ? coarse_pot[(size_t)Q * kP3 + (wz * kP + wy) * kP + wx] : 0.0f;
  }
  __syncthreads();
  for (int l = 0; l < 8; l++) {
    const int p = 8 * g + l, lx = l & 1, ly = (l >> 1) & 1, lz = l >> 2;
    float* r = rho + (size_t)p * R3;
    for (int i = t; i < R3; i += kPackThreads) {
      const int x = i % R, y = (i / R) % R, z = i / (R * R);
      r[i] = fblk[((kP * lz + z) * B + kP * ly + y) * B + kP * lx + x];
    }
    float* c = pot_in + (size_t)p * C3;
    for (int i = t; i < C3; i += kPackThreads) {
      const int x = i % C, y = (i / C) % C, z = i / (C * C);
      c[i] = cblk[((4 * lz + z) * CB + 4 * ly + y) * CB + 4 * lx + x];
    }
    if (flu) {
      float4* o = reinterpret_cast<float4*>(flu + (size_t)p * kNComp * kP3);
      const float4* s = reinterpret_cast<const float4*>(fine_flu + (size_t)(first + l) * kNComp * kP3);
      for (int i = t; i < kNComp * kP3 / 4; i += kPackThreads) o[i] = s[i];
    }
  }
  // per patch: a slot without a patch among those its cells come from (fine: the patch and, with ghost cells,
  // its 26 neighbours; coarse: per axis the slots -1..0 for the lower half of the father, 0..1 for the upper)
  if (status && t < 8) {
    const int l = t, lx = l & 1, ly = (l >> 1) & 1, lz = l >> 2;
    constexpr int e = gf > 0 ? 1 : 0;
    int mf = 0, mc = 0;
    for (int sz = lz - e; sz <= lz + e; sz++)
      for (int sy = ly - e; sy <= ly + e; sy++)
        for (int sx = lx - e; sx <= lx + e; sx++) mf |= fslot[(sz + 1) * 16 + (sy + 1) * 4 + (sx + 1)] < 0;
    for (int sz = lz - 1; sz <= lz; sz++)
      for (int sy = ly - 1; sy <= ly; sy++)
        for (int sx = lx - 1; sx <= lx; sx++) mc |= cslot[nb_dev(sz, sy, sx)] < 0;
    status[8 * g + l] = (mf ? PGAMD_PACK_MISSING_FINE : 0) | (mc ? PGAMD_PACK_MISSING_COARSE : 0);
  }
}

__global__ void __launch_bounds__(kScatterThreads) scatter_groups_kernel(
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
    for (int i = t; i < kP3; i += kScatterThreads) {
      const int x = i & 7, y = (i >> 3) & 7, z = i >> 6;
      fine_pot[(size_t)P * kP3 + i] = w[((z + gg) * G + (y + gg)) * G + (x + gg)];
    }
  }
  if (fine_flu) {
    float4* o = reinterpret_cast<float4*>(fine_flu + (size_t)P * kNComp * kP3);
    const float4* s = reinterpret_cast<const float4*>(flu + (size_t)p * kNComp * kP3);
    for (int i = t; i < kNComp * kP3 / 4; i += kScatterThreads) o[i] = s[i];
  }
}

template <int GS>
void launch_pack(int ngroup, const int* gfirst, const int* gfather, int nfine, const float* fine_flu,
                 const int* fine_sib, int ncoarse, const float* coarse_pot, const int* coarse_sib, float* rho,
                 float* pot_in, float* flu, int* status, hipStream_t s) {
  hipLaunchKernelGGL(pack_groups_kernel<GS>, dim3(ngroup), dim3(kPackThreads), 0, s, ngroup, gfirst, gfather, nfine,
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
  const int n = ngroup;
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
  hipLaunchKernelGGL(scatter_groups_kernel, dim3(8 * ngroup), dim3(kScatterThreads), 0, (hipStream_t)stream, 8 * ngroup,
                     gra_ghost_size, d_group_first, d_pot_out, d_flu, nfine, d_fine_pot, d_fine_flu);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PGAMD_OK : set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
}
