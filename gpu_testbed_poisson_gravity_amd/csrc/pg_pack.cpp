// Patch-group preparation for the batched Poisson-gravity solve, and the reverse scatter: the GAMER-side step the
// reference testbed stands in for ("REPLACE in the actual implementation": GPU_API/CUAPI_Asyn_PoissonGravitySolver.cu:
// 37-41, CPU_Solver/CPU_PoissonSolver_SOR.cpp:5-7, 70-76).  Host code, like GAMER's own preparation: it fills the
// host arrays pgamd_asyn_poisson_gravity_solver takes (the pinned buffers of pgamd_mem_allocate_poisson_gravity) and
// writes the results back into the level.
//
// Level format (include/pg_amd.h, "patch groups"):
//   * fine level: npatch patches of 8^3 cells, fluid [npatch][5][8][8][8] (rho, px, py, pz, E), potential
//     [npatch][8][8][8]; patch group g is the 8 consecutive patches group_first[g] + l, l = (lz*2 + ly)*2 + lx,
//     the children of coarse patch group_father[g] at offset (lx, ly, lz) * 8 fine cells;
//   * coarse level: the potential [ncoarse][8][8][8] at spacing 2h;
//   * neighbours: sib[p][27], the patch at offset (dz, dy, dx) in {-1,0,1}^3 from p, entry (dz+1)*9 + (dy+1)*3 +
//     (dx+1) (entry 13 is p itself), -1 where there is none.
// Packed layout per patch (the solver's): rho [R^3] = the fluid density of the solve interior (the patch and GS-1
// cells of its neighbours), pot_in [C^3] = the coarse potential around it (the patch's 4^3 coarse cells and
// (C-4)/2 more on each side), flu [5][8^3] = the patch's fluid.

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/pg_amd.h"
#include "pg_geometry.h"
#include "pg_internal.h"

using namespace pgamd;

namespace {

constexpr int kP = kPatchSize;           // 8
constexpr int kP3 = kP * kP * kP;

inline int nb_index(int dz, int dy, int dx) { return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); }

// Split a coordinate relative to a patch's origin into (neighbour offset, coordinate inside that patch);
// valid for -8 <= u < 16.
inline void split(int u, int& d, int& w) {
  d = u < 0 ? -1 : (u >= kP ? 1 : 0);
  w = u - kP * d;
}

// One x row of a packed array: n cells starting at x = u0 relative to the patch origin (-8 <= u0, u0 + n <= 16),
// taken in up to three runs from the patches nb[0], nb[1], nb[2] (offsets dx = -1, 0, 1); row_of(Q) is the start
// of the same (z, y) row in patch Q.  Runs without a patch are zero-filled and return `missing`; an index at or
// beyond `count` zero-fills its run, returns PGAMD_PACK_BAD_INDEX and marks the group bad.
template <typename RowOf>
int gather_row(float* dst, int u0, int n, const int* nb, int count, int& bad, int missing, RowOf&& row_of) {
  int flag = 0;
  for (int dx = -1; dx <= 1; dx++) {
    const int lo = std::max(u0, kP * dx), hi = std::min(u0 + n, kP * (dx + 1));
    if (lo >= hi) continue;
    float* out = dst + (lo - u0);
    const int Q = nb[dx + 1];
    if (Q < 0) {
      std::fill(out, out + (hi - lo), 0.0f);
      flag |= missing;
    } else if (Q >= count) {
      std::fill(out, out + (hi - lo), 0.0f);
      flag |= PGAMD_PACK_BAD_INDEX;
      bad = 1;
    } else {
      std::memcpy(out, row_of(Q) + (lo - kP * dx), sizeof(float) * (hi - lo));
    }
  }
  return flag;
}

template <typename F>
void parallel_for(int n, int nthreads, F&& f) {
  const int nt = std::max(1, std::min(nthreads, n));
  if (nt == 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++)
    pool.emplace_back([&, t] {
      for (int i = t; i < n; i += nt) f(i);
    });
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

int pgamd_prepare_patch_groups(int pot_ghost_size, int ngroup, const int* group_first, const int* group_father,
                               int nfine, const float* fine_flu, const int* fine_sib, int ncoarse,
                               const float* coarse_pot, const int* coarse_sib, float* h_rho, float* h_pot_in,
                               float* h_flu, int* status, int nthreads) {
  if (pot_ghost_size < 1 || pot_ghost_size > kMaxGhost) return set_error(PGAMD_ERR_ARG, "POT_GHOST_SIZE outside [1,5]");
  if (ngroup < 0 || nfine < 0 || ncoarse < 0) return set_error(PGAMD_ERR_ARG, "negative count");
  if (ngroup == 0) return PGAMD_OK;
  if (!group_first || !group_father || !fine_flu || !fine_sib || !coarse_pot || !coarse_sib || !h_rho || !h_pot_in ||
      !h_flu)
    return set_error(PGAMD_ERR_ARG, "null array");
  for (int g = 0; g < ngroup; g++)
    if (group_first[g] < 0 || group_first[g] + 8 > nfine || group_father[g] < 0 || group_father[g] >= ncoarse)
      return set_error(PGAMD_ERR_ARG, "patch group refers to a patch outside the level");
  const Geo geo = geo_of(pot_ghost_size);
  const int R = geo.R, C = geo.C, R3 = R * R * R, C3 = C * C * C;
  const int gf = pot_ghost_size - 1;   // fine ghost cells of the density on each side
  const int gc = (C - 4) / 2;          // coarse cells around the patch's own 4^3 (the solve cube plus one layer)
  std::vector<int> bad(ngroup, 0);     // group g met a neighbour index beyond the level (reported as an error)
  parallel_for(ngroup, nthreads, [&](int g) {
    const int F = group_father[g];
    for (int l = 0; l < 8; l++) {
      const int P = group_first[g] + l, p = 8 * g + l;
      const int lx = l & 1, ly = (l >> 1) & 1, lz = l >> 2;
      int st = 0;
      // density of the solve interior: fine cells -gf .. 8+gf-1 of the patch, from the patch and its siblings
      float* rho = h_rho + (size_t)p * R3;
      for (int z = 0; z < R; z++)
        for (int y = 0; y < R; y++) {
          int dz, dy, wz, wy;
          split(z - gf, dz, wz);
          split(y - gf, dy, wy);
          const int* nb = fine_sib + (size_t)P * 27 + nb_index(dz, dy, -1);   // dx = -1, 0, 1 follow
          st |= gather_row(rho + (z * R + y) * R, -gf, R, nb, nfine, bad[g], PGAMD_PACK_MISSING_FINE,
                           [&](int Q) { return fine_flu + (size_t)Q * kNComp * kP3 + (wz * kP + wy) * kP; });
        }
      // coarse potential: coarse cells -gc .. 4+gc-1 around the patch's own, from the father and its siblings
      float* cp = h_pot_in + (size_t)p * C3;
      for (int k = 0; k < C; k++)
        for (int j = 0; j < C; j++) {
          int dz, dy, wz, wy;
          split(4 * lz + k - gc, dz, wz);
          split(4 * ly + j - gc, dy, wy);
          const int* nb = coarse_sib + (size_t)F * 27 + nb_index(dz, dy, -1);
          st |= gather_row(cp + (k * C + j) * C, 4 * lx - gc, C, nb, ncoarse, bad[g], PGAMD_PACK_MISSING_COARSE,
                           [&](int Q) { return coarse_pot + (size_t)Q * kP3 + (wz * kP + wy) * kP; });
        }
      std::memcpy(h_flu + (size_t)p * kNComp * kP3, fine_flu + (size_t)P * kNComp * kP3, sizeof(float) * kNComp * kP3);
      if (status) status[p] = st;
    }
  });
  for (int g = 0; g < ngroup; g++)
    if (bad[g]) return set_error(PGAMD_ERR_ARG, "a neighbour table entry points outside the level");
  return PGAMD_OK;
}

int pgamd_scatter_patch_groups(int gra_ghost_size, int ngroup, const int* group_first, const float* h_pot_out,
                               const float* h_flu, int nfine, float* fine_pot, float* fine_flu, int nthreads) {
  if (gra_ghost_size < 1 || gra_ghost_size > kMaxGhost) return set_error(PGAMD_ERR_ARG, "GRA_GHOST_SIZE outside [1,5]");
  if (ngroup < 0 || nfine < 0) return set_error(PGAMD_ERR_ARG, "negative count");
  if (ngroup == 0) return PGAMD_OK;
  if (!group_first || (!h_pot_out && fine_pot) || (!h_flu && fine_flu) || (!fine_pot && !fine_flu))
    return set_error(PGAMD_ERR_ARG, "null array");
  for (int g = 0; g < ngroup; g++)
    if (group_first[g] < 0 || group_first[g] + 8 > nfine)
      return set_error(PGAMD_ERR_ARG, "patch group refers to a patch outside the level");
  const int G = gra_nxt(gra_ghost_size), gg = gra_ghost_size;
  parallel_for(ngroup, nthreads, [&](int g) {
    for (int l = 0; l < 8; l++) {
      const int P = group_first[g] + l, p = 8 * g + l;
      if (fine_pot) {   // the patch itself, from the output window (CPU_PoissonSolver_SOR.cpp:303-311 offsets)
        const float* w = h_pot_out + (size_t)p * G * G * G;
        float* o = fine_pot + (size_t)P * kP3;
        for (int z = 0; z < kP; z++)
          for (int y = 0; y < kP; y++)
            for (int x = 0; x < kP; x++) o[(z * kP + y) * kP + x] = w[((z + gg) * G + (y + gg)) * G + (x + gg)];
      }
      if (fine_flu)
        std::memcpy(fine_flu + (size_t)P * kNComp * kP3, h_flu + (size_t)p * kNComp * kP3, sizeof(float) * kNComp * kP3);
    }
  });
  return PGAMD_OK;
}

}  // extern "C"
