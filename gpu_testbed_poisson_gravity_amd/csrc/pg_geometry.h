// Size algebra of one patch for the batched Poisson-gravity path, shared by host and device code.
//
// Follows Header/Macro.h:47-61 and CPU_PoissonSolver_SOR.cpp:11-12 of the reference, with the ghost
// sizes as run-time (host) or template (device) values instead of compile-time macros.
#pragma once

namespace pgamd {

constexpr int kPatchSize = 8;  // PATCH_SIZE
constexpr int kNComp = 5;      // NCOMP: rho, px, py, pz, E
constexpr int kPatchCells = kPatchSize * kPatchSize * kPatchSize;
constexpr int kMaxGhost = 5;   // POT_GHOST_SIZE <= 5 (Main.cpp:89-91)

// Interpolation schemes keep the reference's enum values (Header/Typedef.h:7-8).
constexpr int kIntCQuad = 4;
constexpr int kIntQuad = 5;

constexpr int kSchemeSOR = 1;  // Header/Macro.h:47-48
constexpr int kSchemeMG = 2;

struct Geo {
  int gs;  // POT_GHOST_SIZE
  int N;   // solve cube edge including the Dirichlet layer
  int R;   // RHO_NXT, interior edge
  int C;   // POT_NXT, coarse edge
  int NI;  // POT_NXT_INT
  int U;   // POT_USELESS
};

constexpr Geo geo_of(int gs) {
  return Geo{gs,
             kPatchSize + 2 * gs,
             kPatchSize + 2 * gs - 2,
             kPatchSize / 2 + 2 * ((gs + 3) / 2),
             2 * (kPatchSize / 2 + 2 * ((gs + 3) / 2) - 2),
             gs % 2};
}

constexpr int gra_nxt(int gg) { return kPatchSize + 2 * gg; }

}  // namespace pgamd
