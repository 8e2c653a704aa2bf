// Red-black SOR on gfx950 with the whole solve cube resident in LDS, one wavefront per patch (or per 2-4
// small patches), plus the coarse-to-fine interpolation, the output window and the fused gravity update.
//
// Reference behaviour reproduced here (bit for bit in FP32 when the same iteration count is run):
//   interpolation  CPU_PoissonSolver_SOR.cpp:110-203   (GPU: CUPOT_PoissonSolver_SOR_10to14cube.cu:185-343)
//   SOR sweep      CPU_PoissonSolver_SOR.cpp:215-270   (GPU: CUPOT_PoissonSolver_SOR_10to14cube.cu:349-405)
//   stall rule     CPU_PoissonSolver_SOR.cpp:273-283
//   window         CPU_PoissonSolver_SOR.cpp:303-311
//   gravity        CPU_GravitySolver.cpp:34-79         (GPU: CUPOT_GravitySolver.cu:24-115)
//
// Layout.  The N^3 cube lives in LDS split by colour: cell (x,y,z) is stored at
//     lds[((c*N + z)*N + y)*HP + (x>>1)],  c = (x+y+z)&1,  HP = N/2,
// so the cells of one colour along a row are contiguous.  A lane owns a 2x2 block of z-columns
// (x in {2m+1, 2m+2}, y in {2yb+1, 2yb+2}) over the whole interior height and keeps those cells' phi and
// Const*rho in registers.  Updating one cell then needs only two LDS reads (the neighbour across the
// x-pair and the one across the y-pair, both at row offset m+xx in the other colour's array) and one
// LDS write; z-neighbours and the other half of the x/y neighbours come from the lane's own registers.
// Each patch needs (R/2)^2 lanes: 16, 25, 36, 49, 64 for ghost sizes 1..5, so 4 (gs 1) or 2 (gs 2)
// patches can share a wavefront.  One workgroup is one wavefront; barriers are single-wave barriers.
//
// Relative-residual rule (no extra sweep): at a check the first-colour residuals are the pre-update
// residuals of the pass being run, the second-colour ones are |1-omega| times the pre-update residuals
// of the previous second pass; the pass is only committed if the patch has not converged.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "pg_geometry.h"
#include "pg_kernels.h"
#include "pg_interp.h"
#include "pg_internal.h"

#pragma clang fp contract(off)

namespace pgamd {

// Compile-time layout: L = gs + 8*pad selects the ghost size and the padding of the split rows
// (HP = N/2 + pad floats per row).  Padding trades LDS capacity for fewer bank conflicts.
template <int L>
struct SorGeo {
  static constexpr int gs = L & 7;
  static constexpr int PAD = L >> 3;
  static constexpr Geo g = geo_of(gs);
  static constexpr int N = g.N;
  static constexpr int R = g.R;
  static constexpr int C = g.C;
  static constexpr int U = g.U;
  static constexpr int HP = N / 2 + PAD;        // split row stride in floats
  static constexpr int MX = R / 2;              // x-pairs per row
  static constexpr int LP = MX * (R / 2);       // lanes per patch
  static constexpr bool COMPACT = PAD == kPadCompact;   // single-buffer layout (ghost size 5 only)
  static constexpr int FLOATS = COMPACT ? 16 * 256 + 16 * 64 : 2 * N * N * HP;   // LDS floats per patch
  // Lane order within a patch.  Row-major (lane = yb*MX + m) puts consecutive lanes on consecutive words of a
  // split row; x-major (lane = m*MX + yb) puts them 2*HP words apart.  An LDS access serves 32 lanes per
  // cycle over 32 four-byte banks, so x-major at HP = 10 (ghost size 5, one float of padding) sends the 32
  // lanes of a half-wave (yb = 0..7, four values of m) to words 20*yb + m: 32 distinct banks.  Measured on
  // MI355X (16384 patches): x-major is faster at ghost sizes 4 and 5, equal or slower below.
  static constexpr bool XMAJOR = gs >= 4;
};

constexpr int layout_code(int gs, int pad) { return gs + 8 * pad; }

__device__ __forceinline__ int cube_index(int N, int HP, int x, int y, int z) {
  return ((((x + y + z) & 1) * N + z) * N + y) * HP + (x >> 1);
}

// Sum over a power-of-two lane segment; every lane of the segment gets a bitwise-identical result
// (each butterfly step adds the same two operands, only commuted).
template <int SEG>
__device__ __forceinline__ float seg_sum(float v) {
#pragma unroll
  for (int off = SEG / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum for the relative-residual test.  A whole wavefront (SEG = 64) reduces inside the VALU: DPP exchanges
// within each row of 16 lanes (partners 1 and 2 apart in a quad, then the mirrored halves of 8 and of 16
// lanes; each step adds the same two operands in every lane of a pair, only commuted), then the four row sums
// read from lanes 0, 16, 32 and 48 are added as (r0 + r1) + (r2 + r3) from scalar registers, so every lane gets
// the same bits.  seg_sum's butterfly crosses lanes through the LDS crossbar six times in a row; the stall
// rule and the guess check keep that order.
template <int SEG>
__device__ __forceinline__ float check_sum(float v) {
  if constexpr (SEG != 64) {
    return seg_sum<SEG>(v);
  } else {
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));  // quad [1,0,3,2]
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad [2,3,0,1]
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false)); // half-row mirror
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false)); // row mirror
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (r0 + r1) + (r2 + r3);
  }
}

// Any lane of the wavefront (a workgroup is exactly one wavefront here).
__device__ __forceinline__ bool wave_any(bool pred) { return __ballot(pred) != 0; }

// ---- interpolation into LDS --------------------------------------------------------------------------

// Coarse cell q (1-based i, j, k of the NC^3 block that has children in the cube) of the interpolation
// loop: every cell, or (SHELL) only the 2 NC^2 + 4 (NC-1)(NC-2) cells on the outer layer of the block, whose
// children include every cell of the cube's Dirichlet layer (the only cells a zero start or a caller's guess
// leaves to the interpolation).
template <int NC, bool SHELL>
__device__ __forceinline__ void coarse_cell(int q, int& i, int& j, int& k) {
  if (!SHELL) {
    i = q % NC + 1;
    j = (q / NC) % NC + 1;
    k = q / (NC * NC) + 1;
    return;
  }
  constexpr int F = NC * NC, RING = 4 * (NC - 1);
  if (q < 2 * F) {   // the bottom and top faces
    const int t = q % F;
    k = q < F ? 1 : NC;
    i = t % NC + 1;
    j = t / NC + 1;
    return;
  }
  const int r = q - 2 * F, t = r % RING;   // the ring of each layer in between
  k = r / RING + 2;
  if (t < 2 * NC) {
    j = t < NC ? 1 : NC;
    i = t % NC + 1;
  } else {
    const int u = t - 2 * NC;
    i = u < NC - 2 ? 1 : NC;
    j = u % (NC - 2) + 2;
  }
}

// Interpolate a patch's coarse potential into its LDS cube (split layout).  `lane`/`nlanes` index the
// lanes serving this patch.  Two coarse cells per lane and step: both 27-value neighbourhoods are
// requested before either is interpolated, halving the memory round trips of the phase.  SHELL: only the
// cells that reach the Dirichlet layer (zero start or a caller's guess: the interior is overwritten).
template <int GS, bool SHELL>
__device__ __forceinline__ void interp_to_lds(const float* __restrict__ cp, float* __restrict__ lds, int lane, int nlanes,
                              int int_scheme) {
  using S = SorGeo<GS>;
  constexpr int C = S::C, N = S::N, U = S::U, HP = S::HP, NC = C - 2;
  constexpr int NCELL = SHELL ? 2 * NC * NC + 4 * (NC - 1) * (NC - 2) : NC * NC * NC;
  for (int q0 = lane; q0 < NCELL; q0 += 2 * nlanes) {
    float v[2][3][3][3];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int q = q0 + u * nlanes < NCELL ? q0 + u * nlanes : q0;
      int i, j, k;
      coarse_cell<NC, SHELL>(q, i, j, k);
#pragma unroll
      for (int dk = 0; dk < 3; dk++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++)
#pragma unroll
          for (int di = 0; di < 3; di++)
            v[u][dk][dj][di] = cp[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int q = q0 + u * nlanes;
      if (q >= NCELL) break;
      int i, j, k;
      coarse_cell<NC, SHELL>(q, i, j, k);
      float ch[2][2][2];
      interp_children(v[u], int_scheme, ch);
#pragma unroll
      for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
          for (int cx = 0; cx < 2; cx++) {
            const int x = 2 * (i - 1) + cx - U, y = 2 * (j - 1) + cy - U, z = 2 * (k - 1) + cz - U;
            if (x >= 0 && x < N && y >= 0 && y < N && z >= 0 && z < N)
              lds[cube_index(N, HP, x, y, z)] = ch[cz][cy][cx];
          }
    }
  }
}

// The Dirichlet layer of a patch from a caller's whole-cube guess [N^3] into its LDS cube: every load of the
// lane is issued before any value is stored, one memory round trip for the phase.
template <int GS, int SEG>
__device__ __forceinline__ void shell_to_lds(const float* __restrict__ g, float* __restrict__ lds, int lane) {
  using S = SorGeo<GS>;
  constexpr int N = S::N, HP = S::HP;
  constexpr int NSHELL = 2 * N * N + 4 * (N - 1) * (N - 2);
  constexpr int PER = (NSHELL + SEG - 1) / SEG;
  float v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int q = lane + k * SEG;
    int x, y, z;
    shell_cell<N>(q < NSHELL ? q : 0, x, y, z);
    v[k] = g[(z * N + y) * N + x];
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int q = lane + k * SEG;
    if (q < NSHELL) {
      int x, y, z;
      shell_cell<N>(q, x, y, z);
      lds[cube_index(N, HP, x, y, z)] = v[k];
    }
  }
}

// ---- per-lane state and sweeps -----------------------------------------------------------------------

template <int GS>
struct LaneCells {
  static constexpr int R = SorGeo<GS>::R;
  float phi[2][2][R];  // [yy][xx][z-1]
  float cr[2][2][R];   // Const*rho
  float bz[2][2][2];   // Dirichlet values below/above each owned column
};

// LDS index of cell (x = 2m+1+xx, y = 2yb+1+yy, z) relative to the lane base lb = 2*yb*HP + m.
template <int GS>
__device__ __forceinline__ constexpr int own_off(int c, int z, int yy, int xx) {
  using S = SorGeo<GS>;
  return ((c * S::N + z) * S::N + (yy + 1)) * S::HP + xx;
}

// Software pipeline depth of the LDS loads in a half-sweep: the two neighbour values of cell q+kLookahead
// are requested before cell q is computed, so a wavefront keeps ~2*kLookahead LDS reads in flight while
// holding only that many extra registers.
constexpr int kLookahead = 8;

// Cells of colour COL owned by a lane, enumerated z-major: q = 2*zi + yy.
template <int GS, int COL>
struct HalfSweep {
  static constexpr int R = SorGeo<GS>::R;
  static constexpr int NQ = 2 * R;
  static constexpr int yy(int q) { return q & 1; }
  static constexpr int zi(int q) { return q >> 1; }
  static constexpr int xx(int q) { return (COL + (q & 1) + (q >> 1) + 1) & 1; }
};

// LDS addressing of a lane's cells.  load_nb gives the two LDS operands of owned cell q (v[0]: the neighbour
// across the y-pair, v[1]: the one across the x-pair); own gives where the lane publishes cell q.
//
// Split layout: both colours resident, the lane's operands at row offset m+xx of the other colour's array.
template <int GS>
struct SplitAddr {
  static constexpr int kLook = kLookahead;
  int lb;   // 2*yb*HP + m
  template <int COL>
  __device__ __forceinline__ void load_nb(const float* __restrict__ lds, int q, float (&v)[2]) const {
    using H = HalfSweep<GS, COL>;
    const int yy = H::yy(q), zi = H::zi(q), xx = H::xx(q);
    v[0] = lds[lb + own_off<GS>(1 - COL, zi + 1, yy == 0 ? -1 : 2, xx)];
    v[1] = lds[lb + own_off<GS>(1 - COL, zi + 1, yy, xx)];
  }
  template <int COL>
  __device__ __forceinline__ int own(int q) const {
    using H = HalfSweep<GS, COL>;
    return lb + own_off<GS>(COL, H::zi(q) + 1, H::yy(q), H::xx(q));
  }
  template <int COL>
  __device__ __forceinline__ float crv(const LaneCells<GS>& L, const float* __restrict__, int q) const {
    using H = HalfSweep<GS, COL>;
    return L.cr[H::yy(q)][H::xx(q)][H::zi(q)];
  }
};

// Compact layout (ghost size 5, one patch per wavefront).  A red-black half-sweep reads only the other
// colour and then overwrites nothing but its own colour, so one buffer can hold, for each interior cell pair
// x in {2s-1, 2s} of a row, whichever of the two cells was written last: a half-sweep reads the other colour
// there and then publishes its own colour into the same slots.  Plane z (1..16) starts at (z-1)*kPS floats;
// interior row y (1..16) at (y-1)*kRS with slot 0 the Dirichlet cell x = 0, slots 1..8 the pairs and slot 9
// the Dirichlet cell x = 17; the Dirichlet rows y = 0 and 17 keep their 16 interior cells (x = 2m+1 at A+m,
// x = 2m+2 at B+m).  The z = 0 and 17 planes stay in the lanes' registers.  16 KB per patch instead of the
// split layout's 25.9 KB: eight resident patches per CU instead of six.  The 4 KB after the cube hold the
// lane's Const*rho of its top four planes ([cell][lane], conflict-free), and eight 20 KB patches fill the 160 KB
// of LDS exactly.  With those 16 values out of the registers the kernel still uses all 256 VGPRs and spills
// 140 bytes per lane (set-up values, and a few values reloaded once per iteration), but no longer reloads
// spilled Const*rho values all through each half-sweep (9% faster, measured).
// Banks: the 32 lanes of an LDS cycle (yb = 0..7, four values of m) read rows 20*yb + m + const, 28 of the 32
// four-byte banks for a y-neighbour read; A and B sit on the 4 banks left free (23..26 mod 32 for y = 0, 1..4
// for y = 17).  kPS = 256 is a multiple of 64 floats, so two planes pair into one ds_read2st64 from the
// lane's base: the sweeps need no address arithmetic.
struct Compact {
  static constexpr int kRS = 10, kPS = 256, kPlanes = 16;
  static constexpr int kCr = kPlanes * kPS;   // Const*rho of the lane's cells at z - 1 >= kCrZ, 16 per lane
  static constexpr int kCrZ = 12;
  static constexpr int kFloats = kCr + 16 * 64;
  static constexpr int kY0A = 183, kY0B = 215, kY17A = 161, kY17B = 193;
  // slot of cube cell (x, y, z), x, y, z in 1..16 plus the Dirichlet cells x or y in {0, 17}, z in 1..16
  __device__ static constexpr int slot(int x, int y, int z) {
    return (z - 1) * kPS + (y == 0 ? (((x - 1) & 1) ? kY0B : kY0A) + ((x - 1) >> 1)
                            : y == 17 ? (((x - 1) & 1) ? kY17B : kY17A) + ((x - 1) >> 1)
                                      : (y - 1) * kRS + ((x + 1) >> 1));
  }
};

// Six lane registers address every operand: the lane's row base B = 20*yb + m (x-neighbours at B + 10*yy +
// 2*xx, its own slots at B + 10*yy + 1, compile-time offsets from one register), the y-neighbour bases of the
// two rows, which move to the Dirichlet rows for the lanes at yb = 0 (below) and yb = 7 (above), and the base
// of the lane's Const*rho slots.
template <int GS>
struct CompactAddr {
  static constexpr int kLook = 4;
  int b;          // 20*yb + m
  int ylo[2];     // y-neighbour below (yy = 0) of the lane's cell xx, plane 1
  int yhi[2];     // y-neighbour above (yy = 1)
  int crb;        // kCr + lane
  __device__ __forceinline__ void init(int m, int yb, int lane) {
    crb = Compact::kCr + lane;
    asm volatile("" : "+v"(crb));
    b = 2 * yb * Compact::kRS + m;
#pragma unroll
    for (int xx = 0; xx < 2; xx++) {
      const int x = 2 * m + 1 + xx;
      ylo[xx] = Compact::slot(x, 2 * yb, 1);
      yhi[xx] = Compact::slot(x, 2 * yb + 3, 1);
    }
    asm volatile("" : "+v"(b), "+v"(ylo[0]), "+v"(ylo[1]), "+v"(yhi[0]), "+v"(yhi[1]));
  }
  template <int COL>
  __device__ __forceinline__ void load_nb(const float* __restrict__ lds, int q, float (&v)[2]) const {
    using H = HalfSweep<GS, COL>;
    const int yy = H::yy(q), zi = H::zi(q), xx = H::xx(q);
    v[0] = lds[(yy == 0 ? ylo[xx] : yhi[xx]) + Compact::kPS * zi];
    v[1] = lds[b + Compact::kRS * yy + 2 * xx + Compact::kPS * zi];
  }
  template <int COL>
  __device__ __forceinline__ int own(int q) const {
    using H = HalfSweep<GS, COL>;
    return b + Compact::kRS * H::yy(q) + 1 + Compact::kPS * H::zi(q);
  }
  // LDS slot of Const*rho of the lane's cell (yy, xx, zi), zi >= kCrZ, relative to crb
  __device__ __forceinline__ static constexpr int cr_slot(int yy, int xx, int zi) {
    return (((zi - Compact::kCrZ) * 2 + yy) * 2 + xx) * 64;
  }
  template <int COL>
  __device__ __forceinline__ float crv(const LaneCells<GS>& L, const float* __restrict__ lds, int q) const {
    using H = HalfSweep<GS, COL>;
    const int yy = H::yy(q), zi = H::zi(q), xx = H::xx(q);
    return zi >= Compact::kCrZ ? lds[crb + cr_slot(yy, xx, zi)] : L.cr[yy][xx][zi];
  }
};

// Pre-update residual of owned cell q of colour COL, reference operand order
// (CPU_PoissonSolver_SOR.cpp:247-250): ((((((kp+km)+jp)+jm)+ip)+im) - 6 phi) - Const*rho.
template <int GS, int COL>
__device__ __forceinline__ float residual_at(const LaneCells<GS>& L, const float (&v)[2], int q, float cr) {
  using H = HalfSweep<GS, COL>;
  constexpr int R = H::R;
  const int yy = H::yy(q), zi = H::zi(q), xx = H::xx(q);
  const float kp = zi + 1 < R ? L.phi[yy][xx][zi + 1] : L.bz[yy][xx][1];
  const float km = zi > 0 ? L.phi[yy][xx][zi - 1] : L.bz[yy][xx][0];
  const float jp = yy == 0 ? L.phi[1][xx][zi] : v[0];
  const float jm = yy == 1 ? L.phi[0][xx][zi] : v[0];
  const float ip = xx == 0 ? L.phi[yy][1][zi] : v[1];
  const float im = xx == 1 ? L.phi[yy][0][zi] : v[1];
  return kp + km + jp + jm + ip + im - 6.0f * L.phi[yy][xx][zi] - cr;
}

// Publish the lane's cells of colour COL to LDS (one exec-masked block per half-sweep).
template <int GS, int COL, class A>
__device__ __forceinline__ void store_colour(const LaneCells<GS>& L, float* __restrict__ lds, const A& a, bool active) {
  using H = HalfSweep<GS, COL>;
  if (active) {
#pragma unroll
    for (int q = 0; q < H::NQ; q++) lds[a.template own<COL>(q)] = L.phi[H::yy(q)][H::xx(q)][H::zi(q)];
  }
}

// Residuals of all owned cells of colour COL with pipelined loads.  MODE 0: phi += w*R in place (w is
// omega/6 for running patches and 0 for converged ones: phi + 0*R == phi for the finite residuals of a
// converged patch, so no branch is needed).  MODE 1: keep R in res[] and leave phi unchanged.
// Returns sum|R| when ACC.
template <int GS, int COL, int MODE, bool ACC, class A>
__device__ __forceinline__ float half_sweep(LaneCells<GS>& L, const float* __restrict__ lds, const A& a, float w,
                                            float (&res)[2 * SorGeo<GS>::R]) {
  using H = HalfSweep<GS, COL>;
  constexpr int NQ = H::NQ;
  float nb[NQ][2];
  constexpr int LA = A::kLook;
#pragma unroll
  for (int q = 0; q < LA && q < NQ; q++) a.template load_nb<COL>(lds, q, nb[q]);
  float acc = 0.0f;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    if (q + LA < NQ) a.template load_nb<COL>(lds, q + LA, nb[q + LA]);
    const float r = residual_at<GS, COL>(L, nb[q], q, a.template crv<COL>(L, lds, q));
    if (ACC) acc += fabsf(r);
    if (MODE == 0) {
      float& ph = L.phi[H::yy(q)][H::xx(q)][H::zi(q)];
      ph = ph + w * r;
    } else {
      res[q] = r;
    }
  }
  return acc;
}

template <int GS, int COL>
__device__ __forceinline__ void apply_res(LaneCells<GS>& L, const float (&res)[2 * SorGeo<GS>::R], float w) {
  using H = HalfSweep<GS, COL>;
#pragma unroll
  for (int q = 0; q < H::NQ; q++) {
    float& ph = L.phi[H::yy(q)][H::xx(q)][H::zi(q)];
    ph = ph + w * res[q];
  }
}

template <int GS>
__device__ __forceinline__ float abs_sum(const float (&res)[2 * SorGeo<GS>::R]) {
  float s = 0.0f;
#pragma unroll
  for (int q = 0; q < 2 * SorGeo<GS>::R; q++) s += fabsf(res[q]);
  return s;
}

// Per-lane view of the patch a slot of the wavefront works on.
struct LanePatch {
  const float* rho;
  const float* cpot;
  const float* guess;
  float* out;
  float* flu;
  int* iters;
  int* status;
  bool valid;
};

// Gravity update of one patch from the potential held in an LDS cube (CPU_GravitySolver.cpp:34-79).
// The SEG lanes serving the patch take cells lane, lane+SEG, ...; the fluid of up to 8 cells per lane is
// requested before any of them is updated, so a lane waits for memory once per chunk, not once per cell.
template <int GS, int SEG, class Idx>
__device__ void gravity_from_lds(const float* __restrict__ lds, Idx idx, float* __restrict__ flu, int lane,
                                 float gra_const, int p5) {
  using S = SorGeo<GS>;
  constexpr int NT = kPatchCells / SEG;          // cells per lane
  constexpr int CHUNK = NT < 8 ? NT : 8;
  static_assert(kPatchCells % SEG == 0 && NT % CHUNK == 0, "lane segment must divide the patch");
  for (int c0 = 0; c0 < NT; c0 += CHUNK) {
    float f[CHUNK][kNComp];
#pragma unroll
    for (int u = 0; u < CHUNK; u++) {
      const int t = lane + (c0 + u) * SEG;
#pragma unroll
      for (int v = 0; v < kNComp; v++) f[u][v] = flu[v * kPatchCells + t];
    }
#pragma unroll
    for (int u = 0; u < CHUNK; u++) {
      const int t = lane + (c0 + u) * SEG;
      const int i = t % kPatchSize, j = (t / kPatchSize) % kPatchSize, k = t / (kPatchSize * kPatchSize);
      const int x = i + S::gs, y = j + S::gs, z = k + S::gs;
      auto P = [&](int dz, int dy, int dx) { return lds[idx(x + dx, y + dy, z + dz)]; };
      float ax, ay, az;
      if (p5) {
        ax = gra_const * (-P(0, 0, 2) + P(0, 0, -2) + 8.0f * P(0, 0, 1) - 8.0f * P(0, 0, -1));
        ay = gra_const * (-P(0, 2, 0) + P(0, -2, 0) + 8.0f * P(0, 1, 0) - 8.0f * P(0, -1, 0));
        az = gra_const * (-P(2, 0, 0) + P(-2, 0, 0) + 8.0f * P(1, 0, 0) - 8.0f * P(-1, 0, 0));
      } else {
        ax = gra_const * (P(0, 0, 1) - P(0, 0, -1));
        ay = gra_const * (P(0, 1, 0) - P(0, -1, 0));
        az = gra_const * (P(1, 0, 0) - P(-1, 0, 0));
      }
      const float rho = f[u][0];
      float px = f[u][1], py = f[u][2], pz = f[u][3];
      const float temp = 0.5f / rho;
      const float eint = f[u][4] - temp * (px * px + py * py + pz * pz);
      px = px - rho * ax;
      py = py - rho * ay;
      pz = pz - rho * az;
      flu[kPatchCells + t] = px;
      flu[2 * kPatchCells + t] = py;
      flu[3 * kPatchCells + t] = pz;
      flu[4 * kPatchCells + t] = eint + temp * (px * px + py * py + pz * pz);
    }
  }
}

// Owned density (and a caller's guess) of a lane's 2x2 columns.  The lane's two x-neighbours (xx = 0, 1) are
// adjacent in memory: one 8-byte load reads both, so one load instruction of a patch reads R/2 whole rows of
// the cube (the same rows in either lane order).  All loads are issued before any is used: one memory round
// trip for the phase.
template <int GS>
__device__ __forceinline__ void load_owned(const LanePatch& lp, const SolveParams& P, int m, int yb, bool active,
                                           float2 (&r2)[2][SorGeo<GS>::R], float2 (&g2)[2][SorGeo<GS>::R]) {
  constexpr int N = SorGeo<GS>::N, R = SorGeo<GS>::R;
#pragma unroll
  for (int yy = 0; yy < 2; yy++)
#pragma unroll
    for (int zi = 0; zi < R; zi++) r2[yy][zi] = g2[yy][zi] = make_float2(0.0f, 0.0f);
  if (active) {
#pragma unroll
    for (int yy = 0; yy < 2; yy++)
#pragma unroll
      for (int zi = 0; zi < R; zi++)
        r2[yy][zi] = *reinterpret_cast<const float2*>(lp.rho + (zi * R + 2 * yb + yy) * R + 2 * m);
    if (P.init_mode == 2) {
#pragma unroll
      for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int zi = 0; zi < R; zi++)
          g2[yy][zi] = *reinterpret_cast<const float2*>(lp.guess + (zi * R + 2 * yb + yy) * R + 2 * m);
    } else if (P.init_mode == 3) {   // whole-cube guess: the owned cells sit at cube x = 2m+1, 2m+2 (odd start)
#pragma unroll
      for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int zi = 0; zi < R; zi++) {
          const float* gp = lp.guess + ((zi + 1) * N + 2 * yb + yy + 1) * N + 2 * m + 1;
          g2[yy][zi] = make_float2(gp[0], gp[1]);
        }
    }
  }
}

// Reject a caller guess whose relative residual exceeds guess_tol: restart from the zero guess.  Both colours'
// residuals are taken from the same field.  The compact layout holds one colour at a time: the first colour is
// published between the two residual passes, and the second colour (the zero guess after a rejection) is
// published again afterwards, as the first sweep expects.
template <int GS, int SEG, bool COMPACT, class A>
__device__ __forceinline__ bool guess_check(LaneCells<GS>& L, float* __restrict__ lds, const A& a, const SolveParams& P,
                                            bool valid, bool active, float (&res)[2 * SorGeo<GS>::R]) {
  constexpr int R = SorGeo<GS>::R;
  float num = half_sweep<GS, 1, 1, true>(L, lds, a, 0.0f, res);
  if (COMPACT) {
    store_colour<GS, 1>(L, lds, a, active);
    __syncthreads();
  }
  num += half_sweep<GS, 0, 1, true>(L, lds, a, 0.0f, res);
  float den = 0.0f;
#pragma unroll
  for (int yy = 0; yy < 2; yy++)
#pragma unroll
    for (int xx = 0; xx < 2; xx++)
#pragma unroll
      for (int zi = 0; zi < R; zi++) den += fabsf(L.phi[yy][xx][zi]);
  if (!active) num = den = 0.0f;
  num = seg_sum<SEG>(num);
  den = seg_sum<SEG>(den);
  const bool fell_back = valid && !(num <= P.guess_tol * den);
  if (fell_back) {
#pragma unroll
    for (int yy = 0; yy < 2; yy++)
#pragma unroll
      for (int xx = 0; xx < 2; xx++)
#pragma unroll
        for (int zi = 0; zi < R; zi++) L.phi[yy][xx][zi] = 0.0f;
  }
  if (COMPACT) {
    store_colour<GS, 0>(L, lds, a, active);
  } else {
    store_colour<GS, 1>(L, lds, a, active && fell_back);
    store_colour<GS, 0>(L, lds, a, active && fell_back);
  }
  __syncthreads();
  return fell_back;
}

// Red-black iterations of one patch per lane segment.  Sets it_done and st (converged or max_iter).
template <int GS, int SEG, class A>
__device__ __forceinline__ void sor_iterate(LaneCells<GS>& L, float* __restrict__ lds, const A& a, const SolveParams& P,
                                            bool active, float (&res)[2 * SorGeo<GS>::R], int& it_done, int& st) {
  constexpr int R = SorGeo<GS>::R;
  bool running = active;
  it_done = P.max_iter;
  st = kStatusMaxIter;
  const bool relres = P.term_mode == 1, stall = P.term_mode == 0;
  const int first_check = P.min_iter > 1 ? P.min_iter : 1;
  float sum_b = 0.0f, old = FLT_MAX;
  for (int t = 0; t < P.max_iter; ++t) {
    const bool check = relres && t >= first_check && (t - first_check) % P.check_every == 0;
    const bool pre_check = relres && t + 1 >= first_check && (t + 1 - first_check) % P.check_every == 0;
    float tot = 0.0f;
    if (check) {
      float num = half_sweep<GS, 1, 1, true>(L, lds, a, 0.0f, res), den = 0.0f;
#pragma unroll
      for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int xx = 0; xx < 2; xx++)
#pragma unroll
          for (int zi = 0; zi < R; zi++) den += fabsf(L.phi[yy][xx][zi]);
      num = num + P.damp * sum_b;
      if (!active) num = den = 0.0f;
      num = check_sum<SEG>(num);
      den = check_sum<SEG>(den);
      if (running && num <= P.rel_tol * den) {
        running = false;
        it_done = t;
        st = kStatusConverged;
      }
      if (!wave_any(running)) break;
      apply_res<GS, 1>(L, res, running ? P.w6 : 0.0f);
    } else if (stall) {
      tot = half_sweep<GS, 1, 0, true>(L, lds, a, running ? P.w6 : 0.0f, res);
    } else {
      half_sweep<GS, 1, 0, false>(L, lds, a, running ? P.w6 : 0.0f, res);
    }
    store_colour<GS, 1>(L, lds, a, active);
    __syncthreads();
    const float w = running ? P.w6 : 0.0f;
    if (stall || pre_check) {
      const float s2 = half_sweep<GS, 0, 0, true>(L, lds, a, w, res);
      sum_b = s2;
      tot += s2;
    } else {
      half_sweep<GS, 0, 0, false>(L, lds, a, w, res);
    }
    store_colour<GS, 0>(L, lds, a, active);
    __syncthreads();
    if (stall) {
      tot = seg_sum<SEG>(active ? tot : 0.0f);
      if (running && t + 1 >= P.min_iter && tot > old) {
        running = false;
        it_done = t + 1;
        st = kStatusConverged;
      }
      old = tot;
      if (!wave_any(running)) break;
    }
  }
}

// Non-finite check over the owned cells of each patch, then the patch's iteration count and status.
template <int GS, int SEG>
__device__ __forceinline__ void report(const LaneCells<GS>& L, const LanePatch& lp, bool active, bool fell_back,
                                       int it_done, int st, int slot, int pl) {
  constexpr int R = SorGeo<GS>::R;
  bool bad = false;
#pragma unroll
  for (int yy = 0; yy < 2; yy++)
#pragma unroll
    for (int xx = 0; xx < 2; xx++)
#pragma unroll
      for (int zi = 0; zi < R; zi++) bad |= !isfinite(L.phi[yy][xx][zi]);
  const uint64_t bad_mask = __ballot(active && bad);
  uint64_t seg_mask = ~0ull;
  if constexpr (SEG < 64) seg_mask = ((1ull << SEG) - 1) << (slot * SEG);
  if (bad_mask & seg_mask) st |= kStatusNonFinite;
  if (fell_back) st |= kStatusFallback;
  if (lp.valid && pl == 0) {
    if (lp.iters) *lp.iters = it_done;
    if (lp.status) *lp.status = st;
  }
}

// Compact layout, phase A/B: the lane's Dirichlet values below and above its columns, its cells (interpolated
// start or a caller's guess; zero otherwise), and the Dirichlet cells of the x and y faces in their slots.
// The interpolation is the reference's per coarse cell (interp_children, all 8 children, the same operation
// order), distributed by coarse cell: with U = 1 the lane's cells x = 2m+1+cx, y = 2yb+1+cy are the children
// of coarse column (m+2, yb+2); coarse cells k = 1 and 10 give its z = 0 and 17 values, k = 2..9 its cells.
// The 256 coarse cells under the four side faces (8 x 8 each) give the x and y Dirichlet cells, four per lane.
// The coarse cube is staged in the (still unused) LDS buffer first; the face cells reach their slots after a
// barrier, once every lane has read the staged cube.
template <int GS>
__device__ __forceinline__ void compact_start(const LanePatch& lp, float* __restrict__ lds, const SolveParams& P,
                                              int m, int yb, int pl, bool active, LaneCells<GS>& L) {
  using S = SorGeo<GS>;
  constexpr int N = S::N, R = S::R, C = S::C;
  static_assert(S::gs == 5 && S::U == 1 && R == 16, "compact layout: ghost size 5");
  static_assert(C * C * C <= Compact::kFloats, "coarse staging fits the buffer");
#pragma unroll
  for (int yy = 0; yy < 2; yy++)
#pragma unroll
    for (int xx = 0; xx < 2; xx++) {
      L.bz[yy][xx][0] = L.bz[yy][xx][1] = 0.0f;
#pragma unroll
      for (int zi = 0; zi < R; zi++) L.phi[yy][xx][zi] = 0.0f;
    }
  constexpr int PER = 4 * R * R / 64;   // face cells per lane (whole-cube guess)
  float face[PER];                      // this lane's face cells, published after the barrier
  if (P.init_mode == 3) {
    // the whole-cube guess brings the Dirichlet layer
    if (active) {
#pragma unroll
      for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int xx = 0; xx < 2; xx++) {
          const int col = (2 * yb + 1 + yy) * N + 2 * m + 1 + xx;
          L.bz[yy][xx][0] = lp.guess[col];
          L.bz[yy][xx][1] = lp.guess[(N - 1) * N * N + col];
        }
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int q = pl + 64 * k, f = q / (R * R), a = q % R, z = (q / R) % R + 1;
      const int x = f < 2 ? (f == 0 ? 0 : N - 1) : a + 1, y = f < 2 ? a + 1 : (f == 2 ? 0 : N - 1);
      face[k] = lp.valid ? lp.guess[(z * N + y) * N + x] : 0.0f;
    }
  } else {
    // stage the coarse cube (C^3 floats, lane-contiguous loads, all issued before any store)
    constexpr int CPL = (C * C * C + 63) / 64;
    float cv[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) {
      const int t = pl + 64 * k;
      cv[k] = lp.valid && t < C * C * C ? lp.cpot[t] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < CPL; k++) {
      const int t = pl + 64 * k;
      if (t < C * C * C) lds[t] = cv[k];
    }
    __syncthreads();
    auto coarse = [&](int i, int j, int k, float (&v)[3][3][3]) {
#pragma unroll
      for (int dk = 0; dk < 3; dk++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++)
#pragma unroll
          for (int di = 0; di < 3; di++) v[dk][dj][di] = lds[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
    };
    // the side faces: coarse cells (f, a, b) -> i or j = 1 / 10 and the other two 2..9; four children each
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = pl + 64 * u, f = q >> 6, a = (q & 7) + 2, b = ((q >> 3) & 7) + 2;
      const int ci = f < 2 ? (f == 0 ? 1 : 10) : a, cj = f < 2 ? a : (f == 2 ? 1 : 10);
      float v[3][3][3], ch[2][2][2];
      coarse(ci, cj, b, v);
      interp_children(v, P.int_scheme, ch);
#pragma unroll
      for (int c1 = 0; c1 < 2; c1++)
#pragma unroll
        for (int cz = 0; cz < 2; cz++)
          face[4 * u + 2 * c1 + cz] = f < 2 ? ch[cz][c1][f == 0 ? 1 : 0] : ch[cz][f == 2 ? 1 : 0][c1];
    }
    // the lane's column: k = 1 and 10 always, k = 2..9 for the interpolated start
    const int i = m + 2, j = yb + 2;
#pragma unroll
    for (int k = 1; k <= 10; k++) {
      if (k != 1 && k != 10 && P.init_mode != 0) continue;
      float v[3][3][3], ch[2][2][2];
      coarse(i, j, k, v);
      interp_children(v, P.int_scheme, ch);
#pragma unroll
      for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int xx = 0; xx < 2; xx++) {
          if (k == 1) {
            L.bz[yy][xx][0] = ch[1][yy][xx];
          } else if (k == 10) {
            L.bz[yy][xx][1] = ch[0][yy][xx];
          } else {
            L.phi[yy][xx][2 * k - 4] = ch[0][yy][xx];
            L.phi[yy][xx][2 * k - 3] = ch[1][yy][xx];
          }
        }
    }
    __syncthreads();   // every lane has read the staged cube
  }
  // publish the face cells
  if (lp.valid) {
    if (P.init_mode == 3) {
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int q = pl + 64 * k, f = q / (R * R), a = q % R, z = (q / R) % R + 1;
        const int x = f < 2 ? (f == 0 ? 0 : N - 1) : a + 1, y = f < 2 ? a + 1 : (f == 2 ? 0 : N - 1);
        lds[Compact::slot(x, y, z)] = face[k];
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = pl + 64 * u, f = q >> 6, a = (q & 7) + 2, b = ((q >> 3) & 7) + 2;
        const int ci = f < 2 ? (f == 0 ? 1 : 10) : a, cj = f < 2 ? a : (f == 2 ? 1 : 10);
#pragma unroll
        for (int c1 = 0; c1 < 2; c1++)
#pragma unroll
          for (int cz = 0; cz < 2; cz++) {
            const int z = 2 * (b - 1) + cz - 1;
            if (f < 2)   // x = 0 (child cx = 1) or x = 17 (cx = 0); c1 = cy
              lds[Compact::slot(f == 0 ? 0 : N - 1, 2 * (cj - 1) + c1 - 1, z)] = face[4 * u + 2 * c1 + cz];
            else         // y = 0 (cy = 1) or y = 17 (cy = 0); c1 = cx
              lds[Compact::slot(2 * (ci - 1) + c1 - 1, f == 2 ? 0 : N - 1, z)] = face[4 * u + 2 * c1 + cz];
          }
      }
    }
  }
  // owned density (and a caller's guess)
  float2 r2[2][R], g2[2][R];
  load_owned<GS>(lp, P, m, yb, active, r2, g2);
#pragma unroll
  for (int yy = 0; yy < 2; yy++)
#pragma unroll
    for (int xx = 0; xx < 2; xx++)
#pragma unroll
      for (int zi = 0; zi < R; zi++) {
        if (P.init_mode >= 2) L.phi[yy][xx][zi] = xx ? g2[yy][zi].y : g2[yy][zi].x;
        L.cr[yy][xx][zi] = active ? P.cst * (xx ? r2[yy][zi].y : r2[yy][zi].x) : 0.0f;
        if (zi >= Compact::kCrZ) lds[Compact::kCr + CompactAddr<GS>::cr_slot(yy, xx, zi) + pl] = L.cr[yy][xx][zi];
      }
}

// Compact layout (ghost size 5, one patch per wavefront): the window needs gra_ghost <= 4, all of it interior.
template <int GS, int PPW>
__device__ __forceinline__ void sor_task_compact(float* __restrict__ lds, const LanePatch& lp, const SolveParams& P) {
  using S = SorGeo<GS>;
  constexpr int R = S::R, MX = S::MX;
  static_assert(PPW == 1 && S::LP == 64, "compact layout: one ghost-size-5 patch per wavefront");
  int lane = threadIdx.x;
  asm volatile("" : "+v"(lane));
  const int pl = lane;
  const bool active = lp.valid;
  const int m = pl / MX, yb = pl % MX;
  CompactAddr<GS> a;
  a.init(m, yb, pl);
  LaneCells<GS> L;
  compact_start<GS>(lp, lds, P, m, yb, pl, active, L);
  store_colour<GS, 0>(L, lds, a, active);   // the first sweep (colour 1) reads colour 0
  __syncthreads();

  float res[2 * R];
  bool fell_back = false;
  if (P.init_mode >= 2 && P.guess_tol > 0.0f) fell_back = guess_check<GS, 64, true>(L, lds, a, P, lp.valid, active, res);
  int it_done, st;
  sor_iterate<GS, 64>(L, lds, a, P, active, res, it_done, st);
  report<GS, 64>(L, lp, active, fell_back, it_done, st, 0, pl);
  __syncthreads();

  // D. output window from a dense G^3 copy of the window (interior cells only, o = 5 - gra_ghost >= 1)
  const int G = gra_nxt(P.gra_ghost), o = S::gs - P.gra_ghost;
  if (active) {
#pragma unroll
    for (int yy = 0; yy < 2; yy++)
#pragma unroll
      for (int xx = 0; xx < 2; xx++)
#pragma unroll
        for (int zi = 0; zi < R; zi++) {
          const int x = 2 * m + 1 + xx - o, y = 2 * yb + 1 + yy - o, z = zi + 1 - o;
          if ((unsigned)x < (unsigned)G && (unsigned)y < (unsigned)G && (unsigned)z < (unsigned)G)
            lds[(z * G + y) * G + x] = L.phi[yy][xx][zi];
        }
  }
  __syncthreads();
  if (lp.valid) {
    for (int t = pl; t < G * G * G; t += 64) lp.out[t] = lds[t];
    if (P.fuse_gravity && lp.flu)
      gravity_from_lds<GS, 64>(lds, [=](int x, int y, int z) { return ((z - o) * G + (y - o)) * G + (x - o); },
                               lp.flu, pl, P.gra_const, P.p5);
  }
  __syncthreads();
}

// Solve up to PPW patches of ghost size GS with one wavefront (blockDim.x == 64).
template <int GS, int PPW>
__device__ __forceinline__ void sor_task(float* __restrict__ lds_wave, const LanePatch& lp, const SolveParams& P) {
  using S = SorGeo<GS>;
  if constexpr (S::COMPACT) {
    sor_task_compact<GS, PPW>(lds_wave, lp, P);
  } else {
  constexpr int N = S::N, R = S::R, HP = S::HP, MX = S::MX, LP = S::LP;
  constexpr int SEG = 64 / PPW;
  static_assert(LP <= SEG, "patch does not fit its lane segment");
  int lane = threadIdx.x;
  // opaque per patch: inside the persistent kernel's task loop the compiler would otherwise hoist every
  // lane-derived address out of the loop and keep them all live (measured: 2 KB of spills per lane)
  asm volatile("" : "+v"(lane));
  const int slot = lane / SEG, pl = lane % SEG;
  float* lds = lds_wave + slot * S::FLOATS;
  const bool active = lp.valid && pl < LP;
  const int plc = pl < LP ? pl : LP - 1;  // idle lanes mirror a real lane's addresses and never write
  const int m = S::XMAJOR ? plc / MX : plc % MX, yb = S::XMAJOR ? plc % MX : plc / MX;
  const int lb = 2 * yb * HP + m;
  const SplitAddr<GS> a{lb};

  // A. coarse -> fine interpolation of the whole cube (initial guess and Dirichlet layer); a zero start or a
  // caller's guess needs only the Dirichlet layer (a rejected guess restarts from zero, not from the
  // interpolation), and a whole-cube guess brings its own Dirichlet layer
  if (lp.valid) {
    if (P.init_mode == 3)
      shell_to_lds<GS, SEG>(lp.guess, lds, pl);
    else if (P.init_mode != 0)
      interp_to_lds<GS, true>(lp.cpot, lds, pl, SEG, P.int_scheme);
    else
      interp_to_lds<GS, false>(lp.cpot, lds, pl, SEG, P.int_scheme);
  }
  __syncthreads();

  // B. owned cells into registers
  LaneCells<GS> L;
  float2 r2[2][R], g2[2][R];
  load_owned<GS>(lp, P, m, yb, active, r2, g2);
#pragma unroll
  for (int yy = 0; yy < 2; yy++) {
#pragma unroll
    for (int xx = 0; xx < 2; xx++) {
      L.bz[yy][xx][0] = lds[lb + own_off<GS>((xx + yy) & 1, 0, yy, xx)];
      L.bz[yy][xx][1] = lds[lb + own_off<GS>((xx + yy + N - 1) & 1, N - 1, yy, xx)];
    }
#pragma unroll
    for (int zi = 0; zi < R; zi++) {
#pragma unroll
      for (int xx = 0; xx < 2; xx++) {
        const int c = (xx + yy + zi + 1) & 1;
        float v;
        if (P.init_mode == 0)
          v = lds[lb + own_off<GS>(c, zi + 1, yy, xx)];
        else if (P.init_mode == 1)
          v = 0.0f;
        else
          v = xx ? g2[yy][zi].y : g2[yy][zi].x;
        L.phi[yy][xx][zi] = v;
        L.cr[yy][xx][zi] = active ? P.cst * (xx ? r2[yy][zi].y : r2[yy][zi].x) : 0.0f;
        if (P.init_mode != 0 && active) lds[lb + own_off<GS>(c, zi + 1, yy, xx)] = v;
      }
    }
  }
  __syncthreads();

  // B'. guess check, C. iterations
  float res[2 * R];
  bool fell_back = false;
  if (P.init_mode >= 2 && P.guess_tol > 0.0f) fell_back = guess_check<GS, SEG, false>(L, lds, a, P, lp.valid, active, res);
  int it_done, st;
  sor_iterate<GS, SEG>(L, lds, a, P, active, res, it_done, st);
  report<GS, SEG>(L, lp, active, fell_back, it_done, st, slot, pl);
  __syncthreads();

  // D. output window: out[k][j][i] = cube[k+o][j+o][i+o], o = GS - GRA_GHOST_SIZE
  if (lp.valid) {
    const int G = gra_nxt(P.gra_ghost), o = S::gs - P.gra_ghost;
    for (int t = pl; t < G * G * G; t += SEG) {
      const int i = t % G, j = (t / G) % G, k = t / (G * G);
      lp.out[t] = lds[cube_index(N, HP, i + o, j + o, k + o)];
    }
    if (P.fuse_gravity && lp.flu)
      gravity_from_lds<GS, SEG>(lds, [](int x, int y, int z) { return cube_index(N, HP, x, y, z); }, lp.flu, pl,
                                P.gra_const, P.p5);
  }
  __syncthreads();
  }
}

// ---- kernels -----------------------------------------------------------------------------------------

template <int GS>
__device__ __forceinline__ LanePatch uniform_patch(const UniformBatch& b, int patch, int init_mode) {
  using S = SorGeo<GS>;
  constexpr int R3 = S::R * S::R * S::R, C3 = S::C * S::C * S::C, N3 = S::N * S::N * S::N;
  LanePatch lp;
  lp.valid = patch < b.npatch;
  const int p = lp.valid ? patch : 0;
  lp.rho = b.rho + (size_t)p * R3;
  lp.cpot = b.cpot + (size_t)p * C3;
  lp.guess = b.guess ? b.guess + (size_t)p * (init_mode == 3 ? N3 : R3) : nullptr;
  lp.out = nullptr;  // set by the caller (depends on the window size)
  lp.flu = b.flu ? b.flu + (size_t)p * kNComp * kPatchCells : nullptr;
  lp.iters = b.iters ? b.iters + p : nullptr;
  lp.status = b.status ? b.status + p : nullptr;
  return lp;
}

template <int GS, int PPW>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) sor_uniform_kernel(UniformBatch b, SolveParams P) {
  extern __shared__ float lds[];
  constexpr int SEG = 64 / PPW;
  const int slot = threadIdx.x / SEG;
  const int patch = blockIdx.x * PPW + slot;
  LanePatch lp = uniform_patch<GS>(b, patch, P.init_mode);
  const int G = gra_nxt(P.gra_ghost);
  lp.out = b.pot_out + (size_t)(lp.valid ? patch : 0) * G * G * G;
  sor_task<GS, PPW>(lds, lp, P);
}

// Standalone gravity kernel for the Poisson=false path (CUAPI_Asyn_PoissonGravitySolver.cu:192-200,
// 241-246): one workgroup of 256 threads per patch, potential read from HBM.
__global__ void __launch_bounds__(256) gravity_kernel(const float* __restrict__ pot, float* __restrict__ flu,
                                                      int npatch, int gg, float gra_const, int p5) {
  const int P = blockIdx.x;
  if (P >= npatch) return;
  const int G = gra_nxt(gg);
  const float* pp = pot + (size_t)P * G * G * G;
  float* f = flu + (size_t)P * kNComp * kPatchCells;
  for (int t = threadIdx.x; t < kPatchCells; t += blockDim.x) {
    const int i = t % kPatchSize, j = (t / kPatchSize) % kPatchSize, k = t / (kPatchSize * kPatchSize);
    const int x = i + gg, y = j + gg, z = k + gg;
    auto Q = [&](int dz, int dy, int dx) { return pp[((z + dz) * G + (y + dy)) * G + (x + dx)]; };
    float ax, ay, az;
    if (p5) {
      ax = gra_const * (-Q(0, 0, 2) + Q(0, 0, -2) + 8.0f * Q(0, 0, 1) - 8.0f * Q(0, 0, -1));
      ay = gra_const * (-Q(0, 2, 0) + Q(0, -2, 0) + 8.0f * Q(0, 1, 0) - 8.0f * Q(0, -1, 0));
      az = gra_const * (-Q(2, 0, 0) + Q(-2, 0, 0) + 8.0f * Q(1, 0, 0) - 8.0f * Q(-1, 0, 0));
    } else {
      ax = gra_const * (Q(0, 0, 1) - Q(0, 0, -1));
      ay = gra_const * (Q(0, 1, 0) - Q(0, -1, 0));
      az = gra_const * (Q(1, 0, 0) - Q(-1, 0, 0));
    }
    const float rho = f[t];
    float px = f[kPatchCells + t], py = f[2 * kPatchCells + t], pz = f[3 * kPatchCells + t];
    const float temp = 0.5f / rho;
    const float eint = f[4 * kPatchCells + t] - temp * (px * px + py * py + pz * pz);
    px = px - rho * ax;
    py = py - rho * ay;
    pz = pz - rho * az;
    f[kPatchCells + t] = px;
    f[2 * kPatchCells + t] = py;
    f[3 * kPatchCells + t] = pz;
    f[4 * kPatchCells + t] = eint + temp * (px * px + py * py + pz * pz);
  }
}

// ---- persistent kernel over a mixed batch ----------------------------------------------------------------

template <int L, int PPW>
__device__ __forceinline__ void mixed_sor_task(float* __restrict__ lds, const MixedBatch& b, const Task& tk,
                                               const SolveParams& P) {
  constexpr int SEG = 64 / PPW;
  const int slot = threadIdx.x / SEG;
  LanePatch lp;
  int patch = b.perm[tk.first + (slot < tk.count ? slot : 0)];
  lp.valid = slot < tk.count && patch >= 0 && patch < b.npatch;  // a malformed plan never writes out of range
  if (!lp.valid) patch = 0;
  const PatchDesc d = b.desc[patch];
  const int G = gra_nxt(P.gra_ghost);
  lp.rho = b.rho + d.rho_off;
  lp.cpot = b.cpot + d.cpot_off;
  lp.guess = b.guess ? b.guess + d.rho_off : nullptr;
  lp.out = b.pot_out + (size_t)patch * G * G * G;
  lp.flu = b.flu ? b.flu + (size_t)patch * kNComp * kPatchCells : nullptr;
  lp.iters = b.iters ? b.iters + patch : nullptr;
  lp.status = b.status ? b.status + patch : nullptr;
  sor_task<L, PPW>(lds, lp, P);
}

// Every wavefront-sized workgroup pulls tasks from one queue until it is empty (tasks are ordered largest
// first on the host, so the tail of the queue is cheap).  LDS is sized for the largest ghost size present.
template <int PPW1, int PPW2>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    sor_persistent_kernel(MixedBatch b, ParamsByGs P) {
  extern __shared__ float lds[];
  for (;;) {
    int t = 0;
    if (threadIdx.x == 0) t = (int)atomicAdd(b.counter, 1u);
    t = __shfl(t, 0, 64);
    if (t >= b.ntasks) break;
    const Task tk = b.tasks[t];
    switch (tk.gs) {
      case 1: mixed_sor_task<layout_code(1, 0), PPW1>(lds, b, tk, P.p[1]); break;
      case 2: mixed_sor_task<layout_code(2, 0), PPW2>(lds, b, tk, P.p[2]); break;
      case 3: mixed_sor_task<layout_code(3, 0), 1>(lds, b, tk, P.p[3]); break;
      case 4: mixed_sor_task<layout_code(4, 2), 1>(lds, b, tk, P.p[4]); break;
      default: mixed_sor_task<layout_code(5, 0), 1>(lds, b, tk, P.p[5]); break;
    }
  }
}

// The SOR tasks [first, first + gridDim.x) of one ghost size, one task per wavefront-sized workgroup (the
// plan lists each ghost size's tasks contiguously).  Without the persistent kernel's task loop the patch
// solve compiles as in the uniform kernel, with no spills, and the hardware refills a CU's wave slots as
// patches converge, which balances the ragged per-patch iteration counts without a queue.
template <int L, int PPW>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    sor_tasks_kernel(MixedBatch b, int first, int gs, SolveParams P) {
  extern __shared__ float lds[];
  const Task tk = b.tasks[first + blockIdx.x];
  if (tk.gs != gs) return;   // a malformed plan never runs a task with another ghost size's layout
  mixed_sor_task<L, PPW>(lds, b, tk, P);
}

template <int GS, int PPW, int PAD>
static hipError_t launch_sor_tasks_as(const MixedBatch& b, const SolveParams& p, int first, int count, hipStream_t s) {
  hipLaunchKernelGGL((sor_tasks_kernel<layout_code(GS, PAD), PPW>), dim3(count), dim3(64), lds_bytes_sor(GS, PPW, PAD),
                     s, b, first, GS, p);
  return hipGetLastError();
}

template <int GS, int PPW>
static hipError_t launch_sor_tasks_pad(const MixedBatch& b, const SolveParams& p, int pad, int first, int count,
                                       hipStream_t s) {
  if constexpr (GS == 5 && PPW == 1) {
    if (pad == kPadCompact) return p.gra_ghost <= 4 ? launch_sor_tasks_as<5, 1, kPadCompact>(b, p, first, count, s)
                                                    : launch_sor_tasks_as<5, 1, 1>(b, p, first, count, s);
  }
  switch (pad) {
    case 0: return launch_sor_tasks_as<GS, PPW, 0>(b, p, first, count, s);
    case 1: return launch_sor_tasks_as<GS, PPW, 1>(b, p, first, count, s);
    case 2: return launch_sor_tasks_as<GS, PPW, 2>(b, p, first, count, s);
    case 3: return launch_sor_tasks_as<GS, PPW, 3>(b, p, first, count, s);
    default: return hipErrorInvalidValue;
  }
}

// The same (gs, patches per wavefront, row padding) layouts as the uniform kernel; the padding only moves LDS
// rows, so results do not depend on it.
hipError_t launch_sor_tasks(const MixedBatch& b, const SolveParams& p, int gs, int ppw, int pad, int first, int count,
                            hipStream_t s) {
  if (count <= 0) return hipSuccess;
  switch (gs * 8 + ppw) {
    case 1 * 8 + 1: return launch_sor_tasks_pad<1, 1>(b, p, pad, first, count, s);
    case 1 * 8 + 2: return launch_sor_tasks_pad<1, 2>(b, p, pad, first, count, s);
    case 1 * 8 + 4: return launch_sor_tasks_pad<1, 4>(b, p, pad, first, count, s);
    case 2 * 8 + 1: return launch_sor_tasks_pad<2, 1>(b, p, pad, first, count, s);
    case 2 * 8 + 2: return launch_sor_tasks_pad<2, 2>(b, p, pad, first, count, s);
    case 3 * 8 + 1: return launch_sor_tasks_pad<3, 1>(b, p, pad, first, count, s);
    case 4 * 8 + 1: return launch_sor_tasks_pad<4, 1>(b, p, pad, first, count, s);
    case 5 * 8 + 1: return launch_sor_tasks_pad<5, 1>(b, p, pad, first, count, s);
    default: return hipErrorInvalidValue;
  }
}

__global__ void reset_counters_kernel(unsigned int* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

hipError_t launch_reset_counters(unsigned int* counters, int n, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, counters, n);
  return hipGetLastError();
}

template <int PPW1, int PPW2>
static const void* sor_persistent_fn() {
  return reinterpret_cast<const void*>(&sor_persistent_kernel<PPW1, PPW2>);
}

static const void* sor_persistent_select(int ppw1, int ppw2) {
  if (ppw1 == 4 && ppw2 == 2) return sor_persistent_fn<4, 2>();
  if (ppw1 == 2 && ppw2 == 1) return sor_persistent_fn<2, 1>();
  return nullptr;
}

int sor_persistent_blocks_per_cu(int ppw1, int ppw2, size_t lds) {
  const void* fn = sor_persistent_select(ppw1, ppw2);
  int n = 0;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, lds) != hipSuccess) return 0;
  return n;
}

hipError_t launch_sor_persistent(const MixedBatch& b, const ParamsByGs& p, int ppw1, int ppw2, int grid, size_t lds,
                                 hipStream_t stream) {
  if (b.ntasks <= 0) return hipSuccess;
  if (ppw1 == 4 && ppw2 == 2)
    hipLaunchKernelGGL((sor_persistent_kernel<4, 2>), dim3(grid), dim3(64), lds, stream, b, p);
  else if (ppw1 == 2 && ppw2 == 1)
    hipLaunchKernelGGL((sor_persistent_kernel<2, 1>), dim3(grid), dim3(64), lds, stream, b, p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---- host launchers ------------------------------------------------------------------------------------

int sor_lanes_per_patch(int gs) {
  const int r = geo_of(gs).R;
  return (r / 2) * (r / 2);
}

int supported_ppw(int gs, int ppw) {
  if (ppw != 1 && ppw != 2 && ppw != 4) return 0;
  return sor_lanes_per_patch(gs) <= 64 / ppw;
}

int default_ppw(int gs) {
  for (int ppw : {4, 2, 1})
    if (supported_ppw(gs, ppw)) return ppw;
  return 1;
}

// Measured on MI355X (16384 patches, 1e-5, plain and interpolated starts, every padding 0-3): one float of
// row padding is fastest at ghost sizes 1, 4 and 5 (at 5 it turns the x-major lane order conflict-free, +7%
// over the unpadded cube although a CU then holds 6 patches instead of 7; at 1 it gains 9%); ghost sizes 2
// and 3 are best unpadded.
int default_pad(int gs) { return gs == 5 ? kPadCompact : gs == 1 || gs == 4 ? 1 : 0; }

size_t lds_bytes_sor(int gs, int ppw, int pad) {
  const Geo g = geo_of(gs);
  if (pad == kPadCompact) return sizeof(float) * (size_t)ppw * Compact::kFloats;
  return sizeof(float) * (size_t)ppw * 2 * g.N * g.N * (g.N / 2 + pad);
}

bool supported_layout(int gs, int ppw, int pad) {
  if (!supported_ppw(gs, ppw)) return false;
  if (pad == kPadCompact) return gs == 5 && ppw == 1;
  return pad >= 0 && pad <= 3;
}

template <int GS, int PPW, int PAD>
static hipError_t launch_sor_uniform(const UniformBatch& b, const SolveParams& p, hipStream_t s) {
  constexpr int L = layout_code(GS, PAD);
  const int blocks = (b.npatch + PPW - 1) / PPW;
  const size_t lds = lds_bytes_sor(GS, PPW, PAD);
  hipLaunchKernelGGL((sor_uniform_kernel<L, PPW>), dim3(blocks), dim3(64), lds, s, b, p);
  return hipGetLastError();
}

template <int GS, int PPW>
static hipError_t launch_sor_pad(const UniformBatch& b, const SolveParams& p, int pad, hipStream_t s) {
  if constexpr (GS == 5 && PPW == 1) {
    // the compact layout's window holds interior cells only: a window reaching the Dirichlet layer
    // (gra_ghost 5) runs in the split layout
    if (pad == kPadCompact) return p.gra_ghost <= 4 ? launch_sor_uniform<5, 1, kPadCompact>(b, p, s)
                                                    : launch_sor_uniform<5, 1, 1>(b, p, s);
  }
  switch (pad) {
    case 0: return launch_sor_uniform<GS, PPW, 0>(b, p, s);
    case 1: return launch_sor_uniform<GS, PPW, 1>(b, p, s);
    case 2: return launch_sor_uniform<GS, PPW, 2>(b, p, s);
    case 3: return launch_sor_uniform<GS, PPW, 3>(b, p, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sor_uniform_dispatch(const UniformBatch& b, const SolveParams& p, int ppw, int pad, hipStream_t s) {
  switch (b.gs * 8 + ppw) {
    case 1 * 8 + 1: return launch_sor_pad<1, 1>(b, p, pad, s);
    case 1 * 8 + 2: return launch_sor_pad<1, 2>(b, p, pad, s);
    case 1 * 8 + 4: return launch_sor_pad<1, 4>(b, p, pad, s);
    case 2 * 8 + 1: return launch_sor_pad<2, 1>(b, p, pad, s);
    case 2 * 8 + 2: return launch_sor_pad<2, 2>(b, p, pad, s);
    case 3 * 8 + 1: return launch_sor_pad<3, 1>(b, p, pad, s);
    case 4 * 8 + 1: return launch_sor_pad<4, 1>(b, p, pad, s);
    case 5 * 8 + 1: return launch_sor_pad<5, 1>(b, p, pad, s);
    default: return hipErrorInvalidValue;
  }
}

template <int GS, int PPW, int PAD>
static int sor_blocks_as() {
  int n = 0;
  const void* fn = reinterpret_cast<const void*>(&sor_uniform_kernel<layout_code(GS, PAD), PPW>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, lds_bytes_sor(GS, PPW, PAD)) != hipSuccess) return 0;
  return n;
}

template <int GS, int PPW>
static int sor_blocks_pad(int pad) {
  if constexpr (GS == 5 && PPW == 1) {
    if (pad == kPadCompact) return sor_blocks_as<5, 1, kPadCompact>();
  }
  switch (pad) {
    case 0: return sor_blocks_as<GS, PPW, 0>();
    case 1: return sor_blocks_as<GS, PPW, 1>();
    case 2: return sor_blocks_as<GS, PPW, 2>();
    case 3: return sor_blocks_as<GS, PPW, 3>();
  }
  return 0;
}

int sor_uniform_blocks_per_cu(int gs, int ppw, int pad) {
  switch (gs * 8 + ppw) {
    case 1 * 8 + 1: return sor_blocks_pad<1, 1>(pad);
    case 1 * 8 + 2: return sor_blocks_pad<1, 2>(pad);
    case 1 * 8 + 4: return sor_blocks_pad<1, 4>(pad);
    case 2 * 8 + 1: return sor_blocks_pad<2, 1>(pad);
    case 2 * 8 + 2: return sor_blocks_pad<2, 2>(pad);
    case 3 * 8 + 1: return sor_blocks_pad<3, 1>(pad);
    case 4 * 8 + 1: return sor_blocks_pad<4, 1>(pad);
    case 5 * 8 + 1: return sor_blocks_pad<5, 1>(pad);
  }
  return 0;
}

hipError_t launch_uniform(const UniformBatch& b, const SolveParams& p, const SorLayout& lay, hipStream_t s) {
  if (b.npatch <= 0) return hipSuccess;
  if (b.scheme == kSchemeMG) return launch_mg_uniform(b, p, s);
  return launch_sor_uniform_dispatch(b, p, lay.ppw[b.gs], lay.pad[b.gs], s);
}

hipError_t launch_gravity(const float* pot, float* flu, int npatch, int gra_ghost, float gra_const, int p5,
                          hipStream_t stream) {
  if (npatch <= 0) return hipSuccess;
  hipLaunchKernelGGL(gravity_kernel, dim3(npatch), dim3(256), 0, stream, pot, flu, npatch, gra_ghost, gra_const, p5);
  return hipGetLastError();
}

}  // namespace pgamd
