// Per-patch multigrid V-cycle on gfx950: one workgroup of 256 threads per patch, every level resident
// in LDS, the level-0 right-hand side held in registers.
//
// Reference behaviour reproduced (bit for bit in FP32 for the same number of V-cycles):
//   level schedule        CPU_PoissonSolver_MG.cpp:62-87      (GPU: CUPOT_PoissonSolver_MG.cu:16-53)
//   V-cycle driver        CPU_PoissonSolver_MG.cpp:213-258    (GPU: CUPOT_PoissonSolver_MG.cu:414-463)
//   smoother              CPU_PoissonSolver_MG.cpp:392-441
//   defect / error        CPU_PoissonSolver_MG.cpp:460-514
//   restriction           CPU_PoissonSolver_MG.cpp:533-616
//   prolongation          CPU_PoissonSolver_MG.cpp:635-684
//   window                CPU_PoissonSolver_MG.cpp:348-358
//
// Level-0 smoothing enumerates the cells of one colour directly (x = 2h + 1 + parity offset) so that no
// lane idles; coarser levels (at most 9^3) use a plain enumeration with the wrong colour masked.

#include <hip/hip_runtime.h>

#include <cfloat>

#include "pg_geometry.h"
#include "pg_kernels.h"
#include "pg_interp.h"

#pragma clang fp contract(off)

namespace pgamd {

// Level schedule: N_{l+1} = (N_l + 1)/2 while that is >= 3 (at most 4 levels for N <= 18).
struct MgLevels {
  int n[4];
  int nlv;
  int nbottom;
};

constexpr MgLevels mg_levels_of(int n0) {
  MgLevels L{{n0, 0, 0, 0}, 1, 0};
  while ((L.n[L.nlv - 1] + 1) / 2 >= 3 && L.nlv < 4) {
    L.n[L.nlv] = (L.n[L.nlv - 1] + 1) / 2;
    L.nlv++;
  }
  L.nbottom = L.n[L.nlv - 1] == 3 ? 1 : 7;
  return L;
}

template <int GS>
struct MgGeo {
  static constexpr Geo g = geo_of(GS);
  static constexpr int N = g.N, R = g.R, C = g.C, U = g.U;
  static constexpr MgLevels L = mg_levels_of(N);
  static constexpr int NLV = L.nlv;
  static constexpr int n(int l) { return L.n[l]; }
  static constexpr int n3(int l) { return L.n[l] * L.n[l] * L.n[l]; }
  // LDS layout (floats): Sol0, Def0, then Sol/RHS/Def of each coarser level, then 8 reduction slots.
  static constexpr int off_sol(int l) { return l == 0 ? 0 : 2 * n3(0) + 3 * prefix(l); }
  static constexpr int off_rhs(int l) { return 2 * n3(0) + 3 * prefix(l) + n3(l); }
  static constexpr int off_def(int l) { return l == 0 ? n3(0) : 2 * n3(0) + 3 * prefix(l) + 2 * n3(l); }
  static constexpr int prefix(int l) {  // sum of n3 over levels 1..l-1
    int s = 0;
    for (int k = 1; k < l; k++) s += n3(k);
    return s;
  }
  static constexpr int off_red() { return 2 * n3(0) + 3 * prefix(NLV); }
  static constexpr int FLOATS = off_red() + 8;
  static constexpr int CELLS0 = R * R * R / 2;            // level-0 cells per colour
  static constexpr int PER_THREAD = (CELLS0 + 255) / 256;  // level-0 cells per colour per thread
};

constexpr int kMgThreads = 256;

__device__ __forceinline__ int nat(int n, int x, int y, int z) { return (z * n + y) * n + x; }

// Block-wide sums of two values in a fixed order (bitwise reproducible for a given launch shape).
template <typename T>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[4 + w] = b;
  }
  __syncthreads();
  a = ((red[0] + red[1]) + red[2]) + red[3];
  b = ((red[4] + red[5]) + red[6]) + red[7];
  __syncthreads();
}

// Level-0 cell q of colour c (cube parity) in the natural layout.
template <int GS>
__device__ __forceinline__ void cell0(int q, int c, int& x, int& y, int& z) {
  constexpr int R = MgGeo<GS>::R, H = R / 2;
  const int h = q % H;
  y = (q / H) % R + 1;
  z = q / (H * R) + 1;
  x = 2 * h + 1 + ((c + y + z + 1) & 1);
}

// Red-black Gauss-Seidel pass of one colour at level l (CPU_PoissonSolver_MG.cpp:392-441).
template <int GS, typename T>
__device__ void smooth_level0(T* __restrict__ sol, const T (&rhs)[2][MgGeo<GS>::PER_THREAD], T dh2) {
  using M = MgGeo<GS>;
  constexpr int N = M::N;
  const T one_six = T(1) / T(6);
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const int c = 1 - pass;
#pragma unroll
    for (int s = 0; s < M::PER_THREAD; s++) {
      const int q = threadIdx.x + kMgThreads * s;
      if (q < M::CELLS0) {
        int x, y, z;
        cell0<GS>(q, c, x, y, z);
        const int id = nat(N, x, y, z);
        sol[id] = one_six * (sol[id + N * N] + sol[id - N * N] + sol[id + N] + sol[id - N] + sol[id + 1] + sol[id - 1] -
                             dh2 * rhs[c][s]);
      }
    }
    __syncthreads();
  }
}

template <typename T>
__device__ void smooth_coarse(T* __restrict__ sol, const T* __restrict__ rhs, T dh2, int n) {
  const T one_six = T(1) / T(6);
  const int m = n - 2;
  for (int pass = 0; pass < 2; pass++) {
    const int c = 1 - pass;
    for (int q = threadIdx.x; q < m * m * m; q += kMgThreads) {
      const int x = q % m + 1, y = (q / m) % m + 1, z = q / (m * m) + 1;
      if (((x + y + z) & 1) != c) continue;
      const int id = nat(n, x, y, z);
      sol[id] = one_six * (sol[id + n * n] + sol[id - n * n] + sol[id + n] + sol[id - n] + sol[id + 1] + sol[id - 1] -
                           dh2 * rhs[id]);
    }
    __syncthreads();
  }
}

// Defect -(lap(sol) - rhs) on the level-0 interior, with the L1 error estimate when requested
// (CPU_PoissonSolver_MG.cpp:460-514).
template <int GS, typename T>
__device__ T defect_level0(const T* __restrict__ sol, T* __restrict__ def, const T (&rhs)[2][MgGeo<GS>::PER_THREAD],
                           T dh, bool estimate, T* red) {
  using M = MgGeo<GS>;
  constexpr int N = M::N;
  const T m_dh2 = T(-1) / (dh * dh);
  T sd = T(0), ss = T(0);
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int s = 0; s < M::PER_THREAD; s++) {
      const int q = threadIdx.x + kMgThreads * s;
      if (q < M::CELLS0) {
        int x, y, z;
        cell0<GS>(q, c, x, y, z);
        const int id = nat(N, x, y, z);
        const T d = m_dh2 * (sol[id + N * N] + sol[id - N * N] + sol[id + N] + sol[id - N] + sol[id + 1] +
                             sol[id - 1] - T(6) * sol[id]) +
                    rhs[c][s];
        def[id] = d;
        sd += fabs(d);
        ss += fabs(sol[id]);
      }
    }
  __syncthreads();
  if (!estimate) return T(0);
  block_sum2(sd, ss, red);
  return dh * dh * sd / ss;
}

template <typename T>
__device__ void defect_coarse(const T* __restrict__ sol, const T* __restrict__ rhs, T* __restrict__ def, T dh, int n) {
  const T m_dh2 = T(-1) / (dh * dh);
  const int m = n - 2;
  for (int q = threadIdx.x; q < m * m * m; q += kMgThreads) {
    const int x = q % m + 1, y = (q / m) % m + 1, z = q / (m * m) + 1;
    const int id = nat(n, x, y, z);
    def[id] = m_dh2 * (sol[id + n * n] + sol[id - n * n] + sol[id + n] + sol[id - n] + sol[id + 1] + sol[id - 1] -
                       T(6) * sol[id]) +
              rhs[id];
  }
  __syncthreads();
}

// 27-point restriction with ratio (nf-1)/(nc-1) (CPU_PoissonSolver_MG.cpp:533-616).
template <typename T>
__device__ void restrict_level(const T* __restrict__ f, T* __restrict__ c, int nf, int nc) {
  const T ratio = T(nf - 1) / T(nc - 1);
  const int m = nc - 2;
  for (int q = threadIdx.x; q < m * m * m; q += kMgThreads) {
    const int i = q % m + 1, j = (q / m) % m + 1, k = q / (m * m) + 1;
    T w[3][3];
    int ctr[3];
    const int idx[3] = {k, j, i};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const T x = T(idx[a]) * ratio;
      ctr[a] = int(x + T(0.5));
      w[a][0] = T(0.5) * (T(ctr[a]) + T(0.5) - x);
      w[a][2] = T(0.5) - w[a][0];
      w[a][1] = T(0.5);
    }
    T s = T(0);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++)
#pragma unroll
        for (int d = 0; d < 3; d++)
          s = s + w[0][a] * w[1][b] * w[2][d] * f[nat(nf, ctr[2] - 1 + d, ctr[1] - 1 + b, ctr[0] - 1 + a)];
    c[nat(nc, i, j, k)] = s;
  }
  __syncthreads();
}

// Trilinear prolongation added to the fine interior; terms in the order 000,001,010,100,011,110,101,111
// of (z,y,x) (CPU_PoissonSolver_MG.cpp:635-684).
template <typename T>
__device__ void prolongate_add(const T* __restrict__ c, T* __restrict__ f, int nc, int nf) {
  const T ratio = T(nc - 1) / T(nf - 1);
  const int m = nf - 2;
  const int order[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
  for (int q = threadIdx.x; q < m * m * m; q += kMgThreads) {
    const int i = q % m + 1, j = (q / m) % m + 1, k = q / (m * m) + 1;
    T w[3][2];
    int lo[3];
    const int idx[3] = {k, j, i};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const T x = T(idx[a]) * ratio;
      lo[a] = int(x);
      w[a][0] = T(lo[a] + 1) - x;
      w[a][1] = T(1) - w[a][0];
    }
    T s = T(0);
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int a = order[t][0], b = order[t][1], d = order[t][2];
      s = s + w[0][a] * w[1][b] * w[2][d] * c[nat(nc, lo[2] + d, lo[1] + b, lo[0] + a)];
    }
    const int id = nat(nf, i, j, k);
    f[id] = f[id] + s;
  }
  __syncthreads();
}

// One patch: interpolation, V-cycles until Error <= tol or max_iter, window and optional gravity.  PT is the
// parameter type, SolveParams (FP32) or SolveParams64 (FP64); T its element type.
template <int GS, typename PT, typename T = typename PT::real>
__device__ void mg_patch(T* __restrict__ lds, const T* __restrict__ rho, const T* __restrict__ cpot,
                         const T* __restrict__ guess, T* __restrict__ out, T* __restrict__ flu, int* iters,
                         int* status, const PT& P) {
  using M = MgGeo<GS>;
  constexpr int N = M::N, R = M::R, C = M::C, U = M::U, NLV = M::NLV;
  T* sol0 = lds + M::off_sol(0);
  T* red = lds + M::off_red();

  // interpolation into Sol0 (natural layout) and Def boundaries = 0
  for (int t = threadIdx.x; t < M::off_red(); t += kMgThreads) lds[t] = T(0);
  __syncthreads();
  {
    constexpr int NC = C - 2;
    for (int q = threadIdx.x; q < NC * NC * NC; q += kMgThreads) {
      const int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
      T v[3][3][3];
#pragma unroll
      for (int dk = 0; dk < 3; dk++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++)
#pragma unroll
          for (int di = 0; di < 3; di++) v[dk][dj][di] = cpot[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
      T ch[2][2][2];
      interp_children(v, P.int_scheme, ch);
#pragma unroll
      for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
          for (int cx = 0; cx < 2; cx++) {
            const int x = 2 * (i - 1) + cx - U, y = 2 * (j - 1) + cy - U, z = 2 * (k - 1) + cz - U;
            if (x >= 0 && x < N && y >= 0 && y < N && z >= 0 && z < N) sol0[nat(N, x, y, z)] = ch[cz][cy][cx];
          }
    }
  }
  __syncthreads();

  // level-0 right-hand side in registers; initial-guess modes
  T rhs0[2][M::PER_THREAD];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int s = 0; s < M::PER_THREAD; s++) {
      const int q = threadIdx.x + kMgThreads * s;
      rhs0[c][s] = T(0);
      if (q < M::CELLS0) {
        int x, y, z;
        cell0<GS>(q, c, x, y, z);
        const int gi = ((z - 1) * R + (y - 1)) * R + (x - 1);
        rhs0[c][s] = P.poi_coeff * rho[gi];  // CPU_PoissonSolver_MG.cpp:200-207
        if (P.init_mode == 1) sol0[nat(N, x, y, z)] = T(0);
        if (P.init_mode == 2) sol0[nat(N, x, y, z)] = guess[gi];
        if (P.init_mode == 3) sol0[nat(N, x, y, z)] = guess[nat(N, x, y, z)];   // whole-cube guess: its interior
      }
    }
  __syncthreads();

  // Reject a caller guess whose relative residual sum|R| / sum|phi| (the SOR residual in the reference
  // operand order, Const = Poi_Coeff dh^2, over the interior) exceeds guess_tol: start from the zero guess,
  // as the SOR kernel does, and report it in the status.
  bool fell_back = false;
  if (P.init_mode >= 2 && P.guess_tol > T(0)) {
    T num = T(0), den = T(0);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int s = 0; s < M::PER_THREAD; s++) {
        const int q = threadIdx.x + kMgThreads * s;
        if (q < M::CELLS0) {
          int x, y, z;
          cell0<GS>(q, c, x, y, z);
          const T* p = sol0 + nat(N, x, y, z);
          const T r = p[N * N] + p[-N * N] + p[N] + p[-N] + p[1] + p[-1] - T(6) * p[0] -
                      P.cst * rho[((z - 1) * R + (y - 1)) * R + (x - 1)];
          num += fabs(r);
          den += fabs(p[0]);
        }
      }
    block_sum2(num, den, red);
    fell_back = !(num <= P.guess_tol * den);
    if (fell_back) {
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int s = 0; s < M::PER_THREAD; s++) {
          const int q = threadIdx.x + kMgThreads * s;
          if (q < M::CELLS0) {
            int x, y, z;
            cell0<GS>(q, c, x, y, z);
            sol0[nat(N, x, y, z)] = T(0);
          }
        }
    }
    __syncthreads();
  }

  T dh[4];
#pragma unroll
  for (int l = 0; l < 4; l++) dh[l] = l == 0 ? P.dh : P.dh * T(N - 1) / T(l < NLV ? M::n(l) - 1 : 1);

  int iter = 0;
  T error = T(FLT_MAX);
  while (iter < P.mg_max_iter && error > P.mg_tol) {
    // down: pre-smooth, defect, restrict, zero the next correction
#pragma unroll
    for (int l = 0; l < NLV - 1; l++) {
      const int n = M::n(l);
      T* sol = lds + M::off_sol(l);
      T* def = lds + M::off_def(l);
      if (l == 0) {
        for (int s = 0; s < P.mg_npre; s++) smooth_level0<GS>(sol, rhs0, dh[0] * dh[0]);
        defect_level0<GS>(sol, def, rhs0, dh[0], false, red);
      } else {
        for (int s = 0; s < P.mg_npre; s++) smooth_coarse(sol, lds + M::off_rhs(l), dh[l] * dh[l], n);
        defect_coarse(sol, lds + M::off_rhs(l), def, dh[l], n);
      }
      restrict_level(def, lds + M::off_rhs(l + 1), n, M::n(l + 1));
      T* nsol = lds + M::off_sol(l + 1);
      for (int t = threadIdx.x; t < M::n3(l + 1); t += kMgThreads) nsol[t] = T(0);
      __syncthreads();
    }
    // bottom
    {
      constexpr int b = NLV - 1;
      for (int s = 0; s < M::L.nbottom; s++) smooth_coarse(lds + M::off_sol(b), lds + M::off_rhs(b), dh[b] * dh[b], M::n(b));
    }
    // up: prolongate and correct, post-smooth
#pragma unroll
    for (int l = NLV - 2; l >= 0; l--) {
      prolongate_add(lds + M::off_sol(l + 1), lds + M::off_sol(l), M::n(l + 1), M::n(l));
      if (l == 0) {
        for (int s = 0; s < P.mg_npost; s++) smooth_level0<GS>(sol0, rhs0, dh[0] * dh[0]);
      } else {
        for (int s = 0; s < P.mg_npost; s++)
          smooth_coarse(lds + M::off_sol(l), lds + M::off_rhs(l), dh[l] * dh[l], M::n(l));
      }
    }
    error = defect_level0<GS>(sol0, lds + M::off_def(0), rhs0, dh[0], true, red);
    iter++;
  }

  // status: converged / max_iter / non-finite
  if (threadIdx.x == 0) {
    int st = error <= P.mg_tol ? kStatusConverged : kStatusMaxIter;
    if (!(error == error) || error == T(INFINITY)) st |= kStatusNonFinite;
    if (fell_back) st |= kStatusFallback;
    if (iters) *iters = iter;
    if (status) *status = st;
  }

  // window (CPU_PoissonSolver_MG.cpp:348-358) and fused gravity
  const int G = gra_nxt(P.gra_ghost), o = GS - P.gra_ghost;
  for (int t = threadIdx.x; t < G * G * G; t += kMgThreads) {
    const int i = t % G, j = (t / G) % G, k = t / (G * G);
    out[t] = sol0[nat(N, i + o, j + o, k + o)];
  }
  if (P.fuse_gravity && flu) {
    for (int t = threadIdx.x; t < kPatchCells; t += kMgThreads) {
      const int i = t % kPatchSize, j = (t / kPatchSize) % kPatchSize, k = t / (kPatchSize * kPatchSize);
      const int x = i + GS, y = j + GS, z = k + GS;
      auto Q = [&](int dz, int dy, int dx) { return sol0[nat(N, x + dx, y + dy, z + dz)]; };
      T ax, ay, az;
      if (P.p5) {
        ax = P.gra_const * (-Q(0, 0, 2) + Q(0, 0, -2) + T(8) * Q(0, 0, 1) - T(8) * Q(0, 0, -1));
        ay = P.gra_const * (-Q(0, 2, 0) + Q(0, -2, 0) + T(8) * Q(0, 1, 0) - T(8) * Q(0, -1, 0));
        az = P.gra_const * (-Q(2, 0, 0) + Q(-2, 0, 0) + T(8) * Q(1, 0, 0) - T(8) * Q(-1, 0, 0));
      } else {
        ax = P.gra_const * (Q(0, 0, 1) - Q(0, 0, -1));
        ay = P.gra_const * (Q(0, 1, 0) - Q(0, -1, 0));
        az = P.gra_const * (Q(1, 0, 0) - Q(-1, 0, 0));
      }
      const T rho_f = flu[t];
      T px = flu[kPatchCells + t], py = flu[2 * kPatchCells + t], pz = flu[3 * kPatchCells + t];
      const T temp = T(0.5) / rho_f;
      const T eint = flu[4 * kPatchCells + t] - temp * (px * px + py * py + pz * pz);
      px = px - rho_f * ax;
      py = py - rho_f * ay;
      pz = pz - rho_f * az;
      flu[kPatchCells + t] = px;
      flu[2 * kPatchCells + t] = py;
      flu[3 * kPatchCells + t] = pz;
      flu[4 * kPatchCells + t] = eint + temp * (px * px + py * py + pz * pz);
    }
  }
  __syncthreads();
}

template <int GS, typename PT>
__global__ void __launch_bounds__(kMgThreads) mg_uniform_kernel(UniformBatchT<typename PT::real> b, PT P) {
  using T = typename PT::real;
  extern __shared__ __align__(16) unsigned char mg_smem[];
  T* lds = reinterpret_cast<T*>(mg_smem);
  using M = MgGeo<GS>;
  constexpr int R3 = M::R * M::R * M::R, C3 = M::C * M::C * M::C;
  const int p = blockIdx.x;
  if (p >= b.npatch) return;
  const int G = gra_nxt(P.gra_ghost);
  constexpr int N3 = M::N * M::N * M::N;
  mg_patch<GS>(lds, b.rho + (size_t)p * R3, b.cpot + (size_t)p * C3,
               b.guess ? b.guess + (size_t)p * (P.init_mode == 3 ? N3 : R3) : nullptr,
               b.pot_out + (size_t)p * G * G * G, b.flu ? b.flu + (size_t)p * kNComp * kPatchCells : nullptr,
               b.iters ? b.iters + p : nullptr, b.status ? b.status + p : nullptr, P);
}

size_t lds_bytes_mg(int gs) {
  switch (gs) {
    case 1: return sizeof(float) * MgGeo<1>::FLOATS;
    case 2: return sizeof(float) * MgGeo<2>::FLOATS;
    case 3: return sizeof(float) * MgGeo<3>::FLOATS;
    case 4: return sizeof(float) * MgGeo<4>::FLOATS;
    case 5: return sizeof(float) * MgGeo<5>::FLOATS;
  }
  return 0;
}

template <int GS, typename PT>
static hipError_t launch_mg(const UniformBatchT<typename PT::real>& b, const PT& p, hipStream_t s) {
  const size_t lds = lds_bytes_mg(GS) / sizeof(float) * sizeof(typename PT::real);
  if (lds > 64 * 1024)   // the FP64 levels of the larger cubes exceed the default dynamic LDS limit
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&mg_uniform_kernel<GS, PT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((mg_uniform_kernel<GS, PT>), dim3(b.npatch), dim3(kMgThreads), lds, s, b, p);
  return hipGetLastError();
}

// Persistent MG kernel of a mixed batch: each workgroup pulls one patch at a time from the queue.
__global__ void __launch_bounds__(kMgThreads) mg_persistent_kernel(MixedBatch b, ParamsByGs P) {
  extern __shared__ float lds[];
  __shared__ int s_task;
  for (;;) {
    if (threadIdx.x == 0) s_task = (int)atomicAdd(b.counter, 1u);
    __syncthreads();
    const int t = s_task;
    __syncthreads();
    if (t >= b.ntasks) break;
    const Task tk = b.tasks[t];
    const int patch = b.perm[tk.first];
    if (patch < 0 || patch >= b.npatch || tk.gs < 1 || tk.gs > kMaxGhost) continue;  // malformed plan: skip
    const PatchDesc d = b.desc[patch];
    const SolveParams& Q = P.p[tk.gs];
    const int G = gra_nxt(Q.gra_ghost);
    const float* rho = b.rho + d.rho_off;
    const float* cpot = b.cpot + d.cpot_off;
    const float* guess = b.guess ? b.guess + d.rho_off : nullptr;
    float* out = b.pot_out + (size_t)patch * G * G * G;
    float* flu = b.flu ? b.flu + (size_t)patch * kNComp * kPatchCells : nullptr;
    int* it = b.iters ? b.iters + patch : nullptr;
    int* st = b.status ? b.status + patch : nullptr;
    switch (tk.gs) {
      case 1: mg_patch<1>(lds, rho, cpot, guess, out, flu, it, st, Q); break;
      case 2: mg_patch<2>(lds, rho, cpot, guess, out, flu, it, st, Q); break;
      case 3: mg_patch<3>(lds, rho, cpot, guess, out, flu, it, st, Q); break;
      case 4: mg_patch<4>(lds, rho, cpot, guess, out, flu, it, st, Q); break;
      default: mg_patch<5>(lds, rho, cpot, guess, out, flu, it, st, Q); break;
    }
  }
}

// The MG tasks [first, first + gridDim.x) of one ghost size, one patch per workgroup (no task loop: the
// patch solve compiles as in the uniform kernel, without the persistent kernel's spills).
template <int GS>
__global__ void __launch_bounds__(kMgThreads) mg_tasks_kernel(MixedBatch b, int first, SolveParams Q) {
  extern __shared__ float lds[];
  const Task tk = b.tasks[first + blockIdx.x];
  const int patch = b.perm[tk.first];
  if (patch < 0 || patch >= b.npatch || tk.gs != GS) return;   // malformed plan: skip (uniform per workgroup)
  const PatchDesc d = b.desc[patch];
  const int G = gra_nxt(Q.gra_ghost);
  mg_patch<GS>(lds, b.rho + d.rho_off, b.cpot + d.cpot_off, b.guess ? b.guess + d.rho_off : nullptr,
               b.pot_out + (size_t)patch * G * G * G, b.flu ? b.flu + (size_t)patch * kNComp * kPatchCells : nullptr,
               b.iters ? b.iters + patch : nullptr, b.status ? b.status + patch : nullptr, Q);
}

hipError_t launch_mg_tasks(const MixedBatch& b, const SolveParams& p, int gs, int first, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const dim3 grid(count), block(kMgThreads);
  switch (gs) {
    case 1: hipLaunchKernelGGL((mg_tasks_kernel<1>), grid, block, lds_bytes_mg(1), s, b, first, p); break;
    case 2: hipLaunchKernelGGL((mg_tasks_kernel<2>), grid, block, lds_bytes_mg(2), s, b, first, p); break;
    case 3: hipLaunchKernelGGL((mg_tasks_kernel<3>), grid, block, lds_bytes_mg(3), s, b, first, p); break;
    case 4: hipLaunchKernelGGL((mg_tasks_kernel<4>), grid, block, lds_bytes_mg(4), s, b, first, p); break;
    case 5: hipLaunchKernelGGL((mg_tasks_kernel<5>), grid, block, lds_bytes_mg(5), s, b, first, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int mg_persistent_blocks_per_cu(size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(&mg_persistent_kernel),
                                                   kMgThreads, lds) != hipSuccess)
    return 0;
  return n;
}

hipError_t launch_mg_persistent(const MixedBatch& b, const ParamsByGs& p, int grid, size_t lds, hipStream_t s) {
  if (b.ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(mg_persistent_kernel, dim3(grid), dim3(kMgThreads), lds, s, b, p);
  return hipGetLastError();
}

hipError_t launch_mg_uniform(const UniformBatch& b, const SolveParams& p, hipStream_t s) {
  switch (b.gs) {
    case 1: return launch_mg<1>(b, p, s);
    case 2: return launch_mg<2>(b, p, s);
    case 3: return launch_mg<3>(b, p, s);
    case 4: return launch_mg<4>(b, p, s);
    case 5: return launch_mg<5>(b, p, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_mg_uniform64(const UniformBatch64& b, const SolveParams64& p, hipStream_t s) {
  switch (b.gs) {
    case 1: return launch_mg<1>(b, p, s);
    case 2: return launch_mg<2>(b, p, s);
    case 3: return launch_mg<3>(b, p, s);
    case 4: return launch_mg<4>(b, p, s);
    case 5: return launch_mg<5>(b, p, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace pgamd
