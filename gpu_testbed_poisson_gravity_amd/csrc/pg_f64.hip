// FP64 red-black SOR and gravity on gfx950: the reference's FLOAT8 build (Makefile:22, Header/Macro.h:35), in
// which every array and constant is double.  The MG kernel has its FP64 instances in pg_mg.hip.
//
// Reference behaviour reproduced (bit for bit in FP64 when the same iteration count is run):
//   interpolation  CPU_PoissonSolver_SOR.cpp:110-203
//   SOR sweep      CPU_PoissonSolver_SOR.cpp:215-270
//   stall rule     CPU_PoissonSolver_SOR.cpp:273-283
//   window         CPU_PoissonSolver_SOR.cpp:303-311
//   gravity        CPU_GravitySolver.cpp:34-79
//
// Layout.  One workgroup of 256 threads per patch; the N^3 solve cube sits in LDS in the natural layout
// (46.7 KB at ghost size 5, three workgroups per CU).  Thread t updates the interior cells t, t + 256, ... of
// each colour and keeps their Const*rho in registers; every neighbour comes from LDS.  A colour pass ends with
// a barrier, and the sums the stopping rules need are block reductions in a fixed order.

#include <hip/hip_runtime.h>

#include <cfloat>

#include "pg_geometry.h"
#include "pg_interp.h"
#include "pg_kernels.h"

#pragma clang fp contract(off)

namespace pgamd {
namespace {

constexpr int kT64 = 256;

template <int GS>
struct Geo64 {
  static constexpr Geo g = geo_of(GS);
  static constexpr int N = g.N, R = g.R, C = g.C, U = g.U;
  static constexpr int CELLS = R * R * R / 2;            // interior cells per colour (R is even)
  static constexpr int PER = (CELLS + kT64 - 1) / kT64;  // per thread
};

__device__ __forceinline__ int nat3(int n, int x, int y, int z) { return (z * n + y) * n + x; }

// Interior cell q of colour c, i.e. (x + y + z) & 1 == c, in cube coordinates 1..R.
template <int GS>
__device__ __forceinline__ void cell_of(int q, int c, int& x, int& y, int& z) {
  constexpr int R = Geo64<GS>::R, H = R / 2;
  const int h = q % H;
  y = (q / H) % R + 1;
  z = q / (H * R) + 1;
  x = 2 * h + 1 + ((c + y + z + 1) & 1);
}

// Sums of a and b over the workgroup in a fixed order (red: 8 doubles of LDS).
__device__ __forceinline__ void block_sum2_f64(double& a, double& b, double* red) {
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

// SOR residual at cube cell id, reference operand order (CPU_PoissonSolver_SOR.cpp:247-250).
template <int N>
__device__ __forceinline__ double residual64(const double* c, int id, double cr) {
  return c[id + N * N] + c[id - N * N] + c[id + N] + c[id - N] + c[id + 1] + c[id - 1] - 6.0 * c[id] - cr;
}

// One colour pass: updates the thread's cells of colour c, returns the thread's sum of |R| (pre-update).
template <int GS, bool UPDATE>
__device__ __forceinline__ double colour_pass(double* cube, const double (&cr)[2][Geo64<GS>::PER], int c, double w6) {
  using G = Geo64<GS>;
  constexpr int N = G::N;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < G::PER; k++) {
    const int q = threadIdx.x + kT64 * k;
    if (q < G::CELLS) {
      int x, y, z;
      cell_of<GS>(q, c, x, y, z);
      const int id = nat3(N, x, y, z);
      const double r = residual64<N>(cube, id, cr[c][k]);
      if (UPDATE) cube[id] = cube[id] + w6 * r;
      s += fabs(r);
    }
  }
  return s;
}

// Gravity update of one patch from the potential in the LDS cube (CPU_GravitySolver.cpp:34-79).
template <int GS>
__device__ void gravity64_from_cube(const double* cube, double* flu, double gra_const, int p5) {
  constexpr int N = Geo64<GS>::N;
  for (int t = threadIdx.x; t < kPatchCells; t += kT64) {
    const int i = t % kPatchSize, j = (t / kPatchSize) % kPatchSize, k = t / (kPatchSize * kPatchSize);
    const int x = i + GS, y = j + GS, z = k + GS;
    auto Q = [&](int dz, int dy, int dx) { return cube[nat3(N, x + dx, y + dy, z + dz)]; };
    double ax, ay, az;
    if (p5) {
      ax = gra_const * (-Q(0, 0, 2) + Q(0, 0, -2) + 8.0 * Q(0, 0, 1) - 8.0 * Q(0, 0, -1));
      ay = gra_const * (-Q(0, 2, 0) + Q(0, -2, 0) + 8.0 * Q(0, 1, 0) - 8.0 * Q(0, -1, 0));
      az = gra_const * (-Q(2, 0, 0) + Q(-2, 0, 0) + 8.0 * Q(1, 0, 0) - 8.0 * Q(-1, 0, 0));
    } else {
      ax = gra_const * (Q(0, 0, 1) - Q(0, 0, -1));
      ay = gra_const * (Q(0, 1, 0) - Q(0, -1, 0));
      az = gra_const * (Q(1, 0, 0) - Q(-1, 0, 0));
    }
    const double rho = flu[t];
    double px = flu[kPatchCells + t], py = flu[2 * kPatchCells + t], pz = flu[3 * kPatchCells + t];
    const double temp = 0.5 / rho;
    const double eint = flu[4 * kPatchCells + t] - temp * (px * px + py * py + pz * pz);
    px = px - rho * ax;
    py = py - rho * ay;
    pz = pz - rho * az;
    flu[kPatchCells + t] = px;
    flu[2 * kPatchCells + t] = py;
    flu[3 * kPatchCells + t] = pz;
    flu[4 * kPatchCells + t] = eint + temp * (px * px + py * py + pz * pz);
  }
}

template <int GS>
__global__ void __launch_bounds__(kT64) sor64_kernel(UniformBatch64 b, SolveParams64 P) {
  using G = Geo64<GS>;
  constexpr int N = G::N, R = G::R, C = G::C, U = G::U, NC = C - 2;
  constexpr int N3 = N * N * N, R3 = R * R * R, C3 = C * C * C;
  extern __shared__ __align__(16) unsigned char smem64[];
  double* cube = reinterpret_cast<double*>(smem64);
  double* red = cube + N3;
  const int p = blockIdx.x;
  if (p >= b.npatch) return;
  const double* rho = b.rho + (size_t)p * R3;
  const double* cpot = b.cpot + (size_t)p * C3;
  const double* guess = b.guess ? b.guess + (size_t)p * (P.init_mode == 3 ? N3 : R3) : nullptr;

  // A. the solve cube: the interpolated coarse potential (the reference's start and the Dirichlet layer), or
  // the caller's whole cube
  if (P.init_mode == 3) {
    for (int t = threadIdx.x; t < N3; t += kT64) cube[t] = guess[t];
  } else {
    for (int q = threadIdx.x; q < NC * NC * NC; q += kT64) {
      const int i = q % NC + 1, j = (q / NC) % NC + 1, k = q / (NC * NC) + 1;
      double v[3][3][3];
#pragma unroll
      for (int dk = 0; dk < 3; dk++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++)
#pragma unroll
          for (int di = 0; di < 3; di++) v[dk][dj][di] = cpot[((k + dk - 1) * C + (j + dj - 1)) * C + (i + di - 1)];
      double ch[2][2][2];
      interp_children(v, P.int_scheme, ch);
#pragma unroll
      for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
          for (int cx = 0; cx < 2; cx++) {
            const int x = 2 * (i - 1) + cx - U, y = 2 * (j - 1) + cy - U, z = 2 * (k - 1) + cz - U;
            if (x >= 0 && x < N && y >= 0 && y < N && z >= 0 && z < N) cube[nat3(N, x, y, z)] = ch[cz][cy][cx];
          }
    }
  }
  __syncthreads();

  // B. Const*rho of the owned cells; the zero start and the interior guess replace the interior
  double cr[2][G::PER];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int k = 0; k < G::PER; k++) {
      const int q = threadIdx.x + kT64 * k;
      cr[c][k] = 0.0;
      if (q < G::CELLS) {
        int x, y, z;
        cell_of<GS>(q, c, x, y, z);
        const int gi = ((z - 1) * R + (y - 1)) * R + (x - 1);
        cr[c][k] = P.cst * rho[gi];
        if (P.init_mode == 1) cube[nat3(N, x, y, z)] = 0.0;
        if (P.init_mode == 2) cube[nat3(N, x, y, z)] = guess[gi];
      }
    }
  __syncthreads();

  // B'. a caller guess whose relative residual exceeds guess_tol is replaced by the zero guess
  bool fell_back = false;
  if (P.init_mode >= 2 && P.guess_tol > 0.0) {
    double num = colour_pass<GS, false>(cube, cr, 1, 0.0) + colour_pass<GS, false>(cube, cr, 0, 0.0), den = 0.0;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int k = 0; k < G::PER; k++) {
        const int q = threadIdx.x + kT64 * k;
        if (q < G::CELLS) {
          int x, y, z;
          cell_of<GS>(q, c, x, y, z);
          den += fabs(cube[nat3(N, x, y, z)]);
        }
      }
    block_sum2_f64(num, den, red);
    fell_back = !(num <= P.guess_tol * den);
    if (fell_back) {
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int k = 0; k < G::PER; k++) {
          const int q = threadIdx.x + kT64 * k;
          if (q < G::CELLS) {
            int x, y, z;
            cell_of<GS>(q, c, x, y, z);
            cube[nat3(N, x, y, z)] = 0.0;
          }
        }
    }
    __syncthreads();
  }

  // C. red-black iterations: pass 0 updates the cells with x + y + z odd (colour 1), pass 1 the others
  const bool stall = P.term_mode == 0, relres = P.term_mode == 1;
  const int first_check = P.min_iter > 1 ? P.min_iter : 1;
  int it_done = P.max_iter, st = kStatusMaxIter;
  double old = (double)FLT_MAX, sum_b = 0.0;   // Residual_Total_Old = __FLT_MAX__ (CPU_PoissonSolver_SOR.cpp:215)
  for (int t = 0; t < P.max_iter; t++) {
    if (relres && t >= first_check && (t - first_check) % P.check_every == 0) {
      // first-colour residuals before the pass, second-colour ones |1 - omega| times those of the last pass
      double num = colour_pass<GS, false>(cube, cr, 1, 0.0), den = 0.0;
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int k = 0; k < G::PER; k++) {
          const int q = threadIdx.x + kT64 * k;
          if (q < G::CELLS) {
            int x, y, z;
            cell_of<GS>(q, c, x, y, z);
            den += fabs(cube[nat3(N, x, y, z)]);
          }
        }
      block_sum2_f64(num, den, red);
      if (num + P.damp * sum_b <= P.rel_tol * den) {
        it_done = t;
        st = kStatusConverged;
        break;
      }
    }
    double tot = colour_pass<GS, true>(cube, cr, 1, P.w6);
    __syncthreads();
    double tb = colour_pass<GS, true>(cube, cr, 0, P.w6);
    __syncthreads();
    if (relres) {
      const int u = t + 1;
      if (u >= first_check && (u - first_check) % P.check_every == 0) {
        double zero = 0.0;
        block_sum2_f64(tb, zero, red);
        sum_b = tb;
      }
    } else if (stall) {
      double all = tot + tb, zero = 0.0;
      block_sum2_f64(all, zero, red);
      if (t + 1 >= P.min_iter && all > old) {
        it_done = t + 1;
        st = kStatusConverged;
        break;
      }
      old = all;
    }
  }

  // non-finite check over the interior
  bool bad = false;
  for (int t = threadIdx.x; t < R3; t += kT64) {
    const int x = t % R + 1, y = (t / R) % R + 1, z = t / (R * R) + 1;
    bad |= !isfinite(cube[nat3(N, x, y, z)]);
  }
  bad = __syncthreads_or(bad);
  if (bad) st |= kStatusNonFinite;
  if (fell_back) st |= kStatusFallback;
  if (threadIdx.x == 0) {
    if (b.iters) b.iters[p] = it_done;
    if (b.status) b.status[p] = st;
  }

  // D. output window and the fused gravity update
  const int Gw = gra_nxt(P.gra_ghost), o = GS - P.gra_ghost;
  double* out = b.pot_out + (size_t)p * Gw * Gw * Gw;
  for (int t = threadIdx.x; t < Gw * Gw * Gw; t += kT64) {
    const int i = t % Gw, j = (t / Gw) % Gw, k = t / (Gw * Gw);
    out[t] = cube[nat3(N, i + o, j + o, k + o)];
  }
  if (P.fuse_gravity && b.flu) gravity64_from_cube<GS>(cube, b.flu + (size_t)p * kNComp * kPatchCells, P.gra_const, P.p5);
}

// Standalone FP64 gravity update from an output window in HBM (the Poisson = false path).
__global__ void __launch_bounds__(kT64) gravity64_kernel(const double* __restrict__ pot, double* __restrict__ flu,
                                                         int npatch, int gg, double gra_const, int p5) {
  const int P = blockIdx.x;
  if (P >= npatch) return;
  const int G = gra_nxt(gg);
  const double* pp = pot + (size_t)P * G * G * G;
  double* f = flu + (size_t)P * kNComp * kPatchCells;
  for (int t = threadIdx.x; t < kPatchCells; t += kT64) {
    const int i = t % kPatchSize, j = (t / kPatchSize) % kPatchSize, k = t / (kPatchSize * kPatchSize);
    const int x = i + gg, y = j + gg, z = k + gg;
    auto Q = [&](int dz, int dy, int dx) { return pp[((z + dz) * G + (y + dy)) * G + (x + dx)]; };
    double ax, ay, az;
    if (p5) {
      ax = gra_const * (-Q(0, 0, 2) + Q(0, 0, -2) + 8.0 * Q(0, 0, 1) - 8.0 * Q(0, 0, -1));
      ay = gra_const * (-Q(0, 2, 0) + Q(0, -2, 0) + 8.0 * Q(0, 1, 0) - 8.0 * Q(0, -1, 0));
      az = gra_const * (-Q(2, 0, 0) + Q(-2, 0, 0) + 8.0 * Q(1, 0, 0) - 8.0 * Q(-1, 0, 0));
    } else {
      ax = gra_const * (Q(0, 0, 1) - Q(0, 0, -1));
      ay = gra_const * (Q(0, 1, 0) - Q(0, -1, 0));
      az = gra_const * (Q(1, 0, 0) - Q(-1, 0, 0));
    }
    const double rho = f[t];
    double px = f[kPatchCells + t], py = f[2 * kPatchCells + t], pz = f[3 * kPatchCells + t];
    const double temp = 0.5 / rho;
    const double eint = f[4 * kPatchCells + t] - temp * (px * px + py * py + pz * pz);
    px = px - rho * ax;
    py = py - rho * ay;
    pz = pz - rho * az;
    f[kPatchCells + t] = px;
    f[2 * kPatchCells + t] = py;
    f[3 * kPatchCells + t] = pz;
    f[4 * kPatchCells + t] = eint + temp * (px * px + py * py + pz * pz);
  }
}

template <int GS>
hipError_t launch_sor64(const UniformBatch64& b, const SolveParams64& p, hipStream_t s) {
  constexpr int N = Geo64<GS>::N;
  const size_t lds = sizeof(double) * (N * N * N + 8);
  hipLaunchKernelGGL((sor64_kernel<GS>), dim3(b.npatch), dim3(kT64), lds, s, b, p);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_uniform64(const UniformBatch64& b, const SolveParams64& p, hipStream_t s) {
  if (b.npatch <= 0) return hipSuccess;
  if (b.scheme == kSchemeMG) return launch_mg_uniform64(b, p, s);
  switch (b.gs) {
    case 1: return launch_sor64<1>(b, p, s);
    case 2: return launch_sor64<2>(b, p, s);
    case 3: return launch_sor64<3>(b, p, s);
    case 4: return launch_sor64<4>(b, p, s);
    case 5: return launch_sor64<5>(b, p, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_gravity64(const double* pot, double* flu, int npatch, int gra_ghost, double gra_const, int p5,
                            hipStream_t stream) {
  if (npatch <= 0) return hipSuccess;
  hipLaunchKernelGGL(gravity64_kernel, dim3(npatch), dim3(kT64), 0, stream, pot, flu, npatch, gra_ghost, gra_const,
                     p5);
  return hipGetLastError();
}

}  // namespace pgamd
