// Launch-side view of the batched Poisson-gravity kernels (host code includes this; no HIP types in the
// structs so that the C-ABI layer and the autotuner can fill them directly).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace pgamd {

// SOR layout code for the compact single-buffer LDS layout (ghost size 5, one patch per wavefront), in place
// of a split-row padding of 0..3 floats.  Same value as PGAMD_LAYOUT_COMPACT of the C ABI.
constexpr int kPadCompact = 4;

// Uniform per-launch solver settings.  Floating-point constants are formed once on the host, in FP32, the
// way the reference CPU solver forms them (CPU_PoissonSolver_SOR.cpp:51-52), so that device and CPU agree
// bit for bit.
struct SolveParams {
  using real = float;
  float cst;          // Poi_Coeff*dh*dh
  float w6;           // omega/6
  float damp;         // |1 - omega|, second-colour residual factor of the relative-residual rule
  float rel_tol;      // relative-residual tolerance (term_mode 1)
  int min_iter;
  int max_iter;
  int term_mode;      // 0: reference stall rule; 1: relative residual
  int check_every;    // relative-residual cadence
  int init_mode;      // 0: interpolated cube; 1: zero interior; 2: caller guess of the interior [R^3];
                      // 3: caller guess of the whole cube [N^3] (SOR also takes its Dirichlet layer)
  int int_scheme;     // 4 CQUAD, 5 QUAD
  int gra_ghost;      // GRA_GHOST_SIZE of the output window (1..gs; > 2 only without gravity)
  // multigrid
  float poi_coeff;
  float dh;
  float mg_tol;
  int mg_max_iter;
  int mg_npre;
  int mg_npost;
  // gravity (fused after the solve when fuse_gravity != 0)
  int fuse_gravity;
  int p5;
  float gra_const;
  // initial-guess check (init_mode 2, 3): a guess whose relative residual exceeds guess_tol is replaced by
  // the zero guess before the first sweep; 0 disables the check
  float guess_tol;
};

// FP64 settings (the reference's FLOAT8 build, Makefile:22, Header/Macro.h:35): the same fields, with every
// floating-point constant formed in double on the host.
struct SolveParams64 {
  using real = double;
  double cst;
  double w6;
  double damp;
  double rel_tol;
  int min_iter;
  int max_iter;
  int term_mode;
  int check_every;
  int init_mode;
  int int_scheme;
  int gra_ghost;
  double poi_coeff;
  double dh;
  double mg_tol;
  int mg_max_iter;
  int mg_npre;
  int mg_npost;
  int fuse_gravity;
  int p5;
  double gra_const;
  double guess_tol;
};

// Status bits written per patch.
constexpr int kStatusConverged = 1;  // the stopping rule was met before max_iter
constexpr int kStatusMaxIter = 2;    // max_iter reached
constexpr int kStatusNonFinite = 4;  // a non-finite value is present in the returned potential
constexpr int kStatusFallback = 8;   // the caller's guess failed the residual check; the zero guess was used

// One batch of patches sharing a ghost size, with dense per-patch strides (the reference layout); T is the
// element type (float, or double for the FP64 kernels).
template <typename T>
struct UniformBatchT {
  const T* rho;    // [n][R^3]
  const T* cpot;   // [n][C^3]
  const T* guess;  // [n][R^3] (init_mode 2), [n][N^3] (init_mode 3) or null
  T* pot_out;      // [n][G^3]
  T* flu;          // [n][5][8^3] or null
  int* iters;      // [n] or null
  int* status;     // [n] or null
  int npatch;
  int gs;
  int scheme;
};
using UniformBatch = UniformBatchT<float>;
using UniformBatch64 = UniformBatchT<double>;

// A patch of a mixed batch: its ghost size, its method and element offsets into packed arrays.
struct PatchDesc {
  int gs;
  int scheme;
  int64_t rho_off;     // into the packed density (and guess) array
  int64_t cpot_off;    // into the packed coarse-potential array
};

// Work item of the persistent kernel: up to `count` patches of one (gs, scheme) class, listed by index
// in a permutation array.  Built on the host by the scheduler.
struct Task {
  int gs;
  int scheme;
  int first;           // index into the permutation array
  int count;
};

struct MixedBatch {
  const PatchDesc* desc;   // [npatch]
  const int* perm;         // patch indices grouped by task
  const Task* tasks;       // [ntasks], largest work first
  int ntasks;
  int npatch;
  const float* rho;        // packed
  const float* cpot;       // packed
  const float* guess;      // packed like rho, or null
  float* pot_out;          // [npatch][G^3]
  float* flu;              // [npatch][5][8^3] or null
  int* iters;
  int* status;
  unsigned int* counter;   // work-queue head, reset by the launcher
};

// Solver settings per ghost size for a mixed batch (omega may differ per ghost size).
struct ParamsByGs {
  SolveParams p[6];
};

// Per-ghost-size launch layout of the SOR kernel: patches per wavefront.  Chosen by the autotuner;
// the default is the densest packing that fits one wavefront.
struct SorLayout {
  int ppw[6];              // patches per wavefront, indexed by gs
  int pad[6];              // split-row padding in floats (bank-conflict avoidance), or kPadCompact, by gs
  int waves_per_cu;        // persistent kernel occupancy target
};

// Launchers (pg_kernels.hip).  All are asynchronous on `stream`.
hipError_t launch_uniform(const UniformBatch& b, const SolveParams& p, const SorLayout& lay, hipStream_t stream);
hipError_t launch_gravity(const float* pot, float* flu, int npatch, int gra_ghost, float gra_const, int p5,
                          hipStream_t stream);
// Persistent kernels of a mixed batch: SOR tasks (one wavefront each, gs 1/2 packed ppw1/ppw2 per
// wavefront, default row padding) and MG tasks (one 256-thread workgroup per patch).
hipError_t launch_sor_persistent(const MixedBatch& b, const ParamsByGs& p, int ppw1, int ppw2, int grid, size_t lds,
                                 hipStream_t stream);
hipError_t launch_mg_persistent(const MixedBatch& b, const ParamsByGs& p, int grid, size_t lds, hipStream_t stream);
int sor_persistent_blocks_per_cu(int ppw1, int ppw2, size_t lds);
// Resident workgroups per CU of the uniform SOR kernel for a layout (0 if the layout does not exist).
int sor_uniform_blocks_per_cu(int gs, int ppw, int pad);
// Per-ghost-size launches of a mixed batch (the default): the SOR tasks [first, first + count) of ghost size
// gs (ppw patches per wavefront, LDS row padding pad), and the MG tasks likewise.
hipError_t launch_sor_tasks(const MixedBatch& b, const SolveParams& p, int gs, int ppw, int pad, int first, int count,
                            hipStream_t stream);
hipError_t launch_mg_tasks(const MixedBatch& b, const SolveParams& p, int gs, int first, int count, hipStream_t stream);
int mg_persistent_blocks_per_cu(size_t lds);
// Zeroes n work-queue counters with a kernel (graph-capture safe, no memset node).
hipError_t launch_reset_counters(unsigned int* counters, int n, hipStream_t stream);
int sor_lanes_per_patch(int gs);
int default_ppw(int gs);
int default_pad(int gs);
bool supported_layout(int gs, int ppw, int pad);
size_t lds_bytes_sor(int gs, int ppw, int pad);
size_t lds_bytes_mg(int gs);
int supported_ppw(int gs, int ppw);
// FP64 kernels (pg_f64.hip and the FP64 instances of the MG kernel): one workgroup per patch.
hipError_t launch_uniform64(const UniformBatch64& b, const SolveParams64& p, hipStream_t stream);
hipError_t launch_gravity64(const double* pot, double* flu, int npatch, int gra_ghost, double gra_const, int p5,
                            hipStream_t stream);
hipError_t launch_mg_uniform64(const UniformBatch64& b, const SolveParams64& p, hipStream_t s);

}  // namespace pgamd
