/*
 * pg_amd.h - C ABI of the MI355X (gfx950) batched per-patch Poisson-gravity solver.
 *
 * Drop-in for the GPU boundary of the reference testbed (GAMER's Poisson-gravity test program): the entry
 * points below replace the CUAPI_* functions of GPU_API/ and keep their argument order and meaning.  All
 * arrays are FP32, C row-major with x fastest; sizes that the reference fixes at compile time
 * (POT_GHOST_SIZE, GRA_GHOST_SIZE, POT_SCHEME) are context options here.
 *
 *   rho     [npatch][R][R][R]       R = 6 + 2*POT_GHOST_SIZE          (RHO_NXT)
 *   pot_in  [npatch][C][C][C]       C = 4 + 2*((POT_GHOST_SIZE+3)/2)  (POT_NXT)
 *   pot_out [npatch][G][G][G]       G = 8 + 2*GRA_GHOST_SIZE          (GRA_NXT)
 *   flu     [npatch][5][8][8][8]    (rho, px, py, pz, E)
 *
 * Every function returns PGAMD_OK (0) or a negative status; pgamd_last_error() describes the last failure
 * on the calling thread.  Nothing here exits the process (the reference exits on any error,
 * Header/CUAPI.h:10-20).  A context belongs to one host thread at a time.
 */
#ifndef PG_AMD_H
#define PG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGAMD_OK 0
#define PGAMD_ERR_ARG (-1)      /* invalid argument: null pointer, size, enum or combination          */
#define PGAMD_ERR_DEVICE (-2)   /* HIP runtime error                                                  */
#define PGAMD_ERR_STATE (-3)    /* call out of order (e.g. solve before allocate)                     */
#define PGAMD_ERR_NOMEM (-4)    /* allocation failed                                                  */

/* Poisson schemes and interpolation schemes keep the reference's values (Header/Macro.h:47-48,
 * Header/Typedef.h:7-8). */
#define PGAMD_SOR 1
#define PGAMD_MG 2
#define PGAMD_INT_CQUAD 4
#define PGAMD_INT_QUAD 5

/* Initial guess of the interior.  The reference always starts from the interpolated cube
 * (CPU_PoissonSolver_SOR.cpp:65).  INTERP, ZERO, GUESS: the Dirichlet layer comes from the interpolated coarse
 * potential; GUESS takes a caller's interior [npatch][R^3].  GUESS_CUBE takes a caller's whole solve cube
 * [npatch][N^3], N = R + 2: its interior is the start and SOR also takes its Dirichlet layer, which must then
 * be the interpolated coarse potential for the reference problem (pgamd_cnn_warm_start_cube writes exactly
 * that); MG interpolates the layer itself.  GUESS_CUBE is for uniform batches (pgamd_solve_device, graphs). */
#define PGAMD_INIT_INTERP 0
#define PGAMD_INIT_ZERO 1
#define PGAMD_INIT_GUESS 2
#define PGAMD_INIT_GUESS_CUBE 3

/* SOR stopping rule.  REFERENCE: stop once the L1 residual of an iteration grows (after Min_Iter),
 * CPU_PoissonSolver_SOR.cpp:273-283.  The device sums |R| as a wavefront tree, the reference CPU
 * sequentially: where two successive totals are equal to within rounding, the stall can be detected an
 * iteration or a few earlier or later than on the reference CPU (results then differ at FP32 round-off;
 * at equal iteration counts they are bitwise equal).  Measured on the 1320 stall-rule patches of the
 * reference's own outputs (ghost sizes 1..5, CQUAD and QUAD): every patch stops at the reference's iteration.  RELRES: stop when sum|R| <= rel_tol * sum|phi| over the interior,
 * tested after max(1,Min_Iter) iterations and then every check_every iterations.  MG always uses the
 * reference's relative error dh^2 sum|Def| / sum|Sol| <= MG_Tolerated_Error. */
#define PGAMD_TERM_REFERENCE 0
#define PGAMD_TERM_RELRES 1

/* Per-patch status bits (pgamd_get_patch_stats, the *_device entries). */
#define PGAMD_STATUS_CONVERGED 1
#define PGAMD_STATUS_MAX_ITER 2
#define PGAMD_STATUS_NONFINITE 4
#define PGAMD_STATUS_FALLBACK 8   /* init GUESS: guess rejected (relative residual > guess_fallback_tol) */

typedef struct pgamd_context pgamd_context;

/* Options the reference fixes with compile-time macros (Makefile:10-53, Header/CUPOT.h). */
typedef struct {
  int pot_ghost_size;   /* POT_GHOST_SIZE, 1..5                                   */
  int gra_ghost_size;   /* GRA_GHOST_SIZE, 1..2, <= pot_ghost_size; a Poisson-only call (graacc = 0)
                           may use up to pot_ghost_size: pot_out is then the whole solve cube */
  int pot_scheme;       /* PGAMD_SOR or PGAMD_MG                                   */
  int init_mode;        /* PGAMD_INIT_*                                            */
  int term_mode;        /* PGAMD_TERM_*                                            */
  int check_every;      /* >= 1                                                    */
  float rel_tol;        /* RELRES tolerance                                        */
  int fuse_gravity;     /* 1: gravity update inside the solve kernel, phi still in LDS */
  float guess_fallback_tol; /* init GUESS(_CUBE), SOR and MG: use the zero guess instead when the guess's relative
                               residual sum|R| / sum|phi| exceeds this (0: never) */
  int mixed_persistent; /* mixed batches (pgamd_solve_mixed_device, graphs): 0 = one launch per ghost size and
                           scheme, one task per workgroup; 1 = one persistent work-queue launch per scheme.
                           Same results; 0 is faster on gfx950 (the persistent kernels spill registers) */
} pgamd_options;

/* Fills the defaults: ghost 5/1, SOR, interpolated guess, reference rule, check_every 1, rel_tol 1e-5,
 * fused gravity, per-ghost-size launches for mixed batches. */
void pgamd_default_options(pgamd_options* opt);

/* Replaces CUAPI_Set_Diagnose_Device (GPU_API/CUAPI_Set_Diagnose_Device.cu:45-292): selects the device
 * and creates a context.  *npatch_group and *nstream receive defaults when they are <= 0 on input
 * (reference defaults: NPGroup = 32 * number of multiprocessors, GPU_NStream = 48, :236-271; here one
 * stream by default because a single launch already fills the chip). */
int pgamd_set_diagnose_device(int device_id, int* npatch_group, int* nstream, pgamd_context** ctx);
int pgamd_destroy(pgamd_context* ctx);
int pgamd_set_options(pgamd_context* ctx, const pgamd_options* opt);
const char* pgamd_last_error(void);
/* Writes "<device name>;CUs=<n>;LDS/CU=<bytes>;clock_kHz=<n>" into buf. */
int pgamd_device_info(pgamd_context* ctx, char* buf, int buflen);
/* Measured peaks for the roofline report (no reference counterpart): LDS read bandwidth of the whole device
 * (every CU streaming conflict-free 16-byte ds_read_b128 reads) and HBM copy bandwidth (1 GiB device copy,
 * bytes read + written).  Either pointer may be NULL to skip that measurement.  Synchronous; allocates
 * 2 GiB for the copy. */
int pgamd_measure_peaks(pgamd_context* ctx, double* lds_read_bytes_per_s, double* hbm_copy_bytes_per_s);

/* Replaces CUAPI_MemAllocate_PoissonGravity (GPU_API/CUAPI_MemAllocate_PoissonGravity.cu:27-80): device
 * buffers for 8*npatch_group patches at the context's ghost sizes, pinned host buffers handed to the
 * caller, and the streams. */
int pgamd_mem_allocate_poisson_gravity(pgamd_context* ctx, int npatch_group, float** h_rho, float** h_pot_in,
                                       float** h_pot_out, float** h_flu);
/* Replaces CUAPI_MemFree_PoissonGravity (GPU_API/CUAPI_MemFree_PoissonGravity.cu:27-70). */
int pgamd_mem_free_poisson_gravity(pgamd_context* ctx, float** h_rho, float** h_pot_in, float** h_pot_out,
                                   float** h_flu);

/* Replaces CUAPI_Asyn_PoissonGravitySolver (GPU_API/CUAPI_Asyn_PoissonGravitySolver.cu:85-275); same
 * argument order.  Host arrays in (pinned ones from pgamd_mem_allocate_poisson_gravity overlap best);
 * returns before the GPU finishes: h_pot_out and h_flu are valid after pgamd_synchronize.  With
 * poisson == 0 and graacc != 0, h_pot_out is an input (:192-200).  PGAMD_ERR_STATE if the context's ghost
 * sizes differ from those at allocation (the device buffers were sized for those); PGAMD_ERR_ARG for
 * init_mode PGAMD_INIT_GUESS (this call shape has no guess array). */
int pgamd_asyn_poisson_gravity_solver(pgamd_context* ctx, const float* h_rho, float* h_pot_in, float* h_pot_out,
                                      float* h_flu, int npatch_group, float dt, float dh, int sor_min_iter,
                                      int sor_max_iter, float sor_omega, int mg_max_iter, int mg_npre_smooth,
                                      int mg_npost_smooth, float mg_tolerated_error, float poi_coeff,
                                      int int_scheme, int p5_gradient, int poisson, int graacc, int gpu_nstream);
/* Replaces CUAPI_Synchronize (GPU_API/CUAPI_Synchronize.cu:10-13). */
int pgamd_synchronize(pgamd_context* ctx);
/* Iteration count and status of every patch of the last asynchronous solve (valid after synchronize).
 * No reference counterpart: the reference only accumulates an average on the CPU
 * (CPU_PoissonSolver_SOR.cpp:290). */
int pgamd_get_patch_stats(pgamd_context* ctx, int* iters, int* status, int npatch);

/* ---- patch-group preparation (host side; the "REPLACE in the actual implementation" step of
 * GPU_API/CUAPI_Asyn_PoissonGravitySolver.cu:37-41 and CPU_Solver/CPU_PoissonSolver_SOR.cpp:5-7, 70-76: the
 * testbed fills the arrays with synthetic data, Main.cpp:524, where GAMER gathers them from the AMR levels).
 * No GPU involved; nthreads host threads (<= 1: the calling thread only).
 *
 * Level format.  Fine level: nfine patches of 8^3 cells, fine_flu [nfine][5][8][8][8] (rho, px, py, pz, E),
 * fine_pot [nfine][8][8][8].  Patch group g: the 8 patches group_first[g] + l, l = (lz*2 + ly)*2 + lx, the
 * children of coarse patch group_father[g] at offset (lx, ly, lz) * 8 fine cells.  Coarse level: coarse_pot
 * [ncoarse][8][8][8] at twice the spacing.  Neighbour tables fine_sib [nfine][27], coarse_sib [ncoarse][27]:
 * entry (dz+1)*9 + (dy+1)*3 + (dx+1) is the patch at offset (dx, dy, dz) in {-1,0,1}^3 (entry 13: the patch
 * itself), -1 where there is none (outside the domain or the refined region).
 *
 * pgamd_prepare_patch_groups fills, for patch 8g + l of the batch, h_rho [R^3] with the fine density of the
 * solve interior (the patch and pot_ghost_size - 1 cells of its neighbours), h_pot_in [C^3] with the coarse
 * potential around the patch's own 4^3 coarse cells ((C - 4)/2 more on each side, from the father and its
 * neighbours) and h_flu [5][8^3] with the patch's fluid: the layout of pgamd_asyn_poisson_gravity_solver.  Cells
 * with no neighbour to take them from are set to 0 and flagged in status[8g + l] (may be NULL).
 * PGAMD_ERR_ARG for sizes, null arrays, or indices outside the level (a neighbour index beyond the level is
 * also flagged PGAMD_PACK_BAD_INDEX in status). */
#define PGAMD_PACK_MISSING_FINE 1     /* some density cells had no fine neighbour (set to 0) */
#define PGAMD_PACK_MISSING_COARSE 2   /* some coarse potential cells had no coarse neighbour (set to 0) */
#define PGAMD_PACK_BAD_INDEX 4        /* a group or neighbour index outside the level (cells set to 0) */
int pgamd_prepare_patch_groups(int pot_ghost_size, int ngroup, const int* group_first, const int* group_father,
                               int nfine, const float* fine_flu, const int* fine_sib, int ncoarse,
                               const float* coarse_pot, const int* coarse_sib, float* h_rho, float* h_pot_in,
                               float* h_flu, int* status, int nthreads);
/* The reverse step (the reference's results stay in the arrays, Main.cpp:159-220): writes the patch part of
 * h_pot_out [GRA_NXT^3] into fine_pot and h_flu into fine_flu for every patch of the ngroup groups.  Either
 * destination may be NULL (not both). */
int pgamd_scatter_patch_groups(int gra_ghost_size, int ngroup, const int* group_first, const float* h_pot_out,
                               const float* h_flu, int nfine, float* fine_pot, float* fine_flu, int nthreads);
/* The same two steps for levels resident in device memory (no reference counterpart: GAMER keeps its levels
 * on the host).  All pointers are device pointers; one kernel launch each on `stream` (a hipStream_t, NULL =
 * the default stream) of the current device; asynchronous.  The fluid arrays must be 16-byte aligned.  The
 * outputs are bitwise those of the host versions.  Index errors cannot be returned from the device: a patch
 * whose group or neighbour indices lie outside the level gets PGAMD_PACK_BAD_INDEX in d_status (its invalid
 * cells, or the whole patch for a bad group, set to 0), and the scatter skips groups whose group_first is out of
 * range. */
int pgamd_prepare_patch_groups_device(int pot_ghost_size, int ngroup, const int* d_group_first,
                                      const int* d_group_father, int nfine, const float* d_fine_flu,
                                      const int* d_fine_sib, int ncoarse, const float* d_coarse_pot,
                                      const int* d_coarse_sib, float* d_rho, float* d_pot_in, float* d_flu,
                                      int* d_status, void* stream);
int pgamd_scatter_patch_groups_device(int gra_ghost_size, int ngroup, const int* d_group_first,
                                      const float* d_pot_out, const float* d_flu, int nfine, float* d_fine_pot,
                                      float* d_fine_flu, void* stream);

/* Launch layout of the SOR kernel for one ghost size: patches per wavefront (1, 2 or 4 where they fit) and
 * the padding of the colour-split LDS rows (0..3 floats; trades LDS capacity for fewer bank conflicts), or
 * PGAMD_LAYOUT_COMPACT at ghost size 5 with one patch per wavefront: a single-buffer LDS cube of 16 KB instead
 * of 25.9 KB, eight resident patches per CU instead of six (the default there; a solve whose gra_ghost_size
 * is 5 runs that ghost size in the split layout with one float of padding).  Results do not depend on the
 * layout.  This is the autotuner's search space (pgamd_autotune); no reference counterpart: the reference
 * fixes its launch shapes per GPU generation with macros (Header/CUPOT.h:38-152). */
#define PGAMD_LAYOUT_COMPACT 4
int pgamd_set_layout(pgamd_context* ctx, int ghost_size, int patches_per_wave, int row_pad);
int pgamd_get_layout(pgamd_context* ctx, int ghost_size, int* patches_per_wave, int* row_pad);
/* Residency of the SOR kernel at the context's layout for one ghost size: workgroups (one wavefront each) the
 * device keeps resident per CU, and the patches that makes (times patches per wavefront), from the kernel's
 * registers and LDS (hipOccupancyMaxActiveBlocksPerMultiprocessor).  No reference counterpart. */
int pgamd_layout_occupancy(pgamd_context* ctx, int ghost_size, int* workgroups_per_cu, int* patches_per_cu);


/* ---- device-resident entry points (inputs already in HBM; no reference counterpart: the reference
 * always stages through host memory).  All pointers are device pointers; `stream` is a hipStream_t
 * (NULL = the HIP default stream).  iters / status may be NULL. */
typedef struct {
  float dt, dh;
  int sor_min_iter, sor_max_iter;
  float sor_omega;
  int mg_max_iter, mg_npre_smooth, mg_npost_smooth;
  float mg_tolerated_error;
  float poi_coeff;
  int int_scheme;
  int p5_gradient;
  int poisson, graacc;
} pgamd_solver_args;

int pgamd_solve_device(pgamd_context* ctx, const float* d_rho, const float* d_pot_in, const float* d_guess,
                       float* d_pot_out, float* d_flu, int npatch, const pgamd_solver_args* args, int* d_iters,
                       int* d_status, void* stream);

/* ---- FP64: the reference's FLOAT8 build (Makefile:22, Header/Macro.h:35), where every array and constant is
 * double.  Same arguments and options as pgamd_solve_device with double arrays and double constants
 * (Const = Poi_Coeff dh^2, Omega/6 and Gra_Const formed in double).  Uniform batches, SOR or MG, init modes
 * INTERP / ZERO / GUESS / GUESS_CUBE, both stopping rules, fused or separate gravity.  The FP64 kernels keep the
 * reference's operation order: at equal iteration counts the results are bitwise those of the reference's FP64
 * CPU solver. */
typedef struct {
  double dt, dh;
  int sor_min_iter, sor_max_iter;
  double sor_omega;
  int mg_max_iter, mg_npre_smooth, mg_npost_smooth;
  double mg_tolerated_error;
  double poi_coeff;
  int int_scheme;
  int p5_gradient;
  int poisson, graacc;
} pgamd_solver_args_f64;

int pgamd_solve_device_f64(pgamd_context* ctx, const double* d_rho, const double* d_pot_in, const double* d_guess,
                           double* d_pot_out, double* d_flu, int npatch, const pgamd_solver_args_f64* args,
                           int* d_iters, int* d_status, void* stream);

/* Autotuner: times every supported layout of one ghost size on the caller's device batch (SOR with the
 * context's options, gravity off; repeats >= 1 timed solves per candidate after one warm-up), keeps the
 * fastest in the context and returns it.  d_pot_out is overwritten.  The layout table can be saved and
 * reloaded as JSON, keyed by device name. */
int pgamd_autotune(pgamd_context* ctx, int ghost_size, const float* d_rho, const float* d_pot_in, float* d_pot_out,
                   int npatch, const pgamd_solver_args* args, int repeats, int* best_ppw, int* best_pad,
                   float* best_ms);
int pgamd_autotune_save(pgamd_context* ctx, const char* path);
int pgamd_autotune_load(pgamd_context* ctx, const char* path);

/* ---- mixed batches (no reference counterpart: the reference compiles one ghost size and one scheme per
 * binary).  A plan fixes the ghost size (1..5) and scheme (SOR/MG) of every patch, the packed layout of
 * the density and coarse-potential arrays (patch p starts at rho_off[p] / pot_in_off[p] floats) and the
 * work queue of the persistent kernels.  pot_out is [npatch][G^3] and flu [npatch][5][8^3] as usual.
 * Context options other than pot_ghost_size / pot_scheme apply to every patch. */
typedef struct pgamd_plan pgamd_plan;
int pgamd_plan_create(pgamd_context* ctx, const int* ghost_size, const int* scheme, int npatch, pgamd_plan** plan);
int pgamd_plan_sizes(const pgamd_plan* plan, int64_t* rho_floats, int64_t* pot_in_floats);
int pgamd_plan_offsets(const pgamd_plan* plan, int64_t* rho_off, int64_t* pot_in_off);
int pgamd_plan_destroy(pgamd_plan* plan);
/* One persistent launch per scheme on forked streams.  omega_by_ghost[gs-1] (may be NULL) overrides
 * args->sor_omega per ghost size. */
int pgamd_solve_mixed_device(pgamd_context* ctx, pgamd_plan* plan, const float* d_rho, const float* d_pot_in,
                             const float* d_guess, float* d_pot_out, float* d_flu, const pgamd_solver_args* args,
                             const float* omega_by_ghost, int* d_iters, int* d_status, void* stream);

/* ---- HIP graphs: capture a solve once, replay it with one launch (inputs are read from the same device
 * buffers at every replay).  No reference counterpart. */
typedef struct pgamd_graph pgamd_graph;
int pgamd_graph_capture_mixed(pgamd_context* ctx, pgamd_plan* plan, const float* d_rho, const float* d_pot_in,
                              const float* d_guess, float* d_pot_out, float* d_flu, const pgamd_solver_args* args,
                              const float* omega_by_ghost, int* d_iters, int* d_status, pgamd_graph** graph);
int pgamd_graph_capture_device(pgamd_context* ctx, const float* d_rho, const float* d_pot_in, const float* d_guess,
                               float* d_pot_out, float* d_flu, int npatch, const pgamd_solver_args* args,
                               int* d_iters, int* d_status, pgamd_graph** graph);
int pgamd_graph_launch(pgamd_graph* graph, void* stream);
int pgamd_graph_destroy(pgamd_graph* graph);

/* ---- learned warm start (no reference counterpart) ----------------------------------------------------
 * A 3D CNN (BF16 weights, FP32 accumulation, hand-written MFMA convolutions) predicts the correction from
 * the interpolated coarse potential to the converged potential on the whole solve interior.  Frame: the
 * 18^3 solve cube of ghost size 5 (smaller cubes centred, zeros around) with q0 = interp/s and
 * q1 = 64*r/s (interior only), r the SOR residual of the interpolated cube (reference operand order),
 * s = max|interp| of the patch.  Input x0 [n][10^3][16] BF16:
 * space-to-depth of the zero-extended 20^3 frame (fine index f = 2c-1+d per axis, channel
 * ((dz*2+dy)*2+dx)*2 + q).  Six 3^3 convolutions with 16 channels on the 8^3 coarse grid (the first
 * valid on the 10^3 input, the next four residual blocks y += relu(conv(y)+b), the last linear with 8
 * outputs) give d [n][8^3][8] FP32; the guess at interior frame index 2o+p is interp + s*d[o][p].  Solve
 * with init_mode PGAMD_INIT_GUESS and, for a safe fallback, guess_fallback_tol > 0.
 * Parameters: 41568 FP32 values, per layer W [16][432] (W[co][tap*16+ci], tap = dz*9+dy*3+dx) then b [16]. */
typedef struct pgamd_cnn pgamd_cnn;
int pgamd_cnn_num_params(void);
int pgamd_cnn_create(pgamd_context* ctx, uint64_t seed, pgamd_cnn** net);
int pgamd_cnn_destroy(pgamd_cnn* net);
int pgamd_cnn_get_params(pgamd_cnn* net, float* host_params, int count);
int pgamd_cnn_set_params(pgamd_cnn* net, const float* host_params, int count);
/* Network input of a batch of same-ghost-size patches (d_x0: [n][10^3][16] BF16 bits, d_scale: [n]). */
int pgamd_cnn_prepare(pgamd_context* ctx, const float* d_rho, const float* d_pot_in, int npatch, int ghost_size,
                      const pgamd_solver_args* args, uint16_t* d_x0, float* d_scale, void* stream);
int pgamd_cnn_forward(pgamd_cnn* net, const uint16_t* d_x0, int count, float* d_out, void* stream);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same forward as pgamd_cnn_forward, bitwise, with its
 * activations in a fixed device array of the library (10.24 MB per GPU, held until the process ends) instead
 * of an allocated workspace: at most 64 samples per call (PGAMD_ERR_ARG beyond), no allocation, no host
 * synchronisation.  Calls that overlap on one GPU share that array and race. */
int pgamd_cnn_forward_noalloc(pgamd_cnn* net, const uint16_t* d_x0, int count, float* d_out, void* stream);
/* Training loss: mean squared error of the output against d_target ([n][8^3][8]) plus res_weight times the mean
 * squared residual that error leaves in the guess, r = (sum of the six neighbours - 6 e) on the 16^3 frame
 * (e = output - target at frame cell 2o + p, zero outside), both means over the n*4096 outputs.  res_weight = 0
 * (the default) is the plain mean squared error.  Error if res_weight is negative or not finite. */
int pgamd_cnn_set_loss(pgamd_cnn* net, float res_weight);
/* Loss (see pgamd_cnn_set_loss) against d_target ([n][8^3][8]) and its gradient (FP32, 41568 values, may be
 * NULL); *loss (may be NULL) is read back synchronously. */
int pgamd_cnn_loss_grad(pgamd_cnn* net, const uint16_t* d_x0, const float* d_target, int count, float* loss,
                        float* d_grad, void* stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on a batch. */
int pgamd_cnn_train_step(pgamd_cnn* net, const uint16_t* d_x0, const float* d_target, int count, float lr,
                         float* loss, void* stream);
/* Guess of the interior ([n][R^3]) for a batch at the context's ghost size: prepare, forward, assemble. */
int pgamd_cnn_warm_start(pgamd_context* ctx, pgamd_cnn* net, const float* d_rho, const float* d_pot_in, int npatch,
                         const pgamd_solver_args* args, float* d_guess, void* stream);
/* The same guess as the whole solve cube ([n][N^3], N = R + 2) for init_mode PGAMD_INIT_GUESS_CUBE: the interior is
 * the network's guess (bitwise that of pgamd_cnn_warm_start) and the Dirichlet layer the interpolated coarse
 * potential (bitwise that of the solver's own interpolation), so the solve reads its boundary instead of
 * interpolating it. */
int pgamd_cnn_warm_start_cube(pgamd_context* ctx, pgamd_cnn* net, const float* d_rho, const float* d_pot_in,
                              int npatch, const pgamd_solver_args* args, float* d_guess_cube, void* stream);
/* Warm start and solve of a uniform batch in one call, pipelined over nchunk (1..16) slices: the network
 * of slice i+1 runs on the context's second stream while slice i is solved on `stream`.  The context must
 * have init_mode PGAMD_INIT_GUESS or PGAMD_INIT_GUESS_CUBE (guess_fallback_tol applies); d_guess ([n][R^3],
 * or [n][N^3] for the whole cube) is workspace and receives the guesses.  Same results as
 * pgamd_cnn_warm_start (_cube) + pgamd_solve_device. */
int pgamd_cnn_solve_device(pgamd_context* ctx, pgamd_cnn* net, const float* d_rho, const float* d_pot_in,
                           float* d_guess, float* d_pot_out, float* d_flu, int npatch, const pgamd_solver_args* args,
                           int* d_iters, int* d_status, int nchunk, void* stream);
/* Guess for a mixed batch (plan): d_rho / d_pot_in packed as the plan says; d_guess is packed like d_rho
 * (patch p at rho_off[p], R(p)^3 values) and goes to pgamd_solve_mixed_device. */
int pgamd_cnn_warm_start_mixed(pgamd_context* ctx, pgamd_cnn* net, pgamd_plan* plan, const float* d_rho,
                               const float* d_pot_in, const pgamd_solver_args* args, float* d_guess, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PG_AMD_H */
