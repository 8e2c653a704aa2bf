// C-ABI of the batched Poisson-gravity solver (include/pg_amd.h): context, memory, the asynchronous
// host-array solver with the reference call shape, and device-resident entry points.
//
// Host-side behaviour mirrored from the reference GPU_API:
//   parameter checks            CUAPI_Asyn_PoissonGravitySolver.cu:116-140
//   stream split of the batch   CUAPI_Asyn_PoissonGravitySolver.cu:151-164
//   H2D / launch / D2H order    CUAPI_Asyn_PoissonGravitySolver.cu:177-265
//   allocation and free         CUAPI_MemAllocate_PoissonGravity.cu:27-80, CUAPI_MemFree_PoissonGravity.cu:27-70
// Differences: errors are returned, never fatal; the context replaces the reference's file-scope globals.

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pg_amd.h"
#include "pg_geometry.h"
#include "pg_kernels.h"
#include "pg_internal.h"

using namespace pgamd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(PGAMD_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

}  // namespace

namespace pgamd {

int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

int validate_options(const pgamd_options& o) {
  if (o.pot_ghost_size < 1 || o.pot_ghost_size > kMaxGhost)
    return fail(PGAMD_ERR_ARG, "pot_ghost_size %d outside [1,5]", o.pot_ghost_size);
  // GRA_GHOST_SIZE is 1 or 2 for the gravity stencil (the reference's range); a Poisson-only call may ask
  // for a wider output window, up to the whole solve cube (gra_ghost_size = pot_ghost_size).
  if (o.gra_ghost_size < 1 || o.gra_ghost_size > o.pot_ghost_size)
    return fail(PGAMD_ERR_ARG, "gra_ghost_size %d invalid for pot_ghost_size %d", o.gra_ghost_size,
                o.pot_ghost_size);
  if (o.pot_scheme != PGAMD_SOR && o.pot_scheme != PGAMD_MG)
    return fail(PGAMD_ERR_ARG, "pot_scheme %d is neither SOR (1) nor MG (2)", o.pot_scheme);
  if (o.init_mode < 0 || o.init_mode > 3) return fail(PGAMD_ERR_ARG, "init_mode %d", o.init_mode);
  if (o.term_mode < 0 || o.term_mode > 1) return fail(PGAMD_ERR_ARG, "term_mode %d", o.term_mode);
  if (o.check_every < 1) return fail(PGAMD_ERR_ARG, "check_every %d < 1", o.check_every);
  if (!(o.rel_tol >= 0.0f)) return fail(PGAMD_ERR_ARG, "rel_tol must be >= 0");
  if (!(o.guess_fallback_tol >= 0.0f)) return fail(PGAMD_ERR_ARG, "guess_fallback_tol must be >= 0");
  if (o.mixed_persistent != 0 && o.mixed_persistent != 1) return fail(PGAMD_ERR_ARG, "mixed_persistent must be 0 or 1");
  return PGAMD_OK;
}

// Checks of CUAPI_Asyn_PoissonGravitySolver.cu:116-140 plus the ones a run-time API needs.
int validate_args(const pgamd_options& o, const pgamd_solver_args& a, int npatch) {
  if (npatch < 0) return fail(PGAMD_ERR_ARG, "negative patch count %d", npatch);
  if (a.poisson && a.int_scheme != PGAMD_INT_CQUAD && a.int_scheme != PGAMD_INT_QUAD)
    return fail(PGAMD_ERR_ARG, "incorrect parameter IntScheme = %d", a.int_scheme);
  if (a.graacc && o.gra_ghost_size > 2)
    return fail(PGAMD_ERR_ARG, "the gravity update needs GRA_GHOST_SIZE <= 2 (wider windows are Poisson-only)");
  if (a.graacc && a.p5_gradient && o.gra_ghost_size < 2)
    return fail(PGAMD_ERR_ARG, "the 5-point gradient needs GRA_GHOST_SIZE = 2");
  if (a.poisson && o.pot_scheme == PGAMD_SOR && (a.sor_max_iter < 0 || a.sor_min_iter < 0))
    return fail(PGAMD_ERR_ARG, "negative SOR iteration limits");
  if (a.poisson && o.pot_scheme == PGAMD_MG && (a.mg_max_iter < 0 || a.mg_npre_smooth < 0 || a.mg_npost_smooth < 0))
    return fail(PGAMD_ERR_ARG, "negative MG iteration counts");
  return PGAMD_OK;
}

SolveParams make_params(const pgamd_options& o, const pgamd_solver_args& a) {
  SolveParams p{};
  p.cst = a.poi_coeff * a.dh * a.dh;                 // CPU_PoissonSolver_SOR.cpp:51 (FP32)
  p.w6 = a.sor_omega / 6.0f;                         // CPU_PoissonSolver_SOR.cpp:52 (FP32)
  p.damp = fabsf(1.0f - a.sor_omega);
  p.rel_tol = o.rel_tol;
  p.min_iter = a.sor_min_iter;
  p.max_iter = a.sor_max_iter;
  p.term_mode = o.term_mode;
  p.check_every = o.check_every;
  p.init_mode = o.init_mode;
  p.int_scheme = a.int_scheme;
  p.gra_ghost = o.gra_ghost_size;
  p.poi_coeff = a.poi_coeff;
  p.dh = a.dh;
  p.mg_tol = a.mg_tolerated_error;
  p.mg_max_iter = a.mg_max_iter;
  p.mg_npre = a.mg_npre_smooth;
  p.mg_npost = a.mg_npost_smooth;
  p.fuse_gravity = o.fuse_gravity && a.graacc;
  p.p5 = a.p5_gradient;
  // Gra_Const in double, rounded once (CPU_GravitySolver.cpp:23, CUAPI_Asyn_PoissonGravitySolver.cu:112-113)
  p.gra_const = (float)(a.p5_gradient ? (double)a.dt / (12.0 * (double)a.dh) : (double)a.dt / (2.0 * (double)a.dh));
  p.guess_tol = o.guess_fallback_tol;
  return p;
}

}  // namespace pgamd

struct pgamd_context {
  int device = 0;
  hipDeviceProp_t prop{};
  pgamd_options opt{};
  SorLayout layout{};
  std::vector<hipStream_t> streams;
  int cap = 0;  // patches the device buffers hold
  float* d_rho = nullptr;
  float* d_pot_in = nullptr;
  float* d_pot_out = nullptr;
  float* d_flu = nullptr;
  int* d_iters = nullptr;
  int* d_status = nullptr;
  int* h_iters = nullptr;
  int* h_status = nullptr;
  int last_npatch = 0;
  size_t rho_n = 0, cpot_n = 0, out_n = 0;  // floats per patch at allocation time
  int alloc_gs = 0, alloc_gg = 0;           // ghost sizes the device buffers were sized for
  hipStream_t aux = nullptr;                // forked launches (mixed batches)
  hipStream_t capture = nullptr;            // graph capture
  hipEvent_t ev[2] = {nullptr, nullptr};    // fork / join
};

namespace {

void free_device(pgamd_context* c) {
  hipFree(c->d_rho);
  hipFree(c->d_pot_in);
  hipFree(c->d_pot_out);
  hipFree(c->d_flu);
  hipFree(c->d_iters);
  hipFree(c->d_status);
  hipHostFree(c->h_iters);
  hipHostFree(c->h_status);
  c->d_rho = c->d_pot_in = c->d_pot_out = c->d_flu = nullptr;
  c->d_iters = c->d_status = nullptr;
  c->h_iters = c->h_status = nullptr;
  c->cap = 0;
  c->alloc_gs = c->alloc_gg = 0;
}

int ensure_streams(pgamd_context* c, int n) {
  while ((int)c->streams.size() < n) {
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    c->streams.push_back(s);
  }
  return PGAMD_OK;
}

}  // namespace

extern "C" {

void pgamd_default_options(pgamd_options* o) {
  if (!o) return;
  o->pot_ghost_size = 5;
  o->gra_ghost_size = 1;
  o->pot_scheme = PGAMD_SOR;
  o->init_mode = PGAMD_INIT_INTERP;
  o->term_mode = PGAMD_TERM_REFERENCE;
  o->check_every = 1;
  o->rel_tol = 1e-5f;
  o->fuse_gravity = 1;
  o->guess_fallback_tol = 0.0f;
  o->mixed_persistent = 0;
}

const char* pgamd_last_error(void) { return g_last_error.c_str(); }

int pgamd_set_diagnose_device(int device_id, int* npatch_group, int* nstream, pgamd_context** out) {
  if (!out) return fail(PGAMD_ERR_ARG, "null context pointer");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return fail(PGAMD_ERR_DEVICE, "no HIP device");
  if (device_id < 0 || device_id >= ndev) return fail(PGAMD_ERR_ARG, "device id %d outside [0,%d)", device_id, ndev);
  HIP_TRY(hipSetDevice(device_id));
  auto* c = new pgamd_context();
  c->device = device_id;
  hipError_t e = hipGetDeviceProperties(&c->prop, device_id);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipGetDeviceProperties");
  }
  pgamd_default_options(&c->opt);
  for (int g = 1; g <= 5; g++) {
    c->layout.ppw[g] = default_ppw(g);
    c->layout.pad[g] = default_pad(g);
  }
  c->layout.ppw[0] = 1;
  c->layout.pad[0] = 0;
  c->layout.waves_per_cu = 0;
  if (npatch_group && *npatch_group <= 0) *npatch_group = 32 * c->prop.multiProcessorCount;
  if (nstream && *nstream <= 0) *nstream = 1;
  const int rc = ensure_streams(c, nstream && *nstream > 0 ? *nstream : 1);
  if (rc != PGAMD_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return PGAMD_OK;
}

int pgamd_destroy(pgamd_context* c) {
  if (!c) return PGAMD_OK;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  free_device(c);
  pgamd::release_extras(c);
  for (auto s : c->streams) hipStreamDestroy(s);
  if (c->aux) hipStreamDestroy(c->aux);
  if (c->capture) hipStreamDestroy(c->capture);
  for (auto e : c->ev)
    if (e) hipEventDestroy(e);
  delete c;
  return PGAMD_OK;
}

int pgamd_set_options(pgamd_context* c, const pgamd_options* o) {
  if (!c || !o) return fail(PGAMD_ERR_ARG, "null argument");
  const int rc = validate_options(*o);
  if (rc != PGAMD_OK) return rc;
  c->opt = *o;
  return PGAMD_OK;
}

int pgamd_device_info(pgamd_context* c, char* buf, int buflen) {
  if (!c || !buf || buflen <= 0) return fail(PGAMD_ERR_ARG, "null argument");
  snprintf(buf, buflen, "%s;arch=%s;CUs=%d;LDS/CU=%zu;clock_kHz=%d", c->prop.name, c->prop.gcnArchName,
           c->prop.multiProcessorCount, (size_t)c->prop.maxSharedMemoryPerMultiProcessor, c->prop.clockRate);
  return PGAMD_OK;
}

int pgamd_mem_allocate_poisson_gravity(pgamd_context* c, int npatch_group, float** h_rho, float** h_pot_in,
                                       float** h_pot_out, float** h_flu) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  if (npatch_group <= 0) return fail(PGAMD_ERR_ARG, "npatch_group %d <= 0", npatch_group);
  HIP_TRY(hipSetDevice(c->device));
  free_device(c);
  const Geo g = geo_of(c->opt.pot_ghost_size);
  const int G = gra_nxt(c->opt.gra_ghost_size);
  const size_t np = (size_t)8 * npatch_group;
  c->rho_n = (size_t)g.R * g.R * g.R;
  c->cpot_n = (size_t)g.C * g.C * g.C;
  c->out_n = (size_t)G * G * G;
  c->alloc_gs = c->opt.pot_ghost_size;
  c->alloc_gg = c->opt.gra_ghost_size;
  const size_t flu_n = (size_t)kNComp * kPatchCells;
  if (hipMalloc(&c->d_rho, np * c->rho_n * 4) != hipSuccess || hipMalloc(&c->d_pot_in, np * c->cpot_n * 4) != hipSuccess ||
      hipMalloc(&c->d_pot_out, np * c->out_n * 4) != hipSuccess || hipMalloc(&c->d_flu, np * flu_n * 4) != hipSuccess ||
      hipMalloc(&c->d_iters, np * 4) != hipSuccess || hipMalloc(&c->d_status, np * 4) != hipSuccess ||
      hipHostMalloc(&c->h_iters, np * 4) != hipSuccess || hipHostMalloc(&c->h_status, np * 4) != hipSuccess) {
    free_device(c);
    return fail(PGAMD_ERR_NOMEM, "device/pinned allocation for %zu patches failed", np);
  }
  c->cap = (int)np;
  // pinned host buffers handed to the caller (CUAPI_MemAllocate_PoissonGravity.cu:66-69)
  float** outs[4] = {h_rho, h_pot_in, h_pot_out, h_flu};
  const size_t sizes[4] = {np * c->rho_n * 4, np * c->cpot_n * 4, np * c->out_n * 4, np * flu_n * 4};
  for (int i = 0; i < 4; i++) {
    if (!outs[i]) continue;
    *outs[i] = nullptr;
    if (hipHostMalloc((void**)outs[i], sizes[i]) != hipSuccess) return fail(PGAMD_ERR_NOMEM, "pinned host allocation failed");
  }
  return PGAMD_OK;
}

int pgamd_mem_free_poisson_gravity(pgamd_context* c, float** h_rho, float** h_pot_in, float** h_pot_out,
                                   float** h_flu) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  free_device(c);
  float** outs[4] = {h_rho, h_pot_in, h_pot_out, h_flu};
  for (auto p : outs)
    if (p && *p) {
      hipHostFree(*p);
      *p = nullptr;
    }
  return PGAMD_OK;
}

int pgamd_asyn_poisson_gravity_solver(pgamd_context* c, const float* h_rho, float* h_pot_in, float* h_pot_out,
                                      float* h_flu, int npatch_group, float dt, float dh, int sor_min_iter,
                                      int sor_max_iter, float sor_omega, int mg_max_iter, int mg_npre_smooth,
                                      int mg_npost_smooth, float mg_tolerated_error, float poi_coeff,
                                      int int_scheme, int p5_gradient, int poisson, int graacc, int gpu_nstream) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  if (c->cap == 0) return fail(PGAMD_ERR_STATE, "pgamd_mem_allocate_poisson_gravity has not been called");
  // The device buffers and the per-patch copy strides were sized for the ghost sizes current at allocation;
  // the kernels index with the current ones.  The reference fixes both at compile time
  // (CUAPI_MemAllocate_PoissonGravity.cu:27-80), so a change in between is refused, never run.
  if (c->opt.pot_ghost_size != c->alloc_gs || c->opt.gra_ghost_size != c->alloc_gg)
    return fail(PGAMD_ERR_STATE,
                "ghost sizes changed since pgamd_mem_allocate_poisson_gravity (allocated for %d/%d, now %d/%d): "
                "allocate again",
                c->alloc_gs, c->alloc_gg, c->opt.pot_ghost_size, c->opt.gra_ghost_size);
  // the reference call shape has no guess array
  if (poisson && c->opt.init_mode >= PGAMD_INIT_GUESS)
    return fail(PGAMD_ERR_ARG, "init_mode GUESS needs a guess array: use pgamd_solve_device");
  if (gpu_nstream < 1) return fail(PGAMD_ERR_ARG, "GPU_NStream %d < 1", gpu_nstream);
  const int npatch = 8 * npatch_group;
  if (npatch_group < 0 || npatch > c->cap)
    return fail(PGAMD_ERR_ARG, "%d patches exceed the %d allocated", npatch, c->cap);
  const pgamd_solver_args a{dt,          dh,          sor_min_iter,       sor_max_iter, sor_omega,  mg_max_iter,
                            mg_npre_smooth, mg_npost_smooth, mg_tolerated_error, poi_coeff, int_scheme, p5_gradient,
                            poisson,     graacc};
  int rc = validate_args(c->opt, a, npatch);
  if (rc != PGAMD_OK) return rc;
  if ((poisson && (!h_rho || !h_pot_in)) || ((poisson || graacc) && !h_pot_out) || (graacc && !h_flu))
    return fail(PGAMD_ERR_ARG, "null host array");
  if (npatch == 0) return PGAMD_OK;
  HIP_TRY(hipSetDevice(c->device));
  rc = ensure_streams(c, gpu_nstream);
  if (rc != PGAMD_OK) return rc;
  const SolveParams p = make_params(c->opt, a);
  const size_t flu_n = (size_t)kNComp * kPatchCells;

  // patches per stream (CUAPI_Asyn_PoissonGravitySolver.cu:151-164)
  std::vector<int> count(gpu_nstream), first(gpu_nstream, 0);
  for (int s = 0; s < gpu_nstream - 1; s++) {
    count[s] = npatch / gpu_nstream;
    first[s + 1] = first[s] + count[s];
  }
  count[gpu_nstream - 1] = npatch - first[gpu_nstream - 1];

  for (int s = 0; s < gpu_nstream; s++) {  // a. host -> device
    if (count[s] == 0) continue;
    hipStream_t st = c->streams[s];
    const size_t f = first[s], n = count[s];
    if (poisson) {
      HIP_TRY(hipMemcpyAsync(c->d_rho + f * c->rho_n, h_rho + f * c->rho_n, n * c->rho_n * 4, hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(c->d_pot_in + f * c->cpot_n, h_pot_in + f * c->cpot_n, n * c->cpot_n * 4,
                             hipMemcpyHostToDevice, st));
    }
    if (graacc) {
      if (!poisson)
        HIP_TRY(hipMemcpyAsync(c->d_pot_out + f * c->out_n, h_pot_out + f * c->out_n, n * c->out_n * 4,
                               hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(c->d_flu + f * flu_n, h_flu + f * flu_n, n * flu_n * 4, hipMemcpyHostToDevice, st));
    }
  }
  for (int s = 0; s < gpu_nstream; s++) {  // b. kernels
    if (count[s] == 0) continue;
    const size_t f = first[s];
    rc = pgamd::enqueue_solve(c, c->d_rho + f * c->rho_n, c->d_pot_in + f * c->cpot_n, nullptr,
                              c->d_pot_out + f * c->out_n, graacc ? c->d_flu + f * flu_n : nullptr, count[s], a, p,
                              c->d_iters + f, c->d_status + f, c->streams[s]);
    if (rc != PGAMD_OK) return rc;
  }
  for (int s = 0; s < gpu_nstream; s++) {  // c. device -> host
    if (count[s] == 0) continue;
    hipStream_t st = c->streams[s];
    const size_t f = first[s], n = count[s];
    if (poisson) {
      HIP_TRY(hipMemcpyAsync(h_pot_out + f * c->out_n, c->d_pot_out + f * c->out_n, n * c->out_n * 4,
                             hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(c->h_iters + f, c->d_iters + f, n * 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(c->h_status + f, c->d_status + f, n * 4, hipMemcpyDeviceToHost, st));
    }
    if (graacc)
      HIP_TRY(hipMemcpyAsync(h_flu + f * flu_n, c->d_flu + f * flu_n, n * flu_n * 4, hipMemcpyDeviceToHost, st));
  }
  c->last_npatch = poisson ? npatch : 0;
  return PGAMD_OK;
}

int pgamd_synchronize(pgamd_context* c) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  HIP_TRY(hipSetDevice(c->device));
  for (auto s : c->streams) HIP_TRY(hipStreamSynchronize(s));
  return PGAMD_OK;
}

int pgamd_get_patch_stats(pgamd_context* c, int* iters, int* status, int npatch) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  if (npatch > c->last_npatch) return fail(PGAMD_ERR_ARG, "only %d patches in the last solve", c->last_npatch);
  if (iters) memcpy(iters, c->h_iters, sizeof(int) * npatch);
  if (status) memcpy(status, c->h_status, sizeof(int) * npatch);
  return PGAMD_OK;
}

int pgamd_set_layout(pgamd_context* c, int gs, int ppw, int pad) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  if (gs < 1 || gs > kMaxGhost) return fail(PGAMD_ERR_ARG, "ghost size %d outside [1,5]", gs);
  if (!supported_ppw(gs, ppw)) return fail(PGAMD_ERR_ARG, "%d patches per wavefront do not fit ghost size %d", ppw, gs);
  if (!supported_layout(gs, ppw, pad))
    return fail(PGAMD_ERR_ARG, "row padding %d is neither 0..3 nor the compact layout (ghost size 5, one patch "
                "per wavefront)", pad);
  c->layout.ppw[gs] = ppw;
  c->layout.pad[gs] = pad;
  return PGAMD_OK;
}

int pgamd_get_layout(pgamd_context* c, int gs, int* ppw, int* pad) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  if (gs < 1 || gs > kMaxGhost) return fail(PGAMD_ERR_ARG, "ghost size %d outside [1,5]", gs);
  if (ppw) *ppw = c->layout.ppw[gs];
  if (pad) *pad = c->layout.pad[gs];
  return PGAMD_OK;
}

int pgamd_layout_occupancy(pgamd_context* c, int gs, int* wg, int* patches) {
  if (!c) return fail(PGAMD_ERR_ARG, "null context");
  if (gs < 1 || gs > kMaxGhost) return fail(PGAMD_ERR_ARG, "ghost size %d outside [1,5]", gs);
  HIP_TRY(hipSetDevice(c->device));
  const int n = sor_uniform_blocks_per_cu(gs, c->layout.ppw[gs], c->layout.pad[gs]);
  if (n <= 0) return fail(PGAMD_ERR_DEVICE, "occupancy query failed");
  if (wg) *wg = n;
  if (patches) *patches = n * c->layout.ppw[gs];
  return PGAMD_OK;
}

int pgamd_solve_device(pgamd_context* c, const float* d_rho, const float* d_pot_in, const float* d_guess,
                       float* d_pot_out, float* d_flu, int npatch, const pgamd_solver_args* a, int* d_iters,
                       int* d_status, void* stream) {
  if (!c || !a) return fail(PGAMD_ERR_ARG, "null argument");
  int rc = validate_args(c->opt, *a, npatch);
  if (rc != PGAMD_OK) return rc;
  if (npatch == 0) return PGAMD_OK;
  if ((a->poisson && (!d_rho || !d_pot_in)) || !d_pot_out || (a->graacc && !d_flu))
    return fail(PGAMD_ERR_ARG, "null device array");
  if (a->poisson && c->opt.init_mode >= PGAMD_INIT_GUESS && !d_guess)
    return fail(PGAMD_ERR_ARG, "init_mode GUESS / GUESS_CUBE needs a guess array");
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  const SolveParams p = make_params(c->opt, *a);
  return pgamd::enqueue_solve(c, d_rho, d_pot_in, d_guess, d_pot_out, a->graacc ? d_flu : nullptr, npatch, *a, p,
                              d_iters, d_status, st);
}

int pgamd_solve_device_f64(pgamd_context* c, const double* d_rho, const double* d_pot_in, const double* d_guess,
                           double* d_pot_out, double* d_flu, int npatch, const pgamd_solver_args_f64* a64,
                           int* d_iters, int* d_status, void* stream) {
  if (!c || !a64) return fail(PGAMD_ERR_ARG, "null argument");
  // the checks of the FP32 entry, on the same settings
  const pgamd_solver_args a{(float)a64->dt, (float)a64->dh, a64->sor_min_iter, a64->sor_max_iter,
                            (float)a64->sor_omega, a64->mg_max_iter, a64->mg_npre_smooth, a64->mg_npost_smooth,
                            (float)a64->mg_tolerated_error, (float)a64->poi_coeff, a64->int_scheme, a64->p5_gradient,
                            a64->poisson, a64->graacc};
  int rc = validate_args(c->opt, a, npatch);
  if (rc != PGAMD_OK) return rc;
  if (npatch == 0) return PGAMD_OK;
  if ((a.poisson && (!d_rho || !d_pot_in)) || !d_pot_out || (a.graacc && !d_flu))
    return fail(PGAMD_ERR_ARG, "null device array");
  if (a.poisson && c->opt.init_mode >= PGAMD_INIT_GUESS && !d_guess)
    return fail(PGAMD_ERR_ARG, "init_mode GUESS / GUESS_CUBE needs a guess array");
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  const pgamd_options& o = c->opt;
  SolveParams64 p{};
  p.cst = a64->poi_coeff * a64->dh * a64->dh;         // CPU_PoissonSolver_SOR.cpp:51 (real = double)
  p.w6 = a64->sor_omega / 6.0;                        // CPU_PoissonSolver_SOR.cpp:52
  p.damp = fabs(1.0 - a64->sor_omega);
  p.rel_tol = o.rel_tol;
  p.min_iter = a64->sor_min_iter;
  p.max_iter = a64->sor_max_iter;
  p.term_mode = o.term_mode;
  p.check_every = o.check_every;
  p.init_mode = o.init_mode;
  p.int_scheme = a64->int_scheme;
  p.gra_ghost = o.gra_ghost_size;
  p.poi_coeff = a64->poi_coeff;
  p.dh = a64->dh;
  p.mg_tol = a64->mg_tolerated_error;
  p.mg_max_iter = a64->mg_max_iter;
  p.mg_npre = a64->mg_npre_smooth;
  p.mg_npost = a64->mg_npost_smooth;
  p.fuse_gravity = o.fuse_gravity && a64->graacc;
  p.p5 = a64->p5_gradient;
  p.gra_const = a64->p5_gradient ? a64->dt / (12.0 * a64->dh) : a64->dt / (2.0 * a64->dh);   // CPU_GravitySolver.cpp:23
  p.guess_tol = o.guess_fallback_tol;
  if (a.poisson) {
    const UniformBatch64 b{d_rho,   d_pot_in, d_guess, d_pot_out, p.fuse_gravity ? d_flu : nullptr, d_iters, d_status,
                           npatch,  o.pot_ghost_size, o.pot_scheme};
    hipError_t e = launch_uniform64(b, p, st);
    if (e != hipSuccess) return hip_fail(e, "FP64 solve kernel launch");
  }
  if (a.graacc && !(a.poisson && p.fuse_gravity)) {
    hipError_t e = launch_gravity64(d_pot_out, d_flu, npatch, o.gra_ghost_size, p.gra_const, p.p5, st);
    if (e != hipSuccess) return hip_fail(e, "FP64 gravity kernel launch");
  }
  return PGAMD_OK;
}

}  // extern "C"

namespace pgamd {

const pgamd_options& ctx_options(const pgamd_context* c) { return c->opt; }
SorLayout& ctx_layout(pgamd_context* c) { return c->layout; }
int ctx_cu_count(const pgamd_context* c) { return c->prop.multiProcessorCount; }
int ctx_device(const pgamd_context* c) { return c->device; }
hipStream_t ctx_stream(const pgamd_context* c) { return c->streams[0]; }

hipStream_t ctx_aux_stream(pgamd_context* c) {
  if (!c->aux && hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess) c->aux = nullptr;
  return c->aux;
}

hipStream_t ctx_capture_stream(pgamd_context* c) {
  if (!c->capture && hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking) != hipSuccess) c->capture = nullptr;
  return c->capture;
}

hipEvent_t ctx_event(pgamd_context* c, int i) {
  if (!c->ev[i] && hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming) != hipSuccess) c->ev[i] = nullptr;
  return c->ev[i];
}

// Poisson (SOR or MG) then, unless fused, the standalone gravity kernel: the order of
// CUAPI_Asyn_PoissonGravitySolver.cu:210-246.
int enqueue_solve(pgamd_context* c, const float* rho, const float* cpot, const float* guess, float* pot_out,
                  float* flu, int npatch, const pgamd_solver_args& a, const SolveParams& p, int* iters, int* status,
                  hipStream_t st) {
  const pgamd_options& o = c->opt;
  if (a.poisson) {
    UniformBatch b{rho, cpot, guess, pot_out, p.fuse_gravity ? flu : nullptr, iters, status, npatch,
                   o.pot_ghost_size, o.pot_scheme};
    hipError_t e = launch_uniform(b, p, c->layout, st);
    if (e != hipSuccess) return hip_fail(e, "solve kernel launch");
  }
  if (a.graacc && !(a.poisson && p.fuse_gravity)) {
    hipError_t e = launch_gravity(pot_out, flu, npatch, o.gra_ghost_size, p.gra_const, p.p5, st);
    if (e != hipSuccess) return hip_fail(e, "gravity kernel launch");
  }
  return PGAMD_OK;
}

}  // namespace pgamd
