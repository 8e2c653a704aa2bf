// Runtime pieces beyond the reference call shape: mixed-ghost-size batches with persistent work-queue
// kernels, and HIP graph capture of repeated solves.
//
// Scheduling: patches are grouped by (scheme, ghost size); SOR groups are cut into wavefront tasks of
// patches_per_wave patches, MG patches are one task each; each queue is ordered by decreasing work per
// task (cells x expected sweeps), so the last tasks a queue hands out are the cheapest (LPT order).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/pg_amd.h"
#include "pg_geometry.h"
#include "pg_internal.h"
#include "pg_kernels.h"

using namespace pgamd;

struct pgamd_plan {
  int device = 0;
  int npatch = 0;
  std::vector<PatchDesc> desc;
  std::vector<int64_t> rho_off, cpot_off;
  int64_t rho_floats = 0, cpot_floats = 0;
  int ppw1 = 4, ppw2 = 2;
  int n_sor_tasks = 0, n_mg_tasks = 0;
  size_t lds_sor = 0, lds_mg = 0;
  int grid_sor = 0, grid_mg = 0;
  PatchDesc* d_desc = nullptr;
  int* d_perm = nullptr;          // SOR order then MG order
  Task* d_tasks = nullptr;        // SOR tasks then MG tasks
  unsigned int* d_counters = nullptr;  // [0] SOR queue, [1] MG queue
  bool has_gs[6] = {false, false, false, false, false, false};
  // per ghost size: first task and task count (SOR tasks index d_tasks, MG tasks index d_tasks + n_sor_tasks)
  // and the SOR patches per wavefront, for the per-ghost-size launches
  int sor_first[6] = {}, sor_count[6] = {}, sor_ppw[6] = {}, mg_first[6] = {}, mg_count[6] = {};
};

struct pgamd_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace pgamd {
const PatchDesc* plan_device_desc(const pgamd_plan* pl) { return pl->d_desc; }
int plan_npatch(const pgamd_plan* pl) { return pl->npatch; }
int plan_device(const pgamd_plan* pl) { return pl->device; }
bool plan_has_ghost(const pgamd_plan* pl, int gs) { return gs >= 1 && gs <= kMaxGhost && pl->has_gs[gs]; }
}  // namespace pgamd


namespace {

void free_plan(pgamd_plan* p) {
  hipFree(p->d_desc);
  hipFree(p->d_perm);
  hipFree(p->d_tasks);
  hipFree(p->d_counters);
}

// Build per-gs SolveParams (omega may differ per ghost size).
ParamsByGs params_by_gs(pgamd_context* c, const pgamd_solver_args& a, const float* omega_by_ghost) {
  ParamsByGs P{};
  for (int g = 1; g <= kMaxGhost; g++) {
    pgamd_solver_args ag = a;
    if (omega_by_ghost) ag.sor_omega = omega_by_ghost[g - 1];
    P.p[g] = make_params(ctx_options(c), ag);
  }
  P.p[0] = P.p[1];
  return P;
}

int enqueue_mixed(pgamd_context* c, pgamd_plan* pl, const float* rho, const float* cpot, const float* guess,
                  float* pot_out, float* flu, const pgamd_solver_args& a, const float* omega_by_ghost, int* iters,
                  int* status, hipStream_t st) {
  const pgamd_options& o = ctx_options(c);
  const ParamsByGs P = params_by_gs(c, a, omega_by_ghost);
  const bool fused = a.graacc && P.p[1].fuse_gravity;
  MixedBatch b{};
  b.desc = pl->d_desc;
  b.npatch = pl->npatch;
  b.rho = rho;
  b.cpot = cpot;
  b.guess = o.init_mode == PGAMD_INIT_GUESS ? guess : nullptr;
  b.pot_out = pot_out;
  b.flu = fused ? flu : nullptr;
  b.iters = iters;
  b.status = status;
  const bool persistent = o.mixed_persistent != 0;
  hipError_t e = persistent ? launch_reset_counters(pl->d_counters, 2, st) : hipSuccess;
  if (e != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  if (a.poisson) {
    hipStream_t aux = ctx_aux_stream(c);
    if (!aux) return set_error(PGAMD_ERR_DEVICE, "cannot create the auxiliary stream");
    hipEvent_t ev_fork = ctx_event(c, 0), ev_join = ctx_event(c, 1);
    if (!ev_fork || !ev_join) return set_error(PGAMD_ERR_DEVICE, "cannot create events");
    // MG queue on the auxiliary stream, SOR queue on the caller's stream
    if (pl->n_mg_tasks > 0) {
      if ((e = hipEventRecord(ev_fork, st)) != hipSuccess || (e = hipStreamWaitEvent(aux, ev_fork, 0)) != hipSuccess)
        return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
      MixedBatch bm = b;
      bm.perm = pl->d_perm;  // task.first indexes the whole permutation (SOR part first, then MG)
      bm.tasks = pl->d_tasks + pl->n_sor_tasks;
      bm.ntasks = pl->n_mg_tasks;
      bm.counter = pl->d_counters + 1;
      if (persistent) {
        e = launch_mg_persistent(bm, P, pl->grid_mg, pl->lds_mg, aux);
      } else {
        for (int g = kMaxGhost; g >= 1 && e == hipSuccess; g--)
          e = launch_mg_tasks(bm, P.p[g], g, pl->mg_first[g], pl->mg_count[g], aux);
      }
      if (e != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
      if ((e = hipEventRecord(ev_join, aux)) != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
    }
    MixedBatch bs = b;
    bs.perm = pl->d_perm;
    bs.tasks = pl->d_tasks;
    bs.ntasks = pl->n_sor_tasks;
    bs.counter = pl->d_counters;
    if (persistent) {
      e = launch_sor_persistent(bs, P, pl->ppw1, pl->ppw2, pl->grid_sor, pl->lds_sor, st);
    } else {
      for (int g = kMaxGhost; g >= 1 && e == hipSuccess; g--)
        e = launch_sor_tasks(bs, P.p[g], g, pl->sor_ppw[g], ctx_layout(c).pad[g], pl->sor_first[g],
                             pl->sor_count[g], st);
    }
    if (e != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
    if (pl->n_mg_tasks > 0 && (e = hipStreamWaitEvent(st, ev_join, 0)) != hipSuccess)
      return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  }
  if (a.graacc && !(a.poisson && fused)) {
    e = launch_gravity(pot_out, flu, pl->npatch, o.gra_ghost_size, P.p[1].gra_const, P.p[1].p5, st);
    if (e != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  }
  return PGAMD_OK;
}

int check_mixed_args(pgamd_context* c, pgamd_plan* pl, const float* rho, const float* cpot, const float* guess,
                     float* pot_out, float* flu, const pgamd_solver_args* a) {
  if (!c || !pl || !a) return set_error(PGAMD_ERR_ARG, "null argument");
  if (pl->device != ctx_device(c)) return set_error(PGAMD_ERR_ARG, "plan belongs to another device");
  pgamd_options o = ctx_options(c);
  int rc = validate_args(o, *a, pl->npatch);
  if (rc != PGAMD_OK) return rc;
  if (a->graacc && a->p5_gradient) {
    for (int g = 1; g < 2; g++)
      if (pl->has_gs[g]) return set_error(PGAMD_ERR_ARG, "the 5-point gradient needs every ghost size >= 2");
  }
  for (int g = 1; g < o.gra_ghost_size && g <= kMaxGhost; g++)
    if (pl->has_gs[g]) return set_error(PGAMD_ERR_ARG, "GRA_GHOST_SIZE exceeds the ghost size of some patches");
  if ((a->poisson && (!rho || !cpot)) || !pot_out || (a->graacc && !flu))
    return set_error(PGAMD_ERR_ARG, "null device array");
  if (a->poisson && o.init_mode == PGAMD_INIT_GUESS && !guess)
    return set_error(PGAMD_ERR_ARG, "init_mode GUESS needs a guess array");
  if (o.init_mode == PGAMD_INIT_GUESS_CUBE)
    return set_error(PGAMD_ERR_ARG, "mixed batches take guesses packed like the density (init_mode GUESS)");
  return PGAMD_OK;
}

}  // namespace

namespace pgamd {
void release_extras(pgamd_context*) {}
}  // namespace pgamd

extern "C" {

int pgamd_plan_create(pgamd_context* c, const int* ghost_size, const int* scheme, int npatch, pgamd_plan** out) {
  if (!c || !ghost_size || !out) return set_error(PGAMD_ERR_ARG, "null argument");
  *out = nullptr;
  if (npatch <= 0) return set_error(PGAMD_ERR_ARG, "a plan needs at least one patch");
  auto* pl = new pgamd_plan();
  pl->device = ctx_device(c);
  pl->npatch = npatch;
  pl->desc.resize(npatch);
  pl->rho_off.resize(npatch);
  pl->cpot_off.resize(npatch);
  const SorLayout& lay = ctx_layout(c);
  pl->ppw1 = lay.ppw[1] >= 4 ? 4 : 2;   // persistent kernel instantiations: (4,2) and (2,1)
  pl->ppw2 = pl->ppw1 == 4 ? 2 : 1;
  std::vector<int> sor_by_gs[6], mg_by_gs[6];
  int64_t ro = 0, co = 0;
  for (int p = 0; p < npatch; p++) {
    const int g = ghost_size[p];
    const int s = scheme ? scheme[p] : PGAMD_SOR;
    if (g < 1 || g > kMaxGhost || (s != PGAMD_SOR && s != PGAMD_MG)) {
      delete pl;
      return set_error(PGAMD_ERR_ARG, "invalid ghost size or scheme in the plan");
    }
    const Geo ge = geo_of(g);
    pl->desc[p] = PatchDesc{g, s, ro, co};
    pl->rho_off[p] = ro;
    pl->cpot_off[p] = co;
    ro += (int64_t)ge.R * ge.R * ge.R;
    co += (int64_t)ge.C * ge.C * ge.C;
    (s == PGAMD_SOR ? sor_by_gs : mg_by_gs)[g].push_back(p);
    pl->has_gs[g] = true;
  }
  pl->rho_floats = ro;
  pl->cpot_floats = co;

  // tasks, largest ghost size (most work per task) first
  std::vector<int> perm;
  std::vector<Task> tasks;
  size_t lds_sor = 0, lds_mg = 0;
  for (int g = kMaxGhost; g >= 1; g--) {
    const int ppw = g == 1 ? pl->ppw1 : (g == 2 ? pl->ppw2 : 1);
    const auto& v = sor_by_gs[g];
    pl->sor_first[g] = (int)tasks.size();
    pl->sor_ppw[g] = ppw;
    for (size_t i = 0; i < v.size(); i += ppw) {
      const int cnt = (int)std::min<size_t>(ppw, v.size() - i);
      tasks.push_back(Task{g, PGAMD_SOR, (int)perm.size(), cnt});
      for (int k = 0; k < cnt; k++) perm.push_back(v[i + k]);
    }
    pl->sor_count[g] = (int)tasks.size() - pl->sor_first[g];
    if (!v.empty()) lds_sor = std::max(lds_sor, lds_bytes_sor(g, ppw, g == 4 ? 2 : 0));
  }
  pl->n_sor_tasks = (int)tasks.size();
  for (int g = kMaxGhost; g >= 1; g--) {
    pl->mg_first[g] = (int)tasks.size() - pl->n_sor_tasks;
    pl->mg_count[g] = (int)mg_by_gs[g].size();
    for (int p : mg_by_gs[g]) {
      tasks.push_back(Task{g, PGAMD_MG, (int)perm.size(), 1});
      perm.push_back(p);
    }
    if (!mg_by_gs[g].empty()) lds_mg = std::max(lds_mg, lds_bytes_mg(g));
  }
  pl->n_mg_tasks = (int)tasks.size() - pl->n_sor_tasks;
  pl->lds_sor = lds_sor;
  pl->lds_mg = lds_mg;
  const int cus = ctx_cu_count(c);
  pl->grid_sor = pl->n_sor_tasks ? std::min(pl->n_sor_tasks, cus * std::max(1, sor_persistent_blocks_per_cu(pl->ppw1, pl->ppw2, lds_sor))) : 0;
  pl->grid_mg = pl->n_mg_tasks ? std::min(pl->n_mg_tasks, cus * std::max(1, mg_persistent_blocks_per_cu(lds_mg))) : 0;

  if (hipSetDevice(pl->device) != hipSuccess || hipMalloc(&pl->d_desc, sizeof(PatchDesc) * npatch) != hipSuccess ||
      hipMalloc(&pl->d_perm, sizeof(int) * npatch) != hipSuccess ||
      hipMalloc(&pl->d_tasks, sizeof(Task) * tasks.size()) != hipSuccess ||
      hipMalloc(&pl->d_counters, sizeof(unsigned int) * 2) != hipSuccess ||
      hipMemcpy(pl->d_desc, pl->desc.data(), sizeof(PatchDesc) * npatch, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(pl->d_perm, perm.data(), sizeof(int) * npatch, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(pl->d_tasks, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice) != hipSuccess) {
    free_plan(pl);
    delete pl;
    return set_error(PGAMD_ERR_NOMEM, "plan allocation failed");
  }
  *out = pl;
  return PGAMD_OK;
}

int pgamd_plan_sizes(const pgamd_plan* pl, int64_t* rho_floats, int64_t* pot_in_floats) {
  if (!pl) return set_error(PGAMD_ERR_ARG, "null plan");
  if (rho_floats) *rho_floats = pl->rho_floats;
  if (pot_in_floats) *pot_in_floats = pl->cpot_floats;
  return PGAMD_OK;
}

int pgamd_plan_offsets(const pgamd_plan* pl, int64_t* rho_off, int64_t* pot_in_off) {
  if (!pl) return set_error(PGAMD_ERR_ARG, "null plan");
  if (rho_off) std::memcpy(rho_off, pl->rho_off.data(), sizeof(int64_t) * pl->npatch);
  if (pot_in_off) std::memcpy(pot_in_off, pl->cpot_off.data(), sizeof(int64_t) * pl->npatch);
  return PGAMD_OK;
}

int pgamd_plan_destroy(pgamd_plan* pl) {
  if (!pl) return PGAMD_OK;
  hipSetDevice(pl->device);
  hipDeviceSynchronize();
  free_plan(pl);
  delete pl;
  return PGAMD_OK;
}

int pgamd_solve_mixed_device(pgamd_context* c, pgamd_plan* pl, const float* rho, const float* cpot, const float* guess,
                             float* pot_out, float* flu, const pgamd_solver_args* a, const float* omega_by_ghost,
                             int* iters, int* status, void* stream) {
  int rc = check_mixed_args(c, pl, rho, cpot, guess, pot_out, flu, a);
  if (rc != PGAMD_OK) return rc;
  if (hipSetDevice(ctx_device(c)) != hipSuccess) return set_error(PGAMD_ERR_DEVICE, "hipSetDevice failed");
  return enqueue_mixed(c, pl, rho, cpot, guess, pot_out, flu, *a, omega_by_ghost, iters, status, (hipStream_t)stream);
}

static int capture(pgamd_context* c, pgamd_graph** out, int (*body)(pgamd_context*, void*), void* arg) {
  *out = nullptr;
  if (hipSetDevice(ctx_device(c)) != hipSuccess) return set_error(PGAMD_ERR_DEVICE, "hipSetDevice failed");
  hipStream_t s = ctx_capture_stream(c);
  if (!s) return set_error(PGAMD_ERR_DEVICE, "cannot create the capture stream");
  auto* g = new pgamd_graph();
  g->device = ctx_device(c);
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete g;
    return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  }
  const int rc = body(c, arg);
  e = hipStreamEndCapture(s, &g->graph);
  if (rc != PGAMD_OK || e != hipSuccess) {
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return rc != PGAMD_OK ? rc : set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g->graph);
    delete g;
    return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  }
  *out = g;
  return PGAMD_OK;
}

struct MixedArgs {
  pgamd_plan* pl;
  const float *rho, *cpot, *guess;
  float *pot_out, *flu;
  const pgamd_solver_args* a;
  const float* omega;
  int *iters, *status;
};

struct DeviceArgs {
  const float *rho, *cpot, *guess;
  float *pot_out, *flu;
  int npatch;
  const pgamd_solver_args* a;
  int *iters, *status;
};

int pgamd_graph_capture_mixed(pgamd_context* c, pgamd_plan* pl, const float* rho, const float* cpot,
                              const float* guess, float* pot_out, float* flu, const pgamd_solver_args* a,
                              const float* omega_by_ghost, int* iters, int* status, pgamd_graph** out) {
  if (!out) return set_error(PGAMD_ERR_ARG, "null graph pointer");
  int rc = check_mixed_args(c, pl, rho, cpot, guess, pot_out, flu, a);
  if (rc != PGAMD_OK) return rc;
  MixedArgs m{pl, rho, cpot, guess, pot_out, flu, a, omega_by_ghost, iters, status};
  return capture(
      c, out,
      [](pgamd_context* cc, void* v) {
        auto* m = (MixedArgs*)v;
        return enqueue_mixed(cc, m->pl, m->rho, m->cpot, m->guess, m->pot_out, m->flu, *m->a, m->omega, m->iters,
                             m->status, ctx_capture_stream(cc));
      },
      &m);
}

int pgamd_graph_capture_device(pgamd_context* c, const float* rho, const float* cpot, const float* guess,
                               float* pot_out, float* flu, int npatch, const pgamd_solver_args* a, int* iters,
                               int* status, pgamd_graph** out) {
  if (!c || !a || !out) return set_error(PGAMD_ERR_ARG, "null argument");
  int rc = validate_args(ctx_options(c), *a, npatch);
  if (rc != PGAMD_OK) return rc;
  if (npatch <= 0) return set_error(PGAMD_ERR_ARG, "a graph needs at least one patch");
  if ((a->poisson && (!rho || !cpot)) || !pot_out || (a->graacc && !flu))
    return set_error(PGAMD_ERR_ARG, "null device array");
  if (a->poisson && ctx_options(c).init_mode >= PGAMD_INIT_GUESS && !guess)
    return set_error(PGAMD_ERR_ARG, "init_mode GUESS / GUESS_CUBE needs a guess array");
  DeviceArgs d{rho, cpot, guess, pot_out, flu, npatch, a, iters, status};
  return capture(
      c, out,
      [](pgamd_context* cc, void* v) {
        auto* d = (DeviceArgs*)v;
        const SolveParams p = make_params(ctx_options(cc), *d->a);
        return enqueue_solve(cc, d->rho, d->cpot, d->guess, d->pot_out, d->a->graacc ? d->flu : nullptr, d->npatch,
                             *d->a, p, d->iters, d->status, ctx_capture_stream(cc));
      },
      &d);
}

int pgamd_graph_launch(pgamd_graph* g, void* stream) {
  if (!g || !g->exec) return set_error(PGAMD_ERR_ARG, "null graph");
  hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  return e == hipSuccess ? PGAMD_OK : set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
}

int pgamd_graph_destroy(pgamd_graph* g) {
  if (!g) return PGAMD_OK;
  hipSetDevice(g->device);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return PGAMD_OK;
}

}  // extern "C"

// ---- autotuner ------------------------------------------------------------------------------------------
//
// Search space per ghost size: patches per wavefront in {1, 2, 4} (where they fit the 64 lanes) x row
// padding in {0..3} floats, plus the compact layout at ghost size 5 (one patch per wavefront).  Each candidate runs one untimed solve and `repeats` timed ones on the caller's
// stream-less default queue; the fastest mean wins.  The table is cached as a small JSON document:
//   {"device": "<name>", "layouts": {"1": [ppw, pad], ..., "5": [ppw, pad]}}

#include <cstdio>
#include <string>

extern "C" {

int pgamd_autotune(pgamd_context* c, int gs, const float* rho, const float* cpot, float* pot_out, int npatch,
                   const pgamd_solver_args* a, int repeats, int* best_ppw, int* best_pad, float* best_ms) {
  if (!c || !a || !rho || !cpot || !pot_out) return set_error(PGAMD_ERR_ARG, "null argument");
  if (gs < 1 || gs > kMaxGhost || npatch <= 0 || repeats < 1) return set_error(PGAMD_ERR_ARG, "bad autotune size");
  if (hipSetDevice(ctx_device(c)) != hipSuccess) return set_error(PGAMD_ERR_DEVICE, "hipSetDevice failed");
  pgamd_options o = ctx_options(c);
  o.pot_ghost_size = gs;
  o.pot_scheme = PGAMD_SOR;
  o.fuse_gravity = 0;
  // the candidates are timed without a guess array: a GUESS start is timed as the zero start (same sweeps)
  if (o.init_mode >= PGAMD_INIT_GUESS) o.init_mode = PGAMD_INIT_ZERO;
  o.guess_fallback_tol = 0.0f;
  pgamd_solver_args at = *a;
  at.graacc = 0;
  at.poisson = 1;
  int rc = validate_args(o, at, npatch);
  if (rc != PGAMD_OK) return rc;
  const SolveParams p = make_params(o, at);
  hipStream_t st = ctx_stream(c);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return set_error(PGAMD_ERR_DEVICE, "hipEventCreate failed");
  SorLayout& lay = ctx_layout(c);
  int bp = lay.ppw[gs], bd = lay.pad[gs];
  float bt = 1e30f;
  UniformBatch b{rho, cpot, nullptr, pot_out, nullptr, nullptr, nullptr, npatch, gs, PGAMD_SOR};
  for (int ppw : {1, 2, 4}) {
    if (!supported_ppw(gs, ppw)) continue;
    for (int pad = 0; pad <= kPadCompact; pad++) {
      if (!supported_layout(gs, ppw, pad)) continue;
      hipError_t e = launch_sor_uniform_dispatch(b, p, ppw, pad, st);
      if (e != hipSuccess) continue;  // layout does not launch on this device (e.g. LDS): skip it
      hipEventRecord(e0, st);
      for (int r = 0; r < repeats; r++) launch_sor_uniform_dispatch(b, p, ppw, pad, st);
      hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return set_error(PGAMD_ERR_DEVICE, "autotune synchronisation failed");
      }
      float ms = 0.0f;
      hipEventElapsedTime(&ms, e0, e1);
      ms /= repeats;
      if (ms < bt) {
        bt = ms;
        bp = ppw;
        bd = pad;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  lay.ppw[gs] = bp;
  lay.pad[gs] = bd;
  if (best_ppw) *best_ppw = bp;
  if (best_pad) *best_pad = bd;
  if (best_ms) *best_ms = bt;
  return PGAMD_OK;
}

int pgamd_autotune_save(pgamd_context* c, const char* path) {
  if (!c || !path) return set_error(PGAMD_ERR_ARG, "null argument");
  char name[256];
  pgamd_device_info(c, name, sizeof(name));
  FILE* f = fopen(path, "w");
  if (!f) return set_error(PGAMD_ERR_ARG, "cannot open the autotune cache for writing");
  const SorLayout& lay = ctx_layout(c);
  std::string dev(name);
  dev = dev.substr(0, dev.find(';'));
  fprintf(f, "{\"device\": \"%s\", \"layouts\": {", dev.c_str());
  for (int g = 1; g <= kMaxGhost; g++) fprintf(f, "%s\"%d\": [%d, %d]", g > 1 ? ", " : "", g, lay.ppw[g], lay.pad[g]);
  fprintf(f, "}}\n");
  fclose(f);
  return PGAMD_OK;
}

int pgamd_autotune_load(pgamd_context* c, const char* path) {
  if (!c || !path) return set_error(PGAMD_ERR_ARG, "null argument");
  FILE* f = fopen(path, "r");
  if (!f) return set_error(PGAMD_ERR_ARG, "cannot open the autotune cache");
  std::string s;
  char buf[512];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
  fclose(f);
  char name[256];
  pgamd_device_info(c, name, sizeof(name));
  std::string dev(name);
  dev = dev.substr(0, dev.find(';'));
  if (s.find("\"device\": \"" + dev + "\"") == std::string::npos)
    return set_error(PGAMD_ERR_STATE, "autotune cache was written for another device");
  int ppw[6], pad[6];
  for (int g = 1; g <= kMaxGhost; g++) {
    const std::string key = "\"" + std::to_string(g) + "\": [";
    const size_t k = s.find(key);
    if (k == std::string::npos || sscanf(s.c_str() + k + key.size(), "%d, %d", &ppw[g], &pad[g]) != 2 ||
        !supported_layout(g, ppw[g], pad[g]))
      return set_error(PGAMD_ERR_ARG, "malformed autotune cache");
  }
  SorLayout& lay = ctx_layout(c);
  for (int g = 1; g <= kMaxGhost; g++) {
    lay.ppw[g] = ppw[g];
    lay.pad[g] = pad[g];
  }
  return PGAMD_OK;
}

}  // extern "C"
