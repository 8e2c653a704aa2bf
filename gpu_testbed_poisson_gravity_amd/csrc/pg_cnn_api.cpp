// C-ABI of the warm-start CNN (include/pg_amd.h, pgamd_cnn_*): parameters, training step (forward, MSE,
// backward, Adam) and the inference pipeline that turns a batch of patches into an interior guess.
//
// Parameter vector (FP32 master copy): for each of the 6 layers W [16][432] then b [16] (41568 values);
// W[co][tap*16 + ci] with tap = dz*9 + dy*3 + dx.  Rows 8..15 of the output layer are unused.  Forward and
// input-gradient passes use BF16 copies of the weights; gradients are reduced in a fixed order (bitwise
// reproducible).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/pg_amd.h"
#include "pg_cnn.h"
#include "pg_geometry.h"
#include "pg_internal.h"

using namespace pgamd;

namespace {

constexpr int kW = kCnnChannels * kCnnK;   // weights per layer
constexpr int kWF = kCnnChannels * kCnnKF; // forward-layout BF16 weights per layer
constexpr int kMseBlocks = 64;
constexpr int kMaxReduceBlocks = 4096;
constexpr int kMaxChunks = 16;               // pgamd_cnn_solve_device

int ow(int l) { return (l - 1) * kCnnLayerParams; }   // layer l = 1..6
int ob(int l) { return ow(l) + kW; }

}  // namespace

struct pgamd_cnn {
  int device = 0;
  int step = 0;
  float res_weight = 0.0f;  // weight of the residual term of the loss (pgamd_cnn_set_loss)
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* wb = nullptr;   // BF16 W1..W6 in the forward layout [16][480]
  uint16_t* wt = nullptr;   // BF16 W1..W6 transposed per tap (input-gradient pass)
  uint16_t* wl = nullptr;   // BF16 W1..W6 forward layout, lane-major (the fused warm-start kernel)
  // workspace
  int cap = 0;              // samples
  bool train_ws = false;
  uint16_t* y[6] = {};      // y1..y5, padded [n][10^3][16] with zero borders (y[0] unused: the input)
  uint16_t* h[6] = {};      // pre-activations h1..h5 [n][8^3][16] (training only)
  uint16_t* dh = nullptr;   // [n][8^3][16]
  float *d = nullptr, *dy6 = nullptr, *dyA = nullptr, *dyB = nullptr, *part = nullptr, *lossp = nullptr;
  size_t part_floats = 0;
  // pipelined warm-start solve: fork event and one event per chunk
  hipEvent_t ev[kMaxChunks + 1] = {};
};

namespace {

void free_ws(pgamd_cnn* n) {
  for (int l = 1; l <= 5; l++) {
    hipFree(n->y[l]);
    hipFree(n->h[l]);
    n->y[l] = n->h[l] = nullptr;
  }
  hipFree(n->dh);
  hipFree(n->d);
  hipFree(n->dy6);
  hipFree(n->dyA);
  hipFree(n->dyB);
  hipFree(n->part);
  hipFree(n->lossp);
  n->dh = nullptr;
  n->d = n->dy6 = n->dyA = n->dyB = n->part = n->lossp = nullptr;
  n->cap = 0;
  n->train_ws = false;
}

int reduce_blocks(size_t elems) { return (int)std::min<size_t>((elems + 255) / 256, kMaxReduceBlocks); }

int ensure_ws(pgamd_cnn* n, int count, bool train) {
  if (count <= n->cap && (!train || n->train_ws)) return PGAMD_OK;
  free_ws(n);
  const size_t c = (size_t)count;
  const size_t act = c * kCnnVox * kCnnChannels;
  bool ok = true;
  for (int l = 1; l <= 5; l++) {
    const size_t e = c * kCnnInput;
    ok &= hipMalloc(&n->y[l], e * 2) == hipSuccess;
    if (ok) ok &= hipMemset(n->y[l], 0, e * 2) == hipSuccess;   // borders stay zero for good
    if (train) ok &= hipMalloc(&n->h[l], act * 2) == hipSuccess;
  }
  ok &= hipMalloc(&n->d, c * kCnnOutputs * 4) == hipSuccess;
  if (train) {
    ok &= hipMalloc(&n->dh, act * 2) == hipSuccess;
    ok &= hipMalloc(&n->dy6, act * 4) == hipSuccess;
    ok &= hipMalloc(&n->dyA, act * 4) == hipSuccess;
    ok &= hipMalloc(&n->dyB, act * 4) == hipSuccess;
    const size_t pf = std::max((size_t)cnn_bwd_weight_blocks(count) * kW, (size_t)reduce_blocks(act) * kCnnChannels);
    ok &= hipMalloc(&n->part, pf * 4) == hipSuccess;
    n->part_floats = pf;
    ok &= hipMalloc(&n->lossp, kMseBlocks * 4) == hipSuccess;
  }
  if (!ok) {
    free_ws(n);
    return set_error(PGAMD_ERR_NOMEM, "CNN workspace allocation failed");
  }
  n->cap = count;
  n->train_ws = train;
  return PGAMD_OK;
}

#define CK(call)                                                                     \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e_)); \
  } while (0)

int refresh_copies(pgamd_cnn* n, hipStream_t s) {
  CK(cnn_weight_copies(n->p, n->wb, n->wt, n->wl, s));
  return PGAMD_OK;
}

int forward(pgamd_cnn* n, const uint16_t* x0, int count, bool keep_h, hipStream_t s) {
  const uint16_t* in = x0;
  for (int l = 1; l <= 5; l++) {
    CK(cnn_conv_fwd(in, n->wb + (size_t)(l - 1) * kWF, n->p + ob(l), n->y[l], keep_h ? n->h[l] : nullptr, nullptr,
                    count, l == 1 ? 0 : 1, s));
    in = n->y[l];
  }
  CK(cnn_conv_fwd(in, n->wb + (size_t)5 * kWF, n->p + ob(6), nullptr, nullptr, n->d, count, 2, s));
  return PGAMD_OK;
}

// Forward + backward; gradients (mean squared error, plus the residual term when res_weight > 0) go to n->g,
// loss partials to lossp.
int forward_backward(pgamd_cnn* n, const uint16_t* x0, const float* target, int count, hipStream_t s) {
  int rc = forward(n, x0, count, true, s);
  if (rc != PGAMD_OK) return rc;
  const size_t act = (size_t)count * kCnnVox * kCnnChannels;
  const int nb = reduce_blocks(act);
  const int wblocks = cnn_bwd_weight_blocks(count);
  CK(cnn_mse(n->d, target, n->dy6, n->lossp, count * kCnnVox, kMseBlocks, s));
  if (n->res_weight > 0.0f) CK(cnn_res_loss(n->d, target, n->dy6, n->lossp, count, n->res_weight, kMseBlocks, s));
  const float* dy = n->dy6;
  float* bufs[2] = {n->dyA, n->dyB};
  int which = 0;
  for (int l = 6; l >= 1; l--) {
    CK(cnn_relu_bwd(dy, l == 6 ? nullptr : n->h[l], n->dh, n->part, act, nb, s));
    CK(cnn_reduce_parts(n->part, nb, kCnnChannels, n->g + ob(l), s));
    CK(cnn_conv_bwd_weight(l == 1 ? x0 : n->y[l - 1], n->dh, n->part, count, s));
    CK(cnn_reduce_parts(n->part, wblocks, kW, n->g + ow(l), s));
    if (l > 1) {
      float* dx = bufs[which];
      which ^= 1;
      CK(cnn_conv_bwd_data(n->dh, n->wt + (size_t)(l - 1) * kW, l == 6 ? nullptr : dy, dx, count, s));
      dy = dx;
    }
  }
  return PGAMD_OK;
}

float read_loss(pgamd_cnn* n, int count, hipStream_t s) {
  float h[kMseBlocks];
  if (hipMemcpyAsync(h, n->lossp, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return NAN;
  double t = 0.0;
  for (float x : h) t += x;
  return (float)(t / ((double)count * kCnnOutputs));
}

}  // namespace

namespace pgamd {
int cnn_device(const pgamd_cnn* n) { return n->device; }
const uint16_t* cnn_forward_weights(const pgamd_cnn* n, int l) { return n->wb + (size_t)(l - 1) * kWF; }
const float* cnn_bias(const pgamd_cnn* n, int l) { return n->p + ob(l); }
}  // namespace pgamd

extern "C" {

int pgamd_cnn_num_params(void) { return kCnnParams; }

int pgamd_cnn_set_loss(pgamd_cnn* n, float res_weight) {
  if (!n) return set_error(PGAMD_ERR_ARG, "null network");
  if (!(res_weight >= 0.0f) || !std::isfinite(res_weight))
    return set_error(PGAMD_ERR_ARG, "res_weight must be finite and >= 0");
  n->res_weight = res_weight;
  return PGAMD_OK;
}

int pgamd_cnn_create(pgamd_context* ctx, uint64_t seed, pgamd_cnn** out) {
  if (!ctx || !out) return set_error(PGAMD_ERR_ARG, "null argument");
  *out = nullptr;
  auto* n = new pgamd_cnn();
  n->device = ctx_device(ctx);
  if (hipSetDevice(n->device) != hipSuccess || hipMalloc(&n->p, kCnnParams * 4) != hipSuccess ||
      hipMalloc(&n->g, kCnnParams * 4) != hipSuccess || hipMalloc(&n->m, kCnnParams * 4) != hipSuccess ||
      hipMalloc(&n->v, kCnnParams * 4) != hipSuccess || hipMalloc(&n->wb, kCnnLayers * kWF * 2) != hipSuccess ||
      hipMalloc(&n->wt, kCnnLayers * kW * 2) != hipSuccess || hipMalloc(&n->wl, kCnnLayers * kWF * 2) != hipSuccess) {
    pgamd_cnn_destroy(n);
    return set_error(PGAMD_ERR_NOMEM, "CNN parameter allocation failed");
  }
  // He-normal first layer, residual branches scaled by 0.1, output layer and biases zero: an untrained
  // network returns the interpolated guess unchanged.
  std::vector<float> h(kCnnParams, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  const float he = std::sqrt(2.0f / kCnnK);
  for (int i = 0; i < kW; i++) h[ow(1) + i] = nd(rng) * he;
  for (int l = 2; l <= 5; l++)
    for (int i = 0; i < kW; i++) h[ow(l) + i] = 0.1f * nd(rng) * he;
  if (hipMemcpy(n->p, h.data(), kCnnParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, kCnnParams * 4) != hipSuccess || hipMemset(n->v, 0, kCnnParams * 4) != hipSuccess) {
    pgamd_cnn_destroy(n);
    return set_error(PGAMD_ERR_DEVICE, "CNN initialisation failed");
  }
  const int rc = refresh_copies(n, nullptr);
  if (rc != PGAMD_OK) {
    pgamd_cnn_destroy(n);
    return rc;
  }
  *out = n;
  return PGAMD_OK;
}

int pgamd_cnn_destroy(pgamd_cnn* n) {
  if (!n) return PGAMD_OK;
  hipSetDevice(n->device);
  hipDeviceSynchronize();
  free_ws(n);
  hipFree(n->p);
  hipFree(n->g);
  hipFree(n->m);
  hipFree(n->v);
  hipFree(n->wb);
  hipFree(n->wt);
  hipFree(n->wl);
  for (hipEvent_t e : n->ev)
    if (e) hipEventDestroy(e);
  delete n;
  return PGAMD_OK;
}

int pgamd_cnn_get_params(pgamd_cnn* n, float* host, int count) {
  if (!n || !host || count != kCnnParams) return set_error(PGAMD_ERR_ARG, "parameter buffer must hold 41568 floats");
  hipSetDevice(n->device);
  return hipMemcpy(host, n->p, kCnnParams * 4, hipMemcpyDeviceToHost) == hipSuccess
             ? PGAMD_OK
             : set_error(PGAMD_ERR_DEVICE, "parameter copy failed");
}

// Sets the FP32 master weights, resets the optimiser state and refreshes the BF16 copies.
int pgamd_cnn_set_params(pgamd_cnn* n, const float* host, int count) {
  if (!n || !host || count != kCnnParams) return set_error(PGAMD_ERR_ARG, "parameter buffer must hold 41568 floats");
  hipSetDevice(n->device);
  if (hipMemcpy(n->p, host, kCnnParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, kCnnParams * 4) != hipSuccess || hipMemset(n->v, 0, kCnnParams * 4) != hipSuccess)
    return set_error(PGAMD_ERR_DEVICE, "parameter copy failed");
  n->step = 0;
  return refresh_copies(n, nullptr);
}

int pgamd_cnn_forward(pgamd_cnn* n, const uint16_t* d_x0, int count, float* d_out, void* stream) {
  if (!n || !d_x0 || !d_out || count <= 0) return set_error(PGAMD_ERR_ARG, "bad CNN forward arguments");
  hipSetDevice(n->device);
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_ws(n, count, n->train_ws && count <= n->cap);
  if (rc != PGAMD_OK) return rc;
  rc = forward(n, d_x0, count, false, s);
  if (rc != PGAMD_OK) return rc;
  CK(hipMemcpyAsync(d_out, n->d, (size_t)count * kCnnOutputs * 4, hipMemcpyDeviceToDevice, s));
  return PGAMD_OK;
}

int pgamd_cnn_loss_grad(pgamd_cnn* n, const uint16_t* d_x0, const float* d_target, int count, float* loss,
                        float* d_grad, void* stream) {
  if (!n || !d_x0 || !d_target || count <= 0) return set_error(PGAMD_ERR_ARG, "bad CNN arguments");
  hipSetDevice(n->device);
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_ws(n, count, true);
  if (rc != PGAMD_OK) return rc;
  rc = forward_backward(n, d_x0, d_target, count, s);
  if (rc != PGAMD_OK) return rc;
  if (d_grad) CK(hipMemcpyAsync(d_grad, n->g, kCnnParams * 4, hipMemcpyDeviceToDevice, s));
  if (loss) *loss = read_loss(n, count, s);
  return PGAMD_OK;
}

int pgamd_cnn_train_step(pgamd_cnn* n, const uint16_t* d_x0, const float* d_target, int count, float lr, float* loss,
                         void* stream) {
  if (!n || !d_x0 || !d_target || count <= 0 || !(lr > 0.0f)) return set_error(PGAMD_ERR_ARG, "bad training arguments");
  hipSetDevice(n->device);
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_ws(n, count, true);
  if (rc != PGAMD_OK) return rc;
  rc = forward_backward(n, d_x0, d_target, count, s);
  if (rc != PGAMD_OK) return rc;
  n->step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
  CK(cnn_adam(n->p, n->g, n->m, n->v, kCnnParams, lr, b1, b2, eps, c1, c2, s));
  rc = refresh_copies(n, s);
  if (rc != PGAMD_OK) return rc;
  if (loss) *loss = read_loss(n, count, s);
  return PGAMD_OK;
}

int pgamd_cnn_prepare(pgamd_context* ctx, const float* d_rho, const float* d_pot_in, int npatch, int ghost_size,
                      const pgamd_solver_args* a, uint16_t* d_x0, float* d_scale, void* stream) {
  if (!ctx || !a || !d_rho || !d_pot_in || !d_x0 || !d_scale || npatch <= 0)
    return set_error(PGAMD_ERR_ARG, "bad CNN prepare arguments");
  if (ghost_size < 1 || ghost_size > kMaxGhost) return set_error(PGAMD_ERR_ARG, "ghost size outside [1,5]");
  if (a->int_scheme != PGAMD_INT_CQUAD && a->int_scheme != PGAMD_INT_QUAD)
    return set_error(PGAMD_ERR_ARG, "incorrect parameter IntScheme");
  hipSetDevice(ctx_device(ctx));
  const float cst = a->poi_coeff * a->dh * a->dh;
  CK(cnn_prep(d_rho, d_pot_in, npatch, ghost_size, cst, a->int_scheme, d_x0, d_scale, (hipStream_t)stream));
  return PGAMD_OK;
}

int pgamd_cnn_warm_start(pgamd_context* ctx, pgamd_cnn* n, const float* d_rho, const float* d_pot_in, int npatch,
                         const pgamd_solver_args* a, float* d_guess, void* stream) {
  if (!ctx || !n || !a || !d_rho || !d_pot_in || !d_guess || npatch <= 0)
    return set_error(PGAMD_ERR_ARG, "bad warm-start arguments");
  if (a->int_scheme != PGAMD_INT_CQUAD && a->int_scheme != PGAMD_INT_QUAD)
    return set_error(PGAMD_ERR_ARG, "incorrect parameter IntScheme");
  const int gs = ctx_options(ctx).pot_ghost_size;
  hipSetDevice(n->device);
  const float cst = a->poi_coeff * a->dh * a->dh;
  // one launch: interpolation, packing, the six layers in LDS and the guess
  CK(cnn_fused_guess(d_rho, d_pot_in, n->wl, n->p, npatch, gs, cst, a->int_scheme, d_guess, 0, (hipStream_t)stream));
  return PGAMD_OK;
}

int pgamd_cnn_warm_start_cube(pgamd_context* ctx, pgamd_cnn* n, const float* d_rho, const float* d_pot_in, int npatch,
                              const pgamd_solver_args* a, float* d_guess, void* stream) {
  if (!ctx || !n || !a || !d_rho || !d_pot_in || !d_guess || npatch <= 0)
    return set_error(PGAMD_ERR_ARG, "bad warm-start arguments");
  if (a->int_scheme != PGAMD_INT_CQUAD && a->int_scheme != PGAMD_INT_QUAD)
    return set_error(PGAMD_ERR_ARG, "incorrect parameter IntScheme");
  const int gs = ctx_options(ctx).pot_ghost_size;
  hipSetDevice(n->device);
  const float cst = a->poi_coeff * a->dh * a->dh;
  CK(cnn_fused_guess(d_rho, d_pot_in, n->wl, n->p, npatch, gs, cst, a->int_scheme, d_guess, 1, (hipStream_t)stream));
  return PGAMD_OK;
}

int pgamd_cnn_warm_start_mixed(pgamd_context* ctx, pgamd_cnn* n, pgamd_plan* plan, const float* d_rho,
                               const float* d_pot_in, const pgamd_solver_args* a, float* d_guess, void* stream) {
  if (!ctx || !n || !plan || !a || !d_rho || !d_pot_in || !d_guess)
    return set_error(PGAMD_ERR_ARG, "bad warm-start arguments");
  if (plan_device(plan) != n->device) return set_error(PGAMD_ERR_ARG, "plan and network belong to different devices");
  if (a->int_scheme != PGAMD_INT_CQUAD && a->int_scheme != PGAMD_INT_QUAD)
    return set_error(PGAMD_ERR_ARG, "incorrect parameter IntScheme");
  const int np = plan_npatch(plan);
  if (np == 0) return PGAMD_OK;
  hipSetDevice(n->device);
  const float cst = a->poi_coeff * a->dh * a->dh;
  CK(cnn_fused_guess_mixed(d_rho, d_pot_in, plan_device_desc(plan), n->wl, n->p, np, cst, a->int_scheme, d_guess,
                           (hipStream_t)stream));
  return PGAMD_OK;
}

// Warm start and solve of a uniform batch, pipelined: the batch is cut into nchunk slices; the network of
// slice i+1 runs on the context's second stream while slice i is solved on `stream`, so the matrix-core
// network and the VALU-bound SOR sweeps share the GPU instead of running back to back.  Results are those
// of pgamd_cnn_warm_start followed by pgamd_solve_device (same kernels, same inputs per slice).
int pgamd_cnn_solve_device(pgamd_context* ctx, pgamd_cnn* n, const float* d_rho, const float* d_pot_in, float* d_guess,
                           float* d_pot_out, float* d_flu, int npatch, const pgamd_solver_args* a, int* d_iters,
                           int* d_status, int nchunk, void* stream) {
  if (!ctx || !n || !a || !d_rho || !d_pot_in || !d_guess || !d_pot_out)
    return set_error(PGAMD_ERR_ARG, "null argument");
  if (npatch < 0 || nchunk < 1 || nchunk > kMaxChunks) return set_error(PGAMD_ERR_ARG, "bad patch or chunk count");
  const pgamd_options& o = ctx_options(ctx);
  if ((o.init_mode != PGAMD_INIT_GUESS && o.init_mode != PGAMD_INIT_GUESS_CUBE) || !a->poisson)
    return set_error(PGAMD_ERR_ARG, "the warm-start solve needs init_mode GUESS or GUESS_CUBE and poisson = 1");
  if (a->int_scheme != PGAMD_INT_CQUAD && a->int_scheme != PGAMD_INT_QUAD)
    return set_error(PGAMD_ERR_ARG, "incorrect parameter IntScheme");
  if (npatch == 0) return PGAMD_OK;
  if (hipSetDevice(n->device) != hipSuccess) return set_error(PGAMD_ERR_DEVICE, "hipSetDevice failed");
  const Geo g = geo_of(o.pot_ghost_size);
  const size_t R3 = (size_t)g.R * g.R * g.R, C3 = (size_t)g.C * g.C * g.C;
  const size_t G = gra_nxt(o.gra_ghost_size), G3 = G * G * G, F = (size_t)kNComp * kPatchCells;
  const int cube = o.init_mode == PGAMD_INIT_GUESS_CUBE;
  const size_t Q3 = cube ? (size_t)g.N * g.N * g.N : R3;   // guess per patch
  hipStream_t s0 = (hipStream_t)stream, s1 = ctx_aux_stream(ctx);
  if (!s1) return set_error(PGAMD_ERR_DEVICE, "auxiliary stream creation failed");
  for (hipEvent_t& e : n->ev)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return set_error(PGAMD_ERR_DEVICE, "hipEventCreate failed");
  const float cst = a->poi_coeff * a->dh * a->dh;
  CK(hipEventRecord(n->ev[kMaxChunks], s0));  // the network waits for the caller's earlier work
  CK(hipStreamWaitEvent(s1, n->ev[kMaxChunks], 0));
  const int chunk = (npatch + nchunk - 1) / nchunk;
  for (int i = 0, off = 0; off < npatch; i++, off += chunk) {
    const int m = std::min(chunk, npatch - off);
    CK(cnn_fused_guess(d_rho + off * R3, d_pot_in + off * C3, n->wl, n->p, m, o.pot_ghost_size, cst, a->int_scheme,
                       d_guess + off * Q3, cube, s1));
    CK(hipEventRecord(n->ev[i], s1));
    CK(hipStreamWaitEvent(s0, n->ev[i], 0));
    const int rc = pgamd_solve_device(ctx, d_rho + off * R3, d_pot_in + off * C3, d_guess + off * Q3,
                                      d_pot_out + off * G3, d_flu ? d_flu + off * F : nullptr, m, a,
                                      d_iters ? d_iters + off : nullptr, d_status ? d_status + off : nullptr, s0);
    if (rc != PGAMD_OK) return rc;
  }
  return PGAMD_OK;
}

}  // extern "C"
