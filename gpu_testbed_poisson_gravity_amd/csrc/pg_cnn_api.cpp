// C-ABI of the warm-start CNN (include/pg_amd.h, pgamd_cnn_*): parameters, training step (forward, MSE,
// backward, Adam) and the inference pipeline that turns a batch of patches into an interior guess.
//
// Parameter vector (FP32 master copy), in this order:
//   W1 [16][27*8]  b1[16]   W2..W5 [16][27*16]  b2..b5[16]   w6 [27*16]  b6[1]        (31617 values)
// W* are indexed [co][tap*CIN + ci] with tap = dz*9 + dy*3 + dx.  Forward passes use BF16 copies of W1..W5
// (and of w6, rounded to BF16 values); gradients are reduced in a fixed order (bitwise reproducible).

#include <hip/hip_runtime.h>

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

constexpr int kCin1 = 8;
constexpr int kK1 = 27 * kCin1;   // 216
constexpr int kK = 27 * 16;       // 432
constexpr int oW[6] = {0, 0, 3472, 10400, 17328, 24256};   // oW[l], l = 1..5
constexpr int oB[6] = {0, 3456, 10384, 17312, 24240, 31168};
constexpr int oW6 = 31184, oB6 = 31616, kNParams = 31617;
constexpr int kSin[6] = {0, 18, 16, 14, 12, 10};           // input edge of layer l
constexpr int kMseBlocks = 64;

int sz3(int s) { return s * s * s; }

}  // namespace

struct pgamd_cnn {
  int device = 0;
  int step = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* wb = nullptr;   // bf16 W1 | W2 | W3 | W4 | W5 (forward layout)
  uint16_t* wt = nullptr;   // bf16 transposed W2..W5
  float* w6r = nullptr;     // w6 rounded to bf16 values, then b6
  // workspace
  int cap = 0;              // samples
  bool train_ws = false;
  uint16_t* y[6] = {};      // y1..y5 (y[0] unused: input)
  uint16_t* h[6] = {};      // pre-activations (training only)
  uint16_t* dh = nullptr;
  float *d = nullptr, *dd = nullptr, *dyA = nullptr, *dyB = nullptr, *part = nullptr, *lossp = nullptr;
  size_t part_floats = 0;
  // inference
  int icap = 0;
  uint16_t* x0 = nullptr;
  float* scale = nullptr;
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
  hipFree(n->dd);
  hipFree(n->dyA);
  hipFree(n->dyB);
  hipFree(n->part);
  hipFree(n->lossp);
  n->dh = nullptr;
  n->d = n->dd = n->dyA = n->dyB = n->part = n->lossp = nullptr;
  n->cap = 0;
}

int ensure_ws(pgamd_cnn* n, int count, bool train) {
  if (count <= n->cap && (!train || n->train_ws)) return PGAMD_OK;
  free_ws(n);
  const size_t c = (size_t)count;
  bool ok = true;
  for (int l = 1; l <= 5; l++) {
    const size_t e = c * sz3(kSin[l] - 2) * 16;
    ok &= hipMalloc(&n->y[l], e * 2) == hipSuccess;
    if (train) ok &= hipMalloc(&n->h[l], e * 2) == hipSuccess;
  }
  ok &= hipMalloc(&n->d, c * 512 * 4) == hipSuccess;
  if (train) {
    ok &= hipMalloc(&n->dh, c * sz3(16) * 16 * 2) == hipSuccess;
    ok &= hipMalloc(&n->dd, c * 512 * 4) == hipSuccess;
    ok &= hipMalloc(&n->dyA, c * sz3(16) * 16 * 4) == hipSuccess;
    ok &= hipMalloc(&n->dyB, c * sz3(16) * 16 * 4) == hipSuccess;
    size_t pf = (size_t)cnn_bwd_weight_blocks(count, 16) * 16 * kK;       // largest weight partials
    pf = std::max(pf, (size_t)cnn_bwd_weight_blocks(count, 18) * 16 * kK1);
    pf = std::max(pf, c * 433);
    pf = std::max(pf, (size_t)((c * sz3(16) * 16 + 255) / 256 < 4096 ? (c * sz3(16) * 16 + 255) / 256 : 4096) * 16);
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

// BF16 copies used by the forward and input-gradient passes.
int refresh_copies(pgamd_cnn* n, hipStream_t s) {
  CK(cnn_to_bf16(n->p + oW[1], n->wb, 16 * kK1, s));
  for (int l = 2; l <= 5; l++) {
    CK(cnn_to_bf16(n->p + oW[l], n->wb + 16 * kK1 + (l - 2) * 16 * kK, 16 * kK, s));
    CK(cnn_transpose_w(n->p + oW[l], n->wt + (l - 2) * 16 * kK, s));
  }
  // w6 as BF16-rounded FP32 values, b6 as is
  CK(cnn_round_bf16(n->p + oW6, n->w6r, 432, s));
  CK(hipMemcpyAsync(n->w6r + 432, n->p + oB6, 4, hipMemcpyDeviceToDevice, s));
  return PGAMD_OK;
}

const uint16_t* wb_layer(const pgamd_cnn* n, int l) { return l == 1 ? n->wb : n->wb + 16 * kK1 + (l - 2) * 16 * kK; }

int forward(pgamd_cnn* n, const uint16_t* x0, int count, bool keep_h, hipStream_t s) {
  const uint16_t* in = x0;
  for (int l = 1; l <= 5; l++) {
    CK(cnn_conv_fwd(in, wb_layer(n, l), n->p + oB[l], n->y[l], keep_h ? n->h[l] : nullptr, count, kSin[l],
                    l == 1 ? kCin1 : 16, l > 1, s));
    in = n->y[l];
  }
  CK(cnn_conv6_fwd(n->y[5], n->w6r, n->w6r + 432, n->d, count, s));
  return PGAMD_OK;
}

__attribute__((unused)) int zero_grad(pgamd_cnn* n, hipStream_t s) {
  CK(hipMemsetAsync(n->g, 0, kNParams * 4, s));
  return PGAMD_OK;
}

// Forward + backward; gradients (mean-squared-error loss) accumulate into n->g, loss partials into lossp.
int forward_backward(pgamd_cnn* n, const uint16_t* x0, const float* target, int count, hipStream_t s) {
  int rc = forward(n, x0, count, true, s);
  if (rc != PGAMD_OK) return rc;
  CK(hipMemsetAsync(n->g, 0, kNParams * 4, s));
  CK(cnn_mse(n->d, target, n->dd, n->lossp, count * 512, kMseBlocks, s));
  CK(cnn_conv6_bwd(n->y[5], n->w6r, n->dd, n->dyA, n->part, count, s));
  CK(cnn_reduce_parts(n->part, count, 433, n->g + oW6, s));
  float* dy = n->dyA;
  float* dx = n->dyB;
  for (int l = 5; l >= 1; l--) {
    const int S = kSin[l], So = S - 2;
    const size_t elems = (size_t)count * sz3(So) * 16;
    int nb = (int)std::min<size_t>((elems + 255) / 256, 4096);
    CK(cnn_relu_bwd(dy, n->h[l], n->dh, n->part, elems, nb, s));
    CK(cnn_reduce_parts(n->part, nb, 16, n->g + oB[l], s));
    const uint16_t* xin = l == 1 ? x0 : n->y[l - 1];
    const int cin = l == 1 ? kCin1 : 16;
    CK(cnn_conv_bwd_weight(xin, n->dh, n->part, count, S, cin, s));
    CK(cnn_reduce_parts(n->part, cnn_bwd_weight_blocks(count, S), 16 * 27 * cin, n->g + oW[l], s));
    if (l > 1) {
      CK(cnn_conv_bwd_data(n->dh, n->wt + (l - 2) * 16 * kK, l > 1 ? dy : nullptr, dx, count, S, s));
      std::swap(dy, dx);
    }
  }
  return PGAMD_OK;
}

float read_loss(pgamd_cnn* n, int count, hipStream_t s) {
  float h[kMseBlocks];
  if (hipMemcpyAsync(h, n->lossp, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return NAN;
  double t = 0.0;
  for (float x : h) t += x;
  return (float)(t / ((double)count * 512));
}

}  // namespace

extern "C" {

int pgamd_cnn_num_params(void) { return kNParams; }

int pgamd_cnn_create(pgamd_context* ctx, uint64_t seed, pgamd_cnn** out) {
  if (!ctx || !out) return set_error(PGAMD_ERR_ARG, "null argument");
  *out = nullptr;
  auto* n = new pgamd_cnn();
  n->device = ctx_device(ctx);
  if (hipSetDevice(n->device) != hipSuccess || hipMalloc(&n->p, kNParams * 4) != hipSuccess ||
      hipMalloc(&n->g, kNParams * 4) != hipSuccess || hipMalloc(&n->m, kNParams * 4) != hipSuccess ||
      hipMalloc(&n->v, kNParams * 4) != hipSuccess ||
      hipMalloc(&n->wb, (16 * kK1 + 4 * 16 * kK) * 2) != hipSuccess ||
      hipMalloc(&n->wt, 4 * 16 * kK * 2) != hipSuccess || hipMalloc(&n->w6r, 433 * 4) != hipSuccess) {
    pgamd_cnn_destroy(n);
    return set_error(PGAMD_ERR_NOMEM, "CNN parameter allocation failed");
  }
  // He-normal initialisation; residual branches start small and the output layer at zero, so an untrained
  // network returns the interpolated guess unchanged.
  std::vector<float> h(kNParams, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int co = 0; co < 16; co++)
    for (int tap = 0; tap < 27; tap++)
      for (int ci = 0; ci < 2; ci++) h[oW[1] + co * kK1 + tap * kCin1 + ci] = nd(rng) * std::sqrt(2.0f / 54.0f);
  for (int l = 2; l <= 5; l++)
    for (int i = 0; i < 16 * kK; i++) h[oW[l] + i] = 0.1f * nd(rng) * std::sqrt(2.0f / kK);
  if (hipMemcpy(n->p, h.data(), kNParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, kNParams * 4) != hipSuccess || hipMemset(n->v, 0, kNParams * 4) != hipSuccess) {
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
  hipFree(n->w6r);
  hipFree(n->x0);
  hipFree(n->scale);
  delete n;
  return PGAMD_OK;
}

int pgamd_cnn_get_params(pgamd_cnn* n, float* host, int count) {
  if (!n || !host || count != kNParams) return set_error(PGAMD_ERR_ARG, "parameter buffer must hold 31617 floats");
  hipSetDevice(n->device);
  return hipMemcpy(host, n->p, kNParams * 4, hipMemcpyDeviceToHost) == hipSuccess
             ? PGAMD_OK
             : set_error(PGAMD_ERR_DEVICE, "parameter copy failed");
}

// Sets the FP32 master weights, resets the optimiser state and refreshes the BF16 copies.
int pgamd_cnn_set_params(pgamd_cnn* n, const float* host, int count) {
  if (!n || !host || count != kNParams) return set_error(PGAMD_ERR_ARG, "parameter buffer must hold 31617 floats");
  hipSetDevice(n->device);
  if (hipMemcpy(n->p, host, kNParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, kNParams * 4) != hipSuccess || hipMemset(n->v, 0, kNParams * 4) != hipSuccess)
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
  CK(hipMemcpyAsync(d_out, n->d, (size_t)count * 512 * 4, hipMemcpyDeviceToDevice, s));
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
  if (d_grad) CK(hipMemcpyAsync(d_grad, n->g, kNParams * 4, hipMemcpyDeviceToDevice, s));
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
  CK(cnn_adam(n->p, n->g, n->m, n->v, kNParams, lr, b1, b2, eps, c1, c2, s));
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
  const int gs = ctx_options(ctx).pot_ghost_size;
  hipSetDevice(n->device);
  hipStream_t s = (hipStream_t)stream;
  if (npatch > n->icap) {
    hipFree(n->x0);
    hipFree(n->scale);
    n->x0 = nullptr;
    n->scale = nullptr;
    n->icap = 0;
    if (hipMalloc(&n->x0, (size_t)npatch * sz3(18) * 8 * 2) != hipSuccess ||
        hipMalloc(&n->scale, (size_t)npatch * 4) != hipSuccess)
      return set_error(PGAMD_ERR_NOMEM, "warm-start workspace allocation failed");
    n->icap = npatch;
  }
  int rc = pgamd_cnn_prepare(ctx, d_rho, d_pot_in, npatch, gs, a, n->x0, n->scale, stream);
  if (rc != PGAMD_OK) return rc;
  rc = ensure_ws(n, npatch, n->train_ws && npatch <= n->cap);
  if (rc != PGAMD_OK) return rc;
  rc = forward(n, n->x0, npatch, false, s);
  if (rc != PGAMD_OK) return rc;
  CK(cnn_guess(d_pot_in, n->d, n->scale, npatch, gs, a->int_scheme, d_guess, s));
  return PGAMD_OK;
}

}  // extern "C"
