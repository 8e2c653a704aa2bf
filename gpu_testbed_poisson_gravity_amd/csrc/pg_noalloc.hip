// An allocation-free variant of the layer-by-layer warm-start network forward (pgamd_cnn_forward), for callers
// that cannot allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  No reference counterpart.
//
// pgamd_cnn_forward keeps its activations in a workspace it allocates on first use.  This entry keeps them in
// `scratch`, a file-scope device array of fixed size: five padded activation buffers for up to
// kNoAllocMaxSamples samples (5 x 64 x 32,000 bytes = 10.24 MB; every GPU the library runs on holds one copy until
// the process ends).  The caller passes the input and the output as usual; the last layer writes the output
// directly.  Every call on a GPU shares that GPU's array, so calls that overlap on one GPU race; the part of
// the array a call uses is cleared by a fill kernel at the start of the call (the first layer's padding must
// be zero), so nothing depends on what an earlier call left.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/pg_amd.h"
#include "pg_cnn.h"
#include "pg_internal.h"

namespace pgamd {
namespace {

constexpr int kNoAllocMaxSamples = 64;
constexpr size_t kActBytes = (size_t)kNoAllocMaxSamples * kCnnInput * sizeof(uint16_t);   // one buffer
constexpr size_t kScratchBytes = 5 * kActBytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};   // scratch's device address, per device (each GPU has its own copy)

// Device address of this device's copy of `scratch`: looked up once per device, with that device current
// (the array's name is a host-side stand-in, never a device address).
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(PGAMD_ERR_DEVICE, "device id outside the table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_error(PGAMD_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed");
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return PGAMD_OK;
}

__global__ void zero_kernel(uint4* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

}  // namespace
}  // namespace pgamd

extern "C" int pgamd_cnn_forward_noalloc(pgamd_cnn* net, const uint16_t* d_x0, int count, float* d_out,
                                         void* stream) {
  using namespace pgamd;
  if (!net || !d_x0 || !d_out || count <= 0) return set_error(PGAMD_ERR_ARG, "bad CNN forward arguments");
  if (count > kNoAllocMaxSamples)
    return set_error(PGAMD_ERR_ARG, "pgamd_cnn_forward_noalloc takes at most 64 samples per call");
  const int device = cnn_device(net);
  if (hipSetDevice(device) != hipSuccess) return set_error(PGAMD_ERR_DEVICE, "hipSetDevice failed");
  void* base = nullptr;
  int rc = scratch_address(device, &base);
  if (rc != PGAMD_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  uint16_t* y[6] = {};
  for (int l = 1; l <= 5; l++) y[l] = reinterpret_cast<uint16_t*>(static_cast<unsigned char*>(base) + (l - 1) * kActBytes);
  // the used part of each buffer starts at zero (the padding of the next layer's input)
  const size_t used = (size_t)count * kCnnInput * sizeof(uint16_t) / sizeof(uint4);
  for (int l = 1; l <= 5; l++) {
    hipLaunchKernelGGL(zero_kernel, dim3(256), dim3(256), 0, s, reinterpret_cast<uint4*>(y[l]), used);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
  }
  const uint16_t* in = d_x0;
  for (int l = 1; l <= 6; l++) {
    const hipError_t e = cnn_conv_fwd(in, cnn_forward_weights(net, l), cnn_bias(net, l), l < 6 ? y[l] : nullptr,
                                      nullptr, l < 6 ? nullptr : d_out, count, l == 1 ? 0 : (l < 6 ? 1 : 2), s);
    if (e != hipSuccess) return set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e));
    if (l < 6) in = y[l];
  }
  return PGAMD_OK;
}
