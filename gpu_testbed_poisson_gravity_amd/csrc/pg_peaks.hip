// Peak microbenchmarks for the roofline report: LDS read bandwidth (ds_read_b128 streaming) and HBM copy
// bandwidth, measured on the device the context runs on.  No reference counterpart.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pg_amd.h"
#include "pg_internal.h"

namespace {

constexpr int kLdsThreads = 256;
constexpr int kLdsVec = 2048;        // uint4 per workgroup buffer (32 KB)
constexpr int kLdsUnroll = 8;        // independent reads in flight per thread
constexpr int kLdsStride = 128;      // uint4 between the eight reads of one step (2 KB): with a base below 1024,
                                     // base + 7 * 128 stays inside the buffer, so every address is the base
                                     // register plus a constant offset of the instruction

// Each lane reads one 16-byte word per read; the 64 lanes of a wavefront read 64 consecutive words (one
// contiguous 1 KB span, no bank conflicts).  Eight ds_read_b128 are issued back to back and waited for inside
// one asm statement (s_waitcnt lgkmcnt(0) before the statement ends), so every destination register holds its
// data when the compiler next sees it; the outputs are early-clobber, so no destination shares a register with
// the address.  The base moves by 512 words every step (modulo 1024), so no read can be hoisted, and the
// statement is volatile, so none is dropped; one component of the first read is folded into a running XOR for
// the sink.  The buffer is the kernel's dynamic LDS, which starts at LDS offset 0 (no static LDS here), so a
// word's byte address is 16 times its index.
__global__ void __launch_bounds__(kLdsThreads) lds_read_kernel(int iters, uint32_t* sink) {
  extern __shared__ uint4 buf[];
  static_assert(kLdsUnroll == 8 && kLdsStride * 16 == 2048, "the asm statement issues eight reads 2 KB apart");
  for (int i = threadIdx.x; i < kLdsVec; i += kLdsThreads)
    buf[i] = make_uint4(i, i * 3u + 1u, i * 5u + 2u, i * 7u + 3u);
  __syncthreads();
  uint32_t a = 0;
  for (int it = 0; it < iters; it++) {
    const uint32_t base = (uint32_t)(((it * kLdsUnroll * 64 + (int)threadIdx.x) & 1023) * 16);
    uint4 r0, r1, r2, r3, r4, r5, r6, r7;
    asm volatile(
        "ds_read_b128 %0, %8\n\t"
        "ds_read_b128 %1, %8 offset:2048\n\t"
        "ds_read_b128 %2, %8 offset:4096\n\t"
        "ds_read_b128 %3, %8 offset:6144\n\t"
        "ds_read_b128 %4, %8 offset:8192\n\t"
        "ds_read_b128 %5, %8 offset:10240\n\t"
        "ds_read_b128 %6, %8 offset:12288\n\t"
        "ds_read_b128 %7, %8 offset:14336\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6), "=&v"(r7)
        : "v"(base)
        : "memory");
    a ^= r0.x;
  }
  if (a == 0x9e3779b9u) sink[blockIdx.x] = a;   // practically never true
}

__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

#define PK(call)                                                                     \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) { rc = set_error(PGAMD_ERR_DEVICE, hipGetErrorString(e_)); goto done; } \
  } while (0)

}  // namespace

extern "C" int pgamd_measure_peaks(pgamd_context* ctx, double* lds_read_bytes_per_s, double* hbm_copy_bytes_per_s) {
  using namespace pgamd;
  if (!ctx) return set_error(PGAMD_ERR_ARG, "null context");
  int rc = PGAMD_OK;
  const int cus = ctx_cu_count(ctx);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  uint32_t* sink = nullptr;
  float4 *src = nullptr, *dst = nullptr;
  const size_t copy_bytes = size_t(1) << 30;
  const size_t n4 = copy_bytes / sizeof(float4);
  float ms = 0.0f;
  const int lds_blocks = cus * 8, lds_iters = 4096;
  PK(hipSetDevice(ctx_device(ctx)));
  PK(hipEventCreate(&e0));
  PK(hipEventCreate(&e1));
  if (lds_read_bytes_per_s) {
    PK(hipMalloc(&sink, sizeof(uint32_t) * lds_blocks));
    for (int rep = 0; rep < 2; rep++) {   // the first launch warms up, the second is timed
      PK(hipEventRecord(e0, nullptr));
      hipLaunchKernelGGL(lds_read_kernel, dim3(lds_blocks), dim3(kLdsThreads), kLdsVec * sizeof(uint4), nullptr,
                         lds_iters, sink);
      PK(hipGetLastError());
      PK(hipEventRecord(e1, nullptr));
      PK(hipEventSynchronize(e1));
    }
    PK(hipEventElapsedTime(&ms, e0, e1));
    *lds_read_bytes_per_s = (double)lds_blocks * kLdsThreads * lds_iters * kLdsUnroll * 16.0 / (ms * 1e-3);
  }
  if (hbm_copy_bytes_per_s) {
    PK(hipMalloc(&src, copy_bytes));
    PK(hipMalloc(&dst, copy_bytes));
    PK(hipMemset(src, 0, copy_bytes));
    for (int rep = 0; rep < 3; rep++) {
      PK(hipEventRecord(e0, nullptr));
      hipLaunchKernelGGL(copy_kernel, dim3(cus * 16), dim3(256), 0, nullptr, src, dst, n4);
      PK(hipGetLastError());
      PK(hipEventRecord(e1, nullptr));
      PK(hipEventSynchronize(e1));
    }
    PK(hipEventElapsedTime(&ms, e0, e1));
    *hbm_copy_bytes_per_s = 2.0 * (double)copy_bytes / (ms * 1e-3);   // read + write
  }
done:
  if (sink) (void)hipFree(sink);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}
