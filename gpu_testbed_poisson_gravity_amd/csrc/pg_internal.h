// Declarations shared between the C-ABI layer and the solver modules (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pg_amd.h"
#include "pg_kernels.h"

namespace pgamd {

int set_error(int code, const char* msg);
int validate_options(const pgamd_options& o);
int validate_args(const pgamd_options& o, const pgamd_solver_args& a, int npatch);
SolveParams make_params(const pgamd_options& o, const pgamd_solver_args& a);
int enqueue_solve(pgamd_context* c, const float* rho, const float* cpot, const float* guess, float* pot_out,
                  float* flu, int npatch, const pgamd_solver_args& a, const SolveParams& p, int* iters, int* status,
                  hipStream_t st);
const pgamd_options& ctx_options(const pgamd_context* c);
SorLayout& ctx_layout(pgamd_context* c);
int ctx_cu_count(const pgamd_context* c);
int ctx_device(const pgamd_context* c);
hipStream_t ctx_stream(const pgamd_context* c);
// Created on first use and owned by the context: a second stream for forked launches, a stream used
// only for graph capture, and two events for fork/join.
hipStream_t ctx_aux_stream(pgamd_context* c);
hipStream_t ctx_capture_stream(pgamd_context* c);
hipEvent_t ctx_event(pgamd_context* c, int i);
// Plan internals for the other modules.
const PatchDesc* plan_device_desc(const pgamd_plan* pl);
int plan_npatch(const pgamd_plan* pl);
int plan_device(const pgamd_plan* pl);
bool plan_has_ghost(const pgamd_plan* pl, int gs);
// Network internals for the allocation-free forward (layer l = 1..6): the device, the BF16 forward-layout
// weights and the FP32 biases.
int cnn_device(const pgamd_cnn* n);
const uint16_t* cnn_forward_weights(const pgamd_cnn* n, int l);
const float* cnn_bias(const pgamd_cnn* n, int l);
// Frees whatever the optional modules (graphs, mixed batches, CNN) attached to a context.
void release_extras(pgamd_context* c);

hipError_t launch_sor_uniform_dispatch(const UniformBatch& b, const SolveParams& p, int ppw, int pad, hipStream_t s);
hipError_t launch_mg_uniform(const UniformBatch& b, const SolveParams& p, hipStream_t s);

}  // namespace pgamd
