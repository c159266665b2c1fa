// Host <-> kernel interface inside libeuler2d_hip (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>

#include "../../include/euler2d_hip.h"

namespace e2d {

// Host-side trace range (roctx; rocprofv3 --marker-trace records it), named
// after the reference's NVTX ranges (serial_ref/m_run.f90:195,264,336,387,502,571:
// godunov_compute, compute_dt, convertToPrimitives, computeAndStoreFluxes,
// updateHydro, make_boundaries).  It spans the enqueue of the kernels, which
// the trace attributes to it.
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

constexpr int kNumStages = 9;  // E2D_STAGE_Q .. E2D_STAGE_QYP

// Device running-invDt slots.  Slot s holds kInvdtParts partial maxima, one
// per 128-byte line, and its value is their maximum: the atomic maxima that
// every wavefront of a step folds in spread over independent lines instead of
// serialising on one address (at 2048^2 a step folds 6,000 of them).  Slots 0
// and 1 ping-pong between steps; slot 2 serves compute_dt.
constexpr int kInvdtParts = 32;
constexpr int kInvdtStride = 16;                          // doubles between parts
constexpr int kInvdtSlot = kInvdtParts * kInvdtStride;    // doubles per slot
constexpr int kInvdtSlots = 3;
constexpr int kScalDoubles = kInvdtSlots * kInvdtSlot;  // doubles in the allocation

// Device memory of one run.  State element (i, j, v), 0-based, of buffer b
// sits at u[b][v * plane + j * pitch + i].
struct DevBuffers {
  void* u[2] = {nullptr, nullptr};
  int64_t pitch = 0;  // elements per row (>= isize, padded)
  int64_t plane = 0;  // elements per variable plane (pitch * jsize)
  void* stage[kNumStages] = {};
  double* scal = nullptr;  // running-invDt slots (kInvdtSlots x kInvdtSlot doubles)
  double* partial = nullptr;  // per-block partial sums (kSumBlocks * 4)
};

constexpr int kSumBlocks = 240;

// How a step obtains dt.
struct StepCfg {
  int buf_in = 0, buf_out = 1;
  bool dt_from_device = false;  // dt = cfl / (invDt slot slot_cur); else dt_host
  double dt_host = 0.0;
  int slot_cur = 0;
  bool track_dt = false;  // fold the new state's CFL rate into invDt slot slot_next (fused
                          // step) or slot_cur (materialised steps, after their flux kernels)
  int slot_next = 1;
};

size_t state_elem_bytes(int precision);   // storage type of the state
size_t stage_elem_bytes(int precision, int stage);

hipError_t launch_fused_step(const e2d_params& p, const DevBuffers& b, const StepCfg& c, hipStream_t s);
// prim .. update of the materialised version (boundaries launched separately)
hipError_t launch_materialised_step(const e2d_params& p, const DevBuffers& b, const StepCfg& c,
                                    hipStream_t s);
hipError_t launch_boundaries(const e2d_params& p, const DevBuffers& b, int buf, hipStream_t s);
// atomic max of the CFL rate over the compute_dt window of buffer buf into invDt slot `slot`
hipError_t launch_max_invdt(const e2d_params& p, const DevBuffers& b, int buf, int slot, hipStream_t s);
hipError_t launch_fill(double* dst, int n, double v, hipStream_t s);
// dst[0..n) = *src (src a device address)
hipError_t launch_fill_from(double* dst, int n, const double* src, hipStream_t s);
// *dst = value of the running-invDt slot starting at `slot` (the maximum of its parts)
hipError_t launch_slot_value(const double* slot, double* dst, hipStream_t s);
// *flag (device) = 1 when the two device arrays of `bytes` bytes differ in any bit, else 0
hipError_t launch_words_differ(const void* a, const void* b, size_t bytes, double* flag, hipStream_t s);
hipError_t launch_init_problem(const e2d_params& p, const DevBuffers& b, int problem,
                               const double* bumps /* 16 x 5 host values */, hipStream_t s);
hipError_t launch_zero_state(const e2d_params& p, const DevBuffers& b, int buf, hipStream_t s);
hipError_t launch_sums(const e2d_params& p, const DevBuffers& b, int buf, double* d_out4, hipStream_t s);
// conversion between the float64 host layout (isize*jsize*4, Fortran order)
// and the pitched device state, through a device staging array of doubles
hipError_t launch_pack(const e2d_params& p, const DevBuffers& b, int buf, const double* d_src, hipStream_t s);
// host-layout FP64 staging -> FP64 field in the device layout
hipError_t launch_pack_field(const e2d_params& p, const DevBuffers& b, const double* d_src, double* d_field,
                             hipStream_t s);
hipError_t launch_unpack(const e2d_params& p, const DevBuffers& b, const void* d_state, int elem_kind,
                         double* d_dst, hipStream_t s);
hipError_t launch_pointwise(const e2d_params& p, int func, int n, const double* d_x, double* d_y,
                            hipStream_t s);
// tangent-linear step at the state of buffer buf (FP64): v_out = J v_in, dt frozen;
// dual_in/dual_out are two scratch state arrays of 16-byte dual numbers
hipError_t launch_tangent_step(const e2d_params& p, const DevBuffers& b, int buf, double dt, const double* v_in,
                               double* v_out, void* dual_in, void* dual_out, hipStream_t s);
// adjoint step at the state of buffer buf (FP64): lam_out = J^T lam_in, dt frozen
// per-cell scratch of the adjoint step, in the step's arithmetic type
// (adjoint_scratch_elem_bytes bytes per value, plane * 4 values per array)
struct AdjointScratch {
  void *q = nullptr, *xminb = nullptr, *xmaxb = nullptr, *yminb = nullptr, *ymaxb = nullptr;
  void *qb = nullptr, *dqxb = nullptr, *dqyb = nullptr, *ub = nullptr;
  void* seam = nullptr;  // 16 planes: partial sums where adjoint row segments meet
};
size_t adjoint_scratch_elem_bytes(int precision);
hipError_t launch_adjoint_step(const e2d_params& p, const DevBuffers& b, int buf, double dt,
                               const double* lam_in, double* lam_out, const AdjointScratch& a, hipStream_t s);

}  // namespace e2d
