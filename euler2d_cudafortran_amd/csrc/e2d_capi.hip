// C-ABI of libeuler2d_hip: contexts, validation, status codes, and the
// host-side flow of the reference's HydroRun module (cuf/m_HydroRun_gpu.cuf,
// serial_ref/m_run.f90) over the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/euler2d_hip.h"
#include "e2d_adjoint.hpp"
#include "e2d_internal.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(E2D_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// SplitMix64 (Steele, Lea, Flood 2014): seeds the Gaussian bumps.
uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return static_cast<double>(splitmix64(s) >> 11) * 0x1.0p-53; }

int validate(const e2d_params* p) {
  if (!p) return fail(E2D_ERR_ARG, "params is NULL");
  if (p->nx < 4 || p->ny < 4) return fail(E2D_ERR_ARG, "nx and ny must be >= 4 (got %d, %d)", p->nx, p->ny);
  if (static_cast<int64_t>(p->nx + 4) * (p->ny + 4) > (int64_t(1) << 31) / 4)
    return fail(E2D_ERR_ARG, "grid %d x %d too large", p->nx, p->ny);
  if (p->implementation_version != E2D_IMPL_FUSED && p->implementation_version != E2D_IMPL_MATERIALISED &&
      p->implementation_version != E2D_IMPL_FLUX_STORE)
    return fail(E2D_ERR_ARG,
                "The implementation version parameter should be 0, 1 or 2 (got %d); check the namelist, section OTHER",
                p->implementation_version);
  if (p->precision < E2D_PREC_FP64 || p->precision > E2D_PREC_MIXED)
    return fail(E2D_ERR_ARG, "precision must be 0 (fp64), 1 (fp32), 2 (fp32_native) or 3 (mixed), got %d",
                p->precision);
  if (p->riemann_solver != E2D_RIEMANN_APPROX && p->riemann_solver != E2D_RIEMANN_HLLC)
    return fail(E2D_ERR_ARG, "riemann_solver must be 0 (approx) or 2 (hllc), got %d", p->riemann_solver);
  if (p->niter_riemann < 0) return fail(E2D_ERR_ARG, "niter_riemann must be >= 0");
  if (!(p->gamma0 > 1.0)) return fail(E2D_ERR_ARG, "gamma0 must be > 1");
  if (!(p->cfl > 0.0)) return fail(E2D_ERR_ARG, "cfl must be > 0");
  if (!(p->smallr > 0.0) || !(p->smallc > 0.0)) return fail(E2D_ERR_ARG, "smallr and smallc must be > 0");
  // The Riemann solver's radicands are bounded below by gamma0 smallr^2 smallp
  // (1 - gamma6) (densities and pressures are floored at smallr and
  // rho smallp, and p* >= 0); the step takes their square roots without a
  // zero case, so that bound must be a positive normal double.
  const double smallp = p->smallc * p->smallc / p->gamma0;
  const double gamma6 = (p->gamma0 + 1.0) / (2.0 * p->gamma0);
  const double radicand_floor = p->gamma0 * p->smallr * (p->smallr * smallp) * (1.0 - gamma6);
  if (!(gamma6 < 1.0) || !(radicand_floor >= 1e-290))
    return fail(E2D_ERR_ARG,
                "gamma0 %g, smallr %g and smallc %g put the pressure and sound-speed floors below the "
                "double range (gamma0 * smallr^2 * smallc^2 / gamma0 * (1 - (gamma0 + 1) / (2 gamma0)) = %g, "
                "must be >= 1e-290)",
                p->gamma0, p->smallr, p->smallc, radicand_floor);
  // The Newton stop test compares |delp| - conv_tol d with half an ulp of
  // conv_tol times d, d >= smallr smallp; that product must stay a normal
  // double (half an ulp of the FP32 literal 1e-6 is 2^-73).
  if (!(p->smallr * smallp >= 0x1p-949))
    return fail(E2D_ERR_ARG, "smallr %g and smallc %g put the Newton stop test's pressure floor "
                "smallr * smallc^2 / gamma0 = %g below 2^-949", p->smallr, p->smallc, p->smallr * smallp);
  return E2D_OK;
}

}  // namespace

struct e2d_ctx {
  e2d_params p{};
  e2d_ctx() = default;
  e2d::DevBuffers b;
  hipStream_t stream = nullptr;
  int device = 0;
  int isize = 0, jsize = 0;
  double* staging = nullptr;  // float64 host-layout staging array on the device
  double* d_sums = nullptr;
  // running invDt of compute_dt: authoritative on the host when dev_slot < 0,
  // else in device slot dev_slot of b.scal (the maximum of its parts)
  double invdt = 0.0;
  int dev_slot = -1;
  bool stages_valid = false;
  // GPU timers: pending event pairs per category, accumulated on read
  struct Pending {
    int cat;
    hipEvent_t start, stop;
  };
  bool timers_on = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  double timer_ms[E2D_NUM_TIMERS] = {0, 0, 0};
  int64_t timer_count[E2D_NUM_TIMERS] = {0, 0, 0};
  // tangent / adjoint scratch (allocated on first use)
  void* dual[2] = {nullptr, nullptr};
  e2d::AdjointScratch adj;
  std::vector<void*> snapshots;
  double* lam_tmp = nullptr;  // the rollout's second adjoint field
  double* d_flag = nullptr;   // one device word for comparisons
  // states of the last e2d_rollout_forward, kept in the snapshots for the
  // next e2d_adjoint_rollout with the same arguments
  struct Recording {
    bool valid = false;
    int n0 = 0, nsteps = 0, nsnap = 0;
    std::vector<double> dts;
  } rec;
};

namespace {

// Every call on a context runs on the device that was current at e2d_create
// (its stream and buffers live there): make it current on this thread.
int check_ctx(e2d_ctx* c) {
  if (!c) return fail(E2D_ERR_ARG, "context is NULL");
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != c->device)
    if (hipSetDevice(c->device) != hipSuccess) return fail(E2D_ERR_HIP, "hipSetDevice(%d) failed", c->device);
  return E2D_OK;
}
int check_buf(int buf) {
  if (buf != 0 && buf != 1) return fail(E2D_ERR_ARG, "buffer index must be 0 (u) or 1 (u2), got %d", buf);
  return E2D_OK;
}

size_t host_elems(const e2d_ctx* c) { return static_cast<size_t>(c->isize) * c->jsize * 4; }

// set every part of running-invDt slot `slot` to v
int fill_slot(e2d_ctx* c, int slot, double v) {
  HIP_TRY(e2d::launch_fill(c->b.scal + slot * e2d::kInvdtSlot, e2d::kInvdtSlot, v, c->stream));
  return E2D_OK;
}

// value of running-invDt slot `slot` (the maximum of its parts); synchronous
int read_slot(e2d_ctx* c, int slot, double* v) {
  double parts[e2d::kInvdtSlot];
  HIP_TRY(hipMemcpyAsync(parts, c->b.scal + slot * e2d::kInvdtSlot, sizeof(parts), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double m = parts[0];
  for (int p = 1; p < e2d::kInvdtParts; ++p) m = m > parts[p * e2d::kInvdtStride] ? m : parts[p * e2d::kInvdtStride];
  *v = m;
  return E2D_OK;
}

// bring the running invDt to the host
int pull_invdt(e2d_ctx* c) {
  if (c->dev_slot >= 0) {
    if (int rc = read_slot(c, c->dev_slot, &c->invdt)) return rc;
    c->dev_slot = -1;
  }
  return E2D_OK;
}

// dt = cfl / invDt in the run's precision (serial_ref/m_run.f90:308)
double dt_of(const e2d_params& p, double invdt) {
  if (p.precision == E2D_PREC_FP32) return static_cast<double>(static_cast<float>(p.cfl) / static_cast<float>(invdt));
  return p.cfl / invdt;
}

hipEvent_t take_event(e2d_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Launch `fn` on the context stream, bracketed by events when timers are on.
template <class F>
hipError_t timed(e2d_ctx* c, int cat, F&& fn) {
  if (!c->timers_on) return fn();
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) {
    if (a) c->event_pool.push_back(a);
    if (b) c->event_pool.push_back(b);
    return hipErrorOutOfMemory;
  }
  hipError_t e = hipEventRecord(a, c->stream);
  if (e == hipSuccess) e = fn();
  if (e == hipSuccess) e = hipEventRecord(b, c->stream);
  c->pending.push_back({cat, a, b});
  return e;
}

int collect_timers(e2d_ctx* c) {
  if (c->pending.empty()) return E2D_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto& pe : c->pending) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, pe.start, pe.stop));
    c->timer_ms[pe.cat] += ms;
    c->timer_count[pe.cat] += 1;
    c->event_pool.push_back(pe.start);
    c->event_pool.push_back(pe.stop);
  }
  c->pending.clear();
  return E2D_OK;
}

size_t state_bytes(const e2d_ctx* c) {
  return static_cast<size_t>(c->b.plane) * 4 * e2d::state_elem_bytes(c->p.precision);
}
size_t field_bytes(const e2d_ctx* c) { return static_cast<size_t>(c->b.plane) * 4 * sizeof(double); }

void free_adjoint(e2d_ctx* c) {
  for (auto& d : c->dual)
    if (d) (void)hipFree(d), d = nullptr;
  void** arrs[] = {&c->adj.q, &c->adj.xminb, &c->adj.xmaxb, &c->adj.yminb, &c->adj.ymaxb,
                   &c->adj.qb, &c->adj.dqxb, &c->adj.dqyb, &c->adj.ub, &c->adj.seam};
  for (void** a : arrs)
    if (*a) (void)hipFree(*a), *a = nullptr;
  for (void* sn : c->snapshots) (void)hipFree(sn);
  c->snapshots.clear();
  c->rec.valid = false;
  if (c->lam_tmp) (void)hipFree(c->lam_tmp), c->lam_tmp = nullptr;
  if (c->d_flag) (void)hipFree(c->d_flag), c->d_flag = nullptr;
}

void free_all(e2d_ctx* c) {
  free_adjoint(c);
  for (auto& pe : c->pending) {
    (void)hipEventDestroy(pe.start);
    (void)hipEventDestroy(pe.stop);
  }
  c->pending.clear();
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  c->event_pool.clear();
  for (auto& u : c->b.u)
    if (u) (void)hipFree(u), u = nullptr;
  for (auto& s : c->b.stage)
    if (s) (void)hipFree(s), s = nullptr;
  if (c->b.scal) (void)hipFree(c->b.scal), c->b.scal = nullptr;
  if (c->b.partial) (void)hipFree(c->b.partial), c->b.partial = nullptr;
  if (c->staging) (void)hipFree(c->staging), c->staging = nullptr;
  if (c->d_sums) (void)hipFree(c->d_sums), c->d_sums = nullptr;
  if (c->stream) (void)hipStreamDestroy(c->stream), c->stream = nullptr;
}

// Binomial checkpointing (Griewank and Walther, Algorithm 799: revolve, ACM
// TOMS 26 (2000)).
// beta(c, r) = C(c+r, c): the most steps that c kept states (the first one
// included) reverse when no step is advanced more than r times.
double beta(int c, int r) {
  if (r < 0) return 0.0;
  double v = 1.0;
  for (int k = 1; k <= c; ++k) v = v * (r + k) / k;
  return v;
}

// Where to store the next snapshot when reversing l steps whose first state is
// kept, with s snapshot slots free: m steps on.  With c = s + 1 states and r
// the least repetition count with beta(c, r) >= l, every m with
//   max(l - beta(c-1, r), beta(c, r-2) + 1) <= m <= min(beta(c, r-1), l - beta(c-1, r-1) - 1)
// leaves the right part l - m to c - 1 states at r repetitions and the left
// part m to c states at r - 1, so the whole reversal advances the least
// number of steps, r l - beta(c+1, r-1); this takes the largest such m.
int split_point(int l, int s) {
  const int c = s + 1;
  int r = 1;
  while (beta(c, r) < l) ++r;
  const double lo = std::max(l - beta(c - 1, r), beta(c, r - 2) + 1.0);
  const double hi = std::min(beta(c, r - 1), l - beta(c - 1, r - 1) - 1.0);
  double m = hi >= lo ? hi : lo;
  if (m > l - 1) m = l - 1;
  if (m < 1) m = 1;
  return static_cast<int>(m);
}

// (step, slot) of the snapshots that Rollout::reverse stores on its first
// descent from (0, nsteps) with nsnap slots: the same split and slot order
std::vector<std::pair<int, int>> first_descent(int nsteps, int nsnap) {
  std::vector<std::pair<int, int>> out;
  std::vector<int> free_slots;
  for (int k = nsnap - 1; k >= 1; --k) free_slots.push_back(k);
  int a = 0;
  const int b = nsteps;
  for (;;) {
    const int l = b - a;
    const int s = static_cast<int>(free_slots.size());
    if (l == 1 || s == 0) break;
    const int m = split_point(l, s);
    const int mine = free_slots.back();
    free_slots.pop_back();
    out.emplace_back(a + m, mine);
    a += m;
  }
  return out;
}

}  // namespace

extern "C" {

int e2d_version(void) { return E2D_VERSION; }

int e2d_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

const char* e2d_last_error(void) { return g_err; }

int e2d_create(const e2d_params* p, e2d_ctx** out) {
  g_err[0] = '\0';
  if (!out) return fail(E2D_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = validate(p);
  if (rc) return rc;
  if (!e2d_device_available()) return fail(E2D_ERR_HIP, "no HIP device available");
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  e2d_ctx* c = new (std::nothrow) e2d_ctx();
  if (!c) return fail(E2D_ERR_NOMEM, "host allocation failed");
  c->p = *p;
  c->isize = p->nx + 4;
  c->jsize = p->ny + 4;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(E2D_ERR_HIP, "hipStreamCreate failed");
  }
  // rows padded to a multiple of 64 elements (256 B for FP32, 512 B for FP64)
  c->b.pitch = (static_cast<int64_t>(c->isize) + 63) / 64 * 64;
  c->b.plane = c->b.pitch * c->jsize;
  const size_t n = static_cast<size_t>(c->b.plane) * 4;
  const size_t es = e2d::state_elem_bytes(p->precision);
  bool ok = hipMalloc(&c->b.u[0], n * es) == hipSuccess && hipMalloc(&c->b.u[1], n * es) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&c->b.scal), e2d::kScalDoubles * sizeof(double)) ==
                hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&c->b.partial), e2d::kSumBlocks * 4 * sizeof(double)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&c->d_sums), 4 * sizeof(double)) == hipSuccess;
  if (ok && p->implementation_version != E2D_IMPL_FUSED) {
    // version 1 keeps every stage; version 2 only q, fx and fy (cuf/m_HydroRun_gpu.cuf:93-105)
    const int nstages = p->implementation_version == E2D_IMPL_MATERIALISED ? e2d::kNumStages : E2D_STAGE_FY + 1;
    for (int s = 0; s < nstages && ok; ++s) {
      const size_t bytes = n * e2d::stage_elem_bytes(p->precision, s);
      ok = hipMalloc(&c->b.stage[s], bytes) == hipSuccess &&
           hipMemsetAsync(c->b.stage[s], 0, bytes, c->stream) == hipSuccess;
    }
  }
  if (!ok) {
    free_all(c);
    delete c;
    return fail(E2D_ERR_NOMEM, "device allocation failed (%zu bytes per state buffer)", n * es);
  }
  // both buffers start at zero (u2 as the reference leaves it), padding included
  if (hipMemsetAsync(c->b.u[0], 0, n * es, c->stream) != hipSuccess ||
      hipMemsetAsync(c->b.u[1], 0, n * es, c->stream) != hipSuccess ||
      e2d::launch_fill(c->b.scal, e2d::kScalDoubles, 0.0, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    free_all(c);
    delete c;
    return fail(E2D_ERR_HIP, "device initialisation failed");
  }
  *out = c;
  return E2D_OK;
}

int e2d_destroy(e2d_ctx* c) {
  if (!c) return E2D_OK;
  (void)check_ctx(c);
  (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return E2D_OK;
}

int e2d_init_problem(e2d_ctx* c, int problem, uint64_t seed) {
  if (int rc = check_ctx(c)) return rc;
  if (problem < E2D_PROBLEM_IMPLODE || problem > E2D_PROBLEM_BUMPS)
    return fail(E2D_ERR_ARG, "unknown problem %d", problem);
  double bumps[16 * 5];
  uint64_t s = seed;
  for (int k = 0; k < 16; ++k) {
    bumps[5 * k + 0] = uniform01(s);               // centre x
    bumps[5 * k + 1] = uniform01(s);               // centre y
    bumps[5 * k + 2] = 0.1 + 1.9 * uniform01(s);   // amplitude on rho, U[0.1, 2]
    bumps[5 * k + 3] = 0.1 + 1.9 * uniform01(s);   // amplitude on p,   U[0.1, 2]
    bumps[5 * k + 4] = 0.02 + 0.08 * uniform01(s); // sigma,            U[0.02, 0.1]
  }
  HIP_TRY(e2d::launch_init_problem(c->p, c->b, problem, bumps, c->stream));
  c->invdt = 0.0;
  c->dev_slot = -1;
  c->stages_valid = false;
  return E2D_OK;
}

int e2d_upload(e2d_ctx* c, int buf, const double* h) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (!h) return fail(E2D_ERR_ARG, "host array is NULL");
  if (!c->staging)
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->staging), host_elems(c) * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(c->staging, h, host_elems(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(e2d::launch_pack(c->p, c->b, buf, c->staging, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

static int download_array(e2d_ctx* c, const void* d_state, int elem_bytes, double* h) {
  if (!c->staging)
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->staging), host_elems(c) * sizeof(double)));
  HIP_TRY(e2d::launch_unpack(c->p, c->b, d_state, elem_bytes, c->staging, c->stream));
  HIP_TRY(hipMemcpyAsync(h, c->staging, host_elems(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

int e2d_download(e2d_ctx* c, int buf, double* h) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (!h) return fail(E2D_ERR_ARG, "host array is NULL");
  return download_array(c, c->b.u[buf], static_cast<int>(e2d::state_elem_bytes(c->p.precision)), h);
}

int e2d_download_stage(e2d_ctx* c, int stage, double* h) {
  if (int rc = check_ctx(c)) return rc;
  if (!h) return fail(E2D_ERR_ARG, "host array is NULL");
  if (stage < 0 || stage >= e2d::kNumStages) return fail(E2D_ERR_ARG, "unknown stage %d", stage);
  if (c->p.implementation_version == E2D_IMPL_FUSED)
    return fail(E2D_ERR_STATE, "stage arrays exist only with implementationVersion 1 or 2");
  if (!c->b.stage[stage])
    return fail(E2D_ERR_STATE, "implementationVersion 2 keeps only the q, fx and fy stages");
  return download_array(c, c->b.stage[stage], static_cast<int>(e2d::stage_elem_bytes(c->p.precision, stage)), h);
}

int e2d_compute_dt(e2d_ctx* c, int useU, double* dt) {
  if (int rc = check_ctx(c)) return rc;
  const e2d::TraceRange range("compute_dt");
  if (!dt) return fail(E2D_ERR_ARG, "dt is NULL");
  if (int rc = pull_invdt(c)) return rc;
  const int buf = (useU == 0) ? 0 : 1;
  if (int rc = fill_slot(c, 2, c->invdt)) return rc;
  HIP_TRY(timed(c, E2D_TIMER_DT, [&] { return e2d::launch_max_invdt(c->p, c->b, buf, 2, c->stream); }));
  if (int rc = read_slot(c, 2, &c->invdt)) return rc;
  *dt = dt_of(c->p, c->invdt);
  return E2D_OK;
}

int e2d_make_boundaries(e2d_ctx* c, int buf) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  HIP_TRY(timed(c, E2D_TIMER_BOUNDARIES, [&] { return e2d::launch_boundaries(c->p, c->b, buf, c->stream); }));
  return E2D_OK;
}

int e2d_godunov_unsplit(e2d_ctx* c, int nStep, double dt) {
  if (int rc = check_ctx(c)) return rc;
  if (nStep < 0) return fail(E2D_ERR_ARG, "nStep must be >= 0");
  const e2d::TraceRange range("godunov_compute");
  if (int rc = pull_invdt(c)) return rc;
  e2d::StepCfg cfg;
  cfg.buf_in = nStep % 2;
  cfg.buf_out = 1 - cfg.buf_in;
  cfg.dt_from_device = false;
  cfg.dt_host = dt;
  cfg.track_dt = false;
  if (c->p.implementation_version == E2D_IMPL_FUSED) {
    HIP_TRY(timed(c, E2D_TIMER_GODUNOV, [&] { return e2d::launch_fused_step(c->p, c->b, cfg, c->stream); }));
  } else {
    HIP_TRY(timed(c, E2D_TIMER_BOUNDARIES,
                  [&] { return e2d::launch_boundaries(c->p, c->b, cfg.buf_in, c->stream); }));
    HIP_TRY(timed(c, E2D_TIMER_GODUNOV,
                  [&] { return e2d::launch_materialised_step(c->p, c->b, cfg, c->stream); }));
    c->stages_valid = true;
  }
  return E2D_OK;
}

int e2d_run_steps(e2d_ctx* c, int nsteps, int* nStep) {
  if (int rc = check_ctx(c)) return rc;
  if (!nStep) return fail(E2D_ERR_ARG, "nStep is NULL");
  if (nsteps < 0 || *nStep < 0) return fail(E2D_ERR_ARG, "nsteps and *nStep must be >= 0");
  if (nsteps == 0) return E2D_OK;
  int n = *nStep;
  int k = c->dev_slot;
  if (k < 0) {
    // compute_dt of the current input, folded into the running maximum
    k = 0;
    if (int rc = fill_slot(c, k, c->invdt)) return rc;
    HIP_TRY(timed(c, E2D_TIMER_DT, [&] { return e2d::launch_max_invdt(c->p, c->b, n % 2, k, c->stream); }));
  }
  e2d::StepCfg cfg;
  cfg.dt_from_device = true;
  if (c->p.implementation_version == E2D_IMPL_FUSED) {
    // slot k holds the running invDt including the current input; the other
    // slot must not exceed it before the first step writes there
    if (int rc = fill_slot(c, 1 - k, 0.0)) return rc;
    for (int s = 0; s < nsteps; ++s, ++n) {
      cfg.buf_in = n % 2;
      cfg.buf_out = 1 - cfg.buf_in;
      cfg.slot_cur = k;
      cfg.slot_next = 1 - k;
      cfg.track_dt = true;
      const e2d::TraceRange range("godunov_compute");
      HIP_TRY(timed(c, E2D_TIMER_GODUNOV, [&] { return e2d::launch_fused_step(c->p, c->b, cfg, c->stream); }));
      k = 1 - k;
    }
  } else {
    // one slot: each step reads dt from slot k in its flux kernels, and its
    // update kernel then folds the new state's CFL rates into slot k
    for (int s = 0; s < nsteps; ++s, ++n) {
      cfg.buf_in = n % 2;
      cfg.buf_out = 1 - cfg.buf_in;
      cfg.slot_cur = k;
      cfg.track_dt = true;
      const e2d::TraceRange range("godunov_compute");
      HIP_TRY(timed(c, E2D_TIMER_BOUNDARIES,
                    [&] { return e2d::launch_boundaries(c->p, c->b, cfg.buf_in, c->stream); }));
      HIP_TRY(timed(c, E2D_TIMER_GODUNOV,
                    [&] { return e2d::launch_materialised_step(c->p, c->b, cfg, c->stream); }));
    }
    c->stages_valid = true;
  }
  c->dev_slot = k;
  *nStep = n;
  return E2D_OK;
}

int e2d_get_dt(e2d_ctx* c, double* dt) {
  if (int rc = check_ctx(c)) return rc;
  if (!dt) return fail(E2D_ERR_ARG, "dt is NULL");
  if (int rc = pull_invdt(c)) return rc;
  *dt = dt_of(c->p, c->invdt);
  return E2D_OK;
}

int e2d_get_invdt(e2d_ctx* c, double* invdt) {
  if (int rc = check_ctx(c)) return rc;
  if (!invdt) return fail(E2D_ERR_ARG, "invdt is NULL");
  if (int rc = pull_invdt(c)) return rc;
  *invdt = c->invdt;
  return E2D_OK;
}

int e2d_set_invdt(e2d_ctx* c, double invdt) {
  if (int rc = check_ctx(c)) return rc;
  if (!(invdt >= 0.0)) return fail(E2D_ERR_ARG, "invdt must be >= 0");
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->invdt = invdt;
  c->dev_slot = -1;
  return E2D_OK;
}

int e2d_sums(e2d_ctx* c, int buf, double out[4]) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (!out) return fail(E2D_ERR_ARG, "out is NULL");
  HIP_TRY(e2d::launch_sums(c->p, c->b, buf, c->d_sums, c->stream));
  HIP_TRY(hipMemcpyAsync(out, c->d_sums, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

int e2d_synchronize(e2d_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

int e2d_enable_timers(e2d_ctx* c, int on) {
  if (int rc = check_ctx(c)) return rc;
  c->timers_on = (on != 0);
  return E2D_OK;
}

int e2d_get_timers(e2d_ctx* c, double ms[E2D_NUM_TIMERS], int64_t count[E2D_NUM_TIMERS]) {
  if (int rc = check_ctx(c)) return rc;
  if (!ms || !count) return fail(E2D_ERR_ARG, "output array is NULL");
  if (int rc = collect_timers(c)) return rc;
  for (int k = 0; k < E2D_NUM_TIMERS; ++k) {
    ms[k] = c->timer_ms[k];
    count[k] = c->timer_count[k];
    c->timer_ms[k] = 0.0;
    c->timer_count[k] = 0;
  }
  return E2D_OK;
}

int e2d_device_buffer(e2d_ctx* c, int buf, void** dptr, int64_t* pitch) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (!dptr || !pitch) return fail(E2D_ERR_ARG, "output pointer is NULL");
  *dptr = c->b.u[buf];
  *pitch = c->b.pitch;
  return E2D_OK;
}

namespace {

// the adjoint exists for every precision, the tangent (dual numbers) for fp64
int check_linearisable(const e2d_ctx* c, bool tangent) {
  if (tangent && c->p.precision != E2D_PREC_FP64)
    return fail(E2D_ERR_STATE, "the tangent step needs precision 0 (fp64)");
  if (c->p.riemann_solver != E2D_RIEMANN_APPROX)
    return fail(E2D_ERR_STATE, "tangent and adjoint need the two-shock solver (riemann_solver 0)");
  return E2D_OK;
}

int ensure_adjoint_scratch(e2d_ctx* c) {
  if (c->adj.q) return E2D_OK;
  const size_t bytes = static_cast<size_t>(c->b.plane) * 4 * e2d::adjoint_scratch_elem_bytes(c->p.precision);
  void** arrs[] = {&c->adj.q, &c->adj.xminb, &c->adj.xmaxb, &c->adj.yminb, &c->adj.ymaxb,
                   &c->adj.qb, &c->adj.dqxb, &c->adj.dqyb, &c->adj.ub, &c->adj.seam};
  for (void** a : arrs) {
    // the seam buffer holds 16 planes, the others 4
    const size_t n = a == &c->adj.seam ? 4 * bytes : bytes;
    if (hipMalloc(a, n) != hipSuccess) {
      free_adjoint(c);
      return fail(E2D_ERR_NOMEM, "adjoint scratch allocation failed");
    }
    // slots never written by a pass must read as zero
    HIP_TRY(hipMemsetAsync(*a, 0, n, c->stream));
  }
  return E2D_OK;
}

int adjoint_step(e2d_ctx* c, int buf, double dt, const double* lam_in, double* lam_out) {
  HIP_TRY(timed(c, E2D_TIMER_GODUNOV,
                [&] { return e2d::launch_adjoint_step(c->p, c->b, buf, dt, lam_in, lam_out, c->adj, c->stream); }));
  return E2D_OK;
}

// steps n -> n+1 with a given dt (the forward of godunov_unsplit)
int forward_step(e2d_ctx* c, int n, double dt) {
  e2d::StepCfg cfg;
  cfg.buf_in = n % 2;
  cfg.buf_out = 1 - cfg.buf_in;
  cfg.dt_host = dt;
  HIP_TRY(e2d::launch_fused_step(c->p, c->b, cfg, c->stream));
  return E2D_OK;
}

struct Rollout {
  e2d_ctx* c;
  int n0;
  const double* dts;
  // lam ping-pongs between two fields (lam[cur] holds the current one), so
  // that no adjoint step runs in place; both start as lam_n, and ghost cells,
  // which no step writes, keep lam_n's values in either
  double* lam[2];
  int cur = 0;
  std::vector<int> free_slots;
  int64_t nfwd = 0;
  // step whose state each buffer holds (-1: none of this rollout).  A step
  // leaves its input in place (only ghost cells, which every consumer
  // rebuilds from the interior, are rewritten), so after advancing to k the
  // other buffer still holds k-1 and a restore of a held state is skipped.
  int held[2] = {-1, -1};

  // snapshots already stored by e2d_rollout_forward: (step, slot) of the
  // first descent, consumed in order
  std::vector<std::pair<int, int>> pre;

  int buf_of(int k) const { return (n0 + k) % 2; }
  int save(int slot, int k) {
    HIP_TRY(hipMemcpyAsync(c->snapshots[slot], c->b.u[buf_of(k)], state_bytes(c), hipMemcpyDeviceToDevice,
                           c->stream));
    return E2D_OK;
  }
  int restore(int slot, int k) {
    if (held[buf_of(k)] == k) return E2D_OK;
    HIP_TRY(hipMemcpyAsync(c->b.u[buf_of(k)], c->snapshots[slot], state_bytes(c), hipMemcpyDeviceToDevice,
                           c->stream));
    held[buf_of(k)] = k;
    return E2D_OK;
  }
  int advance(int from, int to) {
    for (int k = from; k < to; ++k) {
      if (int rc = forward_step(c, n0 + k, dts[k])) return rc;
      held[buf_of(k + 1)] = k + 1;
      ++nfwd;
    }
    return E2D_OK;
  }
  // lam <- J_a^T ... J_{b-1}^T lam, with the state of step a in snapshot `slot`
  int reverse(int a, int b, int slot) {
    const int l = b - a;
    const int s = static_cast<int>(free_slots.size());
    if (l == 1 || s == 0) {
      for (int k = b - 1; k >= a; --k) {
        if (int rc = restore(slot, a)) return rc;
        if (int rc = advance(a, k)) return rc;
        if (int rc = adjoint_step(c, buf_of(k), dts[k], lam[cur], lam[1 - cur])) return rc;
        cur = 1 - cur;
      }
      return E2D_OK;
    }
    const int m = split_point(l, s);
    const int mine = free_slots.back();
    free_slots.pop_back();
    if (!pre.empty() && pre.front() == std::make_pair(a + m, mine)) {
      pre.erase(pre.begin());  // recorded by the forward
    } else {
      if (int rc = restore(slot, a)) return rc;
      if (int rc = advance(a, a + m)) return rc;
      if (int rc = save(mine, a + m)) return rc;
    }
    if (int rc = reverse(a + m, b, mine)) return rc;
    free_slots.push_back(mine);
    return reverse(a, a + m, slot);
  }
};

}  // namespace

int e2d_tangent_step(e2d_ctx* c, int buf, double dt, const double* v_in, double* v_out) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (int rc = check_linearisable(c, true)) return rc;
  if (!v_in || !v_out) return fail(E2D_ERR_ARG, "field pointer is NULL");
  if (!(dt > 0.0)) return fail(E2D_ERR_ARG, "dt must be > 0");
  for (auto& d : c->dual)
    if (!d && hipMalloc(&d, static_cast<size_t>(c->b.plane) * 4 * sizeof(e2d::Dual)) != hipSuccess)
      return fail(E2D_ERR_NOMEM, "tangent scratch allocation failed");
  HIP_TRY(timed(c, E2D_TIMER_GODUNOV, [&] {
    return e2d::launch_tangent_step(c->p, c->b, buf, dt, v_in, v_out, c->dual[0], c->dual[1], c->stream);
  }));
  return E2D_OK;
}

int e2d_adjoint_step(e2d_ctx* c, int buf, double dt, const double* lam_in, double* lam_out) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (int rc = check_linearisable(c, false)) return rc;
  if (!lam_in || !lam_out) return fail(E2D_ERR_ARG, "field pointer is NULL");
  if (!(dt > 0.0)) return fail(E2D_ERR_ARG, "dt must be > 0");
  if (int rc = ensure_adjoint_scratch(c)) return rc;
  return adjoint_step(c, buf, dt, lam_in, lam_out);
}

int e2d_adjoint_rollout(e2d_ctx* c, int nStep0, int nsteps, const double* dts, int nsnap, const double* lam_n,
                        double* lam_0, int64_t* n_forward) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_linearisable(c, false)) return rc;
  if (nStep0 < 0 || nsteps < 1) return fail(E2D_ERR_ARG, "need nStep0 >= 0 and nsteps >= 1");
  if (nsnap < 1) return fail(E2D_ERR_ARG, "need at least one snapshot");
  if (!dts || !lam_n || !lam_0) return fail(E2D_ERR_ARG, "NULL argument");
  for (int k = 0; k < nsteps; ++k)
    if (!(dts[k] > 0.0)) return fail(E2D_ERR_ARG, "dts[%d] must be > 0", k);
  if (int rc = ensure_adjoint_scratch(c)) return rc;
  while (static_cast<int>(c->snapshots.size()) < nsnap) {
    void* sn = nullptr;
    if (hipMalloc(&sn, state_bytes(c)) != hipSuccess)
      return fail(E2D_ERR_NOMEM, "snapshot %zu allocation failed", c->snapshots.size());
    c->snapshots.push_back(sn);
  }
  if (!c->lam_tmp && hipMalloc(reinterpret_cast<void**>(&c->lam_tmp), field_bytes(c)) != hipSuccess)
    return fail(E2D_ERR_NOMEM, "adjoint field allocation failed");
  if (lam_0 != lam_n)
    HIP_TRY(hipMemcpyAsync(lam_0, lam_n, field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->lam_tmp, lam_n, field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  // a recording of this rollout by e2d_rollout_forward, from the same starting state?
  bool recorded = c->rec.valid && c->rec.n0 == nStep0 && c->rec.nsteps == nsteps && c->rec.nsnap == nsnap &&
                  std::equal(c->rec.dts.begin(), c->rec.dts.end(), dts);
  c->rec.valid = false;
  if (recorded) {
    if (!c->d_flag && hipMalloc(reinterpret_cast<void**>(&c->d_flag), sizeof(double)) != hipSuccess)
      return fail(E2D_ERR_NOMEM, "flag allocation failed");
    HIP_TRY(e2d::launch_words_differ(c->snapshots[0], c->b.u[nStep0 % 2], state_bytes(c), c->d_flag, c->stream));
    double differ = 1.0;
    HIP_TRY(hipMemcpyAsync(&differ, c->d_flag, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    recorded = differ == 0.0;
  }
  Rollout r{c, nStep0, dts, {lam_0, c->lam_tmp}};
  for (int k = nsnap - 1; k >= 1; --k) r.free_slots.push_back(k);
  if (recorded) {
    // snapshot 0 and the first descent's snapshots hold their states; the
    // buffers are not trusted (the caller may have stepped since)
    r.pre = first_descent(nsteps, nsnap);
  } else {
    r.held[r.buf_of(0)] = 0;
    if (int rc = r.save(0, 0)) return rc;
  }
  if (int rc = r.reverse(0, nsteps, 0)) return rc;
  if (r.cur == 1) HIP_TRY(hipMemcpyAsync(lam_0, c->lam_tmp, field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  if (int rc = r.restore(0, 0)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (n_forward) *n_forward = r.nfwd;
  return E2D_OK;
}

int e2d_rollout_forward(e2d_ctx* c, int nStep0, int nsteps, const double* dts, int nsnap) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_linearisable(c, false)) return rc;
  if (nStep0 < 0 || nsteps < 1) return fail(E2D_ERR_ARG, "need nStep0 >= 0 and nsteps >= 1");
  if (nsnap < 1) return fail(E2D_ERR_ARG, "need at least one snapshot");
  if (!dts) return fail(E2D_ERR_ARG, "NULL argument");
  for (int k = 0; k < nsteps; ++k)
    if (!(dts[k] > 0.0)) return fail(E2D_ERR_ARG, "dts[%d] must be > 0", k);
  c->rec.valid = false;
  while (static_cast<int>(c->snapshots.size()) < nsnap) {
    void* sn = nullptr;
    if (hipMalloc(&sn, state_bytes(c)) != hipSuccess)
      return fail(E2D_ERR_NOMEM, "snapshot %zu allocation failed", c->snapshots.size());
    c->snapshots.push_back(sn);
  }
  const std::vector<std::pair<int, int>> plan = first_descent(nsteps, nsnap);
  Rollout r{c, nStep0, dts, {nullptr, nullptr}};
  if (int rc = r.save(0, 0)) return rc;
  size_t next = 0;
  for (int k = 0; k < nsteps; ++k) {
    if (int rc = forward_step(c, nStep0 + k, dts[k])) return rc;
    if (next < plan.size() && plan[next].first == k + 1) {
      if (int rc = r.save(plan[next].second, k + 1)) return rc;
      ++next;
    }
  }
  c->rec.valid = true;
  c->rec.n0 = nStep0;
  c->rec.nsteps = nsteps;
  c->rec.nsnap = nsnap;
  c->rec.dts.assign(dts, dts + nsteps);
  return E2D_OK;
}

int e2d_field_alloc(e2d_ctx* c, double** d_field) {
  if (int rc = check_ctx(c)) return rc;
  if (!d_field) return fail(E2D_ERR_ARG, "NULL argument");
  *d_field = nullptr;
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(d_field), field_bytes(c)));
  HIP_TRY(hipMemsetAsync(*d_field, 0, field_bytes(c), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

int e2d_field_free(e2d_ctx* c, double* d_field) {
  if (int rc = check_ctx(c)) return rc;
  if (d_field) HIP_TRY(hipFree(d_field));
  return E2D_OK;
}

int e2d_field_upload(e2d_ctx* c, double* d_field, const double* h) {
  if (int rc = check_ctx(c)) return rc;
  if (!d_field || !h) return fail(E2D_ERR_ARG, "NULL argument");
  if (!c->staging)
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->staging), host_elems(c) * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(c->staging, h, host_elems(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(e2d::launch_pack_field(c->p, c->b, c->staging, d_field, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

int e2d_field_download(e2d_ctx* c, const double* d_field, double* h) {
  if (int rc = check_ctx(c)) return rc;
  if (!d_field || !h) return fail(E2D_ERR_ARG, "NULL argument");
  return download_array(c, d_field, 8, h);
}

int e2d_copy_to_buffer(e2d_ctx* c, int buf, const void* d_src) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (!d_src) return fail(E2D_ERR_ARG, "NULL argument");
  HIP_TRY(hipMemcpyAsync(c->b.u[buf], d_src, state_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

int e2d_copy_from_buffer(e2d_ctx* c, int buf, void* d_dst) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_buf(buf)) return rc;
  if (!d_dst) return fail(E2D_ERR_ARG, "NULL argument");
  HIP_TRY(hipMemcpyAsync(d_dst, c->b.u[buf], state_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return E2D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the device-resident step loop on caller memory, allocating nothing

namespace {

// Scratch of e2d_run_steps_noalloc: the two running-invDt slots that the
// step loop ping-pongs between (a context keeps them in its own allocation),
// 2 x 32 partial maxima on separate 128-byte lines, 8 KB.  Each device holds
// its own copy from the first call on it until the process ends; every call
// on a device uses that copy, so calls that overlap on one device race.
constexpr int kNoallocSlots = 2;
static __device__ __align__(256) double scratch[kNoallocSlots * e2d::kInvdtSlot];

// Host code never passes `scratch` itself (its name is a host-side stand-in,
// not a device address): it uses the address hipGetSymbolAddress returns,
// fetched once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(double** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(E2D_ERR_HIP, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(E2D_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return E2D_OK;
}

}  // namespace

extern "C" {

int e2d_run_steps_noalloc(const e2d_params* p, void* d_u, void* d_u2, int64_t pitch, int nsteps, int nStep,
                          double* d_invdt, void* stream) {
  g_err[0] = '\0';
  if (int rc = validate(p)) return rc;
  if (p->implementation_version != E2D_IMPL_FUSED)
    return fail(E2D_ERR_ARG,
                "e2d_run_steps_noalloc runs implementationVersion 0 only (versions 1 and 2 keep stage arrays "
                "in device memory); got %d", p->implementation_version);
  if (!d_u || !d_u2 || !d_invdt) return fail(E2D_ERR_ARG, "NULL device pointer");
  if (d_u == d_u2) return fail(E2D_ERR_ARG, "the two state buffers must be distinct");
  const int64_t isize = p->nx + 4, jsize = p->ny + 4;
  if (pitch < isize) return fail(E2D_ERR_ARG, "pitch %lld is below isize %lld", (long long)pitch, (long long)isize);
  if (pitch * jsize > (int64_t(1) << 31) / 4)
    return fail(E2D_ERR_ARG, "pitch %lld x jsize %lld too large", (long long)pitch, (long long)jsize);
  if (nsteps < 0 || nStep < 0) return fail(E2D_ERR_ARG, "nsteps and nStep must be >= 0");
  if (nsteps == 0) return E2D_OK;
  double* scr = nullptr;
  if (int rc = scratch_of_current_device(&scr)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  e2d::DevBuffers b;
  b.u[0] = d_u;
  b.u[1] = d_u2;
  b.pitch = pitch;
  b.plane = pitch * jsize;
  b.scal = scr;
  // slot 0: the running invDt folded with the current input's CFL rate
  // (compute_dt); slot 1: zero until the first step folds into it
  HIP_TRY(e2d::launch_fill_from(scr, e2d::kInvdtSlot, d_invdt, st));
  HIP_TRY(e2d::launch_fill(scr + e2d::kInvdtSlot, e2d::kInvdtSlot, 0.0, st));
  HIP_TRY(e2d::launch_max_invdt(*p, b, nStep % 2, 0, st));
  e2d::StepCfg cfg;
  cfg.dt_from_device = true;
  cfg.track_dt = true;
  int k = 0;
  for (int s = 0, n = nStep; s < nsteps; ++s, ++n) {
    cfg.buf_in = n % 2;
    cfg.buf_out = 1 - cfg.buf_in;
    cfg.slot_cur = k;
    cfg.slot_next = 1 - k;
    HIP_TRY(e2d::launch_fused_step(*p, b, cfg, st));
    k = 1 - k;
  }
  HIP_TRY(e2d::launch_slot_value(scr + k * e2d::kInvdtSlot, d_invdt, st));
  return E2D_OK;
}

int e2d_pointwise(const e2d_params* p, int func, int n, const double* x, double* y) {
  g_err[0] = '\0';
  if (!p) return fail(E2D_ERR_ARG, "params is NULL");
  if (func < E2D_FN_PRIM || func > E2D_FN_CONV_GT) return fail(E2D_ERR_ARG, "unknown function %d", func);
  if (n < 0 || (n > 0 && (!x || !y))) return fail(E2D_ERR_ARG, "bad batch");
  // the device routines assume the floors e2d_create checks
  if (int rc = validate(p)) return rc;
  if (n == 0) return E2D_OK;
  if (!e2d_device_available()) return fail(E2D_ERR_HIP, "no HIP device available");
  static const int nin[E2D_FN_CONV_GT + 1] = {0, 4, 20, 15, 8, 8, 4, 2, 2, 1, 2, 3, 2};
  static const int nout[E2D_FN_CONV_GT + 1] = {0, 5, 8, 4, 8, 8, 4, 2, 1, 1, 2, 1, 1};
  double *dx = nullptr, *dy = nullptr;
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&dx), sizeof(double) * n * nin[func]));
  if (hipMalloc(reinterpret_cast<void**>(&dy), sizeof(double) * n * nout[func]) != hipSuccess) {
    (void)hipFree(dx);
    return fail(E2D_ERR_NOMEM, "device allocation failed");
  }
  int rc = E2D_OK;
  if (hipMemcpy(dx, x, sizeof(double) * n * nin[func], hipMemcpyHostToDevice) != hipSuccess ||
      e2d::launch_pointwise(*p, func, n, dx, dy, nullptr) != hipSuccess ||
      hipMemcpy(y, dy, sizeof(double) * n * nout[func], hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(E2D_ERR_HIP, "pointwise batch failed: %s", hipGetErrorString(hipGetLastError()));
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

}  // extern "C"
