// gfx950 kernels of the euler2d hot path.
//
// Fused step (implementationVersion 0): one kernel per time step.  Each
// wavefront owns a strip of 64 consecutive columns (60 output columns and a
// 2-column halo each side) and marches upward through a segment of rows.  The
// y-neighbourhood of a cell stays in the lane's registers (rolling rows of
// primitive variables); x-neighbours come from adjacent lanes through
// cross-lane shuffles, so no LDS tile is needed.  Per cell the primitive
// variables, slopes and trace source terms are computed once; per face one
// Riemann problem is solved.  Ghost cells are produced on load from their
// interior source (make_boundaries, serial_ref/m_run.f90:561-648) and written
// back to the input buffer, so buffers end up exactly as the reference leaves
// them.  The update epilogue also folds the CFL rate of the new state into the
// running invDt of the next step (compute_dt, serial_ref/m_run.f90:248-312).
//
// Materialised step (implementationVersion 1): boundaries (launched by the
// caller), primitive variables, slopes, face states, fluxes and update as
// separate kernels with
// every stage stored in HBM (cuf/m_HydroRun_gpu.cuf:393-418 structure, with
// serial_ref arithmetic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>

#include "e2d_adjoint.hpp"
#include "e2d_common.hpp"
#include "e2d_internal.hpp"
#include "e2d_numerics.hpp"

namespace e2d {

// ---------------------------------------------------------------------------
// fused step

constexpr int kStripOut = 60;  // output columns per wavefront
// Independent wavefronts per workgroup of the fused and v2 flux kernels (their
// wavefronts share nothing).  One per workgroup: measured on MI355X, the FP64
// step takes 196.0-197.2 us at 2048^2 against 197.9-198.6 with four, and
// 680.6-683.4 against 689.5-692.5 at 4096^2, with the same results; the float
// kernels and the v2 flux kernel time the same either way, eight is slower.
#ifndef E2D_WAVES_PER_BLOCK
#define E2D_WAVES_PER_BLOCK 1
#endif
constexpr int kWavesPerBlock = E2D_WAVES_PER_BLOCK;

// 1: the fused step takes the update's input rows from an LDS ring (0: loads
// them again from memory); 1: wavefronts that load only interior cells skip
// the boundary map (0: every load goes through it).  Both give the same bits.
#ifndef E2D_FUSED_UOLD_LDS
#define E2D_FUSED_UOLD_LDS 1
#endif
#ifndef E2D_FUSED_PLAIN
#define E2D_FUSED_PLAIN 1
#endif

// optional register cap for the fused kernel (minimum resident waves per SIMD)
#if defined(E2D_FUSED_MIN_WAVES) && E2D_FUSED_MIN_WAVES > 0
#define E2D_FUSED_ATTR __attribute__((amdgpu_waves_per_eu(E2D_FUSED_MIN_WAVES)))
#else
#define E2D_FUSED_ATTR
#endif

template <class P>
struct StepArgs {
  Hydro<typename P::T> hT;
  Hydro<typename P::U> hU;
  Grid g;
  typename P::S* uin;   // read (interior, corners) and ghost write-back
  typename P::S* uout;  // interior written
  double* scal;
  int dt_from_device, slot_cur, track_dt, slot_next;
  double dt_host;
  // Row segments, one per wavefront: items [0, nlong) are nsegs segments of
  // seg_rows rows per strip over rows [2, short_base); items [nlong, nitems)
  // are nshort segments of short_rows rows per strip over [short_base, ny + 2);
  // each range is numbered by tile_item.
  int nsegs, seg_rows, nlong;
  int nshort, short_rows, short_base, nitems;
  int chunk;  // items per chunk dealt to one XCD (see fused_segment)
  int nstrips, tile;  // items are numbered over tiles of `tile` adjacent strips (see tile_item)
  // Packed last strip: when the last strip has at most 12 (28) output columns,
  // its wavefronts are split into groups of pk_w = 16 (32) lanes, each group on
  // its own segment of pk_rows rows, so the strip takes a quarter (half) of the
  // wavefronts.  Items [nlong - npk, nlong) are its wavefronts; segment k of the
  // strip covers rows [2 + k pk_rows, 2 + (k+1) pk_rows), the last one moved
  // back to end at ny + 2 (it overlaps its neighbour, whose rows it then
  // computes again with the same bits).  pk_w == 0: no packing; the last
  // strip's index is nstrips (the regular strips are 0 .. nstrips - 1).
  int pk_w, pk_rows, pk_nseg, npk;
};

// Item k of a tiled numbering of nstrips strips x nper segments each: tiles of
// `tile` adjacent strips (the last one narrower), numbered tile by tile and
// within a tile segment by segment, so that consecutive items are horizontal
// neighbours first.  tile == 1 is the strip-major order.
__device__ __forceinline__ void tile_item(int k, int nper, int nstrips, int tile, int& strip, int& seg) {
  const int span = tile * nper;
  const int tc = k / span;
  const int w = min(tile, nstrips - tc * tile);
  const int in = k - tc * span;
  seg = in / w;
  strip = tc * tile + (in - seg * w);
}

// Segment (strip, rows [j0, j1)) of wavefront `wave` of block b; false for a
// wavefront without one.  Blocks go to the 8 XCDs round-robin (b % 8) and
// start in order on each.  The long segments are dealt to the XCDs in chunks
// of `chunk` consecutive items (vertical neighbours, which share rows, meet in
// one L2; every XCD gets a spread of the domain, so a region of costly
// segments does not load one XCD), then the short ones the same way: every XCD
// ends its share on short segments, which shortens the tail where the device
// runs out of wavefronts.
template <class P>
__device__ __forceinline__ bool fused_segment(const StepArgs<P>& a, int b, int wave, int lane, int& strip, int& j0,
                                              int& j1, int& group_w) {
  int item;
  bool is_long;
  if (!dealt_item(b, wave, kWavesPerBlock, a.nlong, a.nitems, a.chunk, item, is_long)) return false;
  int seg;
  group_w = 64;
  if (is_long && item >= a.nlong - a.npk) {
    // a wavefront of the packed last strip: group lane / pk_w takes segment
    // (item - first) * (64 / pk_w) + group; a group past the last segment
    // repeats the last one
    group_w = a.pk_w;
    strip = a.nstrips;
    seg = min((item - (a.nlong - a.npk)) * (64 / a.pk_w) + lane / a.pk_w, a.pk_nseg - 1);
    j0 = 2 + seg * a.pk_rows;
    j1 = j0 + a.pk_rows;
    if (j1 > a.g.ny + 2) {
      j1 = a.g.ny + 2;
      j0 = j1 - a.pk_rows;
    }
    return true;
  }
  if (is_long) {
    tile_item(item, a.nsegs, a.nstrips, a.tile, strip, seg);
    j0 = 2 + seg * a.seg_rows;
    j1 = min(j0 + a.seg_rows, a.short_base);
  } else {
    tile_item(item, a.nshort, a.nstrips, a.tile, strip, seg);
    j0 = a.short_base + seg * a.short_rows;
    j1 = min(j0 + a.short_rows, a.g.ny + 2);
  }
  return true;
}

// Loads of cell (ci, r) of the input, taken from its boundary source as
// stored (ghost cells are never read: every ghost maps to an interior cell,
// corners to themselves), so the loads of several rows can be in flight at
// once.  source_cell applies the reflection's sign flip when the values are
// used and writes ghost values back, so that the input buffer ends up as
// make_boundaries leaves it; doing that here, right after the loads, would
// make every load wait for its data.  plain (uniform over the wavefront):
// every cell the wavefront loads is interior, so the boundary map is skipped.
template <class P>
__device__ __forceinline__ void fetch_cell(const StepArgs<P>& a, int ci, int r, bool plain,
                                           typename P::S raw[4]) {
  if (plain) {
#pragma unroll
    for (int v = 0; v < 4; ++v) raw[v] = a.uin[at(a.g, ci, r, v)];
    return;
  }
  const Src s = ghost_source(a.g, ci, r);
#pragma unroll
  for (int v = 0; v < 4; ++v) raw[v] = a.uin[at(a.g, s.i, s.j, v)];
}

template <class P>
__device__ __forceinline__ void source_cell(const StepArgs<P>& a, int ci, bool col_ok, int r, bool plain,
                                            typename P::S raw[4]) {
  if (plain) return;
  const Src s = ghost_source(a.g, ci, r);
#pragma unroll
  for (int v = 0; v < 4; ++v) raw[v] = (v == s.flip) ? -raw[v] : raw[v];
  if (s.ghost && col_ok) {
#pragma unroll
    for (int v = 0; v < 4; ++v) a.uin[at(a.g, ci, r, v)] = raw[v];
  }
}

template <class P>
__device__ __forceinline__ void prim_of(const Hydro<typename P::T>& h, const typename P::S raw[4],
                                        typename P::T q[4], typename P::T& irho) {
  using T = typename P::T;
  T u[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) u[v] = static_cast<T>(raw[v]);
  cons_to_prim<T, typename P::M>(h, u, q, irho);
}
template <class P>
__device__ __forceinline__ void prim_of(const Hydro<typename P::T>& h, const typename P::S raw[4],
                                        typename P::T q[4]) {
  typename P::T irho;
  prim_of<P>(h, raw, q, irho);
}

// kMinmod: instantiated for slope_type == 1 (the benchmark's minmod), whose
// limiter then skips the products by st = 1 and the first-order select
template <class P, int SOLVER, bool kMinmod>
__device__ __forceinline__ void step_fused_body(const StepArgs<P>& a) {
  using T = typename P::T;
  using U = typename P::U;
  using S = typename P::S;
  using M = typename P::M;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const Grid& g = a.g;
  int strip, j0, j1, group_w;
  if (!fused_segment(a, blockIdx.x, wave, lane, strip, j0, j1, group_w)) return;  // whole wavefront: no barrier here
  // lanes per column group: 64, or 16 / 32 in the packed last strip, whose
  // groups march different rows (j0, j1 per lane, j1 - j0 the same for all)
  const int col = lane & (group_w - 1);
  const int ci = strip * kStripOut + col;  // column of this lane
  const bool col_ok = ci < g.isize;
  const int cl = col_ok ? ci : g.isize - 1;  // clamped column for loads
  const bool out_lane = (col >= 2) && (col < group_w - 2) && (ci <= g.nx + 1);
  // columns strip*60 .. strip*60+63 and rows j0-2 .. j1+1 all interior (never
  // in the packed strip, which holds the last columns)
#if E2D_FUSED_PLAIN
  const bool plain = __builtin_amdgcn_readfirstlane(group_w == 64 && strip * kStripOut >= 2 &&
                                                    strip * kStripOut + 63 <= g.nx + 1 && j0 - 2 >= 2 &&
                                                    j1 + 1 <= g.ny + 1 ? 1 : 0) != 0;
#else
  const bool plain = false;
#endif
  // iterations r = j0 - 1 .. j1 of the row loop, the same number on every lane
  const int n_it = __builtin_amdgcn_readfirstlane(j1 - j0 + 2);

  const Hydro<T>& h = a.hT;
  U rate_max = U(0);

  // the running invDt and the first three rows are all requested before any
  // of them is used, so their round trips overlap
  double invdt_cur = 0.0;
  if (a.dt_from_device) invdt_cur = invdt_part(a.scal, a.slot_cur);
  T qa[4], qb[4], qc[4];
  T irb, irc;  // 1/rho of rows r and r+1 as the primitive conversion forms it, for the trace
  S raw[4], raw_next[4], raw_b[4];
  fetch_cell<P>(a, cl, j0 - 2, plain, raw);
  fetch_cell<P>(a, cl, j0 - 1, plain, raw_b);
  fetch_cell<P>(a, cl, j0, plain, raw_next);
  source_cell<P>(a, cl, col_ok, j0 - 2, plain, raw);
  prim_of<P>(h, raw, qa);
  source_cell<P>(a, cl, col_ok, j0 - 1, plain, raw_b);
  prim_of<P>(h, raw_b, qb, irb);

  // dt of this step (serial_ref/m_run.f90:308, 384-385)
  U dt;
  if (a.dt_from_device) {
    invdt_cur = invdt_reduce(invdt_cur);
    dt = a.hU.cfl / static_cast<U>(invdt_cur);
  } else {
    dt = static_cast<U>(a.dt_host);
  }
  const U dtdxU = dt / a.hU.dx;
  const U dtdyU = dt / a.hU.dy;
  const T dtdx = static_cast<T>(dtdxU);
  const T dtdy = static_cast<T>(dtdyU);

  T ymax_prev[4];
  U fy_prev[4] = {U(0), U(0), U(0), U(0)}, fx_prev[4] = {U(0), U(0), U(0), U(0)};
  U fy_cur[4] = {U(0), U(0), U(0), U(0)}, fx_cur[4] = {U(0), U(0), U(0), U(0)};

#if E2D_FUSED_UOLD_LDS
  // The update of row r-1 adds to that row's input, loaded two iterations
  // earlier: each lane keeps its last three input rows in LDS (slot of row k:
  // k mod 3) instead of loading them again from memory, where the line may
  // have left L2 by then.  A lane reads only what it wrote itself.
  // (raw storage: S may have a constructor, which __shared__ does not allow)
  __shared__ __align__(16) unsigned char rows_in_mem[kWavesPerBlock * 3 * 4 * 64 * sizeof(S)];
  S* const rows_in = reinterpret_cast<S*>(rows_in_mem) + wave * (3 * 4 * 64);  // [slot][v][lane]
  int slot_w = 0;  // slot of row r+1; the next slot (mod 3) holds row r-1, written two iterations back
#endif
  // Variable 0 of row r+2 of the input and of row r-1 of the output in the
  // lane's clamped column cl (ci on every lane that stores), moved one row on
  // per iteration: no row-index product per access.
  const S* in_row = a.uin + at(g, cl, j0 + 1, 0);
  S* out_row = a.uout + at(g, cl, j0 - 2, 0);
  for (int it = 0; it < n_it; ++it, in_row += g.pitch, out_row += g.pitch) {
    const int r = j0 - 1 + it;
    // software pipeline: row r+1 was fetched one iteration ago; fetch row r+2
    // now, so its load is in flight during the Riemann solves
#pragma unroll
    for (int v = 0; v < 4; ++v) raw[v] = raw_next[v];
    if (it < n_it - 1) {  // r + 2 <= j1 + 1
      if (plain) {
#pragma unroll
        for (int v = 0; v < 4; ++v) raw_next[v] = in_row[v * g.plane];
      } else {
        fetch_cell<P>(a, cl, r + 2, plain, raw_next);
      }
    }
    S uold[4];
    const bool do_update = it >= 2;  // r >= j0 + 1
#if !E2D_FUSED_UOLD_LDS
    if (do_update && out_lane) {
#pragma unroll
      for (int v = 0; v < 4; ++v) uold[v] = a.uin[at(g, ci, r - 1, v)];
    }
#endif
    source_cell<P>(a, cl, col_ok, r + 1, plain, raw);
#if E2D_FUSED_UOLD_LDS
    // row r+1 as stored (interior cells carry no reflection flip)
#pragma unroll
    for (int v = 0; v < 4; ++v) rows_in[(slot_w * 4 + v) * 64 + lane] = raw[v];
#endif
    prim_of<P>(h, raw, qc, irc);

    // slopes and trace sources of row r
    T qxm[4], qxp[4], dqx[4], dqy[4], s[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qxm[v] = from_left(qb[v]);
      qxp[v] = from_right(qb[v]);
    }
    cell_slopes<kMinmod>(h, qb, qxp, qxm, qc, qa, dqx, dqy);
    trace_sources<T, M>(h, qb, dqx, dqy, s, irb);

    // y-face between rows r-1 and r
    if (it >= 1) {  // r >= j0
      T ymin[4], f[4];
      face_state(h, qb, dqy, s, dtdy, false, ymin);
      riemann_y<T, M, SOLVER>(h, ymax_prev, ymin, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fy_cur[v] = static_cast<U>(f[v]) * dtdyU;
    }
    face_state(h, qb, dqy, s, dtdy, true, ymax_prev);

    // x-face between columns ci-1 and ci, row r
    if (it >= 1 && it < n_it - 1) {  // j0 <= r < j1
      T xmin[4], xmax[4], xl[4], gd[4], f[4];
      face_state(h, qb, dqx, s, dtdx, false, xmin);
      face_state(h, qb, dqx, s, dtdx, true, xmax);
#pragma unroll
      for (int v = 0; v < 4; ++v) xl[v] = from_left(xmax[v]);
      riemann<T, M, SOLVER>(h, xl, xmin, gd, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fx_cur[v] = static_cast<U>(f[v]) * dtdxU;
    }

    // conservative update of row r-1 (updateHydro, serial_ref/m_run.f90:504-513)
    if (do_update) {
      U fxr[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) fxr[v] = from_right(fx_prev[v]);
      if (out_lane) {
        const int jr = r - 1;
#if E2D_FUSED_UOLD_LDS
        const int slot_r = slot_w == 2 ? 0 : slot_w + 1;  // slot of row r - 1
#pragma unroll
        for (int v = 0; v < 4; ++v) uold[v] = rows_in[(slot_r * 4 + v) * 64 + lane];
#endif
        U un[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const U flux_tot = (fy_prev[v] - fy_cur[v]) + (fx_prev[v] - fxr[v]);
          const S sn = static_cast<S>(static_cast<U>(uold[v]) + flux_tot);
          out_row[v * g.plane] = sn;
          un[v] = static_cast<U>(sn);
        }
        // CFL rate of the new state over the compute_dt window [2, n] (0-based)
        if (a.track_dt && ci <= g.nx && jr <= g.ny) {
          const U rate = cfl_rate<U, typename P::MU>(a.hU, un);
          rate_max = rate_max > rate ? rate_max : rate;
        }
      }
    }

#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qa[v] = qb[v];
      qb[v] = qc[v];
      fy_prev[v] = fy_cur[v];
      fx_prev[v] = fx_cur[v];
    }
    irb = irc;
#if E2D_FUSED_UOLD_LDS
    slot_w = slot_w == 2 ? 0 : slot_w + 1;
#endif
  }

  if (a.track_dt) {
    double m = static_cast<double>(rate_max);
    m = m > invdt_cur ? m : invdt_cur;  // running maximum carries over
    m = wave_max(m);
    if (lane == 0) invdt_fold(a.scal, a.slot_next, static_cast<int>(blockIdx.x) * kWavesPerBlock + wave, m);
  }
}

// Resident wavefronts per SIMD the fused step is held at: 3 for FP64
// arithmetic (168 VGPRs), 5 for FP32 (102), 4 for the native-arithmetic
// variants (128), no cap for the tangent's dual numbers.  The packed last
// strip's per-lane rows cost registers the compiler would otherwise take
// (FP64 179, FP32 107: one wavefront fewer); held at these counts neither
// spills.
template <class P>
constexpr int fused_waves() {
  if constexpr (sizeof(typename P::T) > 8) return 1;
  else if constexpr (sizeof(typename P::T) == 8) return 3;
  else if constexpr (std::is_same<typename P::M, RangedMath>::value) return 5;
  else return 4;
}

template <class P, int SOLVER, bool kMinmod = false>
__global__ __launch_bounds__(64 * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(fused_waves<P>())))
E2D_FUSED_ATTR void k_step_fused(const StepArgs<P> a) {
  step_fused_body<P, SOLVER, kMinmod>(a);
}

// The HLLC step in its own kernel, held at 3 resident wavefronts per SIMD
// (168 VGPRs): left alone the FP64 instantiation takes 172 and drops to 2.
template <class P>
__global__ __launch_bounds__(64 * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_step_fused_hllc(
    const StepArgs<P> a) {
  step_fused_body<P, 2, false>(a);
}

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

// Rows per long segment where the whole-round plan does not apply (small
// grids; the v2 flux kernel).  Each segment re-primes its rolling rows (two
// extra row loads and primitive conversions, one extra row of y-faces), so
// long segments are cheaper per row; short ones give more wavefronts to hide
// latency and balance the data-dependent Newton work, and on small grids a
// wavefront's rows are the critical path.  About ny/128 rows at 3 resident
// wavefronts per SIMD (the FP64 kernels), proportionally fewer where more are
// resident, within [1, 32]; E2D_FUSED_SEG_ROWS overrides it.  Measured on
// MI355X with long segments over 75% of the rows: 8 rows at 1024^2 (76 us per
// FP64 step against 86 us for uniform ny/170-row segments), 4 at 512^2 (29.6
// against 35.6).
static int fused_seg_rows(int ny, int waves_per_simd = 3) {
  const int r = env_int("E2D_FUSED_SEG_ROWS", 0);
  if (r >= 1) return r;
  const int d = static_cast<int>(ny / 128.0 * 3.0 / std::max(1, waves_per_simd) + 0.5);
  return d < 1 ? 1 : (d > 32 ? 32 : d);
}

// compute units of the current device (256 on MI355X), per host thread
static int device_cus() {
  thread_local int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

// resident wavefronts per SIMD of a kernel launched with kWavesPerBlock waves per block
static int waves_per_simd(const void* kernel) {
  int blocks_per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, kernel, 64 * kWavesPerBlock, 0) !=
          hipSuccess ||
      blocks_per_cu <= 0)
    return 3;
  return std::max(1, blocks_per_cu * kWavesPerBlock / 4);
}

// Segment plan of a fused launch over nstrips strips.  Long segments cover
// about 80% (FP64) or 88% (FP32, mixed) of each strip's rows, in as many
// segments per strip as fill k whole rounds of the device's resident
// wavefronts (k the least that keeps them within 40 rows); short ones of a
// quarter of their length cover the rest.
// Every long segment then starts in the first k rounds, and what is left when
// they end is short work, so no long segment runs alone at the end of the
// launch.  Measured on MI355X (us per step, dt on the device): 2048^2 FP64
// 19-row long segments (one round) 213.5 against 223 with 16 rows over 75%
// (which leaves 9% of the long segments for a second round); FP32 12 rows
// 134-136 against 139-140; fp32_native 106.3 against 107.1; 4096^2 FP64 37 rows
// (two rounds) 711 against 716, 74 rows (one round) 719.  Where the rounds
// would give segments of fewer than 8 rows (grids up to about 1024^2) the
// plan is long segments of fused_seg_rows rows over 75% of the rows.
// E2D_FUSED_SEG_ROWS or E2D_FUSED_LONG_PCT select that plan with the given
// values, E2D_FUSED_SHORT_ROWS the short length; the environment is read at
// every launch.  Returns the number of blocks.
template <class P>
static int plan_segments(StepArgs<P>& a, int nstrips_all, int ny, int wps, int resident) {
  // the packed last strip (see StepArgs): output columns left for it
  const int last_cols = a.g.nx - kStripOut * (nstrips_all - 1);
  int pk_w = 0;
  if (nstrips_all >= 2 && env_int("E2D_FUSED_PACK", 1) != 0) pk_w = last_cols <= 12 ? 16 : (last_cols <= 28 ? 32 : 0);
  const int nstrips = pk_w ? nstrips_all - 1 : nstrips_all;
  const int groups = pk_w ? 64 / pk_w : 1;  // segments per packed wavefront
  int R = 0, L = 0;
  const bool forced = env_int("E2D_FUSED_SEG_ROWS", 0) >= 1 || env_int("E2D_FUSED_LONG_PCT", -1) >= 0;
  if (!forced && resident >= nstrips) {
    // the long segments' share of the rows: 80% for the FP64 kernels (3
    // resident wavefronts per SIMD), 88% for the float-arithmetic ones (4-5):
    // measured at 2048^2, FP64 210-213 us at 80% against 214-218 at 76, 84 and
    // 88%; FP32 132.5 at 88% against 137 at 80%, fp32_native 104.5 against 107
    const int cover = env_int("E2D_FUSED_COVER_PCT", wps >= 4 ? 88 : 80);
    const double long_rows = std::min(100, std::max(10, cover)) / 100.0 * ny;
    // the packed wavefronts, each as long as a long segment, take their slots
    // out of the rounds first (a second pass once their number is known)
    int npk = 0;
    for (int pass = 0; pass < 2; ++pass) {
      auto per_strip = [&](int k) { return k * ((resident - (npk + k - 1) / k) / nstrips); };
      int k = 1;
      while (per_strip(k) >= 1 && long_rows / per_strip(k) > 40.0) ++k;
      L = per_strip(k);
      if (L < 1) {
        R = 0;
        break;
      }
      R = static_cast<int>(long_rows / L + 0.5);
      if (R < 8) R = 0;  // small grids: the percentage plan below
      while (R > 0 && L * R > ny) --R;
      if (R == 0 || !pk_w) break;
      const int pk = ((ny + R - 1) / R + groups - 1) / groups;
      if (pk == npk) break;
      npk = pk;
    }
  }
  if (R == 0) {
    R = fused_seg_rows(ny, wps);
    const int pct = std::min(100, std::max(0, env_int("E2D_FUSED_LONG_PCT", 75)));
    L = pct >= 100 ? (ny + R - 1) / R : static_cast<int>(static_cast<int64_t>(ny) * pct / 100 / R);
  }
  const int r = std::max(1, env_int("E2D_FUSED_SHORT_ROWS", R / 4));
  const int covered = std::min(L * R, ny);
  a.seg_rows = R;
  a.nsegs = L;
  a.pk_w = pk_w;
  a.pk_rows = std::min(R, ny);
  a.pk_nseg = pk_w ? (ny + a.pk_rows - 1) / a.pk_rows : 0;
  a.npk = pk_w ? (a.pk_nseg + groups - 1) / groups : 0;
  a.nlong = nstrips * L + a.npk;
  a.short_base = 2 + covered;
  a.short_rows = r;
  a.nshort = covered < ny ? (ny - covered + r - 1) / r : 0;
  a.nitems = a.nlong + nstrips * a.nshort;
  a.chunk = std::max(1, env_int("E2D_FUSED_CHUNK", 16));
  a.nstrips = nstrips;
  a.tile = std::min(std::max(1, env_int("E2D_FUSED_TILE", 1)), std::max(1, nstrips));
  return dealt_blocks(a.nlong, a.nitems, a.chunk, kWavesPerBlock);
}

template <class P, int SOLVER, bool kMinmod>
static hipError_t launch_fused(StepArgs<P>& a, int nstrips, hipStream_t st) {
  const void* kernel;
  if constexpr (SOLVER == 2)
    kernel = reinterpret_cast<const void*>(k_step_fused_hllc<P>);
  else
    kernel = reinterpret_cast<const void*>(k_step_fused<P, SOLVER, kMinmod>);
  thread_local const int wps = waves_per_simd(kernel);
  const int nblocks = plan_segments(a, nstrips, a.g.ny, wps, 4 * wps * device_cus());
  if constexpr (SOLVER == 2)
    hipLaunchKernelGGL((k_step_fused_hllc<P>), dim3(nblocks), dim3(64 * kWavesPerBlock), 0, st, a);
  else
    hipLaunchKernelGGL((k_step_fused<P, SOLVER, kMinmod>), dim3(nblocks), dim3(64 * kWavesPerBlock), 0, st, a);
  return hipGetLastError();
}

template <class P>
static hipError_t fused_step(const e2d_params& p, const DevBuffers& b, const StepCfg& c, hipStream_t st) {
  StepArgs<P> a;
  a.hT = make_hydro<typename P::T>(p);
  a.hU = make_hydro<typename P::U>(p);
  a.g = make_grid(p, b);
  a.uin = static_cast<typename P::S*>(b.u[c.buf_in]);
  a.uout = static_cast<typename P::S*>(b.u[c.buf_out]);
  a.scal = b.scal;
  a.dt_from_device = c.dt_from_device ? 1 : 0;
  a.slot_cur = c.slot_cur;
  a.track_dt = c.track_dt ? 1 : 0;
  a.slot_next = c.slot_next;
  a.dt_host = c.dt_host;
  const bool hllc = p.riemann_solver == E2D_RIEMANN_HLLC;
  const bool minmod = !hllc && p.slope_type == 1;
  const int nstrips = (p.nx + kStripOut - 1) / kStripOut;
  if (hllc) return launch_fused<P, 2, false>(a, nstrips, st);
  if (minmod) return launch_fused<P, 0, true>(a, nstrips, st);
  return launch_fused<P, 0, false>(a, nstrips, st);
}

hipError_t launch_fused_step(const e2d_params& p, const DevBuffers& b, const StepCfg& c, hipStream_t s) {
  return with_precision(p.precision, [&](auto pol) { return fused_step<decltype(pol)>(p, b, c, s); });
}

// ---------------------------------------------------------------------------
// tangent-linear step: the fused kernel on dual numbers (value = the forward
// step's state bit for bit, derivative = the tangent), dt frozen

using PDual = Prec<Dual, Dual, Dual, DualMath<RangedMath>, DualMath<RangedMath>>;

// (U, V) -> dual buffer; V enters on interior cells only (ghost directions
// follow from the interior through the boundary map, corners are constants)
__global__ __launch_bounds__(256) void k_pack_dual(const Grid g, const double* __restrict__ u,
                                                   const double* __restrict__ v, Dual* __restrict__ out) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= g.isize || j >= g.jsize) return;
  const bool interior = i >= 2 && i <= g.nx + 1 && j >= 2 && j <= g.ny + 1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t o = at(g, i, j, k);
    out[o] = Dual(u[o], interior ? v[o] : 0.0);
  }
}

__global__ __launch_bounds__(256) void k_unpack_dual(const Grid g, const Dual* __restrict__ in,
                                                     double* __restrict__ v) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= g.isize || j >= g.jsize) return;
  const bool interior = i >= 2 && i <= g.nx + 1 && j >= 2 && j <= g.ny + 1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t o = at(g, i, j, k);
    v[o] = interior ? in[o].d : 0.0;
  }
}

hipError_t launch_tangent_step(const e2d_params& p, const DevBuffers& b, int buf, double dt, const double* v_in,
                               double* v_out, void* dual_in, void* dual_out, hipStream_t st) {
  const Grid g = make_grid(p, b);
  const dim3 grid((g.isize + 63) / 64, (g.jsize + 3) / 4);
  hipLaunchKernelGGL(k_pack_dual, grid, dim3(256), 0, st, g, static_cast<const double*>(b.u[buf]), v_in,
                     static_cast<Dual*>(dual_in));
  StepArgs<PDual> a;
  a.hT = make_hydro<Dual>(p);
  a.hU = a.hT;
  a.g = g;
  a.uin = static_cast<Dual*>(dual_in);
  a.uout = static_cast<Dual*>(dual_out);
  a.scal = b.scal;
  a.dt_from_device = 0;
  a.slot_cur = 0;
  a.track_dt = 0;
  a.slot_next = 1;
  a.dt_host = dt;
  if (hipError_t e = launch_fused<PDual, 0, false>(a, (p.nx + kStripOut - 1) / kStripOut, st)) return e;
  hipLaunchKernelGGL(k_unpack_dual, grid, dim3(256), 0, st, g, static_cast<const Dual*>(dual_out), v_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// boundaries (make_boundaries): one thread per non-corner ghost cell

template <class S>
__global__ __launch_bounds__(256) void k_boundaries(const Grid g, S* u) {
  const int nxg = 4 * g.ny;  // x-face ghosts: 2 columns each side, interior rows
  const int nyg = 4 * g.nx;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int i, j;
  if (t < nxg) {
    const int k = t % 4;
    j = 2 + t / 4;
    i = (k < 2) ? k : g.nx + k;  // 0,1 | nx+2, nx+3
  } else if (t < nxg + nyg) {
    const int t2 = t - nxg;
    const int k = t2 % 4;
    i = 2 + t2 / 4;
    j = (k < 2) ? k : g.ny + k;
  } else {
    return;
  }
  const Src s = ghost_source(g, i, j);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    S x = u[at(g, s.i, s.j, v)];
    if (v == s.flip) x = -x;
    u[at(g, i, j, v)] = x;
  }
}

template <class P>
static hipError_t boundaries(const e2d_params& p, const DevBuffers& b, int buf, hipStream_t st) {
  const Grid g = make_grid(p, b);
  const int n = 4 * (p.nx + p.ny);
  hipLaunchKernelGGL(k_boundaries<typename P::S>, dim3((n + 255) / 256), dim3(256), 0, st, g,
                     static_cast<typename P::S*>(b.u[buf]));
  return hipGetLastError();
}

hipError_t launch_boundaries(const e2d_params& p, const DevBuffers& b, int buf, hipStream_t s) {
  const TraceRange range("make_boundaries");
  return with_precision(p.precision, [&](auto pol) { return boundaries<decltype(pol)>(p, b, buf, s); });
}

// ---------------------------------------------------------------------------
// compute_dt reduction: max CFL rate over [2, nx] x [2, ny] (0-based)

template <class P>
__global__ __launch_bounds__(256) void k_max_invdt(const Hydro<typename P::U> h, const Grid g,
                                                    const typename P::S* __restrict__ u, double* scal, int slot) {
  using U = typename P::U;
  __shared__ double red[4];
  const int w = g.nx - 1, hgt = g.ny - 1;  // window extent
  const int64_t n = static_cast<int64_t>(w) * hgt;
  U m = U(0);
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int i = 2 + static_cast<int>(t % w);
    const int j = 2 + static_cast<int>(t / w);
    U c[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) c[v] = static_cast<U>(u[at(g, i, j, v)]);
    const U r = cfl_rate<U, typename P::MU>(h, c);
    m = m > r ? m : r;
  }
  double md = wave_max(static_cast<double>(m));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = md;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = red[0];
    for (int k = 1; k < 4; ++k) x = x > red[k] ? x : red[k];
    invdt_fold(scal, slot, blockIdx.x, x);
  }
}

template <class P>
static hipError_t max_invdt(const e2d_params& p, const DevBuffers& b, int buf, int slot, hipStream_t st) {
  const Grid g = make_grid(p, b);
  const int64_t n = static_cast<int64_t>(p.nx - 1) * (p.ny - 1);
  int nb = static_cast<int>((n + 255) / 256);
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_max_invdt<P>, dim3(nb), dim3(256), 0, st, make_hydro<typename P::U>(p), g,
                     static_cast<const typename P::S*>(b.u[buf]), b.scal, slot);
  return hipGetLastError();
}

hipError_t launch_max_invdt(const e2d_params& p, const DevBuffers& b, int buf, int slot, hipStream_t s) {
  return with_precision(p.precision, [&](auto pol) { return max_invdt<decltype(pol)>(p, b, buf, slot, s); });
}

__global__ void k_fill(double* dst, int n, double v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = v;
}

hipError_t launch_fill(double* dst, int n, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3((n + 63) / 64), dim3(64), 0, s, dst, n, v);
  return hipGetLastError();
}

__global__ void k_fill_from(double* dst, int n, const double* src) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = *src;
}

hipError_t launch_fill_from(double* dst, int n, const double* src, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_from, dim3((n + 63) / 64), dim3(64), 0, s, dst, n, src);
  return hipGetLastError();
}

// one wavefront: *dst = the value of the running-invDt slot at `slot` (the maximum of its parts)
__global__ __launch_bounds__(64) void k_slot_value(const double* slot, double* dst) {
  const double v = invdt_read_wave(slot, 0);
  if (threadIdx.x == 0) *dst = v;
}

hipError_t launch_slot_value(const double* slot, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_slot_value, dim3(1), dim3(64), 0, s, slot, dst);
  return hipGetLastError();
}

// *flag = 1 when two device arrays of n 64-bit words differ in any bit
__global__ void k_words_differ(const unsigned long long* a, const unsigned long long* b, int64_t n, double* flag) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride)
    if (a[t] != b[t]) *flag = 1.0;
}

hipError_t launch_words_differ(const void* a, const void* b, size_t bytes, double* flag, hipStream_t s) {
  hipError_t e = launch_fill(flag, 1, 0.0, s);
  if (e != hipSuccess) return e;
  const int64_t n = static_cast<int64_t>(bytes / sizeof(unsigned long long));
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_words_differ, dim3(static_cast<unsigned>(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096)),
                     dim3(256), 0, s, static_cast<const unsigned long long*>(a),
                     static_cast<const unsigned long long*>(b), n, flag);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// materialised steps (implementationVersion 1: q, slopes, face states and
// fluxes in HBM; implementationVersion 2: q and fluxes in HBM)

template <class P>
struct MatArgs {
  Hydro<typename P::T> hT;
  Hydro<typename P::U> hU;
  Grid g;
  const typename P::S* uin;
  typename P::S* uout;
  typename P::T *q, *dqx, *dqy, *qxm, *qxp, *qym, *qyp;
  typename P::U *fx, *fy;
  double* scal;
  int dt_from_device, slot_cur;
  double dt_host;
  int nsegs, seg_rows, seg_groups;  // k_flux_store row segments
  int track_dt;  // k_update folds the new state's CFL rates into invDt slot slot_cur
};

// kWave: called with all 64 lanes of the wavefront active (wave-parallel read)
template <bool kWave = false, class P>
__device__ __forceinline__ void step_dt(const MatArgs<P>& a, typename P::U& dtdx, typename P::U& dtdy) {
  using U = typename P::U;
  U dt = static_cast<U>(a.dt_host);
  if (a.dt_from_device)
    dt = a.hU.cfl / static_cast<U>(kWave ? invdt_read_wave(a.scal, a.slot_cur) : invdt_read(a.scal, a.slot_cur));
  dtdx = dt / a.hU.dx;
  dtdy = dt / a.hU.dy;
}

// convertToPrimitives over all cells (serial_ref/m_run.f90:338-354)
template <class P>
__global__ __launch_bounds__(256) void k_prim(const MatArgs<P> a) {
  using T = typename P::T;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= a.g.isize || j >= a.g.jsize) return;
  T u[4], q[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) u[v] = static_cast<T>(a.uin[at(a.g, i, j, v)]);
  cons_to_prim<T, typename P::M>(a.hT, u, q);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.q[at(a.g, i, j, v)] = q[v];
}

// slopes of cells [1, size-2]^2
template <class P>
__global__ __launch_bounds__(256) void k_slopes(const MatArgs<P> a) {
  using T = typename P::T;
  const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i > a.g.isize - 2 || j > a.g.jsize - 2) return;
  T q[4], xp[4], xm[4], yp[4], ym[4], dx[4], dy[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    q[v] = a.q[at(a.g, i, j, v)];
    xp[v] = a.q[at(a.g, i + 1, j, v)];
    xm[v] = a.q[at(a.g, i - 1, j, v)];
    yp[v] = a.q[at(a.g, i, j + 1, v)];
    ym[v] = a.q[at(a.g, i, j - 1, v)];
  }
  cell_slopes(a.hT, q, xp, xm, yp, ym, dx, dy);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    a.dqx[at(a.g, i, j, v)] = dx[v];
    a.dqy[at(a.g, i, j, v)] = dy[v];
  }
}

// the four face states of cells [1, size-2]^2
template <class P>
__global__ __launch_bounds__(256) void k_trace(const MatArgs<P> a) {
  using T = typename P::T;
  using U = typename P::U;
  const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i > a.g.isize - 2 || j > a.g.jsize - 2) return;
  U dtdxU, dtdyU;
  step_dt(a, dtdxU, dtdyU);
  const T dtdx = static_cast<T>(dtdxU), dtdy = static_cast<T>(dtdyU);
  T q[4], dx[4], dy[4], s[4], f[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    q[v] = a.q[at(a.g, i, j, v)];
    dx[v] = a.dqx[at(a.g, i, j, v)];
    dy[v] = a.dqy[at(a.g, i, j, v)];
  }
  trace_sources<T, typename P::M>(a.hT, q, dx, dy, s);
  face_state(a.hT, q, dx, s, dtdx, true, f);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.qxm[at(a.g, i, j, v)] = f[v];
  face_state(a.hT, q, dx, s, dtdx, false, f);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.qxp[at(a.g, i, j, v)] = f[v];
  face_state(a.hT, q, dy, s, dtdy, true, f);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.qym[at(a.g, i, j, v)] = f[v];
  face_state(a.hT, q, dy, s, dtdy, false, f);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.qyp[at(a.g, i, j, v)] = f[v];
}

// fluxes of the faces of computeAndStoreFluxes, (i, j) in [2, size-2]^2
// (serial_ref/m_run.f90:389-480)
template <class P, int SOLVER>
__global__ __launch_bounds__(256) void k_flux(const MatArgs<P> a) {
  using T = typename P::T;
  using U = typename P::U;
  const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i > a.g.isize - 2 || j > a.g.jsize - 2) return;
  U dtdx, dtdy;
  step_dt(a, dtdx, dtdy);
  T l[4], r[4], gd[4], f[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    l[v] = a.qxm[at(a.g, i - 1, j, v)];
    r[v] = a.qxp[at(a.g, i, j, v)];
  }
  riemann<T, typename P::M, SOLVER>(a.hT, l, r, gd, f);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.fx[at(a.g, i, j, v)] = static_cast<U>(f[v]) * dtdx;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    l[v] = a.qym[at(a.g, i, j - 1, v)];
    r[v] = a.qyp[at(a.g, i, j, v)];
  }
  riemann_y<T, typename P::M, SOLVER>(a.hT, l, r, f);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.fy[at(a.g, i, j, v)] = static_cast<U>(f[v]) * dtdy;
}

// computeFluxesAndStore (implementationVersion 2, cuf/m_HydroRun_gpu.cuf:747-884,
// with serial_ref's arithmetic): fx(i, j) (face between i-1 and i) and fy(i, j)
// (face between j-1 and j) for (i, j) in [2, isize-2]^2, times dt/dx and dt/dy,
// from the primitives k_prim stored.  The reference computes slopes and traces
// three times per cell (own, i-1, j-1); here, as in the fused step, a wavefront
// owns a strip of 64 columns (60 stored, lanes 2..61) and marches up a segment
// of rows, forming each cell's slopes, trace sources and face states once: the
// left state of an x-face comes from lane-1, the lower state of a y-face from
// the previous row, in registers.  Same bits as the per-cell formulation.
template <class P, int SOLVER, bool kMinmod = false>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_flux_store(const MatArgs<P> a) {
  using T = typename P::T;
  using U = typename P::U;
  using M = typename P::M;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  int strip, group;
  strip_and_group<false>(lb, a.seg_groups, strip, group);
  const int seg = group * kWavesPerBlock + wave;
  if (seg >= a.nsegs) return;  // whole wavefront leaves: no barrier in this kernel

  const Grid& g = a.g;
  const int ci = strip * kStripOut + lane;  // column of this lane
  const int cl = ci < g.isize ? ci : g.isize - 1;
  const bool out_lane = lane >= 2 && lane < 62 && ci <= g.isize - 2;
  const int j0 = 2 + seg * a.seg_rows;
  const int j1 = min(j0 + a.seg_rows, g.jsize - 1);  // stored rows [j0, j1)
  U dtdxU, dtdyU;
  step_dt<true>(a, dtdxU, dtdyU);
  const T dtdx = static_cast<T>(dtdxU), dtdy = static_cast<T>(dtdyU);
  const Hydro<T>& h = a.hT;

  T qa[4], qb[4], qc[4], ymax_prev[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    qa[v] = a.q[at(g, cl, j0 - 2, v)];
    qb[v] = a.q[at(g, cl, j0 - 1, v)];
  }
  for (int r = j0 - 1; r < j1; ++r) {
#pragma unroll
    for (int v = 0; v < 4; ++v) qc[v] = a.q[at(g, cl, r + 1, v)];
    T qxm[4], qxp[4], dqx[4], dqy[4], s[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qxm[v] = from_left(qb[v]);
      qxp[v] = from_right(qb[v]);
    }
    cell_slopes<kMinmod>(h, qb, qxp, qxm, qc, qa, dqx, dqy);
    trace_sources<T, M>(h, qb, dqx, dqy, s);
    if (r >= j0) {
      // y-face between rows r-1 and r
      T ymin[4], f[4];
      face_state(h, qb, dqy, s, dtdy, false, ymin);
      riemann_y<T, M, SOLVER>(h, ymax_prev, ymin, f);
      if (out_lane) {
#pragma unroll
        for (int v = 0; v < 4; ++v) a.fy[at(g, ci, r, v)] = static_cast<U>(f[v]) * dtdyU;
      }
      // x-face between columns ci-1 and ci
      T xmin[4], xmax[4], xl[4], gd[4];
      face_state(h, qb, dqx, s, dtdx, false, xmin);
      face_state(h, qb, dqx, s, dtdx, true, xmax);
#pragma unroll
      for (int v = 0; v < 4; ++v) xl[v] = from_left(xmax[v]);
      riemann<T, M, SOLVER>(h, xl, xmin, gd, f);
      if (out_lane) {
#pragma unroll
        for (int v = 0; v < 4; ++v) a.fx[at(g, ci, r, v)] = static_cast<U>(f[v]) * dtdxU;
      }
    }
    face_state(h, qb, dqy, s, dtdy, true, ymax_prev);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qa[v] = qb[v];
      qb[v] = qc[v];
    }
  }
}

// updateHydro over the interior (serial_ref/m_run.f90:504-513).  With
// track_dt, the CFL rate of each new cell in the compute_dt window [2, n]
// (0-based) is folded into the running invDt (compute_dt of the new state,
// serial_ref/m_run.f90:248-312): a wave maximum, a block maximum through LDS,
// one atomic maximum per block.  This step's dt was read from that slot by the
// flux kernels, which have completed; the maximum is order-independent.
template <class P>
__global__ __launch_bounds__(256) void k_update(const MatArgs<P> a) {
  using U = typename P::U;
  using S = typename P::S;
  __shared__ double red[4];
  const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const bool cell = i <= a.g.nx + 1 && j <= a.g.ny + 1;
  U rate = U(0);
  if (cell) {
    U un[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      U flux_tot = a.fx[at(a.g, i, j, v)] - a.fx[at(a.g, i + 1, j, v)];
      flux_tot = a.fy[at(a.g, i, j, v)] - a.fy[at(a.g, i, j + 1, v)] + flux_tot;
      const S sn = static_cast<S>(static_cast<U>(a.uin[at(a.g, i, j, v)]) + flux_tot);
      a.uout[at(a.g, i, j, v)] = sn;
      un[v] = static_cast<U>(sn);
    }
    if (a.track_dt && i <= a.g.nx && j <= a.g.ny) rate = cfl_rate<U, typename P::MU>(a.hU, un);
  }
  if (!a.track_dt) return;  // uniform: the whole block leaves before the barrier
  const double m = wave_max(static_cast<double>(rate));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = red[0];
    for (int k = 1; k < 4; ++k) x = x > red[k] ? x : red[k];
    invdt_fold(a.scal, a.slot_cur, static_cast<int>(blockIdx.y * gridDim.x + blockIdx.x), x);
  }
}

template <class P>
static hipError_t materialised_step(const e2d_params& p, const DevBuffers& b, const StepCfg& c,
                                    hipStream_t st) {
  using T = typename P::T;
  using U = typename P::U;
  MatArgs<P> a;
  a.hT = make_hydro<T>(p);
  a.hU = make_hydro<U>(p);
  a.g = make_grid(p, b);
  a.uin = static_cast<const typename P::S*>(b.u[c.buf_in]);
  a.uout = static_cast<typename P::S*>(b.u[c.buf_out]);
  a.q = static_cast<T*>(b.stage[E2D_STAGE_Q]);
  a.fx = static_cast<U*>(b.stage[E2D_STAGE_FX]);
  a.fy = static_cast<U*>(b.stage[E2D_STAGE_FY]);
  a.dqx = static_cast<T*>(b.stage[E2D_STAGE_DQX]);
  a.dqy = static_cast<T*>(b.stage[E2D_STAGE_DQY]);
  a.qxm = static_cast<T*>(b.stage[E2D_STAGE_QXM]);
  a.qxp = static_cast<T*>(b.stage[E2D_STAGE_QXP]);
  a.qym = static_cast<T*>(b.stage[E2D_STAGE_QYM]);
  a.qyp = static_cast<T*>(b.stage[E2D_STAGE_QYP]);
  a.scal = b.scal;
  a.dt_from_device = c.dt_from_device ? 1 : 0;
  a.slot_cur = c.slot_cur;
  a.dt_host = c.dt_host;
  a.track_dt = c.track_dt ? 1 : 0;
  const int S_ = p.nx + 4, J_ = p.ny + 4;
  const dim3 fgrid((S_ - 3 + 63) / 64, (J_ - 3 + 3) / 4);
  {
    const TraceRange range("convertToPrimitives");
    hipLaunchKernelGGL(k_prim<P>, dim3((S_ + 63) / 64, (J_ + 3) / 4), dim3(256), 0, st, a);
  }
  {
    const TraceRange range("computeAndStoreFluxes");
    if (p.implementation_version == E2D_IMPL_FLUX_STORE) {
      // stored rows and columns [2, size-2]; segment length as for the fused step
      const bool hllc = p.riemann_solver == E2D_RIEMANN_HLLC;
      const bool minmod = !hllc && p.slope_type == 1;
      const void* kernel = hllc ? reinterpret_cast<const void*>(k_flux_store<P, 2>)
                                : (minmod ? reinterpret_cast<const void*>(k_flux_store<P, 0, true>)
                                          : reinterpret_cast<const void*>(k_flux_store<P, 0>));
      thread_local const void* cached_kernel = nullptr;
      thread_local int cached_wps = 3;
      if (kernel != cached_kernel) {
        cached_wps = waves_per_simd(kernel);
        cached_kernel = kernel;
      }
      a.seg_rows = fused_seg_rows(p.ny, cached_wps);
      a.nsegs = (J_ - 3 + a.seg_rows - 1) / a.seg_rows;
      a.seg_groups = (a.nsegs + kWavesPerBlock - 1) / kWavesPerBlock;
      const dim3 mgrid(((S_ - 3 + kStripOut - 1) / kStripOut) * a.seg_groups);
      if (hllc)
        hipLaunchKernelGGL((k_flux_store<P, 2>), mgrid, dim3(64 * kWavesPerBlock), 0, st, a);
      else if (minmod)
        hipLaunchKernelGGL((k_flux_store<P, 0, true>), mgrid, dim3(64 * kWavesPerBlock), 0, st, a);
      else
        hipLaunchKernelGGL((k_flux_store<P, 0>), mgrid, dim3(64 * kWavesPerBlock), 0, st, a);
    } else {
      hipLaunchKernelGGL(k_slopes<P>, dim3((S_ - 2 + 63) / 64, (J_ - 2 + 3) / 4), dim3(256), 0, st, a);
      hipLaunchKernelGGL(k_trace<P>, dim3((S_ - 2 + 63) / 64, (J_ - 2 + 3) / 4), dim3(256), 0, st, a);
      if (p.riemann_solver == E2D_RIEMANN_HLLC)
        hipLaunchKernelGGL((k_flux<P, 2>), fgrid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((k_flux<P, 0>), fgrid, dim3(256), 0, st, a);
    }
  }
  {
    const TraceRange range("updateHydro");
    hipLaunchKernelGGL(k_update<P>, dim3((p.nx + 63) / 64, (p.ny + 3) / 4), dim3(256), 0, st, a);
  }
  return hipGetLastError();
}

hipError_t launch_materialised_step(const e2d_params& p, const DevBuffers& b, const StepCfg& c,
                                    hipStream_t s) {
  return with_precision(p.precision, [&](auto pol) { return materialised_step<decltype(pol)>(p, b, c, s); });
}

// ---------------------------------------------------------------------------
// initial conditions (init_implode, serial_ref/m_run.f90:524-555; diag; bumps)

struct Bumps {
  double v[16 * 5];  // cx, cy, amp_rho, amp_p, sigma
};

template <class P>
__global__ __launch_bounds__(256) void k_init(const Grid g, typename P::S* u, typename P::S* u2,
                                              int problem, typename P::U gamma0, Bumps bumps) {
  using U = typename P::U;
  using S = typename P::S;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= g.isize || j >= g.jsize) return;
  U rho, e;
  if (problem == E2D_PROBLEM_IMPLODE) {
    // tmp in default real (FP32) whatever the run's precision
    const float tmp = 1.0f * static_cast<float>(i - 2) / static_cast<float>(g.nx) +
                      1.0f * static_cast<float>(j - 2) / static_cast<float>(g.ny);
    if (static_cast<U>(tmp) > U(0.5)) {
      rho = U(1.0);
      e = U(1.0) / (gamma0 - U(1.0));
    } else {
      rho = U(0.125);
      e = static_cast<U>(0.14f) / (gamma0 - U(1.0));
    }
  } else {
    const U x = static_cast<U>(i - 2) / static_cast<U>(g.nx);
    const U y = static_cast<U>(j - 2) / static_cast<U>(g.ny);
    U pr;
    if (x + y < U(1)) {
      rho = U(1);
      pr = U(1);
    } else {
      rho = U(0.125);
      pr = U(0.1);
    }
    if (problem == E2D_PROBLEM_BUMPS) {
      const double xc = (static_cast<double>(i - 2) + 0.5) / g.nx;
      const double yc = (static_cast<double>(j - 2) + 0.5) / g.ny;
      double dr = 0.0, dp = 0.0;
      for (int k = 0; k < 16; ++k) {
        const double* bk = bumps.v + 5 * k;
        const double ddx = xc - bk[0], ddy = yc - bk[1], sg = bk[4];
        const double w = exp(-(ddx * ddx + ddy * ddy) / (2.0 * sg * sg));
        dr += bk[2] * w;
        dp += bk[3] * w;
      }
      rho = rho + static_cast<U>(dr);
      pr = pr + static_cast<U>(dp);
    }
    e = pr / (gamma0 - U(1));
  }
  const S vals[4] = {static_cast<S>(rho), static_cast<S>(e), S(0), S(0)};
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    u[at(g, i, j, v)] = vals[v];
    u2[at(g, i, j, v)] = S(0);
  }
}

template <class P>
static hipError_t init_problem(const e2d_params& p, const DevBuffers& b, int problem, const double* bumps,
                               hipStream_t st) {
  const Grid g = make_grid(p, b);
  Bumps bb{};
  if (bumps)
    for (int k = 0; k < 16 * 5; ++k) bb.v[k] = bumps[k];
  hipLaunchKernelGGL(k_init<P>, dim3((g.isize + 63) / 64, (g.jsize + 3) / 4), dim3(256), 0, st, g,
                     static_cast<typename P::S*>(b.u[0]), static_cast<typename P::S*>(b.u[1]), problem,
                     static_cast<typename P::U>(p.gamma0), bb);
  return hipGetLastError();
}

hipError_t launch_init_problem(const e2d_params& p, const DevBuffers& b, int problem, const double* bumps,
                               hipStream_t s) {
  return with_precision(p.precision,
                        [&](auto pol) { return init_problem<decltype(pol)>(p, b, problem, bumps, s); });
}

template <class S>
__global__ __launch_bounds__(256) void k_zero(const Grid g, S* u) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= g.isize || j >= g.jsize) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) u[at(g, i, j, v)] = S(0);
}

hipError_t launch_zero_state(const e2d_params& p, const DevBuffers& b, int buf, hipStream_t s) {
  const Grid g = make_grid(p, b);
  const dim3 grid((g.isize + 63) / 64, (g.jsize + 3) / 4);
  if (state_elem_bytes(p.precision) == 8)
    hipLaunchKernelGGL(k_zero<double>, grid, dim3(256), 0, s, g, static_cast<double*>(b.u[buf]));
  else
    hipLaunchKernelGGL(k_zero<float>, grid, dim3(256), 0, s, g, static_cast<float*>(b.u[buf]));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// deterministic FP64 sums over the interior: block k reduces rows k, k+B, ...
// in a fixed order, then one block adds the partials in block order.

template <class S>
__global__ __launch_bounds__(256) void k_sums_partial(const Grid g, const S* __restrict__ u, double* partial) {
  __shared__ double red[4][256];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = 2 + blockIdx.x; j <= g.ny + 1; j += gridDim.x)
    for (int i = 2 + threadIdx.x; i <= g.nx + 1; i += blockDim.x)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] += static_cast<double>(u[at(g, i, j, v)]);
#pragma unroll
  for (int v = 0; v < 4; ++v) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_sums_final(const double* partial, int nb, double* out) {
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int k = 0; k < nb; ++k) s += partial[k * 4 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

hipError_t launch_sums(const e2d_params& p, const DevBuffers& b, int buf, double* d_out4, hipStream_t s) {
  const Grid g = make_grid(p, b);
  if (state_elem_bytes(p.precision) == 8)
    hipLaunchKernelGGL(k_sums_partial<double>, dim3(kSumBlocks), dim3(256), 0, s, g,
                       static_cast<const double*>(b.u[buf]), b.partial);
  else
    hipLaunchKernelGGL(k_sums_partial<float>, dim3(kSumBlocks), dim3(256), 0, s, g,
                       static_cast<const float*>(b.u[buf]), b.partial);
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, s, b.partial, kSumBlocks, d_out4);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host layout <-> pitched device layout

template <class S>
__global__ __launch_bounds__(256) void k_pack(const Grid g, const double* __restrict__ src, S* dst) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= g.isize || j >= g.jsize) return;
  const int64_t hp = static_cast<int64_t>(g.isize) * g.jsize;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    dst[at(g, i, j, v)] = static_cast<S>(src[v * hp + static_cast<int64_t>(j) * g.isize + i]);
}

template <class S>
__global__ __launch_bounds__(256) void k_unpack(const Grid g, const S* __restrict__ src, double* dst) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= g.isize || j >= g.jsize) return;
  const int64_t hp = static_cast<int64_t>(g.isize) * g.jsize;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    dst[v * hp + static_cast<int64_t>(j) * g.isize + i] = static_cast<double>(src[at(g, i, j, v)]);
}

hipError_t launch_pack(const e2d_params& p, const DevBuffers& b, int buf, const double* d_src, hipStream_t s) {
  const Grid g = make_grid(p, b);
  const dim3 grid((g.isize + 63) / 64, (g.jsize + 3) / 4);
  if (state_elem_bytes(p.precision) == 8)
    hipLaunchKernelGGL(k_pack<double>, grid, dim3(256), 0, s, g, d_src, static_cast<double*>(b.u[buf]));
  else
    hipLaunchKernelGGL(k_pack<float>, grid, dim3(256), 0, s, g, d_src, static_cast<float*>(b.u[buf]));
  return hipGetLastError();
}

hipError_t launch_unpack(const e2d_params& p, const DevBuffers& b, const void* d_state, int elem_bytes,
                         double* d_dst, hipStream_t s) {
  const Grid g = make_grid(p, b);
  const dim3 grid((g.isize + 63) / 64, (g.jsize + 3) / 4);
  if (elem_bytes == 8)
    hipLaunchKernelGGL(k_unpack<double>, grid, dim3(256), 0, s, g, static_cast<const double*>(d_state), d_dst);
  else
    hipLaunchKernelGGL(k_unpack<float>, grid, dim3(256), 0, s, g, static_cast<const float*>(d_state), d_dst);
  return hipGetLastError();
}

hipError_t launch_pack_field(const e2d_params& p, const DevBuffers& b, const double* d_src, double* d_field,
                             hipStream_t s) {
  const Grid g = make_grid(p, b);
  hipLaunchKernelGGL(k_pack<double>, dim3((g.isize + 63) / 64, (g.jsize + 3) / 4), dim3(256), 0, s, g, d_src,
                     d_field);
  return hipGetLastError();
}

size_t state_elem_bytes(int precision) {
  return with_precision(precision, [](auto pol) { return sizeof(typename decltype(pol)::S); });
}

// q, slopes and face states in the flux arithmetic type, fx and fy in the update type
size_t stage_elem_bytes(int precision, int stage) {
  return with_precision(precision, [&](auto pol) {
    using P = decltype(pol);
    return (stage == E2D_STAGE_FX || stage == E2D_STAGE_FY) ? sizeof(typename P::U) : sizeof(typename P::T);
  });
}

// ---------------------------------------------------------------------------
// pointwise numerics (known-answer tests of the device routines)

template <class P>
__global__ __launch_bounds__(256) void k_pointwise(const Hydro<typename P::T> h, int func, int n,
                                                    const double* __restrict__ x, double* __restrict__ y) {
  using T = typename P::T;
  using M = typename P::M;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  T a[20], o[8], g[4];
  switch (func) {
    case E2D_FN_PRIM: {
      for (int m = 0; m < 4; ++m) a[m] = static_cast<T>(x[4 * k + m]);
      cons_to_prim<T, M>(h, a, o);
      const T c = sound_speed<T, M>(h, o);
      for (int m = 0; m < 4; ++m) y[5 * k + m] = o[m];
      y[5 * k + 4] = c;
    } break;
    case E2D_FN_SLOPE: {
      for (int m = 0; m < 20; ++m) a[m] = static_cast<T>(x[20 * k + m]);
      cell_slopes(h, a, a + 4, a + 8, a + 12, a + 16, o, o + 4);
      for (int m = 0; m < 8; ++m) y[8 * k + m] = o[m];
    } break;
    case E2D_FN_TRACE: {
      for (int m = 0; m < 14; ++m) a[m] = static_cast<T>(x[15 * k + m]);
      const int face = static_cast<int>(x[15 * k + 14]);
      T s[4];
      trace_sources<T, M>(h, a, a + 4, a + 8, s);
      const bool xdir = face < 2;
      face_state(h, a, xdir ? a + 4 : a + 8, s, xdir ? a[12] : a[13], (face & 1) == 1, o);
      for (int m = 0; m < 4; ++m) y[4 * k + m] = o[m];
    } break;
    case E2D_FN_RIEMANN:
    case E2D_FN_HLLC: {
      for (int m = 0; m < 8; ++m) a[m] = static_cast<T>(x[8 * k + m]);
      if (func == E2D_FN_RIEMANN)
        riemann_two_shock<T, M>(h, a, a + 4, g, o);
      else
        riemann_hllc<T, ExactMath>(h, a, a + 4, g, o);
      for (int m = 0; m < 4; ++m) y[8 * k + m] = g[m];
      for (int m = 0; m < 4; ++m) y[8 * k + 4 + m] = o[m];
    } break;
    case E2D_FN_CMPFLX: {
      for (int m = 0; m < 4; ++m) a[m] = static_cast<T>(x[4 * k + m]);
      godunov_flux(h, a, o);
      for (int m = 0; m < 4; ++m) y[4 * k + m] = o[m];
    } break;
    case E2D_FN_EOS: {
      const T rho = static_cast<T>(x[2 * k]), eint = static_cast<T>(x[2 * k + 1]);
      const T p = fmax2((h.gamma0 - T(1)) * rho * eint, rho * h.smallp);
      const T c = M::sqrt(M::div(h.gamma0 * p, rho));
      y[2 * k] = p;
      y[2 * k + 1] = c;
    } break;
    case E2D_FN_DIV:
      y[k] = M::div(static_cast<T>(x[2 * k]), static_cast<T>(x[2 * k + 1]));
      break;
    case E2D_FN_SQRT:
      y[k] = M::sqrt(static_cast<T>(x[k]));
      break;
    case E2D_FN_DIV_ROOT: {
      T inv;
      const T s = M::sqrt_pos_inv(static_cast<T>(x[2 * k + 1]), inv);
      y[2 * k] = M::div_inv(static_cast<T>(x[2 * k]), s, inv);
      y[2 * k + 1] = s;
    } break;
    case E2D_FN_DIV_NEAR: {
      const T b = static_cast<T>(x[3 * k + 1]);
      y[k] = M::div_inv(static_cast<T>(x[3 * k]), b, M::recip_near(b, static_cast<T>(x[3 * k + 2])));
    } break;
    case E2D_FN_CONV_GT:
      y[k] = M::quot_gt(static_cast<T>(x[2 * k]), static_cast<T>(x[2 * k + 1]), h.conv_tol, h.conv_half_ulp) ? 1.0 : 0.0;
      break;
    default: break;
  }
}

hipError_t launch_pointwise(const e2d_params& p, int func, int n, const double* d_x, double* d_y, hipStream_t s) {
  const dim3 grid((n + 255) / 256);
  with_precision(p.precision, [&](auto pol) {
    using P = decltype(pol);
    hipLaunchKernelGGL(k_pointwise<P>, grid, dim3(256), 0, s, make_hydro<typename P::T>(p), func, n, d_x, d_y);
    return 0;
  });
  return hipGetLastError();
}

}  // namespace e2d
