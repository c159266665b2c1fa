// Adjoint of one Godunov step (frozen dt): lam_out = J^T lam_in, where J is
// the Jacobian of the interior update u' = u + (fy(j)-fy(j+1)) + (fx(i)-fx(i+1))
// (serial_ref/m_run.f90:504-513) with respect to the interior of u; ghost cells
// are functions of the interior (make_boundaries), corners are constants.
//
// Five deterministic passes (every output element is written by one thread,
// which gathers its contributions in a fixed order):
//   1. k_adj_prim   q = prim(u) on all cells (input buffer boundary-filled first)
//   2. k_adj_faces  per face: flux bar from lam, Riemann adjoint -> bars of the
//                   two face states (stored in the XMIN/XMAX/YMIN/YMAX slot of
//                   the cell that owns each state)
//   3. k_adj_cells  per cell: face-state and trace-source adjoints -> bars of
//                   q, dq_x, dq_y of the cell
//   4. k_adj_gather per cell: limiter adjoints of the cell and its 4 neighbours,
//                   then the primitive-variable adjoint -> bar of u; interior
//                   cells write lam_in + that bar, ghost cells keep their bar
//   5. k_adj_fold_frame  per cell of the 2-cell interior frame: + the bars of
//                   the ghost cells whose source it is
//
// The reverse passes run in the arithmetic type T and math policy of the
// forward step's precision (double for fp64; float for fp32 and the mixed
// variant) and recompute the forward values exactly as the step does, so every
// branch decision matches it.  lam_in and lam_out are double for every
// precision: the face flux bars are formed in double and rounded to T, and
// each step's contribution is added to lam in double.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "e2d_adjoint.hpp"
#include "e2d_common.hpp"
#include "e2d_internal.hpp"

namespace e2d {

namespace {

template <class P>
struct AdjArgs {
  using S = typename P::S;
  using T = typename P::T;
  using U = typename P::U;
  Hydro<T> h;
  Hydro<U> hU;
  Grid g;
  const S* u;
  const double* lam;
  double* out;
  T *q, *xminb, *xmaxb, *yminb, *ymaxb, *qb, *dqxb, *dqyb, *ub;
  T* seam;  // k_adj_march_gather: partial sums of the rows where segments meet (16 planes)
  double dt;
  int nsegs, seg_rows, seg_groups;     // k_adj_march decomposition
  int gnsegs, gseg_rows, gseg_groups;  // k_adj_gather_march decomposition
  // k_adj_march_gather / k_adj_seam: per strip, nsegs long segments of
  // seg_rows completed rows from row 1, then nshort short ones of short_rows
  // rows from short_base to jsize - 1; dealt to the XCDs as dealt_item does
  // (nlong = strips * nsegs long items, nitems in all, chunks of `chunk`)
  int nshort, short_rows, short_base, nlong, nitems, chunk;
};

// completed rows [ra, rb) of segment sg of a strip (long ones first)
template <class P>
__device__ __forceinline__ void adj_segment_rows(const AdjArgs<P>& a, int sg, int& ra, int& rb) {
  if (sg < a.nsegs) {
    ra = 1 + sg * a.seg_rows;
    rb = min(ra + a.seg_rows, a.short_base);
  } else {
    ra = a.short_base + (sg - a.nsegs) * a.short_rows;
    rb = min(ra + a.short_rows, a.g.jsize - 1);
  }
}

template <class X, class T>
__device__ __forceinline__ void load4(const X* a, const Grid& g, int i, int j, T x[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) x[v] = static_cast<T>(a[at(g, i, j, v)]);
}

template <class T>
__device__ __forceinline__ void store4(T* a, const Grid& g, int i, int j, const T x[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) a[at(g, i, j, v)] = x[v];
}

// primitives of cell (i, j) from the (boundary-filled) state, exactly as
// k_adj_prim stores them
template <class P>
__device__ __forceinline__ void prim_at(const AdjArgs<P>& a, int i, int j, typename P::T q[4]) {
  typename P::T u[4];
  load4(a.u, a.g, i, j, u);
  cons_to_prim<typename P::T, typename P::M>(a.h, u, q);
}

// dt/dx and dt/dy as the forward step forms them (in the update type U)
template <class P>
__device__ __forceinline__ void step_ratios(const AdjArgs<P>& a, typename P::U& dtdx, typename P::U& dtdy) {
  using U = typename P::U;
  const U dt = static_cast<U>(a.dt);
  dtdx = dt / a.hU.dx;
  dtdy = dt / a.hU.dy;
}

// slopes and trace sources of cell (i, j) from q of the cell and its neighbours
template <class P>
__device__ __forceinline__ void cell_trace(const AdjArgs<P>& a, int i, int j, typename P::T q[4],
                                           typename P::T dqx[4], typename P::T dqy[4], typename P::T s[4]) {
  typename P::T xp[4], xm[4], yp[4], ym[4];
  load4(a.q, a.g, i, j, q);
  load4(a.q, a.g, i + 1, j, xp);
  load4(a.q, a.g, i - 1, j, xm);
  load4(a.q, a.g, i, j + 1, yp);
  load4(a.q, a.g, i, j - 1, ym);
  cell_slopes(a.h, q, xp, xm, yp, ym, dqx, dqy);
  trace_sources<typename P::T, typename P::M>(a.h, q, dqx, dqy, s);
}

template <class P>
__global__ __launch_bounds__(256) void k_adj_prim(const AdjArgs<P> a) {
  using T = typename P::T;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= a.g.isize || j >= a.g.jsize) return;
  T u[4], q[4];
  load4(a.u, a.g, i, j, u);
  cons_to_prim<T, typename P::M>(a.h, u, q);
  store4(a.q, a.g, i, j, q);
}

__device__ __forceinline__ bool interior(const Grid& g, int i, int j) {
  return i >= 2 && i <= g.nx + 1 && j >= 2 && j <= g.ny + 1;
}

// faces: (i, j) in [2, nx+2] x [2, ny+2]; the x-face left of cell (i, j) for
// interior rows, the y-face below cell (i, j) for interior columns
template <class P>
__global__ __launch_bounds__(256) void k_adj_faces(const AdjArgs<P> a) {
  using T = typename P::T;
  using U = typename P::U;
  using M = typename P::M;
  const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const Grid& g = a.g;
  if (i > g.nx + 2 || j > g.ny + 2) return;
  U dtdxU, dtdyU;
  step_ratios(a, dtdxU, dtdyU);
  const T dtdx = static_cast<T>(dtdxU), dtdy = static_cast<T>(dtdyU);
  if (j <= g.ny + 1) {  // x-face between (i-1, j) and (i, j)
    T fb[4];
    bool any = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double x = 0.0;
      if (interior(g, i, j)) x += a.lam[at(g, i, j, v)];
      if (interior(g, i - 1, j)) x -= a.lam[at(g, i - 1, j, v)];
      fb[v] = static_cast<T>(x * static_cast<double>(dtdxU));
      any |= (x != 0.0);
    }
    T lb[4] = {T(0), T(0), T(0), T(0)}, rb[4] = {T(0), T(0), T(0), T(0)};
    if (any) {
      T qL[4], dxL[4], dyL[4], sL[4], qR[4], dxR[4], dyR[4], sR[4], fl[4], fr[4];
      cell_trace(a, i - 1, j, qL, dxL, dyL, sL);
      cell_trace(a, i, j, qR, dxR, dyR, sR);
      face_state(a.h, qL, dxL, sL, dtdx, true, fl);
      face_state(a.h, qR, dxR, sR, dtdx, false, fr);
      riemann_two_shock_adj<M>(a.h, fl, fr, fb, lb, rb);
    }
    store4(a.xmaxb, g, i - 1, j, lb);
    store4(a.xminb, g, i, j, rb);
  }
  if (i <= g.nx + 1) {  // y-face between (i, j-1) and (i, j)
    T fb[4];
    bool any = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double x = 0.0;
      if (interior(g, i, j)) x += a.lam[at(g, i, j, v)];
      if (interior(g, i, j - 1)) x -= a.lam[at(g, i, j - 1, v)];
      fb[v] = static_cast<T>(x * static_cast<double>(dtdyU));
      any |= (x != 0.0);
    }
    T lb[4] = {T(0), T(0), T(0), T(0)}, rb[4] = {T(0), T(0), T(0), T(0)};
    if (any) {
      T qL[4], dxL[4], dyL[4], sL[4], qR[4], dxR[4], dyR[4], sR[4], fl[4], fr[4];
      cell_trace(a, i, j - 1, qL, dxL, dyL, sL);
      cell_trace(a, i, j, qR, dxR, dyR, sR);
      face_state(a.h, qL, dyL, sL, dtdy, true, fl);
      face_state(a.h, qR, dyR, sR, dtdy, false, fr);
      riemann_y_adj<M>(a.h, fl, fr, fb, lb, rb);
    }
    store4(a.ymaxb, g, i, j - 1, lb);
    store4(a.yminb, g, i, j, rb);
  }
}

// cells [1, size-2]^2: face-state and trace-source adjoints
template <class P>
__global__ __launch_bounds__(256) void k_adj_cells(const AdjArgs<P> a) {
  using T = typename P::T;
  using U = typename P::U;
  const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const Grid& g = a.g;
  if (i > g.isize - 2 || j > g.jsize - 2) return;
  U dtdxU, dtdyU;
  step_ratios(a, dtdxU, dtdyU);
  const T dtdx = static_cast<T>(dtdxU), dtdy = static_cast<T>(dtdyU);
  T bxm[4], bxp[4], bym[4], byp[4];
  load4(a.xminb, g, i, j, bxm);
  load4(a.xmaxb, g, i, j, bxp);
  load4(a.yminb, g, i, j, bym);
  load4(a.ymaxb, g, i, j, byp);
  T qb[4] = {T(0), T(0), T(0), T(0)}, dxb[4] = {T(0), T(0), T(0), T(0)}, dyb[4] = {T(0), T(0), T(0), T(0)};
  bool any = false;
#pragma unroll
  for (int v = 0; v < 4; ++v) any |= (bxm[v] != T(0)) | (bxp[v] != T(0)) | (bym[v] != T(0)) | (byp[v] != T(0));
  if (any) {
    T q[4], dqx[4], dqy[4], s[4], sb[4] = {T(0), T(0), T(0), T(0)};
    cell_trace(a, i, j, q, dqx, dqy, s);
    face_state_adj(a.h, q, dqx, s, dtdx, false, bxm, qb, dxb, sb);
    face_state_adj(a.h, q, dqx, s, dtdx, true, bxp, qb, dxb, sb);
    face_state_adj(a.h, q, dqy, s, dtdy, false, bym, qb, dyb, sb);
    face_state_adj(a.h, q, dqy, s, dtdy, true, byp, qb, dyb, sb);
    trace_sources_adj<typename P::M>(a.h, q, dqx, dqy, sb, qb, dxb, dyb);
  }
  store4(a.qb, g, i, j, qb);
  store4(a.dqxb, g, i, j, dxb);
  store4(a.dqyb, g, i, j, dyb);
}

// Passes 2+3 in one marching kernel (the default): each wavefront owns a strip
// of 64 columns (60 output columns, 2 halo columns each side) and marches up
// a segment of rows, as the forward step does.  At row r it forms the slopes,
// trace sources and face states of row r once, solves the Riemann adjoint of
// the y-face r-1/2, completes cell row r-1 (whose four face-state bars are
// then known: x bars from row r-1, the lower y bar from the face r-3/2, the
// upper one just computed), then the Riemann adjoint of the x-faces of row r;
// the bar of the cell's right face comes from lane+1 by a shuffle.  Faces,
// zero-bar skips and the completion order are those of k_adj_faces and
// k_adj_cells, so qb, dqxb and dqyb are bit-identical to theirs.  Outputs are
// the cells [1, size-2]^2.
// At least 2 resident wavefronts per SIMD: left to itself the compiler gives
// the FP64 kernel more than 256 VGPRs (1 wavefront per SIMD); the cap costs a
// 12-byte spill and is 1.36x faster (2048^2, dense lam).
#ifndef E2D_ADJ_MARCH_MIN_WAVES
#define E2D_ADJ_MARCH_MIN_WAVES 2
#endif
#if E2D_ADJ_MARCH_MIN_WAVES > 0
#define E2D_ADJ_MARCH_ATTR __attribute__((amdgpu_waves_per_eu(E2D_ADJ_MARCH_MIN_WAVES)))
#else
#define E2D_ADJ_MARCH_ATTR
#endif
#ifndef E2D_BAND_MARCH
#define E2D_BAND_MARCH 1
#endif
#ifndef E2D_BAND_GATHER
#define E2D_BAND_GATHER 1
#endif
#ifndef E2D_ADJ_WAVES_PER_BLOCK
#define E2D_ADJ_WAVES_PER_BLOCK 4
#endif
constexpr int kAdjWaves = E2D_ADJ_WAVES_PER_BLOCK;  // wavefronts per workgroup
constexpr int kAdjStripOut = 60;  // output columns per wavefront

template <class P>
__global__ __launch_bounds__(64 * kAdjWaves) E2D_ADJ_MARCH_ATTR void k_adj_march(const AdjArgs<P> a) {
  using T = typename P::T;
  using U = typename P::U;
  using M = typename P::M;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  int strip, group;
  strip_and_group<E2D_BAND_MARCH>(lb, a.seg_groups, strip, group);
  const int seg = group * kAdjWaves + wave;
  if (seg >= a.nsegs) return;  // whole wavefront leaves: no barrier in this kernel

  const Grid& g = a.g;
  const int ci = 1 + strip * kAdjStripOut - 2 + lane;  // column of this lane
  const int cl = ci < 0 ? 0 : (ci > g.isize - 1 ? g.isize - 1 : ci);
  const bool out_col = lane >= 2 && lane < 62 && ci <= g.isize - 2;
  const bool xface = lane >= 2 && lane <= 62 && ci >= 2 && ci <= g.nx + 2;  // face ci-1/2
  const bool yface = lane >= 2 && lane < 62 && ci >= 2 && ci <= g.nx + 1;
  const int ra = 1 + seg * a.seg_rows;
  const int rb = min(ra + a.seg_rows, g.jsize - 1);  // output rows [ra, rb)

  U dtdxU, dtdyU;
  step_ratios(a, dtdxU, dtdyU);
  const T dtdx = static_cast<T>(dtdxU), dtdy = static_cast<T>(dtdyU);
  const T zero4[4] = {T(0), T(0), T(0), T(0)};
  auto row = [&](int r) { return r < 0 ? 0 : (r > g.jsize - 1 ? g.jsize - 1 : r); };

  T qa[4], qb[4], qc[4];
  prim_at(a, cl, row(ra - 2), qa);
  prim_at(a, cl, row(ra - 1), qb);
  // row r-1, carried: slopes, trace sources and the bars of its xmin, xmax, ymin face states
  // (its primitives are qa)
  T pdx[4], pdy[4], ps[4], pbxm[4], pbxp[4], pbym[4];
  // lam of this lane's cell in rows r and r-1, zero outside the interior: one
  // load per row; the x-faces take the left column's by a lane shift
  double lam_r[4], lam_p[4] = {0.0, 0.0, 0.0, 0.0};
  T un[4];  // state of row r+1, loaded an iteration ahead
  load4(a.u, g, cl, row(ra), un);

  for (int r = ra - 1; r <= rb; ++r) {
    {
      T uc[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) uc[v] = un[v];
      load4(a.u, g, cl, row(r + 2), un);
      cons_to_prim<T, M>(a.h, uc, qc);
    }
    {
      const bool in = interior(g, ci, r);
#pragma unroll
      for (int v = 0; v < 4; ++v) lam_r[v] = in ? a.lam[at(g, ci, r, v)] : 0.0;
    }
    T qxm[4], qxp[4], dqx[4], dqy[4], s[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qxm[v] = from_left<true>(qb[v]);
      qxp[v] = from_right<true>(qb[v]);
    }
    cell_slopes(a.h, qb, qxp, qxm, qc, qa, dqx, dqy);
    trace_sources<T, M>(a.h, qb, dqx, dqy, s);

    T bym[4], bxm[4], bxp[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) bym[v] = bxm[v] = bxp[v] = T(0);
    if (r >= ra) {
      // y-face between rows r-1 and r
      T ylb[4] = {T(0), T(0), T(0), T(0)};
      if (yface && r >= 2 && r <= g.ny + 2) {
        T fb[4];
        bool any = false;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // 0 + lam and x - 0 give the bits of leaving out the zero terms
          const double x = (0.0 + lam_r[v]) - lam_p[v];
          fb[v] = static_cast<T>(x * static_cast<double>(dtdyU));
          any |= (x != 0.0);
        }
        if (any) {
          T fl[4], fr[4];
          face_state(a.h, qa, pdy, ps, dtdy, true, fl);
          face_state(a.h, qb, dqy, s, dtdy, false, fr);
          riemann_y_adj<M>(a.h, fl, fr, fb, ylb, bym);
        }
      }
      // cell row r-1 is complete
      if (r - 1 >= ra && out_col) {
        T qbar[4] = {T(0), T(0), T(0), T(0)}, dxb[4] = {T(0), T(0), T(0), T(0)}, dyb[4] = {T(0), T(0), T(0), T(0)};
        bool any = false;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          any |= (pbxm[v] != T(0)) | (pbxp[v] != T(0)) | (pbym[v] != T(0)) | (ylb[v] != T(0));
        if (any) {
          T sb[4] = {T(0), T(0), T(0), T(0)};
          face_state_adj(a.h, qa, pdx, ps, dtdx, false, pbxm, qbar, dxb, sb);
          face_state_adj(a.h, qa, pdx, ps, dtdx, true, pbxp, qbar, dxb, sb);
          face_state_adj(a.h, qa, pdy, ps, dtdy, false, pbym, qbar, dyb, sb);
          face_state_adj(a.h, qa, pdy, ps, dtdy, true, ylb, qbar, dyb, sb);
          trace_sources_adj<M>(a.h, qa, pdx, pdy, sb, qbar, dxb, dyb);
        }
        store4(a.qb, g, ci, r - 1, qbar);
        store4(a.dqxb, g, ci, r - 1, dxb);
        store4(a.dqyb, g, ci, r - 1, dyb);
      }
      // x-faces of row r
      if (r < rb) {
        T xmax_own[4], xl[4], xr[4], xlb[4] = {T(0), T(0), T(0), T(0)};
        face_state(a.h, qb, dqx, s, dtdx, true, xmax_own);
        face_state(a.h, qb, dqx, s, dtdx, false, xr);
#pragma unroll
        for (int v = 0; v < 4; ++v) xl[v] = from_left<true>(xmax_own[v]);
        double lam_l[4];  // lam_r of column ci-1
#pragma unroll
        for (int v = 0; v < 4; ++v) lam_l[v] = from_left<true>(lam_r[v]);
        if (xface && r >= 2 && r <= g.ny + 1) {
          T fb[4];
          bool any = false;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double x = (0.0 + lam_r[v]) - lam_l[v];
            fb[v] = static_cast<T>(x * static_cast<double>(dtdxU));
            any |= (x != 0.0);
          }
          if (any) riemann_two_shock_adj<M>(a.h, xl, xr, fb, xlb, bxm);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) bxp[v] = from_right<true>(xlb[v]);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      pdx[v] = dqx[v];
      pdy[v] = dqy[v];
      ps[v] = s[v];
      pbxm[v] = bxm[v];
      pbxp[v] = bxp[v];
      pbym[v] = bym[v];
      qa[v] = qb[v];
      qb[v] = qc[v];
      lam_p[v] = lam_r[v];
    }
  }
  (void)zero4;
}

// Pass 4 as a marching kernel (with k_adj_march): wavefronts own strips of 64
// columns (60 output columns) and march up segments of rows.  Each cell's x-
// and y-limiter adjoints are formed once, from primitives computed from u in
// rolling rows; a cell collects its neighbours' parts by shuffles (x) and from
// the rows before and after (y), in k_adj_gather's order and with its
// zero-bar skips, then applies the primitive adjoint.  Outputs are all cells.
constexpr int kGatherStripOut = 60;

template <class P>
__global__ __launch_bounds__(64 * kAdjWaves) void k_adj_gather_march(const AdjArgs<P> a) {
  using T = typename P::T;
  using M = typename P::M;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  int strip, group;
  strip_and_group<E2D_BAND_GATHER>(lb, a.gseg_groups, strip, group);
  const int seg = group * kAdjWaves + wave;
  if (seg >= a.gnsegs) return;  // whole wavefront leaves: no barrier in this kernel

  const Grid& g = a.g;
  const int ci = strip * kGatherStripOut + lane - 2;  // column of this lane
  const int cl = ci < 0 ? 0 : (ci > g.isize - 1 ? g.isize - 1 : ci);
  const bool out_col = lane >= 2 && lane < 62 && ci < g.isize;
  const bool slope_col = ci >= 1 && ci <= g.isize - 2;  // the lane's cell has slopes
  const int ra = seg * a.gseg_rows;
  const int rb = min(ra + a.gseg_rows, g.jsize);  // output rows [ra, rb)
  const bool limited = slopes_limited(a.h);
  const T st = static_cast<T>(a.h.slope_type);
  auto row = [&](int r) { return r < 0 ? 0 : (r > g.jsize - 1 ? g.jsize - 1 : r); };

  T qm[4], q0[4], qp[4];  // primitives of rows r-1, r, r+1
  prim_at(a, cl, row(ra - 2), qm);
  prim_at(a, cl, row(ra - 1), q0);
  // y-limiter parts of earlier rows: plus-side part of rows r-2 and r-1, own part of row r-1
  T yp2[4], yp1[4], y01[4];
  bool vp2[4], vp1[4], v01[4];
  T qbx1[4];  // bar of q of row r-1 with its own and x-limiter parts
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    yp2[v] = yp1[v] = y01[v] = qbx1[v] = T(0);
    vp2[v] = vp1[v] = v01[v] = false;
  }

  for (int r = ra - 1; r <= rb; ++r) {
    prim_at(a, cl, row(r + 1), qp);
    const bool cell_has = slope_col && r >= 1 && r <= g.jsize - 2;
    // y-limiter adjoint of cell (ci, r)
    T ym[4], y0[4], yp[4];
    bool yv[4];
    T dyb[4] = {T(0), T(0), T(0), T(0)}, dxb[4] = {T(0), T(0), T(0), T(0)};
    if (cell_has) {
      load4(a.dqyb, g, ci, r, dyb);
      load4(a.dqxb, g, ci, r, dxb);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      ym[v] = y0[v] = yp[v] = T(0);
      yv[v] = limited && cell_has && dyb[v] != T(0);
      if (yv[v]) limited_slope_adj(st, qm[v], q0[v], qp[v], dyb[v], ym[v], y0[v], yp[v]);
    }
    // x-limiter adjoint of cell (ci, r) and the x parts of its bar of q
    T qbx[4] = {T(0), T(0), T(0), T(0)};
    if (r >= ra && r < rb) {
      T xm[4], x0[4], xp[4];
      bool xv[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        xm[v] = x0[v] = xp[v] = T(0);
        xv[v] = limited && cell_has && dxb[v] != T(0);
        const T ql = from_left<true>(q0[v]), qr = from_right<true>(q0[v]);
        if (xv[v]) limited_slope_adj(st, ql, q0[v], qr, dxb[v], xm[v], x0[v], xp[v]);
      }
      if (cell_has) load4(a.qb, g, ci, r, qbx);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T lp = from_left<true>(xp[v]), rm = from_right<true>(xm[v]);
        const bool lv = from_left<true>(static_cast<int>(xv[v])) != 0;
        const bool rv = from_right<true>(static_cast<int>(xv[v])) != 0;
        if (limited) {
          if (lv) qbx[v] += lp;  // left neighbour: this cell is its plus side
          if (xv[v]) qbx[v] += x0[v];
          if (rv) qbx[v] += rm;  // right neighbour: this cell is its minus side
        }
      }
    }
    // cell row r-1 is complete
    if (r - 1 >= ra && r - 1 < rb && out_col) {
      const int jr = r - 1;
      T qb[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        qb[v] = qbx1[v];
        if (vp2[v]) qb[v] += yp2[v];
        if (v01[v]) qb[v] += y01[v];
        if (yv[v]) qb[v] += ym[v];
      }
      T u[4], ub[4] = {T(0), T(0), T(0), T(0)};
      load4(a.u, g, ci, jr, u);
      cons_to_prim_adj<M>(a.h, u, qb, ub);
      if (interior(g, ci, jr)) {
#pragma unroll
        for (int v = 0; v < 4; ++v) a.out[at(g, ci, jr, v)] = a.lam[at(g, ci, jr, v)] + static_cast<double>(ub[v]);
      } else {
        store4(a.ub, g, ci, jr, ub);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      yp2[v] = yp1[v];
      vp2[v] = vp1[v];
      yp1[v] = yp[v];
      vp1[v] = yv[v];
      y01[v] = y0[v];
      v01[v] = yv[v];
      qbx1[v] = qbx[v];
      qm[v] = q0[v];
      q0[v] = qp[v];
    }
  }
}

// Passes 2-4 in one marching kernel (the default): k_adj_march with the
// gather folded in.  A wavefront owns a strip of 64 columns and marches up a
// segment of rows [ra, rb).  Completions (face-state and trace-source
// adjoints, as in k_adj_march) are valid on lanes 2..61; the gather of a cell
// needs the completions of its x-neighbours, so gathered columns are lanes
// 3..60 (58 per strip).  When row k is complete, its x- and y-limiter
// adjoints are formed at once (q of rows k-1, k, k+1 are in registers); the
// bar of q of a cell is
//   ((own bar + x-limiter parts) + plus part from row k-1) + own y part) + minus part from row k+1
// in k_adj_gather's order, so row k-1 is final when row k is complete.  Rows
// ra and rb-1 miss a part that the segment below / above forms: their
// partial sums go to the seam buffer and k_adj_seam finishes them (and rows 0
// and jsize-1).  A part that k_adj_gather skips is carried as -0.0, the
// identity of IEEE addition, so the sums are bit-identical to the unfused
// passes.
// E2D_ADJ_LDS_CARRY=1 (default): the fused kernel's row-carried gather values
// live in per-thread LDS slots (12 values, 24 KB per FP64 workgroup); in
// registers the FP64 kernel spills 132 B and is 7% slower (835 vs 778 us).
#ifndef E2D_ADJ_LDS_CARRY
#define E2D_ADJ_LDS_CARRY 1
#endif
constexpr int kFusedStripOut = 58;
// Seam buffer: 4 slots of 4 values per completed row and column, indexed by
// the row (so that the store addresses vary with the marching loop):
//   row ra:   0 own + x-limiter parts; 1 own y part; 2 minus part (for row ra-1)
//   row ra+1: 0 minus part (for row ra)
//   row rb-1: 1 sum up to the own y part (when rb-1 > ra); 3 plus part (for row rb)
__device__ __forceinline__ int64_t seam_at(const Grid& g, int row, int slot, int v, int ci) {
  return (static_cast<int64_t>(row * 4 + slot) * 4 + v) * g.pitch + ci;
}

// bar of u of cell (ci, row) from its final bar of q: lam_in + bar for
// interior cells, the bar alone for ghost cells (folded by k_adj_fold_frame)
template <class P>
__device__ __forceinline__ void finish_cell(const AdjArgs<P>& a, int ci, int row, const typename P::T qb[4]) {
  using T = typename P::T;
  const Grid& g = a.g;
  T u[4], ub[4] = {T(0), T(0), T(0), T(0)};
  load4(a.u, g, ci, row, u);
  cons_to_prim_adj<typename P::M>(a.h, u, qb, ub);
  if (interior(g, ci, row)) {
#pragma unroll
    for (int v = 0; v < 4; ++v) a.out[at(g, ci, row, v)] = a.lam[at(g, ci, row, v)] + static_cast<double>(ub[v]);
  } else {
    store4(a.ub, g, ci, row, ub);
  }
}

// finish_cell with lam of the cell already in registers (zero for ghost cells,
// where it is not used)
template <class P>
__device__ __forceinline__ void finish_cell_lam(const AdjArgs<P>& a, int ci, int row, const typename P::T qb[4],
                                                const double lam[4]) {
  using T = typename P::T;
  const Grid& g = a.g;
  T u[4], ub[4] = {T(0), T(0), T(0), T(0)};
  load4(a.u, g, ci, row, u);
  cons_to_prim_adj<typename P::M>(a.h, u, qb, ub);
  if (interior(g, ci, row)) {
#pragma unroll
    for (int v = 0; v < 4; ++v) a.out[at(g, ci, row, v)] = lam[v] + static_cast<double>(ub[v]);
  } else {
    store4(a.ub, g, ci, row, ub);
  }
}

template <class P>
__device__ __forceinline__ void adj_march_gather_body(const AdjArgs<P>& a) {
  using T = typename P::T;
  using U = typename P::U;
  using M = typename P::M;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  int item;
  bool is_long;
  if (!dealt_item(blockIdx.x, wave, kAdjWaves, a.nlong, a.nitems, a.chunk, item, is_long))
    return;  // whole wavefront leaves: no barrier in this kernel
  int strip, sg;
  if (is_long) {
    strip = item / a.nsegs;
    sg = item - strip * a.nsegs;
  } else {
    strip = item / a.nshort;
    sg = a.nsegs + (item - strip * a.nshort);
  }

  const Grid& g = a.g;
  const int ci = strip * kFusedStripOut + lane - 3;  // column of this lane
  const int cl = ci < 0 ? 0 : (ci > g.isize - 1 ? g.isize - 1 : ci);
  const bool cell_col = ci >= 1 && ci <= g.isize - 2;          // the column has completions
  const bool done_col = lane >= 2 && lane < 62 && cell_col;    // completed on this lane
  const bool xface = lane >= 2 && lane <= 62 && ci >= 2 && ci <= g.nx + 2;  // face ci-1/2
  const bool yface = lane >= 2 && lane < 62 && ci >= 2 && ci <= g.nx + 1;
  const bool gat_col = lane >= 3 && lane <= 60 && ci >= 0 && ci < g.isize;  // gathered on this lane
  int ra, rb;  // completed rows [ra, rb)
  adj_segment_rows(a, sg, ra, rb);
  const bool limited = slopes_limited(a.h);
  const T st = static_cast<T>(a.h.slope_type);
  const T skip = -T(0);

  U dtdxU, dtdyU;
  step_ratios(a, dtdxU, dtdyU);
  const T dtdx = static_cast<T>(dtdxU), dtdy = static_cast<T>(dtdyU);
  auto row = [&](int r) { return r < 0 ? 0 : (r > g.jsize - 1 ? g.jsize - 1 : r); };

  T qa[4], qb[4], qc[4];  // primitives of rows r-1, r, r+1
  prim_at(a, cl, row(ra - 2), qa);
  prim_at(a, cl, row(ra - 1), qb);
  // row r-1, carried: slopes, trace sources and the bars of its xmin, xmax, ymin face states
  T pdx[4], pdy[4], ps[4], pbxm[4], pbxp[4], pbym[4];
  // gather, carried: partial sum of the last completed row, its plus part, and
  // the primitives of row r-2 (in registers, or with E2D_ADJ_LDS_CARRY in
  // per-thread LDS slots, out of the registers during the Riemann adjoints)
#if E2D_ADJ_LDS_CARRY
  constexpr int kThreads = 64 * kAdjWaves;
  __shared__ __align__(16) unsigned char carry_raw[12 * sizeof(T) * kThreads];
  T* const carry = reinterpret_cast<T*>(carry_raw) + threadIdx.x;
  auto acc = [&](int v) -> T& { return carry[v * kThreads]; };
  auto ypf = [&](int v) -> T& { return carry[(4 + v) * kThreads]; };
  auto qz = [&](int v) -> T& { return carry[(8 + v) * kThreads]; };
#else
  T acc_r[4], ypf_r[4], qz_r[4];
  auto acc = [&](int v) -> T& { return acc_r[v]; };
  auto ypf = [&](int v) -> T& { return ypf_r[v]; };
  auto qz = [&](int v) -> T& { return qz_r[v]; };
#endif
#pragma unroll
  for (int v = 0; v < 4; ++v) qz(v) = acc(v) = ypf(v) = T(0);
  // lam of this lane's cell in rows r and r-1, zero outside the interior: one
  // load per row; the x-faces take the left column's by a lane shift
  double lam_r[4], lam_p[4] = {0.0, 0.0, 0.0, 0.0};
  T un[4];  // state of row r+1, loaded an iteration ahead
  load4(a.u, g, cl, row(ra), un);
  // lam of row r-2 for finish_cell, carried where it keeps the kernel within
  // 168 VGPRs (3 waves per SIMD): the mixed instantiation, 167 VGPRs.  FP32
  // would take 173 (2 waves) and FP64 more scratch spills, both slower.
  constexpr bool kLamPP = std::is_same<M, FastMath>::value;
  double lam_pp[4] = {0.0, 0.0, 0.0, 0.0};

  for (int r = ra - 1; r <= rb; ++r) {
    {
      T uc[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) uc[v] = un[v];
      load4(a.u, g, cl, row(r + 2), un);
      cons_to_prim<T, M>(a.h, uc, qc);
    }
    {
      const bool in = interior(g, ci, r);
#pragma unroll
      for (int v = 0; v < 4; ++v) lam_r[v] = in ? a.lam[at(g, ci, r, v)] : 0.0;
    }
    T qxm[4], qxp[4], dqx[4], dqy[4], s[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qxm[v] = from_left<true>(qb[v]);
      qxp[v] = from_right<true>(qb[v]);
    }
    cell_slopes(a.h, qb, qxp, qxm, qc, qa, dqx, dqy);
    trace_sources<T, M>(a.h, qb, dqx, dqy, s);

    T bym[4], bxm[4], bxp[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) bym[v] = bxm[v] = bxp[v] = T(0);
    if (r >= ra) {
      // y-face between rows r-1 and r
      T ylb[4] = {T(0), T(0), T(0), T(0)};
      if (yface && r >= 2 && r <= g.ny + 2) {
        T fb[4];
        bool any = false;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // 0 + lam and x - 0 give the bits of leaving out the zero terms
          const double x = (0.0 + lam_r[v]) - lam_p[v];
          fb[v] = static_cast<T>(x * static_cast<double>(dtdyU));
          any |= (x != 0.0);
        }
        if (any) {
          T fl[4], fr[4];
          face_state(a.h, qa, pdy, ps, dtdy, true, fl);
          face_state(a.h, qb, dqy, s, dtdy, false, fr);
          riemann_y_adj<M>(a.h, fl, fr, fb, ylb, bym);
        }
      }
      if (r - 1 >= ra) {
        // cell row k = r-1 is complete: bars of its q, x-slopes and y-slopes
        const int k = r - 1;
        T qbar[4] = {T(0), T(0), T(0), T(0)}, dxb[4] = {T(0), T(0), T(0), T(0)}, dyb[4] = {T(0), T(0), T(0), T(0)};
        if (done_col) {
          bool any = false;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            any |= (pbxm[v] != T(0)) | (pbxp[v] != T(0)) | (pbym[v] != T(0)) | (ylb[v] != T(0));
          if (any) {
            T sb[4] = {T(0), T(0), T(0), T(0)};
            face_state_adj(a.h, qa, pdx, ps, dtdx, false, pbxm, qbar, dxb, sb);
            face_state_adj(a.h, qa, pdx, ps, dtdx, true, pbxp, qbar, dxb, sb);
            face_state_adj(a.h, qa, pdy, ps, dtdy, false, pbym, qbar, dyb, sb);
            face_state_adj(a.h, qa, pdy, ps, dtdy, true, ylb, qbar, dyb, sb);
            trace_sources_adj<M>(a.h, qa, pdx, pdy, sb, qbar, dxb, dyb);
          }
        }
        // limiter adjoints of row k (q of rows k-1, k, k+1 = qz, qa, qb); the
        // own bar and the x parts of the cell and its x-neighbours, in k_adj_gather's order
        T qbx[4], y0f[4], ymf[4], ypk[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const bool xv = limited && done_col && dxb[v] != T(0);
          const T ql = from_left<true>(qa[v]), qr = from_right<true>(qa[v]);  // all lanes active here
          T xm = T(0), x0 = T(0), xp = T(0);
          if (xv) limited_slope_adj(st, ql, qa[v], qr, dxb[v], xm, x0, xp);
          const T lp = from_left<true>(xp), rm = from_right<true>(xm);
          const bool lv = from_left<true>(static_cast<int>(xv)) != 0;
          const bool rv = from_right<true>(static_cast<int>(xv)) != 0;
          qbx[v] = qbar[v];
          if (limited) {
            if (lv) qbx[v] += lp;  // left neighbour: this cell is its plus side
            if (xv) qbx[v] += x0;
            if (rv) qbx[v] += rm;  // right neighbour: this cell is its minus side
          }
          const bool yv = limited && done_col && dyb[v] != T(0);
          T ym = T(0), y0 = T(0), yp = T(0);
          if (yv) limited_slope_adj(st, qz(v), qa[v], qb[v], dyb[v], ym, y0, yp);
          ymf[v] = yv ? ym : skip;
          y0f[v] = yv ? y0 : skip;
          ypk[v] = yv ? yp : skip;
        }
        // seam row index through readfirstlane (k is wave-uniform): an opaque
        // scalar, so the compiler forms each store address from it when the
        // store runs instead of carrying 16 address registers through the loop
        const int ks = __builtin_amdgcn_readfirstlane(k);
        if (gat_col) {
          if (k >= ra + 2) {
            // row k-1 is final
            T qfin[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) qfin[v] = acc(v) + ymf[v];
            if constexpr (kLamPP)
              finish_cell_lam(a, ci, ks - 1, qfin, lam_pp);
            else
              finish_cell(a, ci, ks - 1, qfin);
          }
          if (k == ra) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              a.seam[seam_at(g, ks, 0, v, ci)] = qbx[v];
              a.seam[seam_at(g, ks, 1, v, ci)] = y0f[v];
              a.seam[seam_at(g, ks, 2, v, ci)] = ymf[v];
            }
          }
          if (k == ra + 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) a.seam[seam_at(g, ks, 0, v, ci)] = ymf[v];
          }
        }
        // partial sum of row k: ((own + x parts) + plus part of row k-1) + own y part
#pragma unroll
        for (int v = 0; v < 4; ++v) acc(v) = (qbx[v] + ypf(v)) + y0f[v];
        if (gat_col && k == rb - 1) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            if (k > ra) a.seam[seam_at(g, ks, 1, v, ci)] = acc(v);
            a.seam[seam_at(g, ks, 3, v, ci)] = ypk[v];
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) ypf(v) = ypk[v];
      }
      // x-faces of row r
      if (r < rb) {
        T xmax_own[4], xl[4], xr[4], xlb[4] = {T(0), T(0), T(0), T(0)};
        face_state(a.h, qb, dqx, s, dtdx, true, xmax_own);
        face_state(a.h, qb, dqx, s, dtdx, false, xr);
#pragma unroll
        for (int v = 0; v < 4; ++v) xl[v] = from_left<true>(xmax_own[v]);
        double lam_l[4];  // lam_r of column ci-1
#pragma unroll
        for (int v = 0; v < 4; ++v) lam_l[v] = from_left<true>(lam_r[v]);
        if (xface && r >= 2 && r <= g.ny + 1) {
          T fb[4];
          bool any = false;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double x = (0.0 + lam_r[v]) - lam_l[v];
            fb[v] = static_cast<T>(x * static_cast<double>(dtdxU));
            any |= (x != 0.0);
          }
          if (any) riemann_two_shock_adj<M>(a.h, xl, xr, fb, xlb, bxm);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) bxp[v] = from_right<true>(xlb[v]);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      pdx[v] = dqx[v];
      pdy[v] = dqy[v];
      ps[v] = s[v];
      pbxm[v] = bxm[v];
      pbxp[v] = bxp[v];
      pbym[v] = bym[v];
      qz(v) = qa[v];
      qa[v] = qb[v];
      qb[v] = qc[v];
      if constexpr (kLamPP) lam_pp[v] = lam_p[v];
      lam_p[v] = lam_r[v];
    }
  }
}

// The FP64 instantiation is held at 2 resident wavefronts per SIMD (the
// E2D_ADJ_MARCH_MIN_WAVES cap).  The float-arithmetic ones are held at 3 (168
// VGPRs): left to itself the compiler gives fp32_native and mixed 173-177
// VGPRs, 2 wavefronts per SIMD, and their adjoint steps take 312 and 323 us
// at 2048^2 against 284 and 303 us at 3 (with a 28-36 B spill); FP32 fits
// in 163 either way.
template <class P>
__global__ __launch_bounds__(64 * kAdjWaves) E2D_ADJ_MARCH_ATTR void k_adj_march_gather(const AdjArgs<P> a) {
  adj_march_gather_body<P>(a);
}
template <class P>
__global__ __launch_bounds__(64 * kAdjWaves) __attribute__((amdgpu_waves_per_eu(3))) void k_adj_march_gather_w3(
    const AdjArgs<P> a) {
  adj_march_gather_body<P>(a);
}

// The rows k_adj_march_gather leaves unfinished: row 0, row jsize-1, and the
// first and last row of every segment, from the seam buffer.  Thread row t:
// 0 -> row 0, 1 -> row jsize-1, 2 + 2s -> first row of segment s, 3 + 2s ->
// last row of segment s (none when the segment has one row).
template <class P>
__global__ __launch_bounds__(256) void k_adj_seam(const AdjArgs<P> a) {
  using T = typename P::T;
  const Grid& g = a.g;
  const int ci = blockIdx.x * 64 + (threadIdx.x & 63);
  const int t = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (ci >= g.isize || t >= 2 + 2 * (a.nsegs + a.nshort)) return;
  const T skip = -T(0);
  auto V = [&](int row, int slot, int v) { return a.seam[seam_at(g, row, slot, v, ci)]; };
  const int last = g.jsize - 2;  // last completed row
  T qb[4];
  int r;
  if (t == 0) {
    r = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) qb[v] = T(0) + V(1, 2, v);
  } else if (t == 1) {
    r = g.jsize - 1;
#pragma unroll
    for (int v = 0; v < 4; ++v) qb[v] = T(0) + V(last, 3, v);
  } else {
    int ra, rb;
    adj_segment_rows(a, (t - 2) >> 1, ra, rb);
    if (((t - 2) & 1) == 0) {
      r = ra;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T below = ra > 1 ? V(ra - 1, 3, v) : skip;
        const T above = ra + 1 < rb ? V(ra + 1, 0, v) : (rb <= last ? V(rb, 2, v) : skip);
        qb[v] = ((V(ra, 0, v) + below) + V(ra, 1, v)) + above;
      }
    } else {
      r = rb - 1;
      if (r == ra) return;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T above = rb <= last ? V(rb, 2, v) : skip;
        qb[v] = V(r, 1, v) + above;
      }
    }
  }
  finish_cell(a, ci, r, qb);
}

// all cells: bar of q from the cell's own trace and the limiters of the cell
// and its four neighbours, then bar of u
template <class P>
__global__ __launch_bounds__(256) void k_adj_gather(const AdjArgs<P> a) {
  using T = typename P::T;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const Grid& g = a.g;
  if (i >= g.isize || j >= g.jsize) return;
  T qb[4] = {T(0), T(0), T(0), T(0)};
  const bool cell_has = i >= 1 && i <= g.isize - 2 && j >= 1 && j <= g.jsize - 2;
  if (cell_has) load4(a.qb, g, i, j, qb);
  if (slopes_limited(a.h)) {
    const T st = static_cast<T>(a.h.slope_type);
    // x-slopes of cells (i-1, j), (i, j), (i+1, j); y-slopes of (i, j-1), (i, j), (i, j+1)
    for (int dir = 0; dir < 2; ++dir) {
      for (int o = -1; o <= 1; ++o) {
        const int ni = dir == 0 ? i + o : i, nj = dir == 0 ? j : j + o;
        if (ni < 1 || ni > g.isize - 2 || nj < 1 || nj > g.jsize - 2) continue;
        const T* bar = dir == 0 ? a.dqxb : a.dqyb;
        const int mi = dir == 0 ? ni - 1 : ni, mj = dir == 0 ? nj : nj - 1;
        const int pi = dir == 0 ? ni + 1 : ni, pj = dir == 0 ? nj : nj + 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const T ob = bar[at(g, ni, nj, v)];
          if (ob == T(0)) continue;
          T bm = T(0), b0 = T(0), bp = T(0);
          limited_slope_adj(st, a.q[at(g, mi, mj, v)], a.q[at(g, ni, nj, v)], a.q[at(g, pi, pj, v)], ob, bm, b0,
                            bp);
          qb[v] += (o == 0) ? b0 : (o == 1 ? bm : bp);  // this cell is the neighbour's minus/plus side
        }
      }
    }
  }
  T u[4], ub[4] = {T(0), T(0), T(0), T(0)};
  load4(a.u, g, i, j, u);
  cons_to_prim_adj<typename P::M>(a.h, u, qb, ub);
  if (interior(g, i, j)) {
    // lam_in + own bar of u, in double; ghost bars are added by k_adj_fold_frame
#pragma unroll
    for (int v = 0; v < 4; ++v) a.out[at(g, i, j, v)] = a.lam[at(g, i, j, v)] + static_cast<double>(ub[v]);
  } else {
    store4(a.ub, g, i, j, ub);
  }
}

// Ghost cells are filled from interior cells of columns {2, 3, nx, nx+1} or
// rows {2, 3, ny, ny+1} (0-based; reflective, absorbing and periodic alike),
// so only this frame receives ghost bars.  Thread t of the frame: rows
// {2, 3, ny, ny+1} over all interior columns first, then columns
// {2, 3, nx, nx+1} over the remaining rows [4, ny-1].
__device__ __forceinline__ bool frame_cell(const Grid& g, int t, int& i, int& j) {
  const int top = 4 * g.nx;
  if (t < top) {
    const int k = t / g.nx;
    i = 2 + t % g.nx;
    j = k < 2 ? 2 + k : g.ny + k - 2;
    return true;
  }
  t -= top;
  const int rows = g.ny - 4;
  if (rows <= 0 || t >= 4 * rows) return false;
  const int k = t / rows;
  j = 4 + t % rows;
  i = k < 2 ? 2 + k : g.nx + k - 2;
  return true;
}

// frame cells: add the bars of the ghost cells filled from this cell to
// lam_in + own bar (written by k_adj_gather), x-face ghosts of the row first,
// then y-face ghosts of the column, all in double
template <class P>
__global__ __launch_bounds__(256) void k_adj_fold_frame(const AdjArgs<P> a) {
  const Grid& g = a.g;
  int i, j;
  if (!frame_cell(g, blockIdx.x * 256 + threadIdx.x, i, j)) return;
  double r[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) r[v] = a.out[at(g, i, j, v)];
  const int gx[4] = {0, 1, g.nx + 2, g.nx + 3};
  for (int k = 0; k < 4; ++k) {  // x-face ghosts of row j
    const Src s = ghost_source(g, gx[k], j);
    if (s.i != i) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double b = static_cast<double>(a.ub[at(g, gx[k], j, v)]);
      r[v] += (v == s.flip) ? -b : b;
    }
  }
  const int gy[4] = {0, 1, g.ny + 2, g.ny + 3};
  for (int k = 0; k < 4; ++k) {  // y-face ghosts of column i
    const Src s = ghost_source(g, i, gy[k]);
    if (s.j != j) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double b = static_cast<double>(a.ub[at(g, i, gy[k], v)]);
      r[v] += (v == s.flip) ? -b : b;
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) a.out[at(g, i, j, v)] = r[v];
}

// Kernel sequence of an adjoint step, all giving the same bits:
//   fused   k_adj_march_gather + k_adj_seam
//   march   k_adj_march + k_adj_gather_march (bars of q and of the slopes
//           through HBM)
//   separate  k_adj_prim, k_adj_faces, k_adj_cells, k_adj_gather (face bars
//           through HBM as well)
// E2D_ADJOINT_PASSES=fused|march|separate chooses; fused is the default
// (2048^2, dense lam: FP64 778 us fused, 796 us march; FP32 477 us fused,
// 527 us march).  The fused kernel writes lam_out while other wavefronts still
// read lam_in, so an aliased call (lam_out == lam_in) runs march, whose gather
// reads and writes each element of lam in the same thread.
enum class Passes { kFused, kMarch, kSeparate };
Passes adjoint_passes(bool aliased) {
  Passes p = Passes::kFused;
  if (const char* env = getenv("E2D_ADJOINT_PASSES")) {
    if (strcmp(env, "separate") == 0) p = Passes::kSeparate;
    if (strcmp(env, "march") == 0) p = Passes::kMarch;
    if (strcmp(env, "fused") == 0) p = Passes::kFused;
  }
  if (p == Passes::kFused && aliased) p = Passes::kMarch;
  return p;
}

// Rows per k_adj_march segment (E2D_ADJOINT_SEG_ROWS overrides): about
// rows/340, clamped to [4, 16].  The kernel holds 2 wavefronts per SIMD, so
// short segments (many wavefronts, several rounds) balance the data-dependent
// Riemann work best: measured at 2048^2 FP64 with a dense lam, 4-7 rows give
// 860-880 us per adjoint step, 12 rows 896 us, 18 rows 950 us, 24 rows 1030 us.
int march_seg_rows(int rows) {
  if (const char* env = getenv("E2D_ADJOINT_SEG_ROWS")) {
    const int r = atoi(env);
    if (r >= 1) return r;
  }
  const int r = static_cast<int>(rows / 340.0 + 0.5);
  return r < 4 ? 4 : (r > 16 ? 16 : r);
}

static int adj_env(const char* name, int dflt) {
  const char* v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

// Segment plan of k_adj_march_gather over nstrips strips of `rows` completed
// rows: long segments of about rows/170 rows (within [4, 16];
// E2D_ADJOINT_SEG_ROWS overrides) over the lower 75% of the rows
// (E2D_ADJOINT_LONG_PCT), short ones of a third of that
// (E2D_ADJOINT_SHORT_ROWS; 4 rows against 3: FP64 720 -> 715 us, FP32
// 441 -> 430 us, mixed 386 -> 380 us) over the rest, dealt to the XCDs in chunks of 16
// segments (E2D_ADJOINT_CHUNK), so every XCD ends on short segments.
// Measured at 2048^2 with a dense lam, us per adjoint step against uniform
// 6-row segments: FP64 765 -> 721, FP32 481 -> 437, mixed 417 -> 384; 16-row
// long segments suit FP64 (715) but not the float kernels, which hold 3
// wavefronts per SIMD (518 and 440).  Returns the number of blocks.
template <class P>
int plan_adjoint_segments(AdjArgs<P>& a, int nstrips, int rows) {
  int R = adj_env("E2D_ADJOINT_SEG_ROWS", 0);
  if (R < 1) {
    R = static_cast<int>(rows / 170.0 + 0.5);
    R = R < 4 ? 4 : (R > 16 ? 16 : R);
  }
  const int r = adj_env("E2D_ADJOINT_SHORT_ROWS", R / 3);
  const int pct = adj_env("E2D_ADJOINT_LONG_PCT", 75);
  const int L = pct >= 100 ? (rows + R - 1) / R
                           : static_cast<int>(static_cast<int64_t>(rows) * (pct < 0 ? 0 : pct) / 100 / R);
  const int covered = L * R < rows ? L * R : rows;
  a.seg_rows = R;
  a.nsegs = L;
  a.short_rows = r < 1 ? 1 : r;
  a.short_base = 1 + covered;
  a.nshort = covered < rows ? (rows - covered + a.short_rows - 1) / a.short_rows : 0;
  a.nlong = nstrips * L;
  a.nitems = a.nlong + nstrips * a.nshort;
  const int c = adj_env("E2D_ADJOINT_CHUNK", 16);
  a.chunk = c < 1 ? 1 : c;
  return dealt_blocks(a.nlong, a.nitems, a.chunk, kAdjWaves);
}

// Rows per k_adj_gather_march segment (E2D_ADJOINT_GATHER_ROWS overrides):
// short, so that the light, memory-bound kernel has many wavefronts in flight.
int gather_seg_rows() {
  if (const char* env = getenv("E2D_ADJOINT_GATHER_ROWS")) {
    const int r = atoi(env);
    if (r >= 1) return r;
  }
  return 8;
}

template <class P>
hipError_t adjoint_step(const e2d_params& p, const DevBuffers& b, int buf, double dt, const double* lam_in,
                        double* lam_out, const AdjointScratch& s, hipStream_t st) {
  using T = typename P::T;
  AdjArgs<P> a;
  a.h = make_hydro<T>(p);
  a.hU = make_hydro<typename P::U>(p);
  a.g = make_grid(p, b);
  a.u = static_cast<const typename P::S*>(b.u[buf]);
  a.lam = lam_in;
  a.out = lam_out;
  a.q = static_cast<T*>(s.q);
  a.xminb = static_cast<T*>(s.xminb);
  a.xmaxb = static_cast<T*>(s.xmaxb);
  a.yminb = static_cast<T*>(s.yminb);
  a.ymaxb = static_cast<T*>(s.ymaxb);
  a.qb = static_cast<T*>(s.qb);
  a.dqxb = static_cast<T*>(s.dqxb);
  a.dqyb = static_cast<T*>(s.dqyb);
  a.ub = static_cast<T*>(s.ub);
  a.seam = static_cast<T*>(s.seam);
  a.dt = dt;
  hipError_t e = launch_boundaries(p, b, buf, st);
  if (e != hipSuccess) return e;
  const Grid& g = a.g;
  const Passes passes = adjoint_passes(lam_in == lam_out);
  if (passes == Passes::kSeparate) {
    hipLaunchKernelGGL(k_adj_prim<P>, dim3((g.isize + 63) / 64, (g.jsize + 3) / 4), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_adj_faces<P>, dim3((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_adj_cells<P>, dim3((g.isize - 2 + 63) / 64, (g.jsize - 2 + 3) / 4), dim3(256), 0, st,
                       a);
    hipLaunchKernelGGL(k_adj_gather<P>, dim3((g.isize + 63) / 64, (g.jsize + 3) / 4), dim3(256), 0, st, a);
  } else if (passes == Passes::kFused) {
    const int nstrips = (g.isize + kFusedStripOut - 1) / kFusedStripOut;
    const int nblocks = plan_adjoint_segments(a, nstrips, g.jsize - 2);
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL(k_adj_march_gather_w3<P>, dim3(nblocks), dim3(64 * kAdjWaves), 0, st, a);
    else
      hipLaunchKernelGGL(k_adj_march_gather<P>, dim3(nblocks), dim3(64 * kAdjWaves), 0, st, a);
    const int per_strip = a.nsegs + a.nshort;
    hipLaunchKernelGGL(k_adj_seam<P>, dim3((g.isize + 63) / 64, (2 + 2 * per_strip + 3) / 4), dim3(256), 0, st, a);
  } else {
    a.seg_rows = march_seg_rows(g.jsize - 2);
    a.nsegs = (g.jsize - 2 + a.seg_rows - 1) / a.seg_rows;
    a.seg_groups = (a.nsegs + kAdjWaves - 1) / kAdjWaves;
    const int nstrips = (g.isize - 2 + kAdjStripOut - 1) / kAdjStripOut;
    hipLaunchKernelGGL(k_adj_march<P>, dim3(nstrips * a.seg_groups), dim3(64 * kAdjWaves), 0, st, a);
    a.gseg_rows = gather_seg_rows();
    a.gnsegs = (g.jsize + a.gseg_rows - 1) / a.gseg_rows;
    a.gseg_groups = (a.gnsegs + kAdjWaves - 1) / kAdjWaves;
    const int gstrips = (g.isize + kGatherStripOut - 1) / kGatherStripOut;
    hipLaunchKernelGGL(k_adj_gather_march<P>, dim3(gstrips * a.gseg_groups), dim3(64 * kAdjWaves), 0, st, a);
  }
  const int frame = 4 * g.nx + 4 * (g.ny > 4 ? g.ny - 4 : 0);
  hipLaunchKernelGGL(k_adj_fold_frame<P>, dim3((frame + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace

// the scratch arrays hold values of the step's arithmetic type
size_t adjoint_scratch_elem_bytes(int precision) {
  return with_precision(precision, [](auto pol) { return sizeof(typename decltype(pol)::T); });
}

hipError_t launch_adjoint_step(const e2d_params& p, const DevBuffers& b, int buf, double dt, const double* lam_in,
                               double* lam_out, const AdjointScratch& s, hipStream_t st) {
  return with_precision(p.precision, [&](auto pol) {
    return adjoint_step<decltype(pol)>(p, b, buf, dt, lam_in, lam_out, s, st);
  });
}

}  // namespace e2d
