// Pieces shared by the step, tangent and adjoint kernels: precision
// policies, derived parameters, grid description, the ghost-cell source map,
// and wave64 cross-lane helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>

#include "e2d_adjoint.hpp"
#include "e2d_internal.hpp"
#include "e2d_numerics.hpp"

namespace e2d {

// ---------------------------------------------------------------------------
// precision policies: storage S, flux arithmetic T, update / dt type U, and
// the division / square-root policies M (for T) and MU (for U)
template <class S_, class T_, class U_, class M_, class MU_>
struct Prec {
  using S = S_;
  using T = T_;
  using U = U_;
  using M = M_;
  using MU = MU_;
};
using P64 = Prec<double, double, double, RangedMath, RangedMath>;
using P32 = Prec<float, float, float, RangedMath, RangedMath>;
// FP32 state; FP32 flux arithmetic with the native reciprocal and square
// root; the update added in FP64 and rounded to the FP32 state
using PNative = Prec<float, float, double, FastMath, RangedMath>;
// FP64 state and update, FP32 slope / trace / Riemann arithmetic with the
// native reciprocal and square root (as fp32_native; the correctly rounded
// binary32 cores cost 33% more time for 5-7% less error at 4096^2)
using PMix = Prec<double, float, double, FastMath, RangedMath>;

// f(P{}) for the policy P of an E2D_PREC_* code (validated by the caller)
template <class F>
inline auto with_precision(int precision, F&& f) {
  switch (precision) {
    case E2D_PREC_FP64: return f(P64{});
    case E2D_PREC_FP32: return f(P32{});
    case E2D_PREC_FP32_NATIVE: return f(PNative{});
    default: return f(PMix{});
  }
}

// Derived parameters in precision T, as initHydroParameters computes them
// (serial_ref/m_run.f90:111-124) plus the cmpflx constant entho (m_utils.f90:521).
template <class T>
inline Hydro<T> make_hydro(const e2d_params& p) {
  Hydro<T> h;
  h.gamma0 = static_cast<T>(p.gamma0);
  h.cfl = static_cast<T>(p.cfl);
  h.smallr = static_cast<T>(p.smallr);
  h.smallc = static_cast<T>(p.smallc);
  h.smallp = h.smallc * h.smallc / h.gamma0;
  h.smallpp = h.smallr * h.smallp;
  h.gamma6 = (h.gamma0 + T(1)) / (T(2) * h.gamma0);
  h.entho = T(1) / (h.gamma0 - T(1));
  h.dx = (T(1) - T(0)) / static_cast<T>(p.nx);
  h.dy = (T(1) - T(0)) / static_cast<T>(p.ny);
  h.dxy_pow2 = p.nx > 0 && p.ny > 0 && (p.nx & (p.nx - 1)) == 0 && (p.ny & (p.ny - 1)) == 0;
  h.inv_dx = static_cast<T>(p.nx);
  h.inv_dy = static_cast<T>(p.ny);
  h.conv_tol = static_cast<T>(1e-6f);
  h.inv_smallr = T(1) / h.smallr;
  {  // in the scalar type of T (double for the tangent's dual numbers)
    using R = std::conditional_t<std::is_same_v<T, float>, float, double>;
    const R tol = static_cast<R>(1e-6f);
    h.conv_half_ulp = static_cast<T>(std::ldexp(R(1), std::ilogb(tol) - std::numeric_limits<R>::digits));
  }
  h.slope_type = p.slope_type;
  h.niter = p.niter_riemann;
  h.riemann_solver = p.riemann_solver;
  return h;
}

struct Grid {
  int nx, ny, isize, jsize;
  int64_t pitch, plane;
  int bc[4];  // xmin, xmax, ymin, ymax
};

inline Grid make_grid(const e2d_params& p, const DevBuffers& b) {
  Grid g;
  g.nx = p.nx;
  g.ny = p.ny;
  g.isize = p.nx + 4;
  g.jsize = p.ny + 4;
  g.pitch = b.pitch;
  g.plane = b.plane;
  g.bc[0] = p.boundary_type_xmin;
  g.bc[1] = p.boundary_type_xmax;
  g.bc[2] = p.boundary_type_ymin;
  g.bc[3] = p.boundary_type_ymax;
  return g;
}

// Where the value of cell (i, j) comes from once boundaries are filled.
// Non-corner ghost cells map to an interior cell (reflective: mirror with the
// normal momentum negated; absorbing: nearest interior cell; other: periodic),
// everything else (interior cells and the never-filled corners) maps to itself.
struct Src {
  int i, j;
  int flip;  // variable whose sign flips (2 = rho*u, 3 = rho*v), or -1
  bool ghost;
};

__device__ __forceinline__ Src ghost_source(const Grid& g, int i, int j) {
  Src s{i, j, -1, false};
  const bool xg = (i < 2) || (i >= g.nx + 2);
  const bool yg = (j < 2) || (j >= g.ny + 2);
  if (xg && !yg) {
    s.ghost = true;
    if (i < 2) {
      const int bc = g.bc[0];
      if (bc == 1) {
        s.i = 3 - i;
        s.flip = 2;
      } else if (bc == 2) {
        s.i = 2;
      } else {
        s.i = g.nx + i;
      }
    } else {
      const int bc = g.bc[1];
      if (bc == 1) {
        s.i = 2 * g.nx + 3 - i;
        s.flip = 2;
      } else if (bc == 2) {
        s.i = g.nx + 1;
      } else {
        s.i = i - g.nx;
      }
    }
  } else if (yg && !xg) {
    s.ghost = true;
    if (j < 2) {
      const int bc = g.bc[2];
      if (bc == 1) {
        s.j = 3 - j;
        s.flip = 3;
      } else if (bc == 2) {
        s.j = 2;
      } else {
        s.j = g.ny + j;
      }
    } else {
      const int bc = g.bc[3];
      if (bc == 1) {
        s.j = 2 * g.ny + 3 - j;
        s.flip = 3;
      } else if (bc == 2) {
        s.j = g.ny + 1;
      } else {
        s.j = j - g.ny;
      }
    }
  }
  return s;
}

__device__ __forceinline__ int64_t at(const Grid& g, int i, int j, int v) {
  return static_cast<int64_t>(v) * g.plane + static_cast<int64_t>(j) * g.pitch + i;
}

// ---------------------------------------------------------------------------
// cross-lane helpers (wave64)
//
// Neighbour exchange: from_left(x) is the value of lane l-1, from_right(x)
// that of lane l+1, and the edge lane that has no neighbour keeps its own
// value.  Two ways to move it: ds_bpermute (__shfl_up / __shfl_down, issued to
// the LDS pipe, with an LDS round trip of latency) or DPP wavefront shifts
// (wave_shr:1 / wave_shl:1 with bound_ctrl off: one v_mov_b32_dpp per 32 bits
// on the VALU, no round trip).  The forward step is bound by VALU issue and
// keeps its VALU slots for arithmetic (kDpp = false; DPP measured 5% slower in
// FP32 at 2048^2); the adjoint kernels wait on latency more than on issue and
// use DPP (kDpp = true; 1% faster in FP64, 6% in FP32).  Callers run them with
// all 64 lanes active.
template <bool kDpp, bool kUp>
__device__ __forceinline__ int lane_shift(int x) {
  if (kDpp) return __builtin_amdgcn_update_dpp(x, x, kUp ? 0x138 : 0x130, 0xf, 0xf, false);
  return kUp ? __shfl_up(x, 1, 64) : __shfl_down(x, 1, 64);
}
template <bool kDpp, bool kUp>
__device__ __forceinline__ float lane_shift(float x) {
  return __builtin_bit_cast(float, lane_shift<kDpp, kUp>(__builtin_bit_cast(int, x)));
}
template <bool kDpp, bool kUp>
__device__ __forceinline__ double lane_shift(double x) {
  if (!kDpp) return kUp ? __shfl_up(x, 1, 64) : __shfl_down(x, 1, 64);
  const long long b = __double_as_longlong(x);
  const int lo = lane_shift<kDpp, kUp>(static_cast<int>(b));
  const int hi = lane_shift<kDpp, kUp>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <bool kDpp = false, class T>
__device__ __forceinline__ T from_left(T x) {  // value of lane-1
  return lane_shift<kDpp, true>(x);
}
template <bool kDpp = false, class T>
__device__ __forceinline__ T from_right(T x) {  // value of lane+1
  return lane_shift<kDpp, false>(x);
}
template <class T>
__device__ __forceinline__ T wave_max(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T y = __shfl_xor(x, o, 64);
    x = x > y ? x : y;
  }
  return x;
}

__device__ __forceinline__ void atomic_max_nonneg(double* dst, double v) {
  // IEEE order of non-negative doubles equals the order of their bit patterns
  atomicMax(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// value of running-invDt slot `slot`: the maximum of its parts (uniform loads)
__device__ __forceinline__ double invdt_read(const double* scal, int slot) {
  const double* s = scal + slot * kInvdtSlot;
  double m = s[0];
#pragma unroll
  for (int p = 1; p < kInvdtParts; ++p) {
    const double x = s[p * kInvdtStride];
    m = m > x ? m : x;
  }
  return m;
}

// The same for a whole wavefront (all 64 lanes active): lane l loads part
// l mod kInvdtParts, one round trip for the slot, then a wave maximum; the
// result is moved to scalar registers, where the uniform dt arithmetic stays.
// In two halves, so that a kernel can issue the load early and reduce later:
__device__ __forceinline__ double invdt_part(const double* scal, int slot) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  return scal[slot * kInvdtSlot + (lane % kInvdtParts) * kInvdtStride];
}
__device__ __forceinline__ double invdt_reduce(double part) {
  const double m = wave_max(part);
  const long long bits = __double_as_longlong(m);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(bits)));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(bits >> 32)));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ double invdt_read_wave(const double* scal, int slot) {
  return invdt_reduce(invdt_part(scal, slot));
}

// fold a non-negative rate into slot `slot`, part chosen by `who`
__device__ __forceinline__ void invdt_fold(double* scal, int slot, int who, double v) {
  atomic_max_nonneg(scal + slot * kInvdtSlot + (who % kInvdtParts) * kInvdtStride, v);
}

// Segments dealt to the XCDs.  Workgroups reach the 8 XCDs round-robin
// (block b to XCD b mod 8) and start in order on each.  nitems items, the
// first nlong of them long segments, are dealt in chunks of c consecutive
// items, chunk i to XCD i mod 8: the long items first, then the short ones
// (chunks counted afresh), so that every XCD ends its share on short items.
// Items of XCD x among n items dealt that way:
__host__ __device__ __forceinline__ int xcd_share(int n, int c, int x) {
  const int full = n / c, rem = n - full * c;
  return (full / 8 + (x < full % 8 ? 1 : 0)) * c + (x == full % 8 ? rem : 0);
}

// The item of wavefront `wave` of block b (wpb wavefronts per block): its
// index among the long items (is_long) or among the short ones; false when
// the wavefront has none.
__device__ __forceinline__ bool dealt_item(int b, int wave, int wpb, int nlong, int nitems, int c, int& item,
                                           bool& is_long) {
  const int x = b % 8;
  const int k = b / 8 * wpb + wave;
  const int nl = xcd_share(nlong, c, x);
  if (k < nl) {
    item = ((k / c) * 8 + x) * c + k % c;
    is_long = true;
    return true;
  }
  const int ks = k - nl;
  if (ks < xcd_share(nitems - nlong, c, x)) {
    item = ((ks / c) * 8 + x) * c + ks % c;
    is_long = false;
    return true;
  }
  return false;
}

// blocks that give every XCD its share
inline int dealt_blocks(int nlong, int nitems, int c, int wpb) {
  int most = 0;
  for (int x = 0; x < 8; ++x) {
    const int n = xcd_share(nlong, c, x) + xcd_share(nitems - nlong, c, x);
    most = most > (n + wpb - 1) / wpb ? most : (n + wpb - 1) / wpb;
  }
  return 8 * most;
}

// XCD-aware block order: hardware hands consecutive workgroups to the 8 XCDs
// in turn; this gives each XCD a contiguous range of logical blocks so that
// neighbouring strips/segments (which share halo rows and columns) meet in the
// same L2.
__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, r = nb % kXcd;
  const int xcd = b % kXcd, k = b / kXcd;
  return (xcd < r) ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// Strip and segment group of logical block lb in a grid of nstrips x
// seg_groups blocks.  Band-major: neighbouring strips of one band of rows are
// consecutive blocks, so the cache lines two strips share (their column halos)
// and the rows two bands share meet in one L2 at about the same time; this
// pays where a kernel is bound by memory traffic.  Strip-major: the segments
// of one strip are consecutive.
template <bool kBandMajor>
__device__ __forceinline__ void strip_and_group(int lb, int seg_groups, int& strip, int& group) {
  if (kBandMajor) {
    const int nstrips = static_cast<int>(gridDim.x) / seg_groups;
    group = lb / nstrips;
    strip = lb % nstrips;
  } else {
    strip = lb / seg_groups;
    group = lb % seg_groups;
  }
}


// cross-lane shuffles of a dual number move both parts
template <bool kDpp, bool kUp>
__device__ __forceinline__ Dual lane_shift(Dual x) {
  return Dual(lane_shift<kDpp, kUp>(x.v), lane_shift<kDpp, kUp>(x.d));
}

}  // namespace e2d
