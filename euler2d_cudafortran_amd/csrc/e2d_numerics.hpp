// Pointwise numerics of the euler2d hot path, for gfx950.
//
// Every routine keeps the reference's operation order (pkestene/euler2d_cudaFortran,
// serial_ref/m_utils.f90) so that, built with -ffp-contract=off and the exact
// math policy, a result is bit-identical to the reference's for the same inputs.
// Reassociations that are exact are used freely: multiplication by +-1 becomes a
// select of the sign, loop-invariant constants are hoisted, and the four face
// states of a cell share one evaluation of the trace source terms.
//
// Variable order inside a 4-vector follows the reference's state planes:
// [0] = rho, [1] = p (E for conservative vectors), [2] = u, [3] = v
// (serial_ref/m_constants.f90:11-14).
#pragma once

#include <type_traits>

// Host and device: the kernels use these routines; the adjoint's checks also
// compile them for the CPU.
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define E2D_HDI __host__ __device__ __forceinline__
#else
#include <cmath>
#define E2D_HDI inline __attribute__((always_inline))
#endif

namespace e2d {

// Per-run constants in the arithmetic type T (serial_ref/m_parameters.f90:10-36,
// derived as in serial_ref/m_run.f90:111-124).
template <class T>
struct Hydro {
  T gamma0, gamma6, cfl, smallr, smallc, smallp, smallpp;
  T entho;     // 1/(gamma0-1), evaluated as the reference evaluates it in cmpflx
  T dx, dy;
  // 1/dx and 1/dy, used when both cell sizes are powers of two (nx and ny
  // powers of two): then x / dx equals x * (1/dx) exactly for every x, the
  // exact quotient being representable whenever the division does not overflow
  T inv_dx, inv_dy;
  int dxy_pow2;
  T conv_tol;  // Newton stop threshold: the FP32 literal 1e-6 widened to T (m_utils.f90:396)
  T inv_smallr;  // 1 / smallr rounded: a starting reciprocal for quotients by a floored density
  T conv_half_ulp;  // half an ulp of conv_tol: conv_tol + conv_half_ulp is the rounding boundary above it
  int slope_type, niter, riemann_solver;
};

// Division / square root policy.  ExactMath keeps IEEE correctly rounded
// operations (bitwise parity).  FastMath (float only) uses the native
// reciprocal and square root (quotients within 2 ulp, roots within 1 ulp), for
// the fp32_native and mixed variants.
// sqrt_pos(x) is sqrt(x) for a radicand that is never zero (positive or NaN):
// a policy may skip its zero case there.  sqrt_pos_inv(x, inv) is
// sqrt_pos(x) that may also leave in inv a reciprocal of the root which
// div_inv(a, s, inv) = div(a, s) may use in place of its own (the same
// quotient); policies without one leave inv unused.  recip(b) is such a
// reciprocal of b, and recip_near(b, r0) one from a start r0 within about
// 2^-50 of 1/b (one Newton step); both return 0 where div_inv ignores them.
// quot_gt(a, d, tol, h) is div(a, d) > tol for a >= 0 (or NaN) and d > 0,
// h being half an ulp of tol; a policy may decide it without the quotient.
struct ExactMath {
  static E2D_HDI double div(double a, double b) { return a / b; }
  static E2D_HDI float div(float a, float b) { return a / b; }
  static E2D_HDI double sqrt(double a) { return __builtin_sqrt(a); }
  static E2D_HDI float sqrt(float a) { return __builtin_sqrtf(a); }
  static E2D_HDI double sqrt_pos(double a) { return __builtin_sqrt(a); }
  static E2D_HDI float sqrt_pos(float a) { return __builtin_sqrtf(a); }
  template <class T>
  static E2D_HDI T sqrt_pos_inv(T a, T& inv) {
    inv = T(0);
    return sqrt_pos(a);
  }
  template <class T>
  static E2D_HDI T div_inv(T a, T b, T) { return div(a, b); }
  template <class T>
  static E2D_HDI T recip(T) { return T(0); }
  template <class T>
  static E2D_HDI T recip_near(T, T) { return T(0); }
  template <class T>
  static E2D_HDI bool quot_gt(T a, T d, T tol, T) { return div(a, d) > tol; }
};
#if defined(__HIPCC__)
// Its quotient is a times the native reciprocal of b: recip(b) is that
// reciprocal (recip_near and sqrt_pos_inv form it too, whatever the start),
// and div_inv(a, b, r) = a r, the same bits as div(a, b).
struct FastMath {
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
  static __device__ __forceinline__ float sqrt_pos(float a) { return __builtin_amdgcn_sqrtf(a); }
  static __device__ __forceinline__ float sqrt_pos_inv(float a, float& inv) {
    const float s = sqrt_pos(a);
    inv = recip(s);
    return s;
  }
  static __device__ __forceinline__ float div_inv(float a, float, float r) { return a * r; }
  static __device__ __forceinline__ float recip(float b) { return __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float recip_near(float b, float) { return recip(b); }
  static __device__ __forceinline__ bool quot_gt(float a, float d, float tol, float) { return div(a, d) > tol; }
};

// RangedMath (FP64): correctly rounded division and square root for finite,
// nonzero operands whose binary exponents stay far from the double limits
// (|log2| below ~700; the solver's values span about 1e-22..1e9).  These are
// the Newton-Raphson (division) and Goldschmidt (square root) cores of the
// IEEE sequences the compiler emits for a / b and sqrt(x) on gfx950, without
// the range scaling (v_div_scale, ldexp by 2^256) and special-value fixup
// (v_div_fixup, class tests) that act only outside that range.  Inside it the
// results are the IEEE results bit for bit, except that a zero quotient is +0
// whatever the sign of the numerator; a zero radicand returns zero (sqrt; the
// core alone, sqrt_pos, returns NaN for it and is the IEEE result for every
// positive radicand in the range).
//
// The FP32 overloads do the same for binary32 (exponents within about +-90,
// the solver's 1e-22..1e9 again): division is the reciprocal seed, one Newton
// step and two remainder corrections of the IEEE sequence without v_div_scale
// / v_div_fixup; square root takes the hardware estimate s (within 1 ulp) and
// picks s - ulp, s or s + ulp from the signs of the exact residuals
// x - (s -+ ulp) s, which decide on which side of the midpoints sqrt(x) lies.
struct RangedMath {
  static __device__ __forceinline__ float div(float a, float b) {
    const float nb = -b;
    float r = __builtin_amdgcn_rcpf(b);
    r = __builtin_fmaf(__builtin_fmaf(nb, r, 1.0f), r, r);
    float q = a * r;
    q = __builtin_fmaf(__builtin_fmaf(nb, q, a), r, q);
    return __builtin_fmaf(__builtin_fmaf(nb, q, a), r, q);
  }
  static __device__ __forceinline__ float sqrt(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const int bits = __builtin_bit_cast(int, s);
    const float dn = __builtin_bit_cast(float, bits - 1);
    const float up = __builtin_bit_cast(float, bits + 1);
    float t = __builtin_fmaf(-dn, s, x) <= 0.0f ? dn : s;
    t = __builtin_fmaf(-up, s, x) > 0.0f ? up : t;
    return t;
  }
  static __device__ __forceinline__ float sqrt_pos(float x) { return sqrt(x); }
  static __device__ __forceinline__ float sqrt_pos_inv(float x, float& inv) {
    inv = 0.0f;
    return sqrt(x);
  }
  static __device__ __forceinline__ float div_inv(float a, float b, float) { return div(a, b); }
  static __device__ __forceinline__ float recip(float) { return 0.0f; }
  static __device__ __forceinline__ float recip_near(float, float) { return 0.0f; }
  // binary32: the same test as the double one below, evaluated in double,
  // where tol d is exact (48 bits), a - tol d is exact wherever it is near
  // h d (about 49 bits), and h d is exact and normal for every binary32 d
  static __device__ __forceinline__ bool quot_gt(float a, float d, float tol, float h) {
    const double dd = d;
    return __builtin_fma(-static_cast<double>(tol), dd, static_cast<double>(a)) > static_cast<double>(h) * dd;
  }
  static __device__ __forceinline__ double div(double a, double b) { return div_inv(a, b, recip(b)); }
  // the reciprocal of the division sequence: seed and two Newton steps
  static __device__ __forceinline__ double recip(double b) {
    const double nb = -b;
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(nb, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(nb, r, 1.0);
    return __builtin_fma(r, e, r);
  }
  // one Newton step from r0 = 1/b (1 + e0): error e0^2 and the step's
  // rounding, as close as recip(b) once |e0| is below about 2^-50
  static __device__ __forceinline__ double recip_near(double b, double r0) {
    return __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
  }
  // Rounded to nearest, a / d exceeds tol exactly when a / d > mu = tol + h,
  // the midpoint between tol and the next double; with d > 0 that is
  // x = a - tol d > h d.  q = fma(-tol, d, a) is x rounded once, and h d is
  // exact (h is a power of two; the caller keeps h d a normal double), so
  // q > h d implies x > h d and q < h d implies x < h d.  q == h d cannot
  // happen: write d = B 2^b and mu = M 2^e (B < 2^53, M odd with 54 bits,
  // h = 2^e, so h d has the ulp 2^(e+b)).  x - h d = a - mu d is not zero
  // (a would need M B's 54+ bits) and, where a is near mu d, a multiple of
  // 2^(e+b), so x lies at least one ulp of h d away from h d and does not
  // round to it.  The order is decided without the quotient.
  static __device__ __forceinline__ bool quot_gt(double a, double d, double tol, double h) {
    return __builtin_fma(-tol, d, a) > h * d;
  }
  // (the core first, then the select: written as x == 0 ? x : sqrt_pos(x) the
  // compiler branches around the core, which costs the adjoint kernel 14%)
  static __device__ __forceinline__ double sqrt(double x) {
    const double s = sqrt_pos(x);
    return x == 0.0 ? x : s;
  }
  static __device__ __forceinline__ double sqrt_pos(double x) {
    double hh;
    return sqrt_core(x, hh);
  }
  // The Goldschmidt step leaves hh = 1/(2 sqrt(x)) to about 2^-45; 2 hh, one
  // Newton step against the rounded root s (error squared to about 2^-90, then
  // the step's rounding), is a reciprocal of s as good as div's two Newton
  // steps from v_rcp_f64, without the quarter-rate v_rcp_f64.
  static __device__ __forceinline__ double sqrt_pos_inv(double x, double& inv) {
    double hh;
    const double s = sqrt_core(x, hh);
    const double r0 = hh + hh;
    inv = __builtin_fma(r0, __builtin_fma(-s, r0, 1.0), r0);
    return s;
  }
  // div's remainder correction with the reciprocal r of b given
  static __device__ __forceinline__ double div_inv(double a, double b, double r) {
    const double q = a * r;
    const double rem = __builtin_fma(-b, q, a);
    return __builtin_fma(rem, r, q);
  }

 private:
  static __device__ __forceinline__ double sqrt_core(double x, double& hh) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    hh = y * 0.5;
    const double r = __builtin_fma(-hh, g, 0.5);
    g = __builtin_fma(g, r, g);
    hh = __builtin_fma(hh, r, hh);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, hh, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, hh, g);
  }
};
#endif

// Fortran max/min as the reference compiler lowers them: a > b ? a : b and
// a < b ? a : b.  In device code they are the hardware v_max_f64 / v_min_f64
// and v_max_f32 / v_min_f32 (IEEE maxNum/minNum), which return the same bits
// except when b is NaN or {a, b} = {+0, -0}.  Neither case changes a result of
// the step, in either precision:
// every floor compares against a positive constant or rho*smallp > 0, the
// limiter compares absolute values, max(x, 0) for the first pressure guess
// never sees x = -0 (a sum with a positive term), and the Newton correction
// max(t, -p) can meet {+0, -0} only to add a zero of either sign to p = +0.
// NaN enters only with a NaN state.  E2D_SELECT_MINMAX=1 keeps the selects.
#ifndef E2D_SELECT_MINMAX
#define E2D_SELECT_MINMAX 0
#endif
template <class T>
E2D_HDI T fmax2(T a, T b) {
#if defined(__HIP_DEVICE_COMPILE__) && !E2D_SELECT_MINMAX
  if constexpr (std::is_same<T, double>::value) return __builtin_fmax(a, b);
  if constexpr (std::is_same<T, float>::value) return __builtin_fmaxf(a, b);
#endif
  return a > b ? a : b;
}
template <class T>
E2D_HDI T fmin2(T a, T b) {
#if defined(__HIP_DEVICE_COMPILE__) && !E2D_SELECT_MINMAX
  if constexpr (std::is_same<T, double>::value) return __builtin_fmin(a, b);
  if constexpr (std::is_same<T, float>::value) return __builtin_fminf(a, b);
#endif
  return a < b ? a : b;
}
template <class T>
E2D_HDI T fabs_(T a) { return __builtin_fabs(a); }
template <>
E2D_HDI float fabs_<float>(float a) { return __builtin_fabsf(a); }
// s * x for s = +-1, exactly
template <class T>
E2D_HDI T signed_by(bool positive, T x) { return positive ? x : -x; }

// Conservative -> primitive (computePrimitives + eos, m_utils.f90:35-64, 15-29).
// irho: M::recip of the density, the reciprocal its three quotients share,
// for later quotients by the same density (div_inv)
template <class T, class M>
E2D_HDI void cons_to_prim(const Hydro<T>& h, const T u[4], T q[4], T& irho) {
  const T rho = fmax2(u[0], h.smallr);
  irho = M::recip(rho);
  const T vx = M::div_inv(u[2], rho, irho);
  const T vy = M::div_inv(u[3], rho, irho);
  const T kinetic = T(0.5) * (vx * vx + vy * vy);
  const T eint = M::div_inv(u[1], rho, irho) - kinetic;
  q[0] = rho;
  q[1] = fmax2((h.gamma0 - T(1)) * rho * eint, rho * h.smallp);
  q[2] = vx;
  q[3] = vy;
}
template <class T, class M>
E2D_HDI void cons_to_prim(const Hydro<T>& h, const T u[4], T q[4]) {
  T irho;
  cons_to_prim<T, M>(h, u, q, irho);
}

// Sound speed of a primitive state (second half of eos) as cons_to_prim
// leaves it: q[0] >= smallr and q[1] >= q[0] * smallp, so the radicand is at
// least gamma0 * smallp > 0.
template <class T, class M>
E2D_HDI T sound_speed(const Hydro<T>& h, const T q[4]) {
  return M::sqrt_pos(M::div(h.gamma0 * q[1], q[0]));
}

// Local CFL rate (c+|u|)/dx + (c+|v|)/dy of a conservative state
// (compute_dt body, serial_ref/m_run.f90:278-281).
template <class T, class M>
E2D_HDI T cfl_rate(const Hydro<T>& h, const T u[4]) {
  T q[4];
  cons_to_prim<T, M>(h, u, q);
  const T c = sound_speed<T, M>(h, q);
  const T ax = c + fabs_(q[2]);
  const T ay = c + fabs_(q[3]);
  if (h.dxy_pow2) return ax * h.inv_dx + ay * h.inv_dy;
  return M::div(ax, h.dx) + M::div(ay, h.dy);
}

// Limited slope of one variable along one direction (slope_unsplit_hydro_2d,
// m_utils.f90:298-312): minmod for slope_type 1, MC for 2.  kUnit: st = 1,
// where st * x is x itself (exactly, signed zeros and NaN included), so the two
// products are skipped.
template <bool kUnit, class T>
E2D_HDI T limited_slope_st(T st, T qm, T q0, T qp) {
  const T dl = kUnit ? (q0 - qm) : st * (q0 - qm);
  const T dr = kUnit ? (qp - q0) : st * (qp - q0);
  const T dc = T(0.5) * (qp - qm);
  T lim = fmin2(fabs_(dl), fabs_(dr));
  if ((dl * dr) <= T(0)) lim = T(0);
  return signed_by(dc >= T(0), fmin2(lim, fabs_(dc)));
}
template <class T>
E2D_HDI T limited_slope(T st, T qm, T q0, T qp) {
  return limited_slope_st<false>(st, qm, q0, qp);
}

// Slopes of a cell in both directions.  slope_type 0 and values other than
// 1/2 give zero slopes (first order), m_utils.f90:289-337.  kMinmod: the
// caller knows slope_type == 1, so the limiter skips the products by st = 1
// and the first-order select.
template <bool kMinmod = false, class T>
E2D_HDI void cell_slopes(const Hydro<T>& h, const T q[4], const T qxp[4],
                                            const T qxm[4], const T qyp[4], const T qym[4],
                                            T dqx[4], T dqy[4]) {
  if (kMinmod) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      dqx[v] = limited_slope_st<true>(T(1), qxm[v], q[v], qxp[v]);
      dqy[v] = limited_slope_st<true>(T(1), qym[v], q[v], qyp[v]);
    }
    return;
  }
  const bool limited = (h.slope_type == 1 || h.slope_type == 2);
  const T st = T(h.slope_type);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const T sx = limited_slope(st, qxm[v], q[v], qxp[v]);
    const T sy = limited_slope(st, qym[v], q[v], qyp[v]);
    dqx[v] = limited ? sx : T(0);
    dqy[v] = limited ? sy : T(0);
  }
}

// Hancock predictor source terms of a cell (trace_unsplit_2d_along_dir,
// m_utils.f90:213-216): s = (sr0, sp0, su0, sv0).
// irho: M::recip(q[0]), as cons_to_prim leaves it
template <class T, class M>
E2D_HDI void trace_sources(const Hydro<T>& h, const T q[4], const T dqx[4],
                                              const T dqy[4], T s[4], T irho) {
  const T r = q[0], p = q[1], u = q[2], v = q[3];
  const T div_vel = dqx[2] + dqy[3];
  s[0] = -u * dqx[0] - v * dqy[0] - div_vel * r;
  s[1] = -u * dqx[1] - v * dqy[1] - div_vel * h.gamma0 * p;
  s[2] = -u * dqx[2] - v * dqy[2] - M::div_inv(dqx[1], r, irho);
  s[3] = -u * dqx[3] - v * dqy[3] - M::div_inv(dqy[1], r, irho);
}
template <class T, class M>
E2D_HDI void trace_sources(const Hydro<T>& h, const T q[4], const T dqx[4],
                                              const T dqy[4], T s[4]) {
  trace_sources<T, M>(h, q, dqx, dqy, s, M::recip(q[0]));
}

// Face state of a cell: q + side*dq/2 + s*dt_dir/2 with only the density
// floored (m_utils.f90:218-252).  side = +1 for the XMAX/YMAX face, -1 for
// XMIN/YMIN; dq and dtd are those of the face's direction.
template <class T>
E2D_HDI void face_state(const Hydro<T>& h, const T q[4], const T dq[4],
                                           const T s[4], T dtd, bool plus_side, T f[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const T half_dq = T(0.5) * dq[v];
    const T edge = plus_side ? q[v] + half_dq : q[v] - half_dq;
    // s (dtd / 2) is (s dtd) / 2 rounded alike (halving is exact in the
    // solver's range), with dtd / 2 formed once
    f[v] = edge + s[v] * (dtd * T(0.5));
  }
  f[0] = fmax2(h.smallr, f[0]);
}

// Flux of a Godunov state (cmpflx, m_utils.f90:498-526).
template <class T>
E2D_HDI void godunov_flux(const Hydro<T>& h, const T g[4], T f[4]) {
  const T mass = g[0] * g[2];
  f[0] = mass;
  f[2] = mass * g[2] + g[1];
  f[3] = mass * g[3];
  const T ekin = T(0.5) * g[0] * (g[2] * g[2] + g[3] * g[3]);
  const T etot = g[1] * h.entho + ekin;
  f[1] = g[2] * (etot + g[1]);
}

// Approximate two-shock Riemann solver (riemann_2d, m_utils.f90:349-492):
// Newton iterations on p*, star state, sampling at x/t = 0, transverse
// velocity upwinded by the contact.  Written as three pieces: floors and
// first guess; one Newton iteration; star state, sampling and flux.
template <class T>
struct TwoShock {
  T rl, ul, pl, rr, ur, pr, cl, cr, p_star;
  bool more;  // the convergence measure |delp| / (p* + smallpp) still above conv_tol
};

template <class T, class M>
E2D_HDI void two_shock_start(const Hydro<T>& h, const T ql[4], const T qr[4], TwoShock<T>& s) {
  s.rl = fmax2(ql[0], h.smallr);
  s.ul = ql[2];
  s.pl = fmax2(ql[1], s.rl * h.smallp);
  s.rr = fmax2(qr[0], h.smallr);
  s.ur = qr[2];
  s.pr = fmax2(qr[1], s.rr * h.smallp);
  // Lagrangian sound speeds squared and first guess
  s.cl = h.gamma0 * s.pl * s.rl;
  s.cr = h.gamma0 * s.pr * s.rr;
  // cl >= gamma0 * smallr^2 * smallp > 0: rl and pl are floored above
  const T wl = M::sqrt_pos(s.cl);
  const T wr = M::sqrt_pos(s.cr);
  s.p_star = fmax2(M::div((wr * s.pl + wl * s.pr) + wl * wr * (s.ul - s.ur), wl + wr), T(0));
  s.more = true;  // the reference starts from conv = 1 > conv_tol
}

// one Newton iteration: the next p* and whether its convergence measure
// |delp / (p* + smallpp)| is still above conv_tol (p* >= 0, so the divisor
// is positive; 1 + gamma6 (p* - p) / p >= 1 - gamma6 > 0 and the radicands
// are at least cl (1 - gamma6) > 0)
template <class T, class M>
E2D_HDI void two_shock_newton(const Hydro<T>& h, const TwoShock<T>& s, T& p_next, bool& more) {
  T inv_l, inv_r;
  const T wwl = M::sqrt_pos_inv(s.cl * (T(1) + M::div(h.gamma6 * (s.p_star - s.pl), s.pl)), inv_l);
  const T wwr = M::sqrt_pos_inv(s.cr * (T(1) + M::div(h.gamma6 * (s.p_star - s.pr), s.pr)), inv_r);
  const T zl = M::div(T(2) * wwl * wwl * wwl, wwl * wwl + s.cl);
  const T zr = M::div(T(2) * wwr * wwr * wwr, wwr * wwr + s.cr);
  const T usl = s.ul - M::div_inv(s.p_star - s.pl, wwl, inv_l);
  const T usr = s.ur + M::div_inv(s.p_star - s.pr, wwr, inv_r);
  const T delp = fmax2(M::div(zr * zl, zr + zl) * (usl - usr), -s.p_star);
  p_next = s.p_star + delp;
  more = M::quot_gt(fabs_(delp), p_next + h.smallpp, h.conv_tol, h.conv_half_ulp);
}

template <class T, class M>
E2D_HDI void two_shock_finish(const Hydro<T>& h, const T ql[4], const T qr[4], const TwoShock<T>& s, T g[4],
                              T f[4]) {
  const T rl = s.rl, ul = s.ul, pl = s.pl, rr = s.rr, ur = s.ur, pr = s.pr, p_star = s.p_star;
  T inv_l, inv_r;
  const T wl = M::sqrt_pos_inv(s.cl * (T(1) + M::div(h.gamma6 * (p_star - pl), pl)), inv_l);
  const T wr = M::sqrt_pos_inv(s.cr * (T(1) + M::div(h.gamma6 * (p_star - pr), pr)), inv_r);
  const T u_star = T(0.5) * (ul + M::div_inv(pl - p_star, wl, inv_l) + ur - M::div_inv(pr - p_star, wr, inv_r));

  // upwind side of the contact (sgnm = +-1)
  const bool right_going = u_star >= T(0);
  const T ro = right_going ? rl : rr;
  const T uo = right_going ? ul : ur;
  const T po = right_going ? pl : pr;
  const T wo = right_going ? wl : wr;
  // 1/ro serves three quotients.  1/wo^2 starts from (1/wo)^2 and 1/r* from
  // (1 + E)/ro (r* = ro / (1 + E) unless floored; 1 + E lies between
  // (gamma0 - 1)/(gamma0 + 1) and its inverse): each start is within a few
  // ulp, and one Newton step makes it a reciprocal as close as recip's, so
  // the quotients are the same bits.
  const T rro = M::recip(ro);
  // po / ro >= smallp (the same side's floors); p* / r* may be zero
  const T co = fmax2(h.smallc, M::sqrt_pos(fabs_(M::div_inv(h.gamma0 * po, ro, rro))));
  const T wo2 = wo * wo;
  const T iwo = right_going ? inv_l : inv_r;
  const T D = T(1) + M::div_inv(ro * (po - p_star), wo2, M::recip_near(wo2, iwo * iwo));
  const T r_raw = M::div(ro, D);
  const T r_star = fmax2(r_raw, h.smallr);
  const T ir0 = r_raw > h.smallr ? D * rro : h.inv_smallr;
  const T c_star = fmax2(h.smallc, M::sqrt(fabs_(M::div_inv(h.gamma0 * p_star, r_star, M::recip_near(r_star, ir0)))));

  const T s_uo = signed_by(right_going, uo);
  T spout = co - s_uo;
  T spin = c_star - signed_by(right_going, u_star);
  const T ushock = M::div_inv(wo, ro, rro) - s_uo;
  if (p_star >= po) {
    spin = ushock;
    spout = ushock;
  }
  const T scr = fmax2(spout - spin, h.smallc + fabs_(spout + spin));
  const T frac = T(0.5) * (T(1) + M::div(spout + spin, scr));

  T gr = frac * r_star + (T(1) - frac) * ro;
  T gu = frac * u_star + (T(1) - frac) * uo;
  T gp = frac * p_star + (T(1) - frac) * po;
  if (spout < T(0)) {
    gr = ro; gu = uo; gp = po;
  }
  if (spin > T(0)) {
    gr = r_star; gu = u_star; gp = p_star;
  }
  g[0] = gr;
  g[1] = gp;
  g[2] = gu;
  g[3] = right_going ? ql[3] : qr[3];
  godunov_flux(h, g, f);
}

template <class T, class M>
E2D_HDI void riemann_two_shock(const Hydro<T>& h, const T ql[4], const T qr[4],
                                                  T g[4], T f[4]) {
  TwoShock<T> s;
  two_shock_start<T, M>(h, ql, qr, s);
  for (int it = 0; it < h.niter && s.more; ++it) two_shock_newton<T, M>(h, s, s.p_star, s.more);
  two_shock_finish<T, M>(h, ql, qr, s, g, f);
}

// HLLC solver (riemann_2d_hllc, m_utils.f90:534-644).  The reference never
// selects it in the step; here riemann_solver == 2 does.
template <class T, class M>
E2D_HDI void riemann_hllc(const Hydro<T>& h, const T ql[4], const T qr[4],
                                             T g[4], T f[4]) {
  const T H = T(0.5);
  const T rl = fmax2(ql[0], h.smallr);
  const T pl = fmax2(ql[1], rl * h.smallp);
  const T ul = ql[2];
  T ekl = H * rl * ul * ul;
  ekl = ekl + H * rl * ql[3] * ql[3];
  const T el = pl * h.entho + ekl;
  const T rr = fmax2(qr[0], h.smallr);
  const T pr = fmax2(qr[1], rr * h.smallp);
  const T ur = qr[2];
  T ekr = H * rr * ur * ur;
  ekr = ekr + H * rr * qr[3] * qr[3];
  const T er = pr * h.entho + ekr;

  const T cfl_ = M::sqrt(fmax2(M::div(h.gamma0 * pl, rl), h.smallc * h.smallc));
  const T cfr = M::sqrt(fmax2(M::div(h.gamma0 * pr, rr), h.smallc * h.smallc));
  const T cmax = fmax2(cfl_, cfr);
  const T SL = fmin2(ul, ur) - cmax;
  const T SR = fmax2(ul, ur) + cmax;
  const T rcl = rl * (ul - SL);
  const T rcr = rr * (SR - ur);
  const T ustar = M::div(rcr * ur + rcl * ul + (pl - pr), rcr + rcl);
  const T pstar = M::div(rcr * pl + rcl * pr + rcl * rcr * (ul - ur), rcr + rcl);
  const T rstl = M::div(rl * (SL - ul), SL - ustar);
  const T estl = M::div((SL - ul) * el - pl * ul + pstar * ustar, SL - ustar);
  const T rstr = M::div(rr * (SR - ur), SR - ustar);
  const T estr = M::div((SR - ur) * er - pr * ur + pstar * ustar, SR - ustar);

  T ro, uo, po, eo;
  if (SL > T(0)) {
    ro = rl; uo = ul; po = pl; eo = el;
  } else if (ustar > T(0)) {
    ro = rstl; uo = ustar; po = pstar; eo = estl;
  } else if (SR > T(0)) {
    ro = rstr; uo = ustar; po = pstar; eo = estr;
  } else {
    ro = rr; uo = ur; po = pr; eo = er;
  }
  g[0] = g[1] = g[2] = g[3] = T(0);
  f[0] = ro * uo;
  f[2] = ro * uo * uo + po;
  f[1] = (eo + po) * uo;
  f[3] = f[0] > T(0) ? f[0] * ql[3] : f[0] * qr[3];
}

// Solver selection at compile time (riemannSolverType, serial_ref/m_constants.f90:21-23).
// HLLC divides by wave-speed differences that can vanish, so it always keeps
// IEEE division (the reference's inf/NaN behaviour included).
template <class T, class M, int SOLVER>
E2D_HDI void riemann(const Hydro<T>& h, const T ql[4], const T qr[4], T g[4], T f[4]) {
  if constexpr (SOLVER == 2)
    riemann_hllc<T, ExactMath>(h, ql, qr, g, f);
  else
    riemann_two_shock<T, M>(h, ql, qr, g, f);
}

// Flux across a y-face: the x-solver with the velocity components exchanged
// on input and output (serial_ref/m_run.f90:463-470).
template <class T, class M, int SOLVER>
E2D_HDI void riemann_y(const Hydro<T>& h, const T ql[4], const T qr[4], T f[4]) {
  const T l[4] = {ql[0], ql[1], ql[3], ql[2]};
  const T r[4] = {qr[0], qr[1], qr[3], qr[2]};
  T g[4], fs[4];
  riemann<T, M, SOLVER>(h, l, r, g, fs);
  f[0] = fs[0];
  f[1] = fs[1];
  f[2] = fs[3];
  f[3] = fs[2];
}

}  // namespace e2d
