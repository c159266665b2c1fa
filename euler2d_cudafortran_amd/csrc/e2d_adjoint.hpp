// Tangent-linear and adjoint building blocks of the Godunov step.
//
// The linearisation is the one of the step as executed: every max/min floor,
// limiter branch, sampling branch and the Newton iteration count of the
// two-shock solver are taken as the forward pass takes them (fixed
// subgradients), and dt is frozen.  |x| has derivative +1 at x = 0.
//
// Tangent: `Dual` carries (value, derivative) through the same templated
// numerics as the forward step (e2d_numerics.hpp), with DualMath<Base> giving
// division and square root; values are computed with Base, so they and all
// branch decisions equal the forward step's.
//
// Adjoint: hand-written reverses of cons_to_prim, limited_slope,
// trace_sources, face_state, godunov_flux and riemann_two_shock.  Each
// recomputes its forward values and accumulates (+=) into the bars of its
// inputs.
#pragma once

#include "e2d_numerics.hpp"

namespace e2d {

// ---------------------------------------------------------------------------
// dual numbers

struct Dual {
  double v, d;
  E2D_HDI Dual() : v(0.0), d(0.0) {}
  E2D_HDI Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote implicitly
  E2D_HDI Dual(double x, double dx) : v(x), d(dx) {}
  E2D_HDI explicit Dual(int x) : v(static_cast<double>(x)), d(0.0) {}
  E2D_HDI explicit operator double() const { return v; }
};

E2D_HDI Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
E2D_HDI Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
E2D_HDI Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
E2D_HDI Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
E2D_HDI Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
E2D_HDI bool operator>(Dual a, Dual b) { return a.v > b.v; }
E2D_HDI bool operator<(Dual a, Dual b) { return a.v < b.v; }
E2D_HDI bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
E2D_HDI bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
E2D_HDI bool operator==(Dual a, Dual b) { return a.v == b.v; }
template <>
E2D_HDI Dual fabs_<Dual>(Dual a) { return a.v < 0.0 ? -a : a; }

template <class Base>
struct DualMath {
  static E2D_HDI Dual div(Dual a, Dual b) {
    const double q = Base::div(a.v, b.v);
    return Dual(q, (a.d - q * b.d) / b.v);
  }
  static E2D_HDI Dual sqrt(Dual a) {
    const double s = Base::sqrt(a.v);
    return Dual(s, s == 0.0 ? 0.0 : a.d / (2.0 * s));
  }
  static E2D_HDI Dual sqrt_pos(Dual a) {
    const double s = Base::sqrt_pos(a.v);
    return Dual(s, a.d / (2.0 * s));
  }
  static E2D_HDI Dual sqrt_pos_inv(Dual a, Dual& inv) {
    inv = Dual(0.0, 0.0);
    return sqrt_pos(a);
  }
  static E2D_HDI Dual div_inv(Dual a, Dual b, Dual) { return div(a, b); }
  static E2D_HDI Dual recip(Dual) { return Dual(0.0, 0.0); }
  static E2D_HDI Dual recip_near(Dual, Dual) { return Dual(0.0, 0.0); }
  static E2D_HDI bool quot_gt(Dual a, Dual d, Dual tol, Dual h) { return Base::quot_gt(a.v, d.v, tol.v, h.v); }
};

template <class T>
E2D_HDI Hydro<Dual> dual_hydro(const Hydro<T>& h) {
  Hydro<Dual> r;
  r.gamma0 = h.gamma0;
  r.gamma6 = h.gamma6;
  r.cfl = h.cfl;
  r.smallr = h.smallr;
  r.smallc = h.smallc;
  r.smallp = h.smallp;
  r.smallpp = h.smallpp;
  r.entho = h.entho;
  r.dx = h.dx;
  r.dy = h.dy;
  r.inv_dx = h.inv_dx;
  r.inv_dy = h.inv_dy;
  r.dxy_pow2 = h.dxy_pow2;
  r.conv_tol = h.conv_tol;
  r.inv_smallr = h.inv_smallr;
  r.conv_half_ulp = h.conv_half_ulp;
  r.slope_type = h.slope_type;
  r.niter = h.niter;
  r.riemann_solver = h.riemann_solver;
  return r;
}

// ---------------------------------------------------------------------------
// adjoints, in the arithmetic type T of the step (double, or float for the
// FP32 variants; the caller accumulates the adjoint state itself in double)

template <class T>
E2D_HDI T sgn(T x) { return x >= T(0.0) ? T(1.0) : -T(1.0); }

// cons_to_prim: u -> q (rho, p, vx, vy)
template <class M, class T>
E2D_HDI void cons_to_prim_adj(const Hydro<T>& h, const T u[4], const T qb[4], T ub[4]) {
  const T rho = fmax2(u[0], h.smallr);
  const T vx = M::div(u[2], rho);
  const T vy = M::div(u[3], rho);
  const T kin = T(0.5) * (vx * vx + vy * vy);
  const T e0 = M::div(u[1], rho);
  const T eint = e0 - kin;
  const T gm1 = h.gamma0 - T(1.0);
  const T pa = gm1 * rho * eint;
  const T pb = rho * h.smallp;
  T rho_b = qb[0], vx_b = qb[2], vy_b = qb[3];
  T pa_b = T(0.0), pb_b = T(0.0);
  if (pa > pb)
    pa_b = qb[1];
  else
    pb_b = qb[1];
  rho_b += pa_b * gm1 * eint + pb_b * h.smallp;
  const T eint_b = pa_b * gm1 * rho;
  const T e0_b = eint_b, kin_b = -eint_b;
  const T irho = M::div(T(1), rho);
  ub[1] += e0_b * irho;
  rho_b -= e0_b * e0 * irho;
  vx_b += kin_b * vx;
  vy_b += kin_b * vy;
  ub[2] += vx_b * irho;
  ub[3] += vy_b * irho;
  rho_b -= (vx_b * vx + vy_b * vy) * irho;
  if (u[0] > h.smallr) ub[0] += rho_b;
}

// limited_slope(st, qm, q0, qp): bars of (qm, q0, qp) from the bar of the slope
template <class T>
E2D_HDI void limited_slope_adj(T st, T qm, T q0, T qp, T ob, T& qm_b,
                               T& q0_b, T& qp_b) {
  const T dl = st * (q0 - qm);
  const T dr = st * (qp - q0);
  const T dc = T(0.5) * (qp - qm);
  const T adl = fabs_(dl), adr = fabs_(dr), adc = fabs_(dc);
  const T lim0 = fmin2(adl, adr);
  const T lim = ((dl * dr) <= T(0.0)) ? T(0.0) : lim0;
  const T m_b = (dc >= T(0.0)) ? ob : -ob;
  T lim_b = T(0.0), adc_b = T(0.0);
  if (lim < adc)
    lim_b = m_b;
  else
    adc_b = m_b;
  const T lim0_b = ((dl * dr) <= T(0.0)) ? T(0.0) : lim_b;
  T adl_b = T(0.0), adr_b = T(0.0);
  if (adl < adr)
    adl_b = lim0_b;
  else
    adr_b = lim0_b;
  const T dl_b = adl_b * sgn(dl), dr_b = adr_b * sgn(dr), dc_b = adc_b * sgn(dc);
  q0_b += st * dl_b - st * dr_b;
  qm_b += -st * dl_b - T(0.5) * dc_b;
  qp_b += st * dr_b + T(0.5) * dc_b;
}

template <class T>
E2D_HDI bool slopes_limited(const Hydro<T>& h) { return h.slope_type == 1 || h.slope_type == 2; }

// trace_sources: (q, dqx, dqy) -> s
template <class M, class T>
E2D_HDI void trace_sources_adj(const Hydro<T>& h, const T q[4], const T dqx[4],
                               const T dqy[4], const T sb[4], T qb[4], T dqxb[4],
                               T dqyb[4]) {
  const T r = q[0], p = q[1], u = q[2], v = q[3];
  const T dv = dqx[2] + dqy[3];
  qb[2] += -(sb[0] * dqx[0] + sb[1] * dqx[1] + sb[2] * dqx[2] + sb[3] * dqx[3]);
  qb[3] += -(sb[0] * dqy[0] + sb[1] * dqy[1] + sb[2] * dqy[2] + sb[3] * dqy[3]);
  for (int k = 0; k < 4; ++k) {
    dqxb[k] += -u * sb[k];
    dqyb[k] += -v * sb[k];
  }
  const T ir = M::div(T(1), r);
  dqxb[1] += -sb[2] * ir;
  dqyb[1] += -sb[3] * ir;
  const T dv_b = -sb[0] * r - sb[1] * h.gamma0 * p;
  dqxb[2] += dv_b;
  dqyb[3] += dv_b;
  qb[0] += -sb[0] * dv + (sb[2] * dqx[1] + sb[3] * dqy[1]) * (ir * ir);
  qb[1] += -sb[1] * dv * h.gamma0;
}

// face_state: (q, dq, s) -> f, density floored
template <class T>
E2D_HDI void face_state_adj(const Hydro<T>& h, const T q[4], const T dq[4], const T s[4],
                            T dtd, bool plus_side, const T fb[4], T qb[4], T dqb[4],
                            T sb[4]) {
  const T half_dq0 = T(0.5) * dq[0];
  const T edge0 = plus_side ? q[0] + half_dq0 : q[0] - half_dq0;
  const T f0raw = edge0 + s[0] * (dtd * T(0.5));  // as face_state forms it
  for (int k = 0; k < 4; ++k) {
    const T b = (k == 0 && h.smallr > f0raw) ? T(0.0) : fb[k];
    qb[k] += b;
    dqb[k] += plus_side ? T(0.5) * b : -T(0.5) * b;
    sb[k] += T(0.5) * dtd * b;
  }
}

// godunov_flux: g -> f
template <class T>
E2D_HDI void godunov_flux_adj(const Hydro<T>& h, const T g[4], const T fb[4], T gb[4]) {
  const T r = g[0], p = g[1], u = g[2], v = g[3];
  const T mass = r * u;
  const T ekin = T(0.5) * r * (u * u + v * v);
  const T etot = p * h.entho + ekin;
  // f1 = u * (etot + p)
  gb[2] += fb[1] * (etot + p);
  const T etot_b = fb[1] * u;
  gb[1] += fb[1] * u + etot_b * h.entho;
  gb[0] += etot_b * T(0.5) * (u * u + v * v);
  gb[2] += etot_b * r * u;
  gb[3] += etot_b * r * v;
  // f0 = mass, f2 = mass*u + p, f3 = mass*v
  const T mass_b = fb[0] + fb[2] * u + fb[3] * v;
  gb[2] += fb[2] * mass;
  gb[1] += fb[2];
  gb[3] += fb[3] * mass;
  gb[0] += mass_b * u;
  gb[2] += mass_b * r;
}

// One Newton iteration of riemann_two_shock from p_k; returns p_{k+1}.  (The
// square roots keep their zero case here: without it, the FP64 adjoint
// kernel is scheduled 1.5% slower.)
template <class M, class T>
E2D_HDI T newton_step(const Hydro<T>& h, T pk, T cl, T cr, T pl, T pr,
                           T ul, T ur, bool& more, bool& take_t) {
  const T wwl = M::sqrt(cl * (T(1.0) + M::div(h.gamma6 * (pk - pl), pl)));
  const T wwr = M::sqrt(cr * (T(1.0) + M::div(h.gamma6 * (pk - pr), pr)));
  const T zl = M::div(T(2.0) * wwl * wwl * wwl, wwl * wwl + cl);
  const T zr = M::div(T(2.0) * wwr * wwr * wwr, wwr * wwr + cr);
  const T usl = ul - M::div(pk - pl, wwl);
  const T usr = ur + M::div(pk - pr, wwr);
  const T t = M::div(zr * zl, zr + zl) * (usl - usr);
  take_t = t > -pk;
  const T delp = take_t ? t : -pk;
  const T pn = pk + delp;
  more = M::quot_gt(fabs_(delp), pn + h.smallpp, h.conv_tol, h.conv_half_ulp);
  return pn;
}

// Reverse of one Newton iteration at p_k: given pn_b (bar of p_{k+1}), returns
// the bar of p_k and accumulates the bars of (cl, cr, pl, pr, ul, ur).
template <class M, class T>
E2D_HDI T newton_step_adj(const Hydro<T>& h, T pk, T cl, T cr, T pl, T pr,
                               T ul, T ur, T pn_b, T& cl_b, T& cr_b, T& pl_b,
                               T& pr_b, T& ul_b, T& ur_b) {
  const T g6 = h.gamma6;
  const T xl = M::div(g6 * (pk - pl), pl), xr = M::div(g6 * (pk - pr), pr);
  const T wwl = M::sqrt(cl * (T(1.0) + xl));
  const T wwr = M::sqrt(cr * (T(1.0) + xr));
  const T dzl = wwl * wwl + cl, dzr = wwr * wwr + cr;
  const T zl = M::div(T(2.0) * wwl * wwl * wwl, dzl);
  const T zr = M::div(T(2.0) * wwr * wwr * wwr, dzr);
  const T c1 = M::div(pk - pl, wwl), c2 = M::div(pk - pr, wwr);
  const T usl = ul - c1, usr = ur + c2;
  const T S = zr + zl;
  const T m = M::div(zr * zl, S);
  const T t = m * (usl - usr);
  T pk_b = pn_b;  // p_{k+1} = p_k + delp
  const T delp_b = pn_b;
  T t_b = T(0.0);
  if (t > -pk)
    t_b = delp_b;
  else
    pk_b -= delp_b;
  const T m_b = t_b * (usl - usr);
  const T usl_b = t_b * m, usr_b = -t_b * m;
  const T iS = M::div(T(1), S), iwwl = M::div(T(1), wwl), iwwr = M::div(T(1), wwr);
  const T idzl = M::div(T(1), dzl), idzr = M::div(T(1), dzr);
  const T ipl = M::div(T(1), pl), ipr = M::div(T(1), pr);
  const T pm_b = m_b * iS, s_b = -m_b * m * iS;
  const T zr_b = pm_b * zl + s_b, zl_b = pm_b * zr + s_b;
  // usl = ul - (pk - pl)/wwl ; usr = ur + (pk - pr)/wwr
  ul_b += usl_b;
  const T c1_b = -usl_b;
  pk_b += c1_b * iwwl;
  pl_b -= c1_b * iwwl;
  T wwl_b = -c1_b * c1 * iwwl;
  ur_b += usr_b;
  const T c2_b = usr_b;
  pk_b += c2_b * iwwr;
  pr_b -= c2_b * iwwr;
  T wwr_b = -c2_b * c2 * iwwr;
  // z = 2 ww^3 / (ww^2 + c)
  wwl_b += zl_b * idzl * T(6.0) * wwl * wwl - zl_b * zl * idzl * T(2.0) * wwl;
  cl_b += -zl_b * zl * idzl;
  wwr_b += zr_b * idzr * T(6.0) * wwr * wwr - zr_b * zr * idzr * T(2.0) * wwr;
  cr_b += -zr_b * zr * idzr;
  // ww = sqrt(c (1 + g6 (pk - p0)/p0))
  const T al_b = wwl_b * T(0.5) * iwwl, ar_b = wwr_b * T(0.5) * iwwr;
  cl_b += al_b * (T(1.0) + xl);
  cr_b += ar_b * (T(1.0) + xr);
  const T xl_b = al_b * cl, xr_b = ar_b * cr;
  pk_b += xl_b * g6 * ipl + xr_b * g6 * ipr;
  pl_b += -xl_b * g6 * pk * (ipl * ipl);
  pr_b += -xr_b * g6 * pk * (ipr * ipr);
  return pk_b;
}

// riemann_two_shock adjoint: bars of (ql, qr) from the bar of the flux
template <class M, class T>
E2D_HDI void riemann_two_shock_adj(const Hydro<T>& h, const T ql[4], const T qr[4],
                                   const T fb[4], T qlb[4], T qrb[4]) {
  const T g = h.gamma0, g6 = h.gamma6;
  // ---- forward
  const T rl = fmax2(ql[0], h.smallr);
  const T ul = ql[2];
  const T plf = rl * h.smallp;
  const T pl = fmax2(ql[1], plf);
  const T rr = fmax2(qr[0], h.smallr);
  const T ur = qr[2];
  const T prf = rr * h.smallp;
  const T pr = fmax2(qr[1], prf);
  const T cl = g * pl * rl;
  const T cr = g * pr * rr;
  const T wl0 = M::sqrt(cl);
  const T wr0 = M::sqrt(cr);
  const T num = (wr0 * pl + wl0 * pr) + wl0 * wr0 * (ul - ur);
  const T den = wl0 + wr0;
  const T p0r = M::div(num, den);
  const T p0 = fmax2(p0r, T(0.0));
  T ps = p0;
  bool more = true;
  int K = 0;
  for (; K < h.niter && more; ++K) {
    bool tt;
    ps = newton_step<M>(h, ps, cl, cr, pl, pr, ul, ur, more, tt);
  }
  const T xlf = M::div(g6 * (ps - pl), pl), xrf = M::div(g6 * (ps - pr), pr);
  const T wl = M::sqrt(cl * (T(1.0) + xlf));
  const T wr = M::sqrt(cr * (T(1.0) + xrf));
  const T a_us = M::div(pl - ps, wl), b_us = M::div(pr - ps, wr);
  const T us = T(0.5) * (ul + a_us + ur - b_us);
  const bool right = us >= T(0.0);
  const T ro = right ? rl : rr, uo = right ? ul : ur, po = right ? pl : pr, wo = right ? wl : wr;
  const T sg = right ? T(1.0) : -T(1.0);
  const T Y = M::div(g * po, ro);
  const T co_r = M::sqrt(fabs_(Y));
  const T co = fmax2(h.smallc, co_r);
  const T E = M::div(ro * (po - ps), wo * wo);
  const T D = T(1.0) + E;
  const T rs_r = M::div(ro, D);
  const T rs = fmax2(rs_r, h.smallr);
  const T X = M::div(g * ps, rs);
  const T cs_r = M::sqrt(fabs_(X));
  const T cs = fmax2(h.smallc, cs_r);
  const T s_uo = sg * uo;
  T spout = co - s_uo;
  T spin = cs - sg * us;
  const T ush = M::div(wo, ro) - s_uo;
  const bool shock = ps >= po;
  if (shock) {
    spin = ush;
    spout = ush;
  }
  const T sum = spout + spin, dif = spout - spin;
  const T scrb = h.smallc + fabs_(sum);
  const bool take_dif = dif > scrb;
  const T scr = take_dif ? dif : scrb;
  const T frac = T(0.5) * (T(1.0) + M::div(sum, scr));
  const bool outside = spout < T(0.0);
  const bool inside = spin > T(0.0);
  T gs[4];
  gs[0] = frac * rs + (T(1.0) - frac) * ro;
  gs[2] = frac * us + (T(1.0) - frac) * uo;
  gs[1] = frac * ps + (T(1.0) - frac) * po;
  if (outside) {
    gs[0] = ro;
    gs[2] = uo;
    gs[1] = po;
  }
  if (inside) {
    gs[0] = rs;
    gs[2] = us;
    gs[1] = ps;
  }
  gs[3] = right ? ql[3] : qr[3];

  // ---- reverse (derivative arithmetic with one reciprocal per denominator)
  const T iro = M::div(T(1), ro), irs = M::div(T(1), rs);
  const T ipl = M::div(T(1), pl), ipr = M::div(T(1), pr);
  T gb[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  godunov_flux_adj(h, gs, fb, gb);
  if (right)
    qlb[3] += gb[3];
  else
    qrb[3] += gb[3];
  T rs_b = T(0.0), us_b = T(0.0), ps_b = T(0.0), ro_b = T(0.0), uo_b = T(0.0), po_b = T(0.0), wo_b = T(0.0);
  if (inside) {
    rs_b += gb[0];
    us_b += gb[2];
    ps_b += gb[1];
  } else if (outside) {
    ro_b += gb[0];
    uo_b += gb[2];
    po_b += gb[1];
  } else {
    const T frac_b = gb[0] * (rs - ro) + gb[2] * (us - uo) + gb[1] * (ps - po);
    rs_b += gb[0] * frac;
    ro_b += gb[0] * (T(1.0) - frac);
    us_b += gb[2] * frac;
    uo_b += gb[2] * (T(1.0) - frac);
    ps_b += gb[1] * frac;
    po_b += gb[1] * (T(1.0) - frac);
    // frac = 0.5 (1 + sum/scr)
    const T iscr = M::div(T(1), scr);
    T sum_b = frac_b * T(0.5) * iscr;
    const T scr_b = -frac_b * T(0.5) * sum * (iscr * iscr);
    T dif_b = T(0.0);
    if (take_dif)
      dif_b = scr_b;
    else
      sum_b += scr_b * sgn(sum);
    const T spout_b = sum_b + dif_b, spin_b = sum_b - dif_b;
    T co_b = T(0.0), cs_b = T(0.0), s_uo_b = T(0.0);
    if (shock) {
      const T ush_b = spout_b + spin_b;
      wo_b += ush_b * iro;
      ro_b -= ush_b * wo * (iro * iro);
      s_uo_b -= ush_b;
    } else {
      co_b += spout_b;
      s_uo_b -= spout_b;
      cs_b += spin_b;
      us_b -= spin_b * sg;
    }
    uo_b += s_uo_b * sg;
    // cs = max(smallc, sqrt|g ps / rs|)
    if (!(h.smallc > cs_r) && cs_r > T(0.0)) {
      const T X_b = cs_b * T(0.5) * M::div(T(1), cs_r) * sgn(X);
      ps_b += X_b * g * irs;
      rs_b -= X_b * g * ps * (irs * irs);
    }
    // co = max(smallc, sqrt|g po / ro|)
    if (!(h.smallc > co_r) && co_r > T(0.0)) {
      const T Y_b = co_b * T(0.5) * M::div(T(1), co_r) * sgn(Y);
      po_b += Y_b * g * iro;
      ro_b -= Y_b * g * po * (iro * iro);
    }
  }
  // rs = max(ro / (1 + ro (po - ps) / wo^2), smallr)
  if (rs_r > h.smallr) {
    const T rsr_b = rs_b;
    const T iD = M::div(T(1), D);
    ro_b += rsr_b * iD;
    const T E_b = -rsr_b * ro * (iD * iD);
    const T iw2 = M::div(T(1), wo * wo);
    ro_b += E_b * (po - ps) * iw2;
    po_b += E_b * ro * iw2;
    ps_b -= E_b * ro * iw2;
    wo_b -= E_b * T(2.0) * E * M::div(T(1), wo);
  }
  T rl_b = T(0.0), ul_b = T(0.0), pl_b = T(0.0), wl_b = T(0.0);
  T rr_b = T(0.0), ur_b = T(0.0), pr_b = T(0.0), wr_b = T(0.0);
  if (right) {
    rl_b += ro_b;
    ul_b += uo_b;
    pl_b += po_b;
    wl_b += wo_b;
  } else {
    rr_b += ro_b;
    ur_b += uo_b;
    pr_b += po_b;
    wr_b += wo_b;
  }
  // us = 0.5 (ul + (pl - ps)/wl + ur - (pr - ps)/wr)
  ul_b += T(0.5) * us_b;
  ur_b += T(0.5) * us_b;
  const T a_b = T(0.5) * us_b, b_b = -T(0.5) * us_b;
  const T iwl = M::div(T(1), wl), iwr = M::div(T(1), wr);
  pl_b += a_b * iwl;
  ps_b -= a_b * iwl;
  wl_b -= a_b * a_us * iwl;
  pr_b += b_b * iwr;
  ps_b -= b_b * iwr;
  wr_b -= b_b * b_us * iwr;
  // wl = sqrt(cl (1 + g6 (ps - pl)/pl))
  T cl_b = T(0.0), cr_b = T(0.0);
  {
    const T W_b = wl_b * T(0.5) * iwl;
    cl_b += W_b * (T(1.0) + xlf);
    const T x_b = W_b * cl;
    ps_b += x_b * g6 * ipl;
    pl_b += -x_b * g6 * ps * (ipl * ipl);
  }
  {
    const T W_b = wr_b * T(0.5) * iwr;
    cr_b += W_b * (T(1.0) + xrf);
    const T x_b = W_b * cr;
    ps_b += x_b * g6 * ipr;
    pr_b += -x_b * g6 * ps * (ipr * ipr);
  }
  // Newton iterations, last to first (p_k recomputed from p_0)
  T p_b = ps_b;
  for (int k = K - 1; k >= 0; --k) {
    T pk = p0;
    bool cv, tt;
    for (int j = 0; j < k; ++j) pk = newton_step<M>(h, pk, cl, cr, pl, pr, ul, ur, cv, tt);
    p_b = newton_step_adj<M>(h, pk, cl, cr, pl, pr, ul, ur, p_b, cl_b, cr_b, pl_b, pr_b, ul_b, ur_b);
  }
  // p0 = max(num/den, 0)
  T wl0_b = T(0.0), wr0_b = T(0.0);
  if (p0r > T(0.0)) {
    const T iden = M::div(T(1), den);
    const T num_b = p_b * iden;
    const T den_b = -p_b * p0r * iden;
    wr0_b += num_b * pl + num_b * wl0 * (ul - ur) + den_b;
    pl_b += num_b * wr0;
    wl0_b += num_b * pr + num_b * wr0 * (ul - ur) + den_b;
    pr_b += num_b * wl0;
    ul_b += num_b * wl0 * wr0;
    ur_b -= num_b * wl0 * wr0;
  }
  cl_b += wl0_b * T(0.5) * M::div(T(1), wl0);
  cr_b += wr0_b * T(0.5) * M::div(T(1), wr0);
  // c = g p r
  pl_b += cl_b * g * rl;
  rl_b += cl_b * g * pl;
  pr_b += cr_b * g * rr;
  rr_b += cr_b * g * pr;
  // p = max(q_p, r smallp); r = max(q_r, smallr)
  if (ql[1] > plf)
    qlb[1] += pl_b;
  else
    rl_b += pl_b * h.smallp;
  if (qr[1] > prf)
    qrb[1] += pr_b;
  else
    rr_b += pr_b * h.smallp;
  if (ql[0] > h.smallr) qlb[0] += rl_b;
  if (qr[0] > h.smallr) qrb[0] += rr_b;
  qlb[2] += ul_b;
  qrb[2] += ur_b;
}

// y-face: velocities exchanged on input and output (serial_ref/m_run.f90:463-470)
template <class M, class T>
E2D_HDI void riemann_y_adj(const Hydro<T>& h, const T ql[4], const T qr[4], const T fb[4],
                           T qlb[4], T qrb[4]) {
  const T l[4] = {ql[0], ql[1], ql[3], ql[2]};
  const T r[4] = {qr[0], qr[1], qr[3], qr[2]};
  const T fbs[4] = {fb[0], fb[1], fb[3], fb[2]};
  T lb[4] = {T(0.0), T(0.0), T(0.0), T(0.0)}, rb[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  riemann_two_shock_adj<M>(h, l, r, fbs, lb, rb);
  qlb[0] += lb[0];
  qlb[1] += lb[1];
  qlb[2] += lb[3];
  qlb[3] += lb[2];
  qrb[0] += rb[0];
  qrb[1] += rb[1];
  qrb[2] += rb[3];
  qrb[3] += rb[2];
}

}  // namespace e2d
