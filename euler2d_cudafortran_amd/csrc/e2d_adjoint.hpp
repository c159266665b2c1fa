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
  r.conv_tol = h.conv_tol;
  r.slope_type = h.slope_type;
  r.niter = h.niter;
  r.riemann_solver = h.riemann_solver;
  return r;
}

// ---------------------------------------------------------------------------
// adjoints (double)

E2D_HDI double sgn(double x) { return x >= 0.0 ? 1.0 : -1.0; }

// cons_to_prim: u -> q (rho, p, vx, vy)
template <class M>
E2D_HDI void cons_to_prim_adj(const Hydro<double>& h, const double u[4], const double qb[4], double ub[4]) {
  const double rho = fmax2(u[0], h.smallr);
  const double vx = M::div(u[2], rho);
  const double vy = M::div(u[3], rho);
  const double kin = 0.5 * (vx * vx + vy * vy);
  const double e0 = M::div(u[1], rho);
  const double eint = e0 - kin;
  const double gm1 = h.gamma0 - 1.0;
  const double pa = gm1 * rho * eint;
  const double pb = rho * h.smallp;
  double rho_b = qb[0], vx_b = qb[2], vy_b = qb[3];
  double pa_b = 0.0, pb_b = 0.0;
  if (pa > pb)
    pa_b = qb[1];
  else
    pb_b = qb[1];
  rho_b += pa_b * gm1 * eint + pb_b * h.smallp;
  const double eint_b = pa_b * gm1 * rho;
  const double e0_b = eint_b, kin_b = -eint_b;
  ub[1] += e0_b / rho;
  rho_b -= e0_b * e0 / rho;
  vx_b += kin_b * vx;
  vy_b += kin_b * vy;
  ub[2] += vx_b / rho;
  ub[3] += vy_b / rho;
  rho_b -= (vx_b * vx + vy_b * vy) / rho;
  if (u[0] > h.smallr) ub[0] += rho_b;
}

// limited_slope(st, qm, q0, qp): bars of (qm, q0, qp) from the bar of the slope
E2D_HDI void limited_slope_adj(double st, double qm, double q0, double qp, double ob, double& qm_b,
                               double& q0_b, double& qp_b) {
  const double dl = st * (q0 - qm);
  const double dr = st * (qp - q0);
  const double dc = 0.5 * (qp - qm);
  const double adl = fabs_(dl), adr = fabs_(dr), adc = fabs_(dc);
  const double lim0 = fmin2(adl, adr);
  const double lim = ((dl * dr) <= 0.0) ? 0.0 : lim0;
  const double m_b = (dc >= 0.0) ? ob : -ob;
  double lim_b = 0.0, adc_b = 0.0;
  if (lim < adc)
    lim_b = m_b;
  else
    adc_b = m_b;
  const double lim0_b = ((dl * dr) <= 0.0) ? 0.0 : lim_b;
  double adl_b = 0.0, adr_b = 0.0;
  if (adl < adr)
    adl_b = lim0_b;
  else
    adr_b = lim0_b;
  const double dl_b = adl_b * sgn(dl), dr_b = adr_b * sgn(dr), dc_b = adc_b * sgn(dc);
  q0_b += st * dl_b - st * dr_b;
  qm_b += -st * dl_b - 0.5 * dc_b;
  qp_b += st * dr_b + 0.5 * dc_b;
}

E2D_HDI bool slopes_limited(const Hydro<double>& h) { return h.slope_type == 1 || h.slope_type == 2; }

// trace_sources: (q, dqx, dqy) -> s
template <class M>
E2D_HDI void trace_sources_adj(const Hydro<double>& h, const double q[4], const double dqx[4],
                               const double dqy[4], const double sb[4], double qb[4], double dqxb[4],
                               double dqyb[4]) {
  const double r = q[0], p = q[1], u = q[2], v = q[3];
  const double dv = dqx[2] + dqy[3];
  qb[2] += -(sb[0] * dqx[0] + sb[1] * dqx[1] + sb[2] * dqx[2] + sb[3] * dqx[3]);
  qb[3] += -(sb[0] * dqy[0] + sb[1] * dqy[1] + sb[2] * dqy[2] + sb[3] * dqy[3]);
  for (int k = 0; k < 4; ++k) {
    dqxb[k] += -u * sb[k];
    dqyb[k] += -v * sb[k];
  }
  dqxb[1] += -sb[2] / r;
  dqyb[1] += -sb[3] / r;
  const double dv_b = -sb[0] * r - sb[1] * h.gamma0 * p;
  dqxb[2] += dv_b;
  dqyb[3] += dv_b;
  qb[0] += -sb[0] * dv + (sb[2] * dqx[1] + sb[3] * dqy[1]) / (r * r);
  qb[1] += -sb[1] * dv * h.gamma0;
}

// face_state: (q, dq, s) -> f, density floored
E2D_HDI void face_state_adj(const Hydro<double>& h, const double q[4], const double dq[4], const double s[4],
                            double dtd, bool plus_side, const double fb[4], double qb[4], double dqb[4],
                            double sb[4]) {
  const double half_dq0 = 0.5 * dq[0];
  const double edge0 = plus_side ? q[0] + half_dq0 : q[0] - half_dq0;
  const double f0raw = edge0 + s[0] * dtd * 0.5;
  for (int k = 0; k < 4; ++k) {
    const double b = (k == 0 && h.smallr > f0raw) ? 0.0 : fb[k];
    qb[k] += b;
    dqb[k] += plus_side ? 0.5 * b : -0.5 * b;
    sb[k] += 0.5 * dtd * b;
  }
}

// godunov_flux: g -> f
E2D_HDI void godunov_flux_adj(const Hydro<double>& h, const double g[4], const double fb[4], double gb[4]) {
  const double r = g[0], p = g[1], u = g[2], v = g[3];
  const double mass = r * u;
  const double ekin = 0.5 * r * (u * u + v * v);
  const double etot = p * h.entho + ekin;
  // f1 = u * (etot + p)
  gb[2] += fb[1] * (etot + p);
  const double etot_b = fb[1] * u;
  gb[1] += fb[1] * u + etot_b * h.entho;
  gb[0] += etot_b * 0.5 * (u * u + v * v);
  gb[2] += etot_b * r * u;
  gb[3] += etot_b * r * v;
  // f0 = mass, f2 = mass*u + p, f3 = mass*v
  const double mass_b = fb[0] + fb[2] * u + fb[3] * v;
  gb[2] += fb[2] * mass;
  gb[1] += fb[2];
  gb[3] += fb[3] * mass;
  gb[0] += mass_b * u;
  gb[2] += mass_b * r;
}

// One Newton iteration of riemann_two_shock from p_k; returns p_{k+1}.
template <class M>
E2D_HDI double newton_step(const Hydro<double>& h, double pk, double cl, double cr, double pl, double pr,
                           double ul, double ur, double& conv, bool& take_t) {
  const double wwl = M::sqrt(cl * (1.0 + M::div(h.gamma6 * (pk - pl), pl)));
  const double wwr = M::sqrt(cr * (1.0 + M::div(h.gamma6 * (pk - pr), pr)));
  const double zl = M::div(2.0 * wwl * wwl * wwl, wwl * wwl + cl);
  const double zr = M::div(2.0 * wwr * wwr * wwr, wwr * wwr + cr);
  const double usl = ul - M::div(pk - pl, wwl);
  const double usr = ur + M::div(pk - pr, wwr);
  const double t = M::div(zr * zl, zr + zl) * (usl - usr);
  take_t = t > -pk;
  const double delp = take_t ? t : -pk;
  const double pn = pk + delp;
  conv = fabs_(M::div(delp, pn + h.smallpp));
  return pn;
}

// Reverse of one Newton iteration at p_k: given pn_b (bar of p_{k+1}), returns
// the bar of p_k and accumulates the bars of (cl, cr, pl, pr, ul, ur).
template <class M>
E2D_HDI double newton_step_adj(const Hydro<double>& h, double pk, double cl, double cr, double pl, double pr,
                               double ul, double ur, double pn_b, double& cl_b, double& cr_b, double& pl_b,
                               double& pr_b, double& ul_b, double& ur_b) {
  const double g6 = h.gamma6;
  const double xl = M::div(g6 * (pk - pl), pl), xr = M::div(g6 * (pk - pr), pr);
  const double wwl = M::sqrt(cl * (1.0 + xl));
  const double wwr = M::sqrt(cr * (1.0 + xr));
  const double dzl = wwl * wwl + cl, dzr = wwr * wwr + cr;
  const double zl = M::div(2.0 * wwl * wwl * wwl, dzl);
  const double zr = M::div(2.0 * wwr * wwr * wwr, dzr);
  const double c1 = M::div(pk - pl, wwl), c2 = M::div(pk - pr, wwr);
  const double usl = ul - c1, usr = ur + c2;
  const double S = zr + zl;
  const double m = M::div(zr * zl, S);
  const double t = m * (usl - usr);
  double pk_b = pn_b;  // p_{k+1} = p_k + delp
  const double delp_b = pn_b;
  double t_b = 0.0;
  if (t > -pk)
    t_b = delp_b;
  else
    pk_b -= delp_b;
  const double m_b = t_b * (usl - usr);
  const double usl_b = t_b * m, usr_b = -t_b * m;
  const double pm_b = m_b / S, s_b = -m_b * m / S;
  const double zr_b = pm_b * zl + s_b, zl_b = pm_b * zr + s_b;
  // usl = ul - (pk - pl)/wwl ; usr = ur + (pk - pr)/wwr
  ul_b += usl_b;
  const double c1_b = -usl_b;
  pk_b += c1_b / wwl;
  pl_b -= c1_b / wwl;
  double wwl_b = -c1_b * c1 / wwl;
  ur_b += usr_b;
  const double c2_b = usr_b;
  pk_b += c2_b / wwr;
  pr_b -= c2_b / wwr;
  double wwr_b = -c2_b * c2 / wwr;
  // z = 2 ww^3 / (ww^2 + c)
  wwl_b += zl_b / dzl * 6.0 * wwl * wwl - zl_b * zl / dzl * 2.0 * wwl;
  cl_b += -zl_b * zl / dzl;
  wwr_b += zr_b / dzr * 6.0 * wwr * wwr - zr_b * zr / dzr * 2.0 * wwr;
  cr_b += -zr_b * zr / dzr;
  // ww = sqrt(c (1 + g6 (pk - p0)/p0))
  const double al_b = wwl_b * 0.5 / wwl, ar_b = wwr_b * 0.5 / wwr;
  cl_b += al_b * (1.0 + xl);
  cr_b += ar_b * (1.0 + xr);
  const double xl_b = al_b * cl, xr_b = ar_b * cr;
  pk_b += xl_b * g6 / pl + xr_b * g6 / pr;
  pl_b += -xl_b * g6 * pk / (pl * pl);
  pr_b += -xr_b * g6 * pk / (pr * pr);
  return pk_b;
}

// riemann_two_shock adjoint: bars of (ql, qr) from the bar of the flux
template <class M>
E2D_HDI void riemann_two_shock_adj(const Hydro<double>& h, const double ql[4], const double qr[4],
                                   const double fb[4], double qlb[4], double qrb[4]) {
  const double g = h.gamma0, g6 = h.gamma6;
  // ---- forward
  const double rl = fmax2(ql[0], h.smallr);
  const double ul = ql[2];
  const double plf = rl * h.smallp;
  const double pl = fmax2(ql[1], plf);
  const double rr = fmax2(qr[0], h.smallr);
  const double ur = qr[2];
  const double prf = rr * h.smallp;
  const double pr = fmax2(qr[1], prf);
  const double cl = g * pl * rl;
  const double cr = g * pr * rr;
  const double wl0 = M::sqrt(cl);
  const double wr0 = M::sqrt(cr);
  const double num = (wr0 * pl + wl0 * pr) + wl0 * wr0 * (ul - ur);
  const double den = wl0 + wr0;
  const double p0r = M::div(num, den);
  const double p0 = fmax2(p0r, 0.0);
  double ps = p0, conv = 1.0;
  int K = 0;
  for (; K < h.niter && conv > h.conv_tol; ++K) {
    bool tt;
    ps = newton_step<M>(h, ps, cl, cr, pl, pr, ul, ur, conv, tt);
  }
  const double xlf = M::div(g6 * (ps - pl), pl), xrf = M::div(g6 * (ps - pr), pr);
  const double wl = M::sqrt(cl * (1.0 + xlf));
  const double wr = M::sqrt(cr * (1.0 + xrf));
  const double a_us = M::div(pl - ps, wl), b_us = M::div(pr - ps, wr);
  const double us = 0.5 * (ul + a_us + ur - b_us);
  const bool right = us >= 0.0;
  const double ro = right ? rl : rr, uo = right ? ul : ur, po = right ? pl : pr, wo = right ? wl : wr;
  const double sg = right ? 1.0 : -1.0;
  const double Y = M::div(g * po, ro);
  const double co_r = M::sqrt(fabs_(Y));
  const double co = fmax2(h.smallc, co_r);
  const double E = M::div(ro * (po - ps), wo * wo);
  const double D = 1.0 + E;
  const double rs_r = M::div(ro, D);
  const double rs = fmax2(rs_r, h.smallr);
  const double X = M::div(g * ps, rs);
  const double cs_r = M::sqrt(fabs_(X));
  const double cs = fmax2(h.smallc, cs_r);
  const double s_uo = sg * uo;
  double spout = co - s_uo;
  double spin = cs - sg * us;
  const double ush = M::div(wo, ro) - s_uo;
  const bool shock = ps >= po;
  if (shock) {
    spin = ush;
    spout = ush;
  }
  const double sum = spout + spin, dif = spout - spin;
  const double scrb = h.smallc + fabs_(sum);
  const bool take_dif = dif > scrb;
  const double scr = take_dif ? dif : scrb;
  const double frac = 0.5 * (1.0 + M::div(sum, scr));
  const bool outside = spout < 0.0;
  const bool inside = spin > 0.0;
  double gs[4];
  gs[0] = frac * rs + (1.0 - frac) * ro;
  gs[2] = frac * us + (1.0 - frac) * uo;
  gs[1] = frac * ps + (1.0 - frac) * po;
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

  // ---- reverse
  double gb[4] = {0.0, 0.0, 0.0, 0.0};
  godunov_flux_adj(h, gs, fb, gb);
  if (right)
    qlb[3] += gb[3];
  else
    qrb[3] += gb[3];
  double rs_b = 0.0, us_b = 0.0, ps_b = 0.0, ro_b = 0.0, uo_b = 0.0, po_b = 0.0, wo_b = 0.0;
  if (inside) {
    rs_b += gb[0];
    us_b += gb[2];
    ps_b += gb[1];
  } else if (outside) {
    ro_b += gb[0];
    uo_b += gb[2];
    po_b += gb[1];
  } else {
    const double frac_b = gb[0] * (rs - ro) + gb[2] * (us - uo) + gb[1] * (ps - po);
    rs_b += gb[0] * frac;
    ro_b += gb[0] * (1.0 - frac);
    us_b += gb[2] * frac;
    uo_b += gb[2] * (1.0 - frac);
    ps_b += gb[1] * frac;
    po_b += gb[1] * (1.0 - frac);
    // frac = 0.5 (1 + sum/scr)
    double sum_b = frac_b * 0.5 / scr;
    const double scr_b = -frac_b * 0.5 * sum / (scr * scr);
    double dif_b = 0.0;
    if (take_dif)
      dif_b = scr_b;
    else
      sum_b += scr_b * sgn(sum);
    const double spout_b = sum_b + dif_b, spin_b = sum_b - dif_b;
    double co_b = 0.0, cs_b = 0.0, s_uo_b = 0.0;
    if (shock) {
      const double ush_b = spout_b + spin_b;
      wo_b += ush_b / ro;
      ro_b -= ush_b * wo / (ro * ro);
      s_uo_b -= ush_b;
    } else {
      co_b += spout_b;
      s_uo_b -= spout_b;
      cs_b += spin_b;
      us_b -= spin_b * sg;
    }
    uo_b += s_uo_b * sg;
    // cs = max(smallc, sqrt|g ps / rs|)
    if (!(h.smallc > cs_r) && cs_r > 0.0) {
      const double X_b = cs_b * 0.5 / cs_r * sgn(X);
      ps_b += X_b * g / rs;
      rs_b -= X_b * g * ps / (rs * rs);
    }
    // co = max(smallc, sqrt|g po / ro|)
    if (!(h.smallc > co_r) && co_r > 0.0) {
      const double Y_b = co_b * 0.5 / co_r * sgn(Y);
      po_b += Y_b * g / ro;
      ro_b -= Y_b * g * po / (ro * ro);
    }
  }
  // rs = max(ro / (1 + ro (po - ps) / wo^2), smallr)
  if (rs_r > h.smallr) {
    const double rsr_b = rs_b;
    ro_b += rsr_b / D;
    const double E_b = -rsr_b * ro / (D * D);
    const double w2 = wo * wo;
    ro_b += E_b * (po - ps) / w2;
    po_b += E_b * ro / w2;
    ps_b -= E_b * ro / w2;
    wo_b -= E_b * 2.0 * E / wo;
  }
  double rl_b = 0.0, ul_b = 0.0, pl_b = 0.0, wl_b = 0.0;
  double rr_b = 0.0, ur_b = 0.0, pr_b = 0.0, wr_b = 0.0;
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
  ul_b += 0.5 * us_b;
  ur_b += 0.5 * us_b;
  const double a_b = 0.5 * us_b, b_b = -0.5 * us_b;
  pl_b += a_b / wl;
  ps_b -= a_b / wl;
  wl_b -= a_b * a_us / wl;
  pr_b += b_b / wr;
  ps_b -= b_b / wr;
  wr_b -= b_b * b_us / wr;
  // wl = sqrt(cl (1 + g6 (ps - pl)/pl))
  double cl_b = 0.0, cr_b = 0.0;
  {
    const double W_b = wl_b * 0.5 / wl;
    cl_b += W_b * (1.0 + xlf);
    const double x_b = W_b * cl;
    ps_b += x_b * g6 / pl;
    pl_b += -x_b * g6 * ps / (pl * pl);
  }
  {
    const double W_b = wr_b * 0.5 / wr;
    cr_b += W_b * (1.0 + xrf);
    const double x_b = W_b * cr;
    ps_b += x_b * g6 / pr;
    pr_b += -x_b * g6 * ps / (pr * pr);
  }
  // Newton iterations, last to first (p_k recomputed from p_0)
  double p_b = ps_b;
  for (int k = K - 1; k >= 0; --k) {
    double pk = p0, cv;
    bool tt;
    for (int j = 0; j < k; ++j) pk = newton_step<M>(h, pk, cl, cr, pl, pr, ul, ur, cv, tt);
    p_b = newton_step_adj<M>(h, pk, cl, cr, pl, pr, ul, ur, p_b, cl_b, cr_b, pl_b, pr_b, ul_b, ur_b);
  }
  // p0 = max(num/den, 0)
  double wl0_b = 0.0, wr0_b = 0.0;
  if (p0r > 0.0) {
    const double num_b = p_b / den;
    const double den_b = -p_b * p0r / den;
    wr0_b += num_b * pl + num_b * wl0 * (ul - ur) + den_b;
    pl_b += num_b * wr0;
    wl0_b += num_b * pr + num_b * wr0 * (ul - ur) + den_b;
    pr_b += num_b * wl0;
    ul_b += num_b * wl0 * wr0;
    ur_b -= num_b * wl0 * wr0;
  }
  cl_b += wl0_b * 0.5 / wl0;
  cr_b += wr0_b * 0.5 / wr0;
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
template <class M>
E2D_HDI void riemann_y_adj(const Hydro<double>& h, const double ql[4], const double qr[4], const double fb[4],
                           double qlb[4], double qrb[4]) {
  const double l[4] = {ql[0], ql[1], ql[3], ql[2]};
  const double r[4] = {qr[0], qr[1], qr[3], qr[2]};
  const double fbs[4] = {fb[0], fb[1], fb[3], fb[2]};
  double lb[4] = {0.0, 0.0, 0.0, 0.0}, rb[4] = {0.0, 0.0, 0.0, 0.0};
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
