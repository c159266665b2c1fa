// CPU checks of the adjoint building blocks (euler2d_cudafortran_amd/csrc/e2d_adjoint.hpp).
// Test infrastructure: built and run by tests/test_adjoint_cpu.py with g++.
//
// For random inputs of each routine: the hand-written adjoint against the
// dual-number tangent (<J v, w> = <v, J^T w>, relative error printed), and the
// tangent against a central finite difference.  Prints one line per routine:
//   <name> n=<cases> dot=<max rel err> fd=<max rel err>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../euler2d_cudafortran_amd/csrc/e2d_adjoint.hpp"

using namespace e2d;
using M = ExactMath;
using DM = DualMath<ExactMath>;

static Hydro<double> hydro(int slope_type) {
  Hydro<double> h;
  h.gamma0 = 1.4;
  h.cfl = 0.8;
  h.smallr = 1e-7;
  h.smallc = 1e-7;
  h.smallp = h.smallc * h.smallc / h.gamma0;
  h.smallpp = h.smallr * h.smallp;
  h.gamma6 = (h.gamma0 + 1.0) / (2.0 * h.gamma0);
  h.entho = 1.0 / (h.gamma0 - 1.0);
  h.dx = h.dy = 1.0 / 64;
  h.inv_dx = h.inv_dy = 64.0;
  h.dxy_pow2 = 1;
  h.conv_tol = static_cast<double>(1e-6f);
  h.inv_smallr = 1.0 / h.smallr;
  h.conv_half_ulp = std::ldexp(1.0, std::ilogb(h.conv_tol) - 53);
  h.slope_type = slope_type;
  h.niter = 10;
  h.riemann_solver = 0;
  return h;
}

static double rel(double a, double b) { return std::fabs(a - b) / std::fmax(1e-300, std::fmax(std::fabs(a), std::fabs(b))); }

struct Stats {
  double dot = 0, fd = 0;
  int n = 0;
};

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> N(0.0, 1.0);
  const Hydro<double> h = hydro(1);
  const Hydro<Dual> hd = dual_hydro(h);

  // ---- riemann_two_shock (x face) and riemann_y
  for (int yface = 0; yface < 2; ++yface) {
    Stats st;
    for (int c = 0; c < 4000; ++c) {
      double ql[4], qr[4];
      ql[0] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      qr[0] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      ql[1] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      qr[1] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      for (int k = 2; k < 4; ++k) {
        ql[k] = 2.0 * N(rng);
        qr[k] = 2.0 * N(rng);
      }
      double v[8], w[4];
      for (double& x : v) x = N(rng);
      for (double& x : w) x = N(rng);
      Dual dl[4], dr[4], g[4], f[4];
      for (int k = 0; k < 4; ++k) {
        dl[k] = Dual(ql[k], v[k]);
        dr[k] = Dual(qr[k], v[4 + k]);
      }
      if (yface)
        riemann_y<Dual, DM, 0>(hd, dl, dr, f);
      else
        riemann_two_shock<Dual, DM>(hd, dl, dr, g, f);
      double jv_w = 0;
      for (int k = 0; k < 4; ++k) jv_w += f[k].d * w[k];
      double lb[4] = {0, 0, 0, 0}, rb[4] = {0, 0, 0, 0};
      if (yface)
        riemann_y_adj<M>(h, ql, qr, w, lb, rb);
      else
        riemann_two_shock_adj<M>(h, ql, qr, w, lb, rb);
      double v_jtw = 0;
      for (int k = 0; k < 4; ++k) v_jtw += v[k] * lb[k] + v[4 + k] * rb[k];
      st.dot = std::fmax(st.dot, rel(jv_w, v_jtw));
      // finite difference of <F, w> along v
      const double eps = 1e-6;
      double fp = 0, fm = 0;
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        double a[4], b[4], gg[4], ff[4];
        for (int k = 0; k < 4; ++k) {
          a[k] = ql[k] + sgn * eps * v[k];
          b[k] = qr[k] + sgn * eps * v[4 + k];
        }
        if (yface)
          riemann_y<double, M, 0>(h, a, b, ff);
        else
          riemann_two_shock<double, M>(h, a, b, gg, ff);
        double s = 0;
        for (int k = 0; k < 4; ++k) s += ff[k] * w[k];
        (sgn > 0 ? fp : fm) = s;
      }
      const double fd = (fp - fm) / (2 * eps);
      const double e = std::fabs(fd - jv_w) / std::fmax(1.0, std::fabs(jv_w));
      st.fd = std::fmax(st.fd, e);
      st.n++;
    }
    std::printf("%s n=%d dot=%.3e fd=%.3e\n", yface ? "riemann_y" : "riemann_two_shock", st.n, st.dot, st.fd);
  }

  // ---- cons_to_prim
  {
    Stats st;
    for (int c = 0; c < 4000; ++c) {
      double u[4];
      u[0] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      u[2] = N(rng) * u[0];
      u[3] = N(rng) * u[0];
      u[1] = 0.5 * (u[2] * u[2] + u[3] * u[3]) / u[0] + std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      double v[4], w[4];
      for (double& x : v) x = N(rng);
      for (double& x : w) x = N(rng);
      Dual du[4], dq[4];
      for (int k = 0; k < 4; ++k) du[k] = Dual(u[k], v[k]);
      cons_to_prim<Dual, DM>(hd, du, dq);
      double jv_w = 0;
      for (int k = 0; k < 4; ++k) jv_w += dq[k].d * w[k];
      double ub[4] = {0, 0, 0, 0};
      cons_to_prim_adj<M>(h, u, w, ub);
      double v_jtw = 0;
      for (int k = 0; k < 4; ++k) v_jtw += v[k] * ub[k];
      st.dot = std::fmax(st.dot, rel(jv_w, v_jtw));
      st.n++;
    }
    std::printf("cons_to_prim n=%d dot=%.3e fd=0\n", st.n, st.dot);
  }

  // ---- limited_slope (minmod and MC)
  for (int stype = 1; stype <= 2; ++stype) {
    Stats st;
    for (int c = 0; c < 20000; ++c) {
      const double qm = N(rng), q0 = N(rng), qp = N(rng);
      const double v0 = N(rng), v1 = N(rng), v2 = N(rng), w = N(rng);
      const Dual o = limited_slope<Dual>(Dual(double(stype)), Dual(qm, v0), Dual(q0, v1), Dual(qp, v2));
      const double jv_w = o.d * w;
      double a = 0, b = 0, cc = 0;
      limited_slope_adj(double(stype), qm, q0, qp, w, a, b, cc);
      const double v_jtw = v0 * a + v1 * b + v2 * cc;
      st.dot = std::fmax(st.dot, std::fabs(jv_w - v_jtw) / std::fmax(1.0, std::fabs(jv_w)));
      st.n++;
    }
    std::printf("limited_slope_st%d n=%d dot=%.3e fd=0\n", stype, st.n, st.dot);
  }

  // ---- trace_sources + the four face states
  {
    Stats st;
    for (int c = 0; c < 4000; ++c) {
      double q[4], dqx[4], dqy[4];
      q[0] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      q[1] = std::exp(std::log(0.05) + U(rng) * std::log(40.0));
      q[2] = N(rng);
      q[3] = N(rng);
      for (int k = 0; k < 4; ++k) {
        dqx[k] = 0.1 * N(rng);
        dqy[k] = 0.1 * N(rng);
      }
      const double dtdx = 0.3 * U(rng), dtdy = 0.3 * U(rng);
      double v[12], w[16];
      for (double& x : v) x = N(rng);
      for (double& x : w) x = N(rng);
      Dual dq[4], ddx[4], ddy[4], s[4], f[4];
      for (int k = 0; k < 4; ++k) {
        dq[k] = Dual(q[k], v[k]);
        ddx[k] = Dual(dqx[k], v[4 + k]);
        ddy[k] = Dual(dqy[k], v[8 + k]);
      }
      trace_sources<Dual, DM>(hd, dq, ddx, ddy, s);
      double jv_w = 0;
      for (int face = 0; face < 4; ++face) {
        const bool xdir = face < 2, plus = face & 1;
        face_state<Dual>(hd, dq, xdir ? ddx : ddy, s, Dual(xdir ? dtdx : dtdy), plus, f);
        for (int k = 0; k < 4; ++k) jv_w += f[k].d * w[4 * face + k];
      }
      double sv[4];
      {
        double sd[4];
        trace_sources<double, M>(h, q, dqx, dqy, sd);
        for (int k = 0; k < 4; ++k) sv[k] = sd[k];
      }
      double qb[4] = {0, 0, 0, 0}, dxb[4] = {0, 0, 0, 0}, dyb[4] = {0, 0, 0, 0}, sb[4] = {0, 0, 0, 0};
      for (int face = 0; face < 4; ++face) {
        const bool xdir = face < 2, plus = face & 1;
        face_state_adj(h, q, xdir ? dqx : dqy, sv, xdir ? dtdx : dtdy, plus, w + 4 * face, qb, xdir ? dxb : dyb, sb);
      }
      trace_sources_adj<M>(h, q, dqx, dqy, sb, qb, dxb, dyb);
      double v_jtw = 0;
      for (int k = 0; k < 4; ++k) v_jtw += v[k] * qb[k] + v[4 + k] * dxb[k] + v[8 + k] * dyb[k];
      st.dot = std::fmax(st.dot, rel(jv_w, v_jtw));
      st.n++;
    }
    std::printf("trace n=%d dot=%.3e fd=0\n", st.n, st.dot);
  }

  // ---- the float instantiation (FP32 and mixed adjoints) against the double
  // one on the same float inputs: 99th percentile of the relative difference
  // of <v, J^T w> (cases where the float forward takes another branch, e.g.
  // one Newton iteration more or less, make up the tail)
  {
    Hydro<float> hf;
    hf.gamma0 = static_cast<float>(h.gamma0);
    hf.gamma6 = static_cast<float>(h.gamma6);
    hf.cfl = static_cast<float>(h.cfl);
    hf.smallr = static_cast<float>(h.smallr);
    hf.smallc = static_cast<float>(h.smallc);
    hf.smallp = static_cast<float>(h.smallp);
    hf.smallpp = static_cast<float>(h.smallpp);
    hf.entho = static_cast<float>(h.entho);
    hf.dx = hf.dy = static_cast<float>(h.dx);
    hf.inv_dx = hf.inv_dy = static_cast<float>(h.inv_dx);
    hf.dxy_pow2 = h.dxy_pow2;
    hf.conv_tol = 1e-6f;
    hf.inv_smallr = 1.0f / hf.smallr;
    hf.conv_half_ulp = std::ldexp(1.0f, std::ilogb(hf.conv_tol) - 24);
    hf.slope_type = h.slope_type;
    hf.niter = h.niter;
    hf.riemann_solver = 0;
    std::vector<double> errs;
    for (int c = 0; c < 4000; ++c) {
      float ql[4], qr[4], w[4];
      double qld[4], qrd[4], wd[4], v[8];
      ql[0] = static_cast<float>(std::exp(std::log(0.05) + U(rng) * std::log(40.0)));
      qr[0] = static_cast<float>(std::exp(std::log(0.05) + U(rng) * std::log(40.0)));
      ql[1] = static_cast<float>(std::exp(std::log(0.05) + U(rng) * std::log(40.0)));
      qr[1] = static_cast<float>(std::exp(std::log(0.05) + U(rng) * std::log(40.0)));
      for (int k = 2; k < 4; ++k) {
        ql[k] = static_cast<float>(2.0 * N(rng));
        qr[k] = static_cast<float>(2.0 * N(rng));
      }
      for (int k = 0; k < 4; ++k) {
        w[k] = static_cast<float>(N(rng));
        qld[k] = ql[k];
        qrd[k] = qr[k];
        wd[k] = w[k];
      }
      for (double& x : v) x = N(rng);
      float lb[4] = {0, 0, 0, 0}, rb[4] = {0, 0, 0, 0};
      double lbd[4] = {0, 0, 0, 0}, rbd[4] = {0, 0, 0, 0};
      riemann_two_shock_adj<M>(hf, ql, qr, w, lb, rb);
      riemann_two_shock_adj<M>(h, qld, qrd, wd, lbd, rbd);
      double a = 0, b = 0;
      for (int k = 0; k < 4; ++k) {
        a += v[k] * lb[k] + v[4 + k] * rb[k];
        b += v[k] * lbd[k] + v[4 + k] * rbd[k];
      }
      errs.push_back(std::fabs(a - b) / std::fmax(1e-30, std::fabs(b)));
    }
    std::sort(errs.begin(), errs.end());
    std::printf("riemann_two_shock_f32 n=%zu dot=%.3e fd=0\n", errs.size(), errs[errs.size() * 99 / 100]);
  }
  return 0;
}
