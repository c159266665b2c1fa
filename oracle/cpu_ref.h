/*
 * Test infrastructure only: the oracle.  Nothing on the product path links,
 * loads or calls this library; tests, the smoke check and the CPU-baseline leg
 * of the benchmark use it as the checker.
 *
 * CPU restatement of the reference serial solver (pkestene/euler2d_cudaFortran,
 * serial_ref/), operation by operation and in source order, in two precisions:
 * suffix _f64 (the reference built with fp_kind = dp_kind) and _f32 (the
 * reference default, serial_ref/m_precision.f90:25), plus _f80 (long double)
 * and _f128 (__float128): the FP64 map's arithmetic with a 64-bit and a
 * 113-bit significand, used only to take finite differences above the FP64
 * roundoff floor.  Built with
 * -ffp-contract=off so every operation rounds as the Fortran source reads.
 *
 * Arrays exchanged with the caller are Fortran column-major u(isize,jsize,4)
 * with isize = nx+4, jsize = ny+4, variable order (rho, E, rho*u, rho*v)
 * (serial_ref/m_constants.f90:11-14), always as double on the interface.
 */
#ifndef EULER2D_ORACLE_CPU_REF_H
#define EULER2D_ORACLE_CPU_REF_H

#ifdef __cplusplus
extern "C" {
#endif

/* which array for oracle_get/oracle_set */
enum { ORACLE_U = 0, ORACLE_U2 = 1, ORACLE_Q = 2, ORACLE_FX = 3, ORACLE_FY = 4 };
/* initial conditions */
enum { ORACLE_IC_IMPLODE = 0, ORACLE_IC_DIAG = 1 };
/* pointwise function ids (same numbering as ref_kat.f90) */
enum {
  ORACLE_KAT_PRIM = 1, ORACLE_KAT_SLOPE = 2, ORACLE_KAT_TRACE = 3, ORACLE_KAT_RIEMANN = 4,
  ORACLE_KAT_HLLC = 5, ORACLE_KAT_CMPFLX = 6, ORACLE_KAT_EOS = 7, ORACLE_KAT_RIEMANN_ITERS = 8
};

#define ORACLE_DECLARE(S)                                                                     \
  void *oracle_create_##S(int nx, int ny, const int bc[4], int slope_type, int niter_riemann,  \
                          int riemann_solver, double gamma0, double cfl, double smallr,       \
                          double smallc);                                                     \
  void oracle_destroy_##S(void *h);                                                           \
  void oracle_init_##S(void *h, int problem);                                                 \
  double oracle_compute_dt_##S(void *h, int useU);                                            \
  double oracle_get_invdt_##S(void *h);                                                       \
  void oracle_set_invdt_##S(void *h, double v);                                               \
  void oracle_make_boundaries_##S(void *h, int which);                                        \
  void oracle_godunov_##S(void *h, int nStep, double dt);                                     \
  int oracle_run_##S(void *h, int nStepmax, double tEnd, double *dts);                        \
  void oracle_get_##S(void *h, int which, double *out);                                       \
  void oracle_set_##S(void *h, int which, const double *in);                                  \
  void oracle_kat_##S(void *h, int func, int n, const double *x, double *y);                   \
  int oracle_fd_##S(void *h, const double *u0, const double *dir, int nsteps, const double *dts, \
                    double eps, const double *weight, int freeze, double *out);

ORACLE_DECLARE(f64)
ORACLE_DECLARE(f32)
ORACLE_DECLARE(f80)  /* x87 extended precision: finite differences of the FP64 map */
ORACLE_DECLARE(f128) /* binary128: the same, where a smaller step is needed */

/* thread count of the grid loops (n > 0 sets it); returns the count in use */
int oracle_threads(int n);

#undef ORACLE_DECLARE

#ifdef __cplusplus
}
#endif
#endif
