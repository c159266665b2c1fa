/*
 * Test infrastructure only: body of the oracle, instantiated once per precision
 * by cpu_ref.c with
 *   R         the reference's fp_kind (double or float)
 *   S         the symbol suffix (f64 or f32)
 *   SQRT/FABS the libm routines of that precision
 *   LIT_F64   an fp_kind literal (used by the benchmark initial condition)
 *   LIT_F32   how a default-real (FP32) literal of the Fortran source enters
 *             R arithmetic: (R)(x##f).  In the FP64 build `1e-6` and `0.14`
 *             are FP32 constants widened to double (serial_ref/m_utils.f90:396,
 *             serial_ref/m_run.f90:548); every other literal of the path is
 *             exact in both precisions.
 * Each function cites the reference lines it restates.  Fortran max/min are
 * restated as (a > b ? a : b) / (a < b ? a : b), which is what the reference
 * compiler emits (checked on signed zeros).
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, S)

typedef struct {
  int nx, ny, isize, jsize;
  int bc[4]; /* xmin, xmax, ymin, ymax */
  int slope_type, niter_riemann, riemann_solver;
  R gamma0, gamma6, cfl, smallr, smallc, smallp, smallpp, dx, dy;
  R *u, *u2, *q, *fx, *fy;
  R invDt; /* SAVEd running max of compute_dt (serial_ref/m_run.f90:257) */
  /* decisions of the two-shock solver per step and face (Newton count and
   * sampling branches, FN(riemann_n)), for oracle_fd only: mode 1 records
   * them, mode 2 replays them (0: unused) */
  unsigned char *tape;
  unsigned char *ltape; /* the limiters' decisions per step and cell (FN(slope_at)) */
  int tape_mode, tape_step;
} FN(run);

#define NBVAR 4
/* 1-based Fortran indexing, as in the reference: a(i,j,v) */
#define AT(a, i, j, v) (a)[((size_t)((v)-1) * (size_t)h->jsize + (size_t)((j)-1)) * (size_t)h->isize + (size_t)((i)-1)]

/* tape entry (2 bytes) of face dir (0: x, 1: y) of cell (i, j) in the current step */
static inline unsigned char *FN(tape_slot)(const FN(run) * h, int i, int j, int dir) {
  if (!h->tape || !h->tape_mode) return NULL;
  return h->tape + ((((size_t)h->tape_step * (size_t)h->jsize + (size_t)(j - 1)) * (size_t)h->isize +
                     (size_t)(i - 1)) * 2 + (size_t)dir) * 2;
}

/* limiter tape entry (8 bytes: X then Y slopes of the 4 variables) of cell (i, j) in the current step */
static inline unsigned char *FN(ltape_slot)(const FN(run) * h, int i, int j) {
  if (!h->ltape || !h->tape_mode) return NULL;
  return h->ltape + (((size_t)h->tape_step * (size_t)h->jsize + (size_t)(j - 1)) * (size_t)h->isize +
                     (size_t)(i - 1)) * 8;
}

static inline R FN(fmax2)(R a, R b) { return a > b ? a : b; }
static inline R FN(fmin2)(R a, R b) { return a < b ? a : b; }
#define MAX2 FN(fmax2)
#define MIN2 FN(fmin2)
#define ABS(x) FABS(x) /* Fortran abs clears the sign bit */

/* eos: serial_ref/m_utils.f90:15-29 */
static inline void FN(eos)(const FN(run) * h, R rho, R eint, R *p, R *c) {
  *p = MAX2((h->gamma0 - (R)1.0) * rho * eint, rho * h->smallp);
  *c = SQRT(h->gamma0 * *p / rho);
}

/* computePrimitives: serial_ref/m_utils.f90:35-64 (order ID, IP, IU, IV) */
static inline void FN(prim)(const FN(run) * h, const R u[4], R q[4], R *c) {
  q[0] = MAX2(u[0], h->smallr);
  q[2] = u[2] / q[0];
  q[3] = u[3] / q[0];
  R eken = (R)0.5 * (q[2] * q[2] + q[3] * q[3]);
  R e = u[1] / q[0] - eken;
  FN(eos)(h, q[0], e, &q[1], c);
}

/* slope_unsplit_hydro_2d: serial_ref/m_utils.f90:268-342.
 * dq[0..3] = X slopes, dq[4..7] = Y slopes. */
static inline void FN(slope)(const FN(run) * h, const R q[4], const R qpx[4], const R qmx[4],
                             const R qpy[4], const R qmy[4], R dq[8]) {
  if (h->slope_type == 1 || h->slope_type == 2) {
    const R st = (R)h->slope_type;
    for (int v = 0; v < NBVAR; ++v) {
      R dlft = st * (q[v] - qmx[v]);
      R drgt = st * (qpx[v] - q[v]);
      R dcen = (R)0.5 * (qpx[v] - qmx[v]);
      R dsgn = (dcen >= (R)0) ? (R)1 : (R)-1;
      R slop = MIN2(ABS(dlft), ABS(drgt));
      R dlim = slop;
      if ((dlft * drgt) <= (R)0) dlim = (R)0;
      dq[v] = dsgn * MIN2(dlim, ABS(dcen));

      dlft = st * (q[v] - qmy[v]);
      drgt = st * (qpy[v] - q[v]);
      dcen = (R)0.5 * (qpy[v] - qmy[v]);
      dsgn = (dcen >= (R)0) ? (R)1 : (R)-1;
      slop = MIN2(ABS(dlft), ABS(drgt));
      dlim = slop;
      if ((dlft * drgt) <= (R)0) dlim = (R)0;
      dq[4 + v] = dsgn * MIN2(dlim, ABS(dcen));
    }
  } else {
    for (int v = 0; v < 8; ++v) dq[v] = (R)0;
  }
}

/* One limiter of FN(slope) whose decisions oracle_fd records (mode 1: the
 * reference's arithmetic, and the byte *t) or replays (mode 2: the same
 * operands combined as *t says).  Bits: 0 dcen >= 0, 1 dlft >= 0, 2 drgt >= 0,
 * 3 |dlft| < |drgt|, 4 dlft*drgt <= 0, 5 dlim < |dcen|. */
static inline R FN(limiter_t)(R st, R qm, R q0, R qp, unsigned char *t, int mode) {
  const R dlft = st * (q0 - qm);
  const R drgt = st * (qp - q0);
  const R dcen = (R)0.5 * (qp - qm);
  if (mode == 1) {
    const R dsgn = (dcen >= (R)0) ? (R)1 : (R)-1;
    const R slop = MIN2(ABS(dlft), ABS(drgt));
    R dlim = slop;
    if ((dlft * drgt) <= (R)0) dlim = (R)0;
    *t = (unsigned char)((dcen >= (R)0) | ((dlft >= (R)0) << 1) | ((drgt >= (R)0) << 2) |
                         ((ABS(dlft) < ABS(drgt)) << 3) | (((dlft * drgt) <= (R)0) << 4) |
                         ((dlim < ABS(dcen)) << 5));
    return dsgn * MIN2(dlim, ABS(dcen));
  }
  const int b = *t;
  const R al = (b & 2) ? dlft : -dlft;
  const R ar = (b & 4) ? drgt : -drgt;
  const R dlim = (b & 16) ? (R)0 : ((b & 8) ? al : ar);
  const R m = (b & 32) ? dlim : ((b & 1) ? dcen : -dcen);
  return (b & 1) ? m : -m;
}

/* FN(slope) of cell (i, j) with its limiter decisions taped (oracle_fd) */
static inline void FN(slope_at)(const FN(run) * h, int i, int j, const R q[4], const R qpx[4], const R qmx[4],
                                const R qpy[4], const R qmy[4], R dq[8]) {
  unsigned char *t = FN(ltape_slot)(h, i, j);
  if (!t || !(h->slope_type == 1 || h->slope_type == 2)) {
    FN(slope)(h, q, qpx, qmx, qpy, qmy, dq);
    return;
  }
  /* a cell's slopes are formed again as a neighbour's, possibly on another
   * thread: every recording of a cell writes the same bytes */
  const R st = (R)h->slope_type;
  for (int v = 0; v < NBVAR; ++v) {
    dq[v] = FN(limiter_t)(st, qmx[v], q[v], qpx[v], t + v, h->tape_mode);
    dq[4 + v] = FN(limiter_t)(st, qmy[v], q[v], qpy[v], t + 4 + v, h->tape_mode);
  }
}

/* trace_unsplit_2d_along_dir: serial_ref/m_utils.f90:177-254 */
static inline void FN(trace)(const FN(run) * h, const R q[4], const R dqX[4], const R dqY[4], R dtdx,
                             R dtdy, int faceId, R qf[4]) {
  const R r = q[0], p = q[1], u = q[2], v = q[3];
  const R drx = dqX[0], dpx = dqX[1], dux = dqX[2], dvx = dqX[3];
  const R dry = dqY[0], dpy = dqY[1], duy = dqY[2], dvy = dqY[3];
  const R sr0 = -u * drx - v * dry - (dux + dvy) * r;
  const R sp0 = -u * dpx - v * dpy - (dux + dvy) * h->gamma0 * p;
  const R su0 = -u * dux - v * duy - (dpx) / r;
  const R sv0 = -u * dvx - v * dvy - (dpy) / r;
  const R H = (R)0.5;
  if (faceId == FACE_XMIN) {
    qf[0] = r - H * drx + sr0 * dtdx * H;
    qf[1] = p - H * dpx + sp0 * dtdx * H;
    qf[2] = u - H * dux + su0 * dtdx * H;
    qf[3] = v - H * dvx + sv0 * dtdx * H;
    qf[0] = MAX2(h->smallr, qf[0]);
  }
  if (faceId == FACE_XMAX) {
    qf[0] = r + H * drx + sr0 * dtdx * H;
    qf[1] = p + H * dpx + sp0 * dtdx * H;
    qf[2] = u + H * dux + su0 * dtdx * H;
    qf[3] = v + H * dvx + sv0 * dtdx * H;
    qf[0] = MAX2(h->smallr, qf[0]);
  }
  if (faceId == FACE_YMIN) {
    qf[0] = r - H * dry + sr0 * dtdy * H;
    qf[1] = p - H * dpy + sp0 * dtdy * H;
    qf[2] = u - H * duy + su0 * dtdy * H;
    qf[3] = v - H * dvy + sv0 * dtdy * H;
    qf[0] = MAX2(h->smallr, qf[0]);
  }
  if (faceId == FACE_YMAX) {
    qf[0] = r + H * dry + sr0 * dtdy * H;
    qf[1] = p + H * dpy + sp0 * dtdy * H;
    qf[2] = u + H * duy + su0 * dtdy * H;
    qf[3] = v + H * dvy + sv0 * dtdy * H;
    qf[0] = MAX2(h->smallr, qf[0]);
  }
}

/* cmpflx: serial_ref/m_utils.f90:498-526 */
static inline void FN(cmpflx)(const FN(run) * h, const R g[4], R f[4]) {
  f[0] = g[0] * g[2];
  f[2] = f[0] * g[2] + g[1];
  f[3] = f[0] * g[3];
  R entho = (R)1.0 / (h->gamma0 - (R)1.0);
  R ekin = (R)0.5 * g[0] * (g[2] * g[2] + g[3] * g[3]);
  R etot = g[1] * entho + ekin;
  f[1] = g[2] * (etot + g[1]);
}

/* max(a, b) and |x| whose outcome oracle_fd records (bit `bit` of rec[1])
 * or replays (from rep[1]); with neither, MAX2 and ABS themselves */
static inline R FN(max_t)(R a, R b, const unsigned char *rep, unsigned char *rec, int bit) {
  if (rep) return ((rep[1] >> bit) & 1) ? a : b;
  if (rec && a > b) rec[1] = (unsigned char)(rec[1] | (1 << bit));
  return MAX2(a, b);
}
static inline R FN(abs_t)(R x, const unsigned char *rep, unsigned char *rec, int bit) {
  if (rep) return ((rep[1] >> bit) & 1) ? x : -x;
  if (rec && x >= (R)0) rec[1] = (unsigned char)(rec[1] | (1 << bit));
  return ABS(x);
}

/* riemann_2d: serial_ref/m_utils.f90:349-492.  Returns the Newton iteration
 * count (not part of the reference interface; used to grade test coverage).
 * For oracle_fd's frozen decisions: `record` (if not NULL) receives two
 * bytes, the first holding the Newton count (bits 0-3) and the outcomes of
 * the four branch tests, u* >= 0 (bit 4), p* >= po (bit 5), spout < 0
 * (bit 6), spin > 0 (bit 7), the second the sides taken by the maxima and
 * the absolute value the solver passes through (first guess 0, c_o and c*
 * floors, r* floor, the sampling's denominator and |spout + spin|);
 * `replay` (if not NULL) imposes two such bytes instead of evaluating the
 * stop rule and the tests.  Both NULL: the reference's solver. */
static inline int FN(riemann_n)(const FN(run) * h, const R ql[4], const R qr[4], R g[4], R f[4],
                                const unsigned char *replay, unsigned char *record) {
  const R rl = MAX2(ql[0], h->smallr);
  const R ul = ql[2];
  const R pl = MAX2(ql[1], rl * h->smallp);
  const R rr = MAX2(qr[0], h->smallr);
  const R ur = qr[2];
  const R pr = MAX2(qr[1], rr * h->smallp);

  const R cl = h->gamma0 * pl * rl;
  const R cr = h->gamma0 * pr * rr;
  if (record) record[1] = 0;

  R wl = SQRT(cl);
  R wr = SQRT(cr);
  R pstar = FN(max_t)(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), (R)0, replay, record, 0);
  R pold = pstar;
  R conv = (R)1;

  int iter = 0;
  while (replay ? iter < (*replay & 15) : (iter < h->niter_riemann && conv > LIT_F32(1e-6))) {
    R wwl = SQRT(cl * ((R)1 + h->gamma6 * (pold - pl) / pl));
    R wwr = SQRT(cr * ((R)1 + h->gamma6 * (pold - pr) / pr));
    R qlw = (R)2.0 * wwl * wwl * wwl / (wwl * wwl + cl);
    R qrw = (R)2.0 * wwr * wwr * wwr / (wwr * wwr + cr);
    R usl = ul - (pold - pl) / wwl;
    R usr = ur + (pold - pr) / wwr;
    R delp = MAX2(qrw * qlw / (qrw + qlw) * (usl - usr), -pold);
    pold = pold + delp;
    conv = ABS(delp / (pold + h->smallpp));
    iter = iter + 1;
  }

  pstar = pold;
  wl = SQRT(cl * ((R)1 + h->gamma6 * (pstar - pl) / pl));
  wr = SQRT(cr * ((R)1 + h->gamma6 * (pstar - pr) / pr));

  const R ustar = (R)0.5 * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr);
  const int b_sg = replay ? (*replay >> 4) & 1 : (ustar >= (R)0);
  const R sgnm = b_sg ? (R)1.0 : (R)-1.0;

  R ro, uo, po, wo;
  if (sgnm > (R)0) {
    ro = rl; uo = ul; po = pl; wo = wl;
  } else {
    ro = rr; uo = ur; po = pr; wo = wr;
  }
  const R co = FN(max_t)(h->smallc, SQRT(ABS(h->gamma0 * po / ro)), replay, record, 1);
  const R rstar = FN(max_t)(ro / ((R)1 + ro * (po - pstar) / (wo * wo)), h->smallr, replay, record, 2);
  const R cstar = FN(max_t)(h->smallc, SQRT(ABS(h->gamma0 * pstar / rstar)), replay, record, 3);

  R spout = co - sgnm * uo;
  R spin = cstar - sgnm * ustar;
  const R ushock = wo / ro - sgnm * uo;
  const int b_shock = replay ? (*replay >> 5) & 1 : (pstar >= po);
  if (b_shock) {
    spin = ushock;
    spout = ushock;
  }
  const R scr = FN(max_t)(spout - spin, h->smallc + FN(abs_t)(spout + spin, replay, record, 5), replay, record, 4);
  const R frac = (R)0.5 * ((R)1 + (spout + spin) / scr);

  g[0] = frac * rstar + ((R)1 - frac) * ro;
  g[2] = frac * ustar + ((R)1 - frac) * uo;
  g[1] = frac * pstar + ((R)1 - frac) * po;
  const int b_out = replay ? (*replay >> 6) & 1 : (spout < (R)0);
  const int b_in = replay ? (*replay >> 7) & 1 : (spin > (R)0);
  if (b_out) {
    g[0] = ro; g[2] = uo; g[1] = po;
  }
  if (b_in) {
    g[0] = rstar; g[2] = ustar; g[1] = pstar;
  }
  g[3] = (sgnm > (R)0) ? ql[3] : qr[3];

  FN(cmpflx)(h, g, f);
  if (record) record[0] = (unsigned char)(iter | (b_sg << 4) | (b_shock << 5) | (b_out << 6) | (b_in << 7));
  return iter;
}

static inline int FN(riemann)(const FN(run) * h, const R ql[4], const R qr[4], R g[4], R f[4]) {
  return FN(riemann_n)(h, ql, qr, g, f, NULL, NULL);
}

/* riemann_2d_hllc: serial_ref/m_utils.f90:534-644 */
static inline void FN(hllc)(const FN(run) * h, const R ql[4], const R qr[4], R g[4], R f[4]) {
  const R H = (R)0.5;
  const R entho = (R)1 / (h->gamma0 - (R)1);
  const R rl = MAX2(ql[0], h->smallr);
  const R pl = MAX2(ql[1], rl * h->smallp);
  const R ul = ql[2];
  R ecinl = H * rl * ul * ul;
  ecinl = ecinl + H * rl * ql[3] * ql[3];
  const R etotl = pl * entho + ecinl;
  const R ptotl = pl;

  const R rr = MAX2(qr[0], h->smallr);
  const R pr = MAX2(qr[1], rr * h->smallp);
  const R ur = qr[2];
  R ecinr = H * rr * ur * ur;
  ecinr = ecinr + H * rr * qr[3] * qr[3];
  const R etotr = pr * entho + ecinr;
  const R ptotr = pr;

  const R cfastl = SQRT(MAX2(h->gamma0 * pl / rl, h->smallc * h->smallc));
  const R cfastr = SQRT(MAX2(h->gamma0 * pr / rr, h->smallc * h->smallc));

  const R SL = MIN2(ul, ur) - MAX2(cfastl, cfastr);
  const R SR = MAX2(ul, ur) + MAX2(cfastl, cfastr);

  const R rcl = rl * (ul - SL);
  const R rcr = rr * (SR - ur);
  const R rcsum = MAX2(rcl + rcr, h->smallr);
  (void)rcsum; /* computed and unused in the reference as well (:594) */

  const R ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const R ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  const R rstarl = rl * (SL - ul) / (SL - ustar);
  const R etotstarl = ((SL - ul) * etotl - ptotl * ul + ptotstar * ustar) / (SL - ustar);
  const R rstarr = rr * (SR - ur) / (SR - ustar);
  const R etotstarr = ((SR - ur) * etotr - ptotr * ur + ptotstar * ustar) / (SR - ustar);

  R ro, uo, ptoto, etoto;
  if (SL > (R)0) {
    ro = rl; uo = ul; ptoto = ptotl; etoto = etotl;
  } else if (ustar > (R)0) {
    ro = rstarl; uo = ustar; ptoto = ptotstar; etoto = etotstarl;
  } else if (SR > (R)0) {
    ro = rstarr; uo = ustar; ptoto = ptotstar; etoto = etotstarr;
  } else {
    ro = rr; uo = ur; ptoto = ptotr; etoto = etotr;
  }
  g[0] = g[1] = g[2] = g[3] = (R)0;
  f[0] = ro * uo;
  f[2] = ro * uo * uo + ptoto;
  f[1] = (etoto + ptoto) * uo;
  f[3] = (f[0] > (R)0) ? f[0] * ql[3] : f[0] * qr[3];
}

/* slot: this face's entry of the decision tape, or NULL */
static inline void FN(solve)(const FN(run) * h, const R ql[4], const R qr[4], R g[4], R f[4],
                             unsigned char *slot) {
  if (h->riemann_solver == 2)
    FN(hllc)(h, ql, qr, g, f);
  else
    (void)FN(riemann_n)(h, ql, qr, g, f, slot && h->tape_mode == 2 ? slot : NULL,
                        slot && h->tape_mode == 1 ? slot : NULL);
}


/* ------------------------------------------------------------------------ */

void *FN(oracle_create)(int nx, int ny, const int bc[4], int slope_type, int niter_riemann,
                        int riemann_solver, double gamma0, double cfl, double smallr, double smallc) {
  FN(run) *h = (FN(run) *)calloc(1, sizeof(FN(run)));
  if (!h) return NULL;
  const int ghostWidth = 2;
  h->nx = nx;
  h->ny = ny;
  /* serial_ref/m_run.f90:103-124 */
  h->isize = nx + 2 * ghostWidth;
  h->jsize = ny + 2 * ghostWidth;
  for (int k = 0; k < 4; ++k) h->bc[k] = bc[k];
  h->slope_type = slope_type;
  h->niter_riemann = niter_riemann;
  h->riemann_solver = riemann_solver;
  h->gamma0 = (R)gamma0;
  h->cfl = (R)cfl;
  h->smallr = (R)smallr;
  h->smallc = (R)smallc;
  h->dx = ((R)1.0 - (R)0.0) / (R)nx;
  h->dy = ((R)1.0 - (R)0.0) / (R)ny;
  h->smallp = h->smallc * h->smallc / h->gamma0;
  h->smallpp = h->smallr * h->smallp;
  h->gamma6 = (h->gamma0 + (R)1) / ((R)2 * h->gamma0);
  size_t n = (size_t)h->isize * (size_t)h->jsize * NBVAR;
  h->u = (R *)calloc(n, sizeof(R));
  h->u2 = (R *)calloc(n, sizeof(R)); /* zero, as the reference's unwritten u2 */
  h->q = (R *)calloc(n, sizeof(R));
  h->fx = (R *)calloc(n, sizeof(R));
  h->fy = (R *)calloc(n, sizeof(R));
  h->invDt = (R)0;
  if (!h->u || !h->u2 || !h->q || !h->fx || !h->fy) {
    FN(oracle_destroy)(h);
    return NULL;
  }
  return h;
}

void FN(oracle_destroy)(void *hv) {
  FN(run) *h = (FN(run) *)hv;
  if (!h) return;
  free(h->u);
  free(h->u2);
  free(h->q);
  free(h->fx);
  free(h->fy);
  free(h);
}

/* init_implode: serial_ref/m_run.f90:524-555 (tmp evaluated in default real);
 * diag: the benchmark initial condition (x+y<1 high), same cell indexing. */
void FN(oracle_init)(void *hv, int problem) {
  FN(run) *h = (FN(run) *)hv;
  const int ghostWidth = 2;
  for (int j = 1; j <= h->jsize; ++j)
    for (int i = 1; i <= h->isize; ++i) {
      if (problem == ORACLE_IC_DIAG) {
        R x = (R)(i - ghostWidth - 1) / (R)h->nx;
        R y = (R)(j - ghostWidth - 1) / (R)h->ny;
        AT(h->u, i, j, IU) = (R)0;
        AT(h->u, i, j, IV) = (R)0;
        if (x + y < (R)1) {
          AT(h->u, i, j, ID) = (R)1;
          AT(h->u, i, j, IP) = (R)1 / (h->gamma0 - (R)1);
        } else {
          AT(h->u, i, j, ID) = (R)0.125;
          AT(h->u, i, j, IP) = LIT_F64(0.1) / (h->gamma0 - (R)1);
        }
      } else {
        float tmp = 1.0f * (float)(i - ghostWidth - 1) / (float)h->nx +
                    1.0f * (float)(j - ghostWidth - 1) / (float)h->ny;
        if ((R)tmp > (R)0.5) {
          AT(h->u, i, j, ID) = (R)1.0;
          AT(h->u, i, j, IP) = (R)1.0 / (h->gamma0 - (R)1.0);
          AT(h->u, i, j, IU) = (R)0.0;
          AT(h->u, i, j, IV) = (R)0.0;
        } else {
          AT(h->u, i, j, ID) = (R)0.125;
          AT(h->u, i, j, IP) = LIT_F32(0.14) / (h->gamma0 - (R)1.0);
          AT(h->u, i, j, IU) = (R)0.0;
          AT(h->u, i, j, IV) = (R)0.0;
        }
      }
    }
  size_t n = (size_t)h->isize * (size_t)h->jsize * NBVAR;
  memset(h->u2, 0, n * sizeof(R));
  h->invDt = (R)0;
}

/* compute_dt: serial_ref/m_run.f90:248-312.  Window [3, size-3], running max. */
double FN(oracle_compute_dt)(void *hv, int useU) {
  FN(run) *h = (FN(run) *)hv;
  const int ghostWidth = 2;
  R *a = (useU == 0) ? h->u : h->u2;
  /* the running max is the largest of invDt and every rate, whatever the
   * order, so rows may run in parallel */
  R m = h->invDt;
#pragma omp parallel for schedule(static) reduction(max : m)
  for (int j = ghostWidth + 1; j <= h->jsize - ghostWidth - 1; ++j)
    for (int i = ghostWidth + 1; i <= h->isize - ghostWidth - 1; ++i) {
      R uLoc[4] = {AT(a, i, j, ID), AT(a, i, j, IP), AT(a, i, j, IU), AT(a, i, j, IV)};
      R qLoc[4], c;
      FN(prim)(h, uLoc, qLoc, &c);
      R vx = c + ABS(qLoc[2]);
      R vy = c + ABS(qLoc[3]);
      m = MAX2(m, vx / h->dx + vy / h->dy);
    }
  h->invDt = m;
  return (double)(h->cfl / h->invDt);
}

double FN(oracle_get_invdt)(void *hv) { return (double)((FN(run) *)hv)->invDt; }
void FN(oracle_set_invdt)(void *hv, double v) { ((FN(run) *)hv)->invDt = (R)v; }

/* make_boundaries: serial_ref/m_run.f90:561-648.  Corners are never written. */
void FN(oracle_make_boundaries)(void *hv, int which) {
  FN(run) *h = (FN(run) *)hv;
  const int gw = 2, nx = h->nx, ny = h->ny;
  const int imin = 1, jmin = 1, imax = nx + 2 * gw, jmax = ny + 2 * gw;
  R *d = (which == 0) ? h->u : h->u2;
  for (int j = jmin + gw; j <= jmax - gw; ++j)
    for (int v = 1; v <= NBVAR; ++v)
      for (int i = 1; i <= gw; ++i) {
        R dir = (R)1;
        int i0;
        if (h->bc[0] == 1) {
          i0 = 2 * gw + 1 - i;
          if (v == IU) dir = (R)-1;
        } else if (h->bc[0] == 2) {
          i0 = gw + 1;
        } else {
          i0 = nx + i;
        }
        AT(d, i, j, v) = AT(d, i0, j, v) * dir;
      }
  for (int j = jmin + gw; j <= jmax - gw; ++j)
    for (int v = 1; v <= NBVAR; ++v)
      for (int i = nx + gw + 1; i <= nx + 2 * gw; ++i) {
        R dir = (R)1;
        int i0;
        if (h->bc[1] == 1) {
          i0 = 2 * nx + 2 * gw + 1 - i;
          if (v == IU) dir = (R)-1;
        } else if (h->bc[1] == 2) {
          i0 = nx + gw;
        } else {
          i0 = i - nx;
        }
        AT(d, i, j, v) = AT(d, i0, j, v) * dir;
      }
  for (int i = imin + gw; i <= imax - gw; ++i)
    for (int v = 1; v <= NBVAR; ++v)
      for (int j = 1; j <= gw; ++j) {
        R dir = (R)1;
        int j0;
        if (h->bc[2] == 1) {
          j0 = 2 * gw + 1 - j;
          if (v == IV) dir = (R)-1;
        } else if (h->bc[2] == 2) {
          j0 = gw + 1;
        } else {
          j0 = ny + j;
        }
        AT(d, i, j, v) = AT(d, i, j0, v) * dir;
      }
  for (int i = imin + gw; i <= imax - gw; ++i)
    for (int v = 1; v <= NBVAR; ++v)
      for (int j = ny + gw + 1; j <= ny + 2 * gw; ++j) {
        R dir = (R)1;
        int j0;
        if (h->bc[3] == 1) {
          j0 = 2 * ny + 2 * gw + 1 - j;
          if (v == IV) dir = (R)-1;
        } else if (h->bc[3] == 2) {
          j0 = ny + gw;
        } else {
          j0 = j - ny;
        }
        AT(d, i, j, v) = AT(d, i, j0, v) * dir;
      }
}

/* convertToPrimitives: serial_ref/m_run.f90:317-358 (all cells) */
static void FN(convert_to_primitives)(FN(run) * h, const R *ud) {
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= h->jsize; ++j)
    for (int i = 1; i <= h->isize; ++i) {
      R uLoc[4] = {AT(ud, i, j, ID), AT(ud, i, j, IP), AT(ud, i, j, IU), AT(ud, i, j, IV)};
      R qLoc[4], c;
      FN(prim)(h, uLoc, qLoc, &c);
      AT(h->q, i, j, ID) = qLoc[0];
      AT(h->q, i, j, IP) = qLoc[1];
      AT(h->q, i, j, IU) = qLoc[2];
      AT(h->q, i, j, IV) = qLoc[3];
    }
}

static inline void FN(load_q)(const FN(run) * h, int i, int j, R out[4]) {
  out[0] = AT(h->q, i, j, ID);
  out[1] = AT(h->q, i, j, IP);
  out[2] = AT(h->q, i, j, IU);
  out[3] = AT(h->q, i, j, IV);
}

/* computeAndStoreFluxes: serial_ref/m_run.f90:363-484 */
static void FN(compute_and_store_fluxes)(FN(run) * h, R dt) {
  const int gw = 2;
  const R dtdx = dt / h->dx;
  const R dtdy = dt / h->dy;
#pragma omp parallel for schedule(static)
  for (int j = gw + 1; j <= h->jsize - gw + 1; ++j)
    for (int i = gw + 1; i <= h->isize - gw + 1; ++i) {
      R qLoc[4], qLocN[4], qN0[4], qN1[4], qN2[4], qN3[4], dq[8], dqN[8];
      R qleft[4], qright[4], qgdnv[4], flux_x[4], flux_y[4], tmp;

      FN(load_q)(h, i, j, qLoc);
      FN(load_q)(h, i + 1, j, qN0);
      FN(load_q)(h, i - 1, j, qN1);
      FN(load_q)(h, i, j + 1, qN2);
      FN(load_q)(h, i, j - 1, qN3);
      FN(slope_at)(h, i, j, qLoc, qN0, qN1, qN2, qN3, dq);

      FN(load_q)(h, i - 1, j, qLocN);
      FN(load_q)(h, i, j, qN0);
      FN(load_q)(h, i - 2, j, qN1);
      FN(load_q)(h, i - 1, j + 1, qN2);
      FN(load_q)(h, i - 1, j - 1, qN3);
      FN(slope_at)(h, i - 1, j, qLocN, qN0, qN1, qN2, qN3, dqN);

      FN(trace)(h, qLoc, dq, dq + 4, dtdx, dtdy, FACE_XMIN, qright);
      FN(trace)(h, qLocN, dqN, dqN + 4, dtdx, dtdy, FACE_XMAX, qleft);
      FN(solve)(h, qleft, qright, qgdnv, flux_x, FN(tape_slot)(h, i, j, 0));
      for (int v = 0; v < NBVAR; ++v) AT(h->fx, i, j, v + 1) = flux_x[v] * dtdx;

      FN(load_q)(h, i, j - 1, qLocN);
      FN(load_q)(h, i + 1, j - 1, qN0);
      FN(load_q)(h, i - 1, j - 1, qN1);
      FN(load_q)(h, i, j, qN2);
      FN(load_q)(h, i, j - 2, qN3);
      FN(slope_at)(h, i, j - 1, qLocN, qN0, qN1, qN2, qN3, dqN);

      FN(trace)(h, qLoc, dq, dq + 4, dtdx, dtdy, FACE_YMIN, qright);
      FN(trace)(h, qLocN, dqN, dqN + 4, dtdx, dtdy, FACE_YMAX, qleft);
      tmp = qleft[2]; qleft[2] = qleft[3]; qleft[3] = tmp;
      tmp = qright[2]; qright[2] = qright[3]; qright[3] = tmp;
      FN(solve)(h, qleft, qright, qgdnv, flux_y, FN(tape_slot)(h, i, j, 1));
      tmp = flux_y[2]; flux_y[2] = flux_y[3]; flux_y[3] = tmp;
      for (int v = 0; v < NBVAR; ++v) AT(h->fy, i, j, v + 1) = flux_y[v] * dtdy;
    }
}

/* updateHydro: serial_ref/m_run.f90:489-518 (interior only) */
static void FN(update_hydro)(FN(run) * h, const R *din, R *dout) {
  const int gw = 2;
#pragma omp parallel for schedule(static)
  for (int j = gw + 1; j <= h->jsize - gw; ++j)
    for (int i = gw + 1; i <= h->isize - gw; ++i)
      for (int v = 1; v <= NBVAR; ++v) {
        R flux_tot = AT(h->fx, i, j, v) - AT(h->fx, i + 1, j, v);
        flux_tot = AT(h->fy, i, j, v) - AT(h->fy, i, j + 1, v) + flux_tot;
        AT(dout, i, j, v) = AT(din, i, j, v) + flux_tot;
      }
}

/* godunov_unsplit + godunov_unsplit_cpu: serial_ref/m_run.f90:187-243 */
void FN(oracle_godunov)(void *hv, int nStep, double dt) {
  FN(run) *h = (FN(run) *)hv;
  R *din = (nStep % 2 == 0) ? h->u : h->u2;
  R *dout = (nStep % 2 == 0) ? h->u2 : h->u;
  FN(oracle_make_boundaries)(h, (nStep % 2 == 0) ? 0 : 1);
  FN(convert_to_primitives)(h, din);
  FN(compute_and_store_fluxes)(h, (R)dt);
  FN(update_hydro)(h, din, dout);
}

/* main program loop: serial_ref/main.f90:36-76 (without output).  Returns the
 * step count; dts[n] receives the dt used by step n when non-NULL. */
int FN(oracle_run)(void *hv, int nStepmax, double tEnd, double *dts) {
  FN(run) *h = (FN(run) *)hv;
  R t = (R)0, dt;
  int nStep = 0;
  dt = (R)FN(oracle_compute_dt)(h, nStep % 2);
  FN(oracle_make_boundaries)(h, 0);
  while (t < (R)tEnd && nStep < nStepmax) {
    dt = (R)FN(oracle_compute_dt)(h, nStep % 2);
    if (dts) dts[nStep] = (double)dt;
    FN(oracle_godunov)(h, nStep, (double)dt);
    nStep = nStep + 1;
  }
  return nStep;
}

static R *FN(pick)(FN(run) * h, int which) {
  switch (which) {
    case ORACLE_U: return h->u;
    case ORACLE_U2: return h->u2;
    case ORACLE_Q: return h->q;
    case ORACLE_FX: return h->fx;
    case ORACLE_FY: return h->fy;
    default: return NULL;
  }
}

void FN(oracle_get)(void *hv, int which, double *out) {
  FN(run) *h = (FN(run) *)hv;
  R *a = FN(pick)(h, which);
  size_t n = (size_t)h->isize * (size_t)h->jsize * NBVAR;
  for (size_t k = 0; k < n; ++k) out[k] = (double)a[k];
}

void FN(oracle_set)(void *hv, int which, const double *in) {
  FN(run) *h = (FN(run) *)hv;
  R *a = FN(pick)(h, which);
  size_t n = (size_t)h->isize * (size_t)h->jsize * NBVAR;
  for (size_t k = 0; k < n; ++k) a[k] = (R)in[k];
}

/* Pointwise known-answer entry: same layouts as ref_kat.f90. */
void FN(oracle_kat)(void *hv, int func, int n, const double *x, double *y) {
  FN(run) *h = (FN(run) *)hv;
  for (int k = 0; k < n; ++k) {
    R a[20], o[8] = {0}, c = 0;
    int nin = (func == ORACLE_KAT_PRIM) ? 4 : (func == ORACLE_KAT_SLOPE) ? 20
            : (func == ORACLE_KAT_TRACE) ? 15 : (func == ORACLE_KAT_CMPFLX) ? 4
            : (func == ORACLE_KAT_EOS) ? 2 : 8;
    const double *xk = x + (size_t)k * (size_t)nin;
    for (int m = 0; m < nin && m < 20; ++m) a[m] = (R)xk[m];
    switch (func) {
      case ORACLE_KAT_PRIM: {
        FN(prim)(h, a, o, &c);
        double *yk = y + (size_t)k * 5;
        for (int m = 0; m < 4; ++m) yk[m] = (double)o[m];
        yk[4] = (double)c;
      } break;
      case ORACLE_KAT_SLOPE: {
        FN(slope)(h, a, a + 4, a + 8, a + 12, a + 16, o);
        double *yk = y + (size_t)k * 8;
        for (int m = 0; m < 8; ++m) yk[m] = (double)o[m];
      } break;
      case ORACLE_KAT_TRACE: {
        FN(trace)(h, a, a + 4, a + 8, a[12], a[13], (int)xk[14], o);
        double *yk = y + (size_t)k * 4;
        for (int m = 0; m < 4; ++m) yk[m] = (double)o[m];
      } break;
      case ORACLE_KAT_RIEMANN:
      case ORACLE_KAT_HLLC: {
        R g[4];
        if (func == ORACLE_KAT_RIEMANN)
          (void)FN(riemann)(h, a, a + 4, g, o);
        else
          FN(hllc)(h, a, a + 4, g, o);
        double *yk = y + (size_t)k * 8;
        for (int m = 0; m < 4; ++m) yk[m] = (double)g[m];
        for (int m = 0; m < 4; ++m) yk[4 + m] = (double)o[m];
      } break;
      case ORACLE_KAT_CMPFLX: {
        FN(cmpflx)(h, a, o);
        double *yk = y + (size_t)k * 4;
        for (int m = 0; m < 4; ++m) yk[m] = (double)o[m];
      } break;
      case ORACLE_KAT_EOS: {
        R p;
        FN(eos)(h, a[0], a[1], &p, &c);
        y[(size_t)k * 2] = (double)p;
        y[(size_t)k * 2 + 1] = (double)c;
      } break;
      case ORACLE_KAT_RIEMANN_ITERS: {
        R g[4];
        y[k] = (double)FN(riemann)(h, a, a + 4, g, o);
      } break;
      default: break;
    }
  }
}

/* For the adjoint checks: the 4-point central difference
 *   (8 (f(+e) - f(-e)) - (f(+2e) - f(-2e))) / (12 e)
 * of the frozen-dt rollout map f: U0 -> U_n (n = nsteps steps of
 * oracle_godunov with dt = dts[0..n-1]) along dir, evaluated in R: the
 * perturbed starts U0 + s e dir are formed in R from the double inputs, and
 * the differences are taken in R, element by element, before the result is
 * rounded to double.  Buffer u2 starts each evaluation at zero (the
 * reference's never-written u2).  freeze: the decisions of every face, cell
 * and step are recorded in an unperturbed run first and the perturbed runs
 * repeat them: the two-shock solver's (Newton count, upwind side, shock or
 * rarefaction, the two sampling tests, its maxima and |spout + spin|) and the
 * limiters' (signs, which of |dl|, |dr|, |dc| is taken, the extremum test).
 * That is the map as executed at U0, smooth in U0, which the tangent and
 * adjoint linearise.  Free, the solver jumps where its decisions switch (the
 * stop rule by the size of the next Newton correction, the sonic-point blend
 * by order smallc / |spin|, the shock test in spin), which a difference
 * quotient divides by e, and every limiter kink within 2 e leaves an O(1)
 * error in that cell's slope derivative.  The floors against smallr, smallp
 * and the Newton step's -p stay free: they act only near vacuum.
 * weight == NULL: out (isize*jsize*4, the Fortran layout) receives the
 * derivative of the final state; else out[0] receives the derivative of
 * sum(weight * U_n).  The run's arrays are overwritten.  Returns 0, or -1
 * when a scratch allocation fails. */
int FN(oracle_fd)(void *hv, const double *u0, const double *dir, int nsteps, const double *dts, double eps,
                  const double *weight, int freeze, double *out) {
  FN(run) *h = (FN(run) *)hv;
  const size_t n = (size_t)h->isize * (size_t)h->jsize * NBVAR;
  static const int mult[4] = {1, -1, 2, -2};
  static const int coef[4] = {8, -8, -1, 1};
  R *acc = (R *)calloc(n, sizeof(R));
  if (!acc) return -1;
  if (freeze && h->niter_riemann > 15) freeze = 0; /* counts above 15 do not fit the tape */
  if (freeze) {
    const size_t cells = (size_t)(nsteps > 0 ? nsteps : 1) * (size_t)h->isize * (size_t)h->jsize;
    h->tape = (unsigned char *)calloc(cells * 4, 1);
    h->ltape = (unsigned char *)calloc(cells * 8, 1);
    if (!h->tape || !h->ltape) {
      free(h->tape);
      free(h->ltape);
      h->tape = h->ltape = NULL;
      free(acc);
      return -1;
    }
  }
  const R e = (R)eps;
  for (int m = freeze ? -1 : 0; m < 4; ++m) {
    /* m = -1: the unperturbed run that records the decisions */
    h->tape_mode = freeze ? (m < 0 ? 1 : 2) : 0;
    for (size_t k = 0; k < n; ++k) {
      h->u[k] = m < 0 ? (R)u0[k] : (R)u0[k] + ((R)mult[m] * e) * (R)dir[k];
      h->u2[k] = (R)0;
    }
    for (int s = 0; s < nsteps; ++s) {
      h->tape_step = s;
      FN(oracle_godunov)(h, s, dts[s]);
    }
    if (m < 0) continue;
    const R *fin = (nsteps % 2 == 0) ? h->u : h->u2;
    for (size_t k = 0; k < n; ++k) acc[k] += (R)coef[m] * fin[k];
  }
  h->tape_mode = 0;
  free(h->tape);
  free(h->ltape);
  h->tape = h->ltape = NULL;
  const R den = (R)12 * e;
  if (weight) {
    R sum = (R)0;
    for (size_t k = 0; k < n; ++k)
      if (weight[k] != 0.0) sum += (R)weight[k] * (acc[k] / den);
    out[0] = (double)sum;
  } else {
    for (size_t k = 0; k < n; ++k) out[k] = (double)(acc[k] / den);
  }
  free(acc);
  return 0;
}

#undef AT
#undef MAX2
#undef MIN2
#undef ABS
#undef NBVAR
