/*
 * Oracle (test infrastructure only): fixed-step BDF1/BDF2 + Newton + left-preconditioned
 * GMRES(m) in FP64, the CPU restatement of the product's integrator, which stands in for the
 * reference's BDFNewtonKrylov over CVODE (src/libfosolvers/numerics/SUNDIALS.f90:745-845).
 * CVODE is third-party and absent: integrator parity with CVODE itself is unpinned.  This oracle is
 * pinned by the reference's integrator KAT (src/test/BDFNewtonKrylov1.f90:20, harmonic oscillator
 * to 1e-6) and by closed forms of the linear scalar ODE y' = l y (tests/test_oracle_solver.py).
 * The Jacobian-vector product is CVODE's difference quotient (sigma = 1/||v||_WRMS), the
 * weights are 1/(rtol|y| + atol + rtol rms(y)) (CVODE's plus a floor, see orc_wrms), the
 * orthogonalisation is modified Gram-Schmidt.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct {
  int n, maxl, maxRestarts, maxNewton, order;
  double rtol, atol, newtonTol, linTol, dt, t;
  double absTol; /* > 0: GMRES to this weighted L2 norm instead of linTol relative */
  orc_f f;
  orc_psetup pset;
  orc_psolve psol;
  void *ud;
  double *yn, *ynm1, *y, *G, *fy, *w, *d, *tmp, *tmp2, *V, *H;
  long long nni, nli;
  int lastNewton, lastLin;
  double lastG;
} orc_bdf;

static double wdot(int n, const double *a, const double *b, const double *w) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * w[i] * w[i] * b[i];
  return s;
}

static void matvec(orc_bdf *s, double gamma, const double *v, double *out) {
  /* out = v - gamma * DQ(J) v; sigma = 1/||v||_WRMS with the weights of the current iterate, as
   * the fused Jacobian-vector product computes it */
  int n = s->n;
  double nv = orc_wrms(n, v, s->y, s->rtol, s->atol);
  if (nv == 0.0) { memcpy(out, v, sizeof(double) * n); return; }
  double sig = 1.0 / nv;
  for (int i = 0; i < n; i++) s->tmp[i] = sig * v[i] + s->y[i];
  s->f(s->t, s->tmp, s->tmp2, s->ud);
  for (int i = 0; i < n; i++) {
    double jv = (s->tmp2[i] - s->fy[i]) / sig;
    out[i] = v[i] - gamma * jv;
  }
}

static void prec(orc_bdf *s, double gamma, const double *r, double *z) {
  if (s->psol) s->psol(r, z, gamma, s->ud);
  else if (r != z) memcpy(z, r, sizeof(double) * s->n);
}

static int gmres(orc_bdf *s, double gamma, const double *b, double *x, int *conv) {
  int n = s->n, m = s->maxl, iters = 0;
  double *V = s->V, *H = s->H;
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m), *g = malloc(sizeof(double) * (m + 1));
  double *yv = malloc(sizeof(double) * m), *hc = malloc(sizeof(double) * (m + 2));
  memset(x, 0, sizeof(double) * n);
  double beta0 = -1.0;
  int ok = 0;
  for (int restart = 0; restart <= s->maxRestarts; restart++) {
    if (restart == 0) {
      prec(s, gamma, b, V);
    } else {
      matvec(s, gamma, x, s->d);
      for (int i = 0; i < n; i++) s->d[i] = b[i] - s->d[i];
      prec(s, gamma, s->d, V);
    }
    double beta = sqrt(wdot(n, V, V, s->w));
    if (beta0 < 0) beta0 = beta;
    const double tol = s->absTol > 0.0 ? s->absTol : s->linTol * beta0;
    if (beta == 0.0 || ((s->absTol > 0.0 || restart > 0) && beta <= tol)) { ok = 1; break; }
    for (int i = 0; i < n; i++) V[i] /= beta;
    for (int k = 0; k <= m; k++) g[k] = 0.0;
    g[0] = beta;
    int j = 0, done = 0;
    for (; j < m; j++) {
      double *vj = V + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
      matvec(s, gamma, vj, s->d);
      prec(s, gamma, s->d, w);
      iters++;
      for (int k = 0; k <= j; k++) { /* modified Gram-Schmidt */
        double h = wdot(n, V + (size_t)k * n, w, s->w);
        hc[k] = h;
        for (int i = 0; i < n; i++) w[i] -= h * V[(size_t)k * n + i];
      }
      double hh = sqrt(wdot(n, w, w, s->w));
      hc[j + 1] = hh;
      if (hh > 0.0) for (int i = 0; i < n; i++) w[i] /= hh;
      for (int k = 0; k < j; k++) {
        double a = cs[k] * hc[k] + sn[k] * hc[k + 1];
        hc[k + 1] = -sn[k] * hc[k] + cs[k] * hc[k + 1];
        hc[k] = a;
      }
      double r = hypot(hc[j], hc[j + 1]);
      cs[j] = r > 0 ? hc[j] / r : 1.0;
      sn[j] = r > 0 ? hc[j + 1] / r : 0.0;
      hc[j] = r;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int k = 0; k <= j; k++) H[(size_t)k * m + j] = hc[k];
      if (fabs(g[j + 1]) <= tol || hh == 0.0) { j++; done = 1; break; }
    }
    for (int k = j - 1; k >= 0; k--) {
      double t = g[k];
      for (int q = k + 1; q < j; q++) t -= H[(size_t)k * m + q] * yv[q];
      yv[k] = t / H[(size_t)k * m + k];
    }
    for (int k = 0; k < j; k++)
      for (int i = 0; i < n; i++) x[i] += yv[k] * V[(size_t)k * n + i];
    if (done) { ok = 1; break; }
  }
  free(cs); free(sn); free(g); free(yv); free(hc);
  if (conv) *conv = ok;
  return iters;
}

/* one fixed step; returns 0 on convergence, 1 otherwise */
static int bdf_step(orc_bdf *s) {
  int n = s->n;
  double a1, a2, gamma;
  if (s->order <= 1) { a1 = 1.0; a2 = 0.0; gamma = s->dt; }
  else { a1 = 4.0 / 3.0; a2 = -1.0 / 3.0; gamma = 2.0 * s->dt / 3.0; }
  double q = 0.0;
  for (int i = 0; i < n; i++) q += s->yn[i] * s->yn[i];
  const double floor_ = s->atol + s->rtol * sqrt(q / n);
  for (int i = 0; i < n; i++) {
    s->y[i] = s->order >= 2 ? 2.0 * s->yn[i] - s->ynm1[i] : s->yn[i];
    s->w[i] = 1.0 / (s->rtol * fabs(s->yn[i]) + floor_);
  }
  double tSave = s->t;
  s->t += s->dt;
  int newton = 0, lin = 0, ok = 0;
  double gn = 0.0;
  for (; newton < s->maxNewton; newton++) {
    s->f(s->t, s->y, s->fy, s->ud);
    for (int i = 0; i < n; i++)
      s->G[i] = s->y[i] - a1 * s->yn[i] - a2 * (s->order >= 2 ? s->ynm1[i] : s->yn[i]) - gamma * s->fy[i];
    gn = sqrt(wdot(n, s->G, s->G, s->w) / n);
    if (!(gn == gn)) break;
    if (gn <= s->newtonTol) { ok = 1; break; }
    if (newton == 0 && s->pset) s->pset(s->t, s->y, gamma, s->ud);
    for (int i = 0; i < n; i++) s->G[i] = -s->G[i];
    double *dx = malloc(sizeof(double) * n);
    lin += gmres(s, gamma, s->G, dx, NULL);
    for (int i = 0; i < n; i++) s->y[i] += dx[i];
    free(dx);
  }
  s->lastNewton = newton;
  s->lastLin = lin;
  s->lastG = gn;
  s->nni += newton;
  s->nli += lin;
  if (!ok) { s->t = tSave; return 1; }
  memcpy(s->ynm1, s->yn, sizeof(double) * n);
  memcpy(s->yn, s->y, sizeof(double) * n);
  s->order = 2;
  return 0;
}

/* Integrate nsteps fixed steps from y0; writes every accepted state to states[k*n] (if non-null)
 * and per-step (newton iterations, gmres iterations) to iters[2k]; returns steps completed. */
int orc_bdf_run(int n, orc_f f, orc_psetup pset, orc_psolve psol, void *ud, const double *y0, double t0,
                double dt, int nsteps, double rtol, double atol, double newtonTol, double linTol, int maxl,
                int maxRestarts, int maxNewton, double *states, int *iters) {
  orc_bdf s;
  memset(&s, 0, sizeof(s));
  s.n = n; s.maxl = maxl; s.maxRestarts = maxRestarts; s.maxNewton = maxNewton; s.order = 1;
  s.rtol = rtol; s.atol = atol; s.newtonTol = newtonTol; s.linTol = linTol; s.dt = dt; s.t = t0;
  s.f = f; s.pset = pset; s.psol = psol; s.ud = ud;
  size_t N = n;
  s.yn = malloc(sizeof(double) * N); s.ynm1 = malloc(sizeof(double) * N); s.y = malloc(sizeof(double) * N);
  s.G = malloc(sizeof(double) * N); s.fy = malloc(sizeof(double) * N); s.w = malloc(sizeof(double) * N);
  s.d = malloc(sizeof(double) * N); s.tmp = malloc(sizeof(double) * N); s.tmp2 = malloc(sizeof(double) * N);
  s.V = malloc(sizeof(double) * N * (maxl + 1)); s.H = malloc(sizeof(double) * (maxl + 1) * maxl);
  memcpy(s.yn, y0, sizeof(double) * N);
  int k = 0;
  for (; k < nsteps; k++) {
    if (bdf_step(&s)) break;
    if (states) memcpy(states + N * k, s.yn, sizeof(double) * N);
    if (iters) { iters[2 * k] = s.lastNewton; iters[2 * k + 1] = s.lastLin; }
  }
  free(s.yn); free(s.ynm1); free(s.y); free(s.G); free(s.fy); free(s.w); free(s.d); free(s.tmp); free(s.tmp2);
  free(s.V); free(s.H);
  return k;
}

/* ---- the Euler problem through the oracle's rhs / Jacobian / block ILU ---- */

typedef struct {
  const orc_grid *g;
  const orc_bc *bc;
  double *scratch;
  double *jac;
  orc_bilu bilu;
  int useIlu;
} euler_ud;

static int euler_f(double t, const double *y, double *ydot, void *ud) {
  (void)t;
  euler_ud *e = (euler_ud *)ud;
  orc_rhs(e->g, e->bc, y, ydot, e->scratch);
  return 0;
}

static int euler_pset(double t, const double *y, double gamma, void *ud) {
  (void)t;
  euler_ud *e = (euler_ud *)ud;
  const orc_grid *g = e->g;
  int nE = g->nE;
  double *rho = e->scratch, *rhou = rho + nE, *rhoE = rhou + 3 * (size_t)nE, *u = rhoE + nE;
  double *p = u + 3 * (size_t)nE, *temp = p + nE, *c = temp + nE;
  orc_set_bc(g, e->bc, y, rho, rhou, rhoE);
  orc_sync_state(g, e->bc->r, e->bc->gamm, y, rho, rhou, rhoE, u, p, temp, c);
  orc_find_euler_jac(g, rho, rhou, rhoE, p, e->bc->gamm, e->jac);
  return orc_bilu_setup(&e->bilu, e->jac, g->v, gamma);
}

static int euler_psol(const double *r, double *z, double gamma, void *ud) {
  (void)gamma;
  euler_ud *e = (euler_ud *)ud;
  orc_bilu_solve(&e->bilu, r, z);
  return 0;
}

int orc_bdf_euler(const orc_grid *g, const orc_bc *bc, const double *y0, double dt, int nsteps, double rtol,
                  double atol, double newtonTol, double linTol, int maxl, int maxRestarts, int maxNewton, int useIlu,
                  double *states, int *iters) {
  euler_ud e;
  memset(&e, 0, sizeof(e));
  e.g = g; e.bc = bc; e.useIlu = useIlu;
  e.scratch = calloc(13 * (size_t)g->nE, sizeof(double));
  e.jac = calloc(125 * (size_t)g->nC, sizeof(double));
  e.bilu.nC = g->nC; e.bilu.maxNF = g->maxNF; e.bilu.neib = g->neib; e.bilu.order = NULL;
  e.bilu.D = calloc(25 * (size_t)g->nC, sizeof(double));
  e.bilu.Off = calloc(25 * (size_t)g->maxNF * g->nC, sizeof(double));
  int k = orc_bdf_run(5 * g->nC, euler_f, useIlu ? euler_pset : NULL, useIlu ? euler_psol : NULL, &e, y0, 0.0, dt,
                      nsteps, rtol, atol, newtonTol, linTol, maxl, maxRestarts, maxNewton, states, iters);
  free(e.scratch); free(e.jac); free(e.bilu.D); free(e.bilu.Off);
  return k;
}

/* ---- adaptive mode: CVODE's variable-order variable-step BDF (restatement) ----
 *
 * The published CVODE method (Hindmarsh et al., ACM TOMS 31 (2005) 363-396; the CVODE user
 * guide's "Mathematical considerations"): fixed-leading-coefficient BDF of orders 1..qmax in
 * Nordsieck form zn[j] = h^j y^(j)/j!, predictor zn A (Pascal matrix), corrector
 * zn += l acor, local error test ||acor|| tq[2] <= 1, step/order choice from the q-1, q and q+1
 * error estimates with biases 6, 6, 10, Newton with the rate test ||dx|| min(1, rate) <= 0.1/tq[2]
 * over at most 3 corrections, GMRES to 0.05 of that, the initial step from ||y''||.  CVODE is
 * absent: parity with CVODE itself is unpinned; this restatement is pinned by the reference's KAT
 * (src/test/BDFNewtonKrylov1.f90:20), the constant-step coefficients of BDF1..5 and the closed
 * form of the first step on y' = l y (tests/test_oracle_solver.py).  Weights 1/(rtol|y|+atol). */

#define CV_QMAX 5

typedef struct {
  orc_bdf b; /* n, maxl, restarts, f, psol, ud, t, y (iterate), fy, w (weights), GMRES work */
  orc_psetup_j pset;
  int qmax, q, qprime, qu, qwait, L, maxcor, msbp, msbj, jcur;
  double h, hprime, hscale, hu, eta, etamax, h0, hmin, hmaxInv, dgmax;
  double tau[CV_QMAX + 2], tq[6], l[CV_QMAX + 1], rl1, gamma, gammap, gamrat, crate, savedTq5, acnrm;
  long long nst, nstP, nstJ;
  double *zn, *acor, *dx;
  orc_cv_stats st;
} orc_cv;

static double *cv_z(orc_cv *c, int j) { return c->zn + (size_t)j * c->b.n; }
static double cv_wrms(orc_cv *c, const double *x) { return sqrt(wdot(c->b.n, x, x, c->b.w) / c->b.n); }

static void cv_ewt(orc_cv *c) {
  const double *y = cv_z(c, 0);
  for (int i = 0; i < c->b.n; i++) c->b.w[i] = 1.0 / (c->b.rtol * fabs(y[i]) + c->b.atol);
}

/* zn <- zn A (sign +1) or its inverse sweep (sign -1) */
static void cv_pascal(orc_cv *c, double sign) {
  int n = c->b.n;
  for (int k = 1; k <= c->q; k++)
    for (int j = c->q; j >= k; j--) {
      double *lo = cv_z(c, j - 1), *hi = cv_z(c, j);
      for (int i = 0; i < n; i++) lo[i] += sign * hi[i];
    }
}

static void cv_rescale(orc_cv *c) {
  double f = 1.0;
  for (int j = 1; j <= c->q; j++) {
    f *= c->eta;
    double *z = cv_z(c, j);
    for (int i = 0; i < c->b.n; i++) z[i] *= f;
  }
  c->h = c->hscale * c->eta;
  c->hscale = c->h;
}

static void cv_restore(orc_cv *c, double savedT) {
  c->b.t = savedT;
  cv_pascal(c, -1.0);
}

static void cv_set_coeffs(orc_cv *c) {
  int q = c->q;
  double *l = c->l, *tq = c->tq, h = c->h;
  for (int i = 0; i <= CV_QMAX; i++) l[i] = 0.0;
  l[0] = l[1] = 1.0;
  double xiInv = 1.0, xistarInv = 1.0, alpha0 = -1.0, alpha0Hat = -1.0, hsum = h;
  if (q > 1) {
    for (int j = 2; j < q; j++) {
      hsum += c->tau[j - 1];
      xiInv = h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; i--) l[i] += l[i - 1] * xiInv;
    }
    alpha0 -= 1.0 / q;
    xistarInv = -l[1] - alpha0;
    hsum += c->tau[q - 1];
    xiInv = h / hsum;
    alpha0Hat = -l[1] - xiInv;
    for (int i = q; i >= 1; i--) l[i] += l[i - 1] * xistarInv;
  }
  double A1 = 1.0 - alpha0Hat + alpha0, A2 = 1.0 + q * A1;
  tq[2] = fabs(A1 / (alpha0 * A2));
  tq[5] = fabs(A2 * xistarInv / (l[q] * xiInv));
  if (c->qwait == 1) {
    if (q > 1) {
      double C = xistarInv / l[q], A3 = alpha0 + 1.0 / q, A4 = alpha0Hat + xiInv;
      tq[1] = fabs(C * (1.0 - A4 + A3) / A3);
    } else {
      tq[1] = 1.0;
    }
    hsum += c->tau[q];
    xiInv = h / hsum;
    double A5 = alpha0 - 1.0 / (q + 1), A6 = alpha0Hat - xiInv;
    tq[3] = fabs((1.0 - A6 + A5) / A2 / (xiInv * (q + 2) * A5));
  }
  tq[4] = 0.1 / tq[2];
  c->rl1 = 1.0 / l[1];
  c->gamma = h * c->rl1;
  if (c->nst == 0) c->gammap = c->gamma;
  c->gamrat = c->nst > 0 ? c->gamma / c->gammap : 1.0;
}

static void cv_adjust_order(orc_cv *c, int dq) {
  int q = c->q, n = c->b.n;
  if (q == 2 && dq != 1) return;
  double w[CV_QMAX + 2] = {0};
  if (dq == 1) {
    double alpha0 = -1.0, alpha1 = 1.0, prod = 1.0, xiold = 1.0, hsum = c->hscale;
    w[2] = 1.0;
    for (int j = 1; j < q; j++) {
      hsum += c->tau[j + 1];
      double xi = hsum / c->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; i--) w[i] = w[i] * xiold + w[i - 1];
      xiold = xi;
    }
    double A1 = (-alpha0 - alpha1) / prod;
    double *top = cv_z(c, q + 1), *sv = cv_z(c, c->qmax);
    for (int i = 0; i < n; i++) top[i] = A1 * sv[i];
    for (int j = 2; j <= q; j++) {
      double *z = cv_z(c, j);
      for (int i = 0; i < n; i++) z[i] += w[j] * top[i];
    }
  } else if (dq == -1) {
    double hsum = 0.0;
    w[2] = 1.0;
    for (int j = 1; j <= q - 2; j++) {
      hsum += c->tau[j];
      double xi = hsum / c->hscale;
      for (int i = j + 2; i >= 2; i--) w[i] = w[i] * xi + w[i - 1];
    }
    double *zq = cv_z(c, q);
    for (int j = 2; j < q; j++) {
      double *z = cv_z(c, j);
      for (int i = 0; i < n; i++) z[i] -= w[j] * zq[i];
    }
  }
}

/* initial step: bounds from tOut and y / y', then h = 0.5 sqrt(2/||y''||) iterated */
static int cv_hin(orc_cv *c, double tOut) {
  int n = c->b.n;
  double t = c->b.t, *y0 = cv_z(c, 0), *f0 = cv_z(c, 1);
  double tdist = fabs(tOut - t), tround = 2.220446049250313e-16 * fmax(fabs(t), fabs(tOut));
  if (!(tOut > t) || tdist < 2.0 * tround) return -1;
  double hlb = 100.0 * tround, hubInv = 0.0;
  for (int i = 0; i < n; i++) hubInv = fmax(hubInv, fabs(f0[i]) / (0.1 * fabs(y0[i]) + c->b.atol));
  double hub = 0.1 * tdist;
  if (hub * hubInv > 1.0) hub = 1.0 / hubInv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) { c->h = hg; return 0; }
  int hnewOK = 0;
  double hs = hg, hnew = hg, yddnrm = 0.0;
  for (int count1 = 1; count1 <= 4; count1++) {
    int hgOK = 0;
    for (int count2 = 1; count2 <= 4; count2++) {
      for (int i = 0; i < n; i++) c->b.tmp[i] = y0[i] + hg * f0[i];
      int rc = c->b.f(t + hg, c->b.tmp, c->b.tmp2, c->b.ud);
      if (rc < 0) return rc;
      if (rc == 0) {
        for (int i = 0; i < n; i++) c->b.tmp2[i] = (c->b.tmp2[i] - f0[i]) / hg;
        yddnrm = cv_wrms(c, c->b.tmp2);
        hgOK = 1;
        break;
      }
      hg *= 0.2;
    }
    if (!hgOK) {
      if (count1 <= 2) return -1;
      hnew = hs;
      break;
    }
    hs = hg;
    if (hnewOK || count1 == 4) { hnew = hg; break; }
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) hnewOK = 1;
    if (count1 > 1 && hrat > 2.0) { hnew = hg; hnewOK = 1; }
    hg = hnew;
  }
  c->h = fmin(fmax(0.5 * hnew, hlb), hub);
  return 0;
}

enum { CV_FIRST, CV_PREV_CONV_FAIL, CV_PREV_ERR_FAIL };
enum { CV_NO_FAIL, CV_FAIL_BAD_J, CV_FAIL_OTHER };
enum { CV_OK = 0, CV_RECOVER = 1, CV_TRY_AGAIN = 2 };

static int cv_newton(orc_cv *c) {
  int n = c->b.n;
  double delp = 0.0, *z0 = cv_z(c, 0), *z1 = cv_z(c, 1);
  c->b.absTol = 0.05 * c->tq[4] * sqrt((double)n);
  for (int m = 0;; m++) {
    for (int i = 0; i < n; i++) c->b.G[i] = c->gamma * c->b.fy[i] - c->rl1 * z1[i] - c->acor[i];
    int conv = 0;
    c->st.nli += gmres(&c->b, c->gamma, c->b.G, c->dx, &conv);
    c->st.nni++;
    if (!conv && m > 0) return (!c->jcur && c->pset) ? CV_TRY_AGAIN : CV_RECOVER;
    double del = cv_wrms(c, c->dx);
    for (int i = 0; i < n; i++) {
      c->acor[i] += c->dx[i];
      c->b.y[i] = z0[i] + c->acor[i];
    }
    if (m > 0) c->crate = fmax(0.3 * c->crate, del / delp);
    double dcon = del * fmin(1.0, c->crate) / c->tq[4];
    if (dcon <= 1.0) {
      c->acnrm = m == 0 ? del : cv_wrms(c, c->acor);
      c->jcur = 0;
      return CV_OK;
    }
    if (m + 1 == c->maxcor || (m >= 1 && del > 2.0 * delp)) return (!c->jcur && c->pset) ? CV_TRY_AGAIN : CV_RECOVER;
    delp = del;
    int rc = c->b.f(c->b.t, c->b.y, c->b.fy, c->b.ud);
    if (rc < 0) return rc;
    if (rc > 0) return CV_RECOVER;
  }
}

static int cv_nls(orc_cv *c, int nflag) {
  int n = c->b.n, callSetup = 0, convfail = CV_NO_FAIL;
  if (c->pset) {
    convfail = (nflag == CV_FIRST || nflag == CV_PREV_ERR_FAIL) ? CV_NO_FAIL : CV_FAIL_OTHER;
    callSetup = nflag == CV_PREV_CONV_FAIL || nflag == CV_PREV_ERR_FAIL || c->nst == 0 || c->nstP < 0 ||
                c->nst >= c->nstP + c->msbp || fabs(c->gamrat - 1.0) > c->dgmax;
  } else {
    c->crate = 1.0;
  }
  for (;;) {
    memcpy(c->b.y, cv_z(c, 0), sizeof(double) * n);
    memset(c->acor, 0, sizeof(double) * n);
    int rc = c->b.f(c->b.t, c->b.y, c->b.fy, c->b.ud);
    if (rc < 0) return rc;
    if (rc > 0) return CV_RECOVER;
    if (callSetup) {
      double dgamma = fabs(c->gamma / c->gammap - 1.0);
      int jbad = c->nst == 0 || c->nstJ < 0 || c->nst >= c->nstJ + c->msbj ||
                 (convfail == CV_FAIL_BAD_J && dgamma < 0.2) || convfail == CV_FAIL_OTHER;
      int jc = 0;
      c->st.npe++;
      rc = c->pset(c->b.t, c->b.y, c->b.fy, jbad ? 0 : 1, &jc, c->gamma, c->b.ud);
      c->jcur = jc != 0;
      if (c->jcur) c->nstJ = c->nst;
      c->crate = 1.0;
      c->gammap = c->gamma;
      c->gamrat = 1.0;
      c->nstP = c->nst;
      if (rc < 0) return rc;
      if (rc > 0) return CV_RECOVER;
    }
    rc = cv_newton(c);
    if (rc != CV_TRY_AGAIN) return rc;
    callSetup = 1;
    convfail = CV_FAIL_BAD_J;
  }
}

static void cv_prepare_next(orc_cv *c, double dsm) {
  int n = c->b.n, q = c->q;
  if (c->etamax == 1.0) {
    if (c->qwait < 2) c->qwait = 2;
    c->qprime = q;
    c->hprime = c->h;
    c->eta = 1.0;
    return;
  }
  double etaq = 1.0 / (pow(6.0 * dsm, 1.0 / c->L) + 1e-6);
  if (c->qwait != 0) {
    c->eta = etaq;
    c->qprime = q;
  } else {
    c->qwait = 2;
    double etaqm1 = 0.0, etaqp1 = 0.0;
    if (q > 1) etaqm1 = 1.0 / (pow(6.0 * cv_wrms(c, cv_z(c, q)) * c->tq[1], 1.0 / q) + 1e-6);
    if (q != c->qmax && c->savedTq5 != 0.0) {
      double cquot = (c->tq[5] / c->savedTq5) * pow(c->h / c->tau[2], (double)c->L);
      double *sv = cv_z(c, c->qmax);
      for (int i = 0; i < n; i++) c->b.tmp[i] = c->acor[i] - cquot * sv[i];
      etaqp1 = 1.0 / (pow(10.0 * cv_wrms(c, c->b.tmp) * c->tq[3], 1.0 / (c->L + 1)) + 1e-6);
    }
    double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam < 1.5) {
      c->eta = 1.0;
      c->qprime = q;
    } else if (etam == etaqm1) {
      c->eta = etaqm1;
      c->qprime = q - 1;
    } else if (etam == etaq) {
      c->eta = etaq;
      c->qprime = q;
    } else {
      c->eta = etaqp1;
      c->qprime = q + 1;
      memcpy(cv_z(c, c->qmax), c->acor, sizeof(double) * n);
    }
  }
  if (c->eta < 1.5) {
    c->eta = 1.0;
    c->hprime = c->h;
  } else {
    c->eta = fmin(c->eta, c->etamax);
    if (c->hmaxInv > 0.0) c->eta /= fmax(1.0, fabs(c->h) * c->hmaxInv * c->eta);
    c->hprime = c->h * c->eta;
  }
}

/* one internal step; 0 ok, 1 failed (state unchanged), <0 fatal */
static int cv_step(orc_cv *c, double tOut) {
  int n = c->b.n, rc;
  if (c->nst == 0 && c->hu == 0.0) { /* first step: f(t0, y0), initial step, history */
    cv_ewt(c);
    rc = c->b.f(c->b.t, cv_z(c, 0), cv_z(c, 1), c->b.ud);
    if (rc) return rc < 0 ? rc : -1;
    if (c->h0 > 0.0) c->h = c->h0;
    else if ((rc = cv_hin(c, tOut))) return rc;
    if (c->hmaxInv > 0.0 && c->h * c->hmaxInv > 1.0) c->h = 1.0 / c->hmaxInv;
    if (c->h < c->hmin) c->h = c->hmin;
    double *z1 = cv_z(c, 1);
    for (int i = 0; i < n; i++) z1[i] *= c->h;
    c->hscale = c->hprime = c->h;
    c->q = c->qprime = 1;
    c->L = 2;
    c->qwait = 2;
    c->etamax = 1e4;
    c->crate = 1.0;
    c->hu = -1.0; /* started */
  } else if (c->hprime != c->h) {
    if (c->qprime != c->q) {
      cv_adjust_order(c, c->qprime - c->q);
      c->q = c->qprime;
      c->L = c->q + 1;
      c->qwait = c->L;
    }
    cv_rescale(c);
  }
  cv_ewt(c);
  double savedT = c->b.t, dsm = 0.0;
  int ncf = 0, nef = 0, nflag = CV_FIRST;
  for (;;) {
    c->b.t = savedT + c->h;
    cv_pascal(c, 1.0);
    cv_set_coeffs(c);
    rc = cv_nls(c, nflag);
    if (rc < 0) { cv_restore(c, savedT); return rc; }
    if (rc != CV_OK) {
      c->st.ncfn++;
      ncf++;
      c->etamax = 1.0;
      cv_restore(c, savedT);
      if (fabs(c->h) <= c->hmin * 1.000001 || ncf == 10) return 1;
      c->eta = fmax(0.25, c->hmin / fabs(c->h));
      nflag = CV_PREV_CONV_FAIL;
      cv_rescale(c);
      continue;
    }
    dsm = c->acnrm * c->tq[2];
    if (dsm <= 1.0) break;
    nef++;
    c->st.netf++;
    nflag = CV_PREV_ERR_FAIL;
    cv_restore(c, savedT);
    if (fabs(c->h) <= c->hmin * 1.000001 || nef == 7) return 1;
    c->etamax = 1.0;
    if (nef <= 3) {
      c->eta = 1.0 / (pow(6.0 * dsm, 1.0 / c->L) + 1e-6);
      c->eta = fmax(0.1, fmax(c->eta, c->hmin / fabs(c->h)));
      if (nef >= 2) c->eta = fmin(c->eta, 0.2);
      cv_rescale(c);
      continue;
    }
    c->eta = fmax(0.1, c->hmin / fabs(c->h));
    if (c->q > 1) { /* one order lower, qwait = L (cvDoErrorTest after MXNEF1 failures) */
      cv_adjust_order(c, -1);
      c->L = c->q;
      c->q--;
      c->qwait = c->L;
      cv_rescale(c);
      continue;
    }
    c->h *= c->eta;
    c->hscale = c->h;
    c->qwait = 10;
    rc = c->b.f(c->b.t, cv_z(c, 0), c->b.tmp, c->b.ud);
    if (rc) return rc < 0 ? rc : 1;
    double *z1 = cv_z(c, 1);
    for (int i = 0; i < n; i++) z1[i] = c->h * c->b.tmp[i];
  }
  c->nst++;
  c->hu = c->h;
  c->qu = c->q;
  for (int i = c->q; i >= 2; i--) c->tau[i] = c->tau[i - 1];
  if (c->q == 1 && c->nst > 1) c->tau[2] = c->tau[1];
  c->tau[1] = c->h;
  for (int j = 0; j <= c->q; j++) {
    double *z = cv_z(c, j);
    for (int i = 0; i < n; i++) z[i] += c->l[j] * c->acor[i];
  }
  c->qwait--;
  if (c->qwait == 1 && c->q != c->qmax) {
    memcpy(cv_z(c, c->qmax), c->acor, sizeof(double) * n);
    c->savedTq5 = c->tq[5];
  }
  cv_prepare_next(c, dsm);
  c->etamax = 10.0;
  return 0;
}

int orc_cv_run(int n, orc_f f, orc_psetup_j pset, orc_psolve psol, void *ud, const double *y0, double t0,
               double tOut, double tEnd, int maxSteps, double rtol, double atol, int maxl, int maxRestarts,
               int qmax, double h0, double *ts, double *hs, int *qs, double *states, orc_cv_stats *stats) {
  orc_cv c;
  memset(&c, 0, sizeof(c));
  c.b.n = n; c.b.maxl = maxl; c.b.maxRestarts = maxRestarts; c.b.rtol = rtol; c.b.atol = atol; c.b.t = t0;
  c.b.f = f; c.b.psol = psol; c.b.ud = ud; c.b.linTol = 1e-2;
  c.pset = pset;
  c.qmax = qmax > 0 && qmax <= CV_QMAX ? qmax : CV_QMAX;
  c.maxcor = 3; c.msbp = 20; c.msbj = 51; c.dgmax = 0.3; c.h0 = h0;
  c.nstP = c.nstJ = -1;
  size_t N = n;
  c.b.y = malloc(sizeof(double) * N); c.b.G = malloc(sizeof(double) * N); c.b.fy = malloc(sizeof(double) * N);
  c.b.w = malloc(sizeof(double) * N); c.b.d = malloc(sizeof(double) * N); c.b.tmp = malloc(sizeof(double) * N);
  c.b.tmp2 = malloc(sizeof(double) * N); c.b.V = malloc(sizeof(double) * N * (maxl + 1));
  c.b.H = malloc(sizeof(double) * (maxl + 1) * maxl);
  c.zn = calloc(N * (CV_QMAX + 1), sizeof(double)); c.acor = calloc(N, sizeof(double)); c.dx = calloc(N, sizeof(double));
  memcpy(c.zn, y0, sizeof(double) * N);
  int k = 0;
  for (; k < maxSteps; k++) {
    int rc = cv_step(&c, tOut);
    if (rc) break;
    if (ts) ts[k] = c.b.t;
    if (hs) hs[k] = c.hu;
    if (qs) qs[k] = c.qu;
    if (states) memcpy(states + N * k, c.zn, sizeof(double) * N);
    if (c.b.t >= tEnd) { k++; break; }
  }
  if (stats) { *stats = c.st; stats->nst = c.nst; }
  free(c.b.y); free(c.b.G); free(c.b.fy); free(c.b.w); free(c.b.d); free(c.b.tmp); free(c.b.tmp2); free(c.b.V);
  free(c.b.H); free(c.zn); free(c.acor); free(c.dx);
  return k;
}

/* constant-coefficient check: l[0..q] and tq[2] for order q after the step history tau[1..q] */
void orc_cv_coeffs(int q, double h, const double *tau, double *l, double *tq2) {
  orc_cv c;
  memset(&c, 0, sizeof(c));
  c.q = q;
  c.h = h;
  for (int i = 1; i <= q && i <= CV_QMAX; i++) c.tau[i] = tau[i - 1];
  cv_set_coeffs(&c);
  for (int i = 0; i <= q; i++) l[i] = c.l[i];
  *tq2 = c.tq[2];
}

/* the Euler problem, with a preconditioner setup that reuses the Jacobian when jok = 1 */
static int euler_pset_j(double t, const double *y, const double *fy, int jok, int *jcur, double gamma, void *ud) {
  (void)fy;
  euler_ud *e = (euler_ud *)ud;
  if (!jok) {
    const orc_grid *g = e->g;
    int nE = g->nE;
    double *rho = e->scratch, *rhou = rho + nE, *rhoE = rhou + 3 * (size_t)nE, *u = rhoE + nE;
    double *p = u + 3 * (size_t)nE, *temp = p + nE, *c = temp + nE;
    orc_set_bc(g, e->bc, y, rho, rhou, rhoE);
    orc_sync_state(g, e->bc->r, e->bc->gamm, y, rho, rhou, rhoE, u, p, temp, c);
    orc_find_euler_jac(g, rho, rhou, rhoE, p, e->bc->gamm, e->jac);
  }
  *jcur = !jok;
  (void)t;
  return orc_bilu_setup(&e->bilu, e->jac, e->g->v, gamma);
}

int orc_cv_euler(const orc_grid *g, const orc_bc *bc, const double *y0, double tOut, double tEnd, int maxSteps,
                 double rtol, double atol, int maxl, int maxRestarts, const int *order, double *ts, double *hs,
                 int *qs, double *states, orc_cv_stats *stats) {
  euler_ud e;
  memset(&e, 0, sizeof(e));
  e.g = g; e.bc = bc; e.useIlu = 1;
  e.scratch = calloc(13 * (size_t)g->nE, sizeof(double));
  e.jac = calloc(125 * (size_t)g->nC, sizeof(double));
  e.bilu.nC = g->nC; e.bilu.maxNF = g->maxNF; e.bilu.neib = g->neib; e.bilu.order = order;
  e.bilu.D = calloc(25 * (size_t)g->nC, sizeof(double));
  e.bilu.Off = calloc(25 * (size_t)g->maxNF * g->nC, sizeof(double));
  int k = orc_cv_run(5 * g->nC, euler_f, euler_pset_j, euler_psol, &e, y0, 0.0, tOut, tEnd, maxSteps, rtol, atol,
                     maxl, maxRestarts, 5, 0.0, ts, hs, qs, states, stats);
  free(e.scratch); free(e.jac); free(e.bilu.D); free(e.bilu.Off);
  return k;
}
