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

static int gmres(orc_bdf *s, double gamma, const double *b, double *x) {
  int n = s->n, m = s->maxl, iters = 0;
  double *V = s->V, *H = s->H;
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m), *g = malloc(sizeof(double) * (m + 1));
  double *yv = malloc(sizeof(double) * m), *hc = malloc(sizeof(double) * (m + 2));
  memset(x, 0, sizeof(double) * n);
  double beta0 = -1.0;
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
    if (beta == 0.0 || (restart > 0 && beta <= s->linTol * beta0)) break;
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
      if (fabs(g[j + 1]) <= s->linTol * beta0 || hh == 0.0) { j++; done = 1; break; }
    }
    for (int k = j - 1; k >= 0; k--) {
      double t = g[k];
      for (int q = k + 1; q < j; q++) t -= H[(size_t)k * m + q] * yv[q];
      yv[k] = t / H[(size_t)k * m + k];
    }
    for (int k = 0; k < j; k++)
      for (int i = 0; i < n; i++) x[i] += yv[k] * V[(size_t)k * n + i];
    if (done) break;
  }
  free(cs); free(sn); free(g); free(yv); free(hc);
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
    lin += gmres(s, gamma, s->G, dx);
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
