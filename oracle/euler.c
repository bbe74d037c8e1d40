/*
 * Oracle (test infrastructure only): Euler residual path restated from the reference.
 *   setBC        src/foeuler/foeuler.f90:209-330
 *   syncState    src/foeuler/foeuler.f90:186-206
 *   findEuler    src/libfosolvers/FVM/Euler.f90:25-156 (face loop :86-153, 1 thread => face order)
 *   rhs          src/foeuler/foeuler.f90:349-375
 *   findEulerJac src/libfosolvers/FVM/Euler.f90:159-249, eulerNFJ :252-276, eulerANFJ :279-333
 *   DQ Jv        SUNDIALS CVODE cvLsDQJtimes (third-party, absent): Jv = (f(y+sig v)-f(y))/sig,
 *                sig = 1/||v||_WRMS, weights 1/(rtol|y|+atol) -- parity unpinned against CVODE
 *                itself; pinned by linearity, Jv(0)=0 and convergence to the analytic J*v.
 * Parity: setBC/syncState/findEuler/rhs/findEulerJac are pinned against the compiled reference
 * (oracle/_ref) to <=1e-14 relative, and by free-stream preservation, discrete conservation,
 * Roe consistency and the NFJ-vs-finite-difference check (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double norm2_3(const double *a) { return sqrt(dot3(a, a)); }

void orc_set_bc(const orc_grid *g, const orc_bc *bc, const double *x,
                double *rho, double *rhou, double *rhoE) {
  const int nC = g->nC;
  const double gamm = bc->gamm, r = bc->r;
  for (int i = g->nPi; i < g->nPi + g->nPb; i++) {
    int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    const double *nrm = g->normP + 3 * i;
    double xm[3] = {x[nC + m], x[2 * nC + m], x[3 * nC + m]};
    int k = bc->iBC[n];
    int t = k < 0 ? ORC_BC_WALL : bc->t[k];
    const double *par = k < 0 ? NULL : bc->par + (size_t)bc->maxPar * k;
    const double *uv = bc->udf ? bc->udf + 5 * (size_t)(i - g->nPi) : NULL;
    double pGst, TGst, ptGst, TtGst, uGst[3], rhoGst, Mach;
    switch (t) {
      case ORC_BC_WALL: {
        /* foeuler.f90:217-228 */
        double dn = dot3(xm, nrm);
        rho[n] = x[m];
        for (int d = 0; d < 3; d++) rhou[3 * n + d] = xm[d] - 2.0 * dn * nrm[d];
        rhoE[n] = x[4 * nC + m];
        break;
      }
      case ORC_BC_IN_STATIC: case ORC_BC_IN_STATIC_UDF: {
        /* foeuler.f90:229-249 */
        const double *q = t == ORC_BC_IN_STATIC ? par : uv;
        pGst = q[0]; TGst = q[1]; uGst[0] = q[2]; uGst[1] = q[3]; uGst[2] = q[4];
        Mach = norm2_3(uGst) / sqrt(gamm * r * TGst);
        rhoGst = pGst / r / TGst;
        if (Mach < 1.0)
          for (int d = 0; d < 3; d++) uGst[d] = xm[d] / rhoGst;
        rho[n] = rhoGst;
        for (int d = 0; d < 3; d++) rhou[3 * n + d] = rhoGst * uGst[d];
        rhoE[n] = rhoGst * (1.0 / (gamm - 1.0) * r * TGst + 0.5 * dot3(uGst, uGst));
        break;
      }
      case ORC_BC_IN_TOTAL: case ORC_BC_IN_TOTAL_UDF: {
        /* foeuler.f90:250-276 */
        const double *q = t == ORC_BC_IN_TOTAL ? par : uv;
        ptGst = q[0]; TtGst = q[1]; uGst[0] = q[2]; uGst[1] = q[3]; uGst[2] = q[4];
        TGst = TtGst - 0.5 * (gamm - 1.0) / gamm / r * dot3(uGst, uGst);
        Mach = norm2_3(uGst) / sqrt(gamm * r * TGst);
        pGst = ptGst * pow(1.0 + 0.5 * (gamm - 1.0) * (Mach * Mach), -gamm / (gamm - 1.0));
        rhoGst = pGst / r / TGst;
        if (Mach < 1.0) {
          double ui[3] = {xm[0] / x[m], xm[1] / x[m], xm[2] / x[m]};
          double s = dot3(ui, uGst);
          double uu = dot3(uGst, uGst);
          for (int d = 0; d < 3; d++) uGst[d] = s * uGst[d] / uu;
          TGst = TtGst - 0.5 * (gamm - 1.0) / gamm / r * dot3(uGst, uGst);
          Mach = norm2_3(uGst) / sqrt(gamm * r * TGst);
          pGst = ptGst * pow(1.0 + 0.5 * (gamm - 1.0) * (Mach * Mach), -gamm / (gamm - 1.0));
          rhoGst = pGst / r / TGst;
        }
        rho[n] = rhoGst;
        for (int d = 0; d < 3; d++) rhou[3 * n + d] = rhoGst * uGst[d];
        rhoE[n] = rhoGst * (1.0 / (gamm - 1.0) * r * TGst + 0.5 * dot3(uGst, uGst));
        break;
      }
      case ORC_BC_OUT: case ORC_BC_OUT_UDF: {
        /* foeuler.f90:277-287 */
        pGst = t == ORC_BC_OUT ? par[0] : uv[0];
        rho[n] = x[m];
        for (int d = 0; d < 3; d++) rhou[3 * n + d] = xm[d];
        rhoE[n] = 1.0 / (gamm - 1.0) * pGst + 0.5 * dot3(xm, xm) / x[m];
        break;
      }
      case ORC_BC_FAR: case ORC_BC_FAR_UDF: {
        /* foeuler.f90:288-325 (Riemann invariants) */
        const double *q = t == ORC_BC_FAR ? par : uv;
        pGst = q[0]; TGst = q[1]; uGst[0] = q[2]; uGst[1] = q[3]; uGst[2] = q[4];
        double charO = dot3(uGst, nrm) - 2.0 * sqrt(gamm * r * TGst) / (gamm - 1.0);
        double ui[3] = {xm[0] / x[m], xm[1] / x[m], xm[2] / x[m]};
        double eint = x[4 * nC + m] - 0.5 * dot3(ui, ui) * x[m];
        double charI = dot3(ui, nrm) + 2.0 * sqrt(gamm / x[m] / (gamm - 1.0) * eint);
        double un[3], ut[3];
        for (int d = 0; d < 3; d++) un[d] = 0.5 * nrm[d] * (charI + charO);
        double a = 0.25 * (gamm - 1.0) * (charI - charO);
        if (dot3(un, nrm) > 0.0) {
          double nu = dot3(nrm, ui);
          for (int d = 0; d < 3; d++) ut[d] = ui[d] - nrm[d] * nu;
          Mach = norm2_3(ui) / sqrt(gamm / x[m] * (gamm - 1.0) * eint);
          ptGst = (gamm - 1.0) * eint * pow(1.0 + 0.5 * (gamm - 1.0) * (Mach * Mach), gamm / (gamm - 1.0));
        } else {
          for (int d = 0; d < 3; d++) ut[d] = uGst[d] - un[d];
          Mach = norm2_3(uGst) / sqrt(gamm * r * TGst);
          ptGst = pGst * pow(1.0 + 0.5 * (gamm - 1.0) * (Mach * Mach), gamm / (gamm - 1.0));
        }
        for (int d = 0; d < 3; d++) uGst[d] = un[d] + ut[d];
        TGst = a * a / gamm / r;
        Mach = norm2_3(uGst) / a;
        pGst = ptGst * pow(1.0 + 0.5 * (gamm - 1.0) * (Mach * Mach), -gamm / (gamm - 1.0));
        rhoGst = pGst / r / TGst;
        rho[n] = rhoGst;
        for (int d = 0; d < 3; d++) rhou[3 * n + d] = rhoGst * uGst[d];
        rhoE[n] = rhoGst * (1.0 / (gamm - 1.0) * r * TGst + 0.5 * dot3(uGst, uGst));
        break;
      }
      default:
        break; /* foeuler.f90:326 case default: ghost slot untouched */
    }
  }
}

void orc_sync_state(const orc_grid *g, double r, double gamm, const double *x,
                    double *rho, double *rhou, double *rhoE, double *u, double *p,
                    double *temp, double *c) {
  const int nC = g->nC;
  for (int i = 0; i < nC; i++) {
    rho[i] = x[i];
    rhou[3 * i] = x[nC + i];
    rhou[3 * i + 1] = x[2 * nC + i];
    rhou[3 * i + 2] = x[3 * nC + i];
    rhoE[i] = x[4 * nC + i];
  }
  for (int i = 0; i < g->nE; i++) {
    for (int d = 0; d < 3; d++) u[3 * i + d] = rhou[3 * i + d] / rho[i];
    p[i] = (gamm - 1.0) * (rhoE[i] - 0.5 * dot3(rhou + 3 * i, rhou + 3 * i) / rho[i]);
    temp[i] = p[i] / rho[i] / r;
    c[i] = sqrt(gamm * r * temp[i]);
  }
}

void orc_find_euler(const orc_grid *g, const double *rho, const double *rhou, const double *rhoE,
                    const double *p, double gamm, double *dRho, double *dRhou, double *dRhoE) {
  const int nC = g->nC, k = g->nE;
  memset(dRho, 0, sizeof(double) * nC);
  memset(dRhou, 0, sizeof(double) * 3 * nC);
  memset(dRhoE, 0, sizeof(double) * nC);
  /* cell auxiliaries (Euler.f90:73-81) */
  double *u = (double *)malloc(sizeof(double) * 3 * (size_t)k);
  double *H = (double *)malloc(sizeof(double) * (size_t)k);
  double *f2c = (double *)malloc(sizeof(double) * 9 * (size_t)k); /* f2c(d,j,i) -> [9i + 3j + d] */
  double *f3c = (double *)malloc(sizeof(double) * 3 * (size_t)k);
  for (int i = 0; i < k; i++) {
    for (int d = 0; d < 3; d++) u[3 * i + d] = rhou[3 * i + d] / rho[i];
    H[i] = (rhoE[i] + p[i]) / rho[i];
    for (int jj = 0; jj < 3; jj++)
      for (int d = 0; d < 3; d++)
        f2c[9 * i + 3 * jj + d] = rhou[3 * i + jj] * u[3 * i + d] + (d == jj ? p[i] : 0.0);
    for (int d = 0; d < 3; d++) f3c[3 * i + d] = (rhoE[i] + p[i]) * u[3 * i + d];
  }
  /* Roe flux-difference splitting, face order (Euler.f90:86-153) */
  for (int i = 0; i < g->nP; i++) {
    int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    if (!(m < k && n < k)) continue;
    const double *nr = g->normP + 3 * i;
    const double *um = u + 3 * m, *un = u + 3 * n;
    double srm = sqrt(rho[m]), srn = sqrt(rho[n]);
    double rhoAvg = sqrt(rho[m] * rho[n]);
    double uAvg[3];
    for (int d = 0; d < 3; d++) uAvg[d] = (srm * um[d] + srn * un[d]) / (srm + srn);
    double uNormAvg = dot3(uAvg, nr);
    double HAvg = (srm * H[m] + srn * H[n]) / (srm + srn);
    double cAvg = sqrt((gamm - 1.0) * (HAvg - 0.5 * dot3(uAvg, uAvg)));
    double rhoJump = rho[n] - rho[m];
    double pJump = p[n] - p[m];
    double uJump[3] = {un[0] - um[0], un[1] - um[1], un[2] - um[2]};
    double uNormJump = dot3(uJump, nr);
    double c2 = cAvg * cAvg;
    double s1 = rhoJump - pJump / c2;
    double ent[5];
    ent[0] = fabs(uNormAvg) * (s1 * 1.0 + rhoAvg * 0.0);
    for (int d = 0; d < 3; d++)
      ent[1 + d] = fabs(uNormAvg) * (s1 * uAvg[d] + rhoAvg * (uJump[d] - nr[d] * uNormJump));
    ent[4] = fabs(uNormAvg) * (s1 * (0.5 * dot3(uAvg, uAvg)) + rhoAvg * (dot3(uAvg, uJump) - uNormAvg * uNormJump));
    double w1 = fabs(uNormAvg - cAvg) * (0.5 * (pJump - rhoAvg * cAvg * uNormJump) / c2);
    double w2 = fabs(uNormAvg + cAvg) * (0.5 * (pJump + rhoAvg * cAvg * uNormJump) / c2);
    double ac1[5], ac2[5];
    ac1[0] = w1 * 1.0;
    ac2[0] = w2 * 1.0;
    for (int d = 0; d < 3; d++) {
      ac1[1 + d] = w1 * (uAvg[d] - nr[d] * cAvg);
      ac2[1 + d] = w2 * (uAvg[d] + nr[d] * cAvg);
    }
    ac1[4] = w1 * (HAvg - uNormAvg * cAvg);
    ac2[4] = w2 * (HAvg + uNormAvg * cAvg);
    double cf[5], tmp[3];
    for (int d = 0; d < 3; d++) tmp[d] = 0.5 * (rhou[3 * m + d] + rhou[3 * n + d]);
    cf[0] = dot3(tmp, nr);
    for (int jj = 0; jj < 3; jj++) {
      for (int d = 0; d < 3; d++) tmp[d] = 0.5 * (f2c[9 * m + 3 * jj + d] + f2c[9 * n + 3 * jj + d]);
      cf[1 + jj] = dot3(tmp, nr);
    }
    for (int d = 0; d < 3; d++) tmp[d] = 0.5 * (f3c[3 * m + d] + f3c[3 * n + d]);
    cf[4] = dot3(tmp, nr);
    double flow[5];
    for (int q = 0; q < 5; q++) flow[q] = g->aP[i] * (cf[q] - 0.5 * (ent[q] + ac1[q] + ac2[q]));
    dRho[m] -= flow[0];
    for (int d = 0; d < 3; d++) dRhou[3 * m + d] -= flow[1 + d];
    dRhoE[m] -= flow[4];
    if (n < nC) {
      dRho[n] += flow[0];
      for (int d = 0; d < 3; d++) dRhou[3 * n + d] += flow[1 + d];
      dRhoE[n] += flow[4];
    }
  }
  free(u); free(H); free(f2c); free(f3c);
}

void orc_rhs(const orc_grid *g, const orc_bc *bc, const double *x, double *dxdt, double *scratch) {
  /* foeuler.f90:349-375; scratch holds the module arrays rho,rhou,rhoE,u,p,temp,c (13 nE)
   * plus dRho,dRhou,dRhoE. The caller keeps scratch between calls like the module globals. */
  const int nE = g->nE, nC = g->nC;
  double *rho = scratch, *rhou = rho + nE, *rhoE = rhou + 3 * (size_t)nE, *u = rhoE + nE;
  double *p = u + 3 * (size_t)nE, *temp = p + nE, *c = temp + nE;
  double *dR = (double *)malloc(sizeof(double) * 5 * (size_t)nC);
  orc_set_bc(g, bc, x, rho, rhou, rhoE);
  orc_sync_state(g, bc->r, bc->gamm, x, rho, rhou, rhoE, u, p, temp, c);
  orc_find_euler(g, rho, rhou, rhoE, p, bc->gamm, dR, dR + nC, dR + 4 * (size_t)nC);
  for (int i = 0; i < nC; i++) {
    dxdt[i] = dR[i] / g->v[i];
    for (int d = 0; d < 3; d++) dxdt[(size_t)(1 + d) * nC + i] = dR[nC + 3 * (size_t)i + d] / g->v[i];
    dxdt[4 * (size_t)nC + i] = dR[4 * (size_t)nC + i] / g->v[i];
  }
  free(dR);
}

/* ---- approximate Jacobian ---- */

/* dger: A := alpha*x*y' + A, A 3x3 column-major (BLAS reference semantics) */
static void dger3(double alpha, const double *x, const double *y, double *A) {
  for (int j = 0; j < 3; j++) {
    double t = alpha * y[j];
    for (int i = 0; i < 3; i++) A[3 * j + i] += x[i] * t;
  }
}

/* J is 5x5 column-major: J[5*col + row] */
void orc_euler_nfj(const double *nrm, const double *u, double H, double gamm, double *J) {
  /* Euler.f90:252-276 */
  double v = norm2_3(u), vn = dot3(u, nrm), tmp[9] = {0};
  dger3(1.0, u, nrm, tmp);
  dger3(1.0 - gamm, nrm, u, tmp);
  for (int i = 0; i < 3; i++) tmp[4 * i] += vn;
  memset(J, 0, sizeof(double) * 25);
  double v2 = v * v;
  for (int d = 0; d < 3; d++) J[1 + d] = 0.5 * (gamm - 1.0) * v2 * nrm[d] - u[d] * vn;
  J[4] = (0.5 * (gamm - 1.0) * v2 - H) * vn;
  for (int d = 0; d < 3; d++) J[5 * (1 + d)] = nrm[d];
  for (int cI = 0; cI < 3; cI++)
    for (int rI = 0; rI < 3; rI++) J[5 * (1 + cI) + 1 + rI] = tmp[3 * cI + rI];
  for (int d = 0; d < 3; d++) J[5 * (1 + d) + 4] = H * nrm[d] - (gamm - 1.0) * u[d] * vn;
  for (int d = 0; d < 3; d++) J[5 * 4 + 1 + d] = (gamm - 1.0) * nrm[d];
  J[24] = gamm * vn;
}

void orc_euler_anfj(const double *nrm, const double *u, double H, double c, double gamm, double *J) {
  /* Euler.f90:279-333 */
  double v = norm2_3(u), vn = dot3(u, nrm), M = v / c, Mn = vn / c, M2 = M * M, v2 = v * v;
  double tmp1[9] = {0}, t2[25];
  double g1 = gamm - 1.0, c2 = c * c;
  dger3(g1 / c2, u, u, tmp1);
  dger3(-1.0, nrm, nrm, tmp1);
  for (int i = 0; i < 3; i++) tmp1[4 * i] += 1.0;
#define T2(rw, cl) t2[5 * ((cl) - 1) + (rw) - 1]
  T2(1, 1) = 1.0 - 0.5 * g1 * M2;
  for (int d = 0; d < 3; d++) T2(2 + d, 1) = -(1.0 + 0.5 * g1 * M2) * u[d] + vn * nrm[d];
  T2(5, 1) = vn * vn - 0.5 * v2 * (1.0 + 0.5 * g1 * M2);
  for (int d = 0; d < 3; d++) T2(1, 2 + d) = g1 / c2 * u[d];
  for (int cI = 0; cI < 3; cI++) for (int rI = 0; rI < 3; rI++) T2(2 + rI, 2 + cI) = tmp1[3 * cI + rI];
  for (int d = 0; d < 3; d++) T2(5, 2 + d) = (1.0 + 0.5 * g1 * M2) * u[d] - vn * nrm[d];
  T2(1, 5) = -g1 / c2;
  for (int d = 0; d < 3; d++) T2(2 + d, 5) = -g1 / c2 * u[d];
  T2(5, 5) = -0.5 * g1 * M2;
  for (int q = 0; q < 25; q++) J[q] = fabs(vn) * t2[q];
  for (int sgn = -1; sgn <= 1; sgn += 2) {
    /* sgn=-1: acoustic wave u-c n (Euler.f90:309-320); sgn=+1: u+c n (:321-332) */
    double a[3], b[3];
    for (int d = 0; d < 3; d++) {
      a[d] = u[d] + sgn * c * nrm[d];
      b[d] = -0.5 * g1 / c2 * u[d] + sgn * 0.5 * nrm[d] / c;
    }
    double hh = H + sgn * vn * c;
    memset(tmp1, 0, sizeof(tmp1));
    dger3(1.0, a, b, tmp1);
    T2(1, 1) = 0.25 * g1 * M2 - sgn * 0.5 * Mn;
    for (int d = 0; d < 3; d++) T2(2 + d, 1) = a[d] * T2(1, 1);
    T2(5, 1) = hh * T2(1, 1);
    for (int d = 0; d < 3; d++) T2(1, 2 + d) = -0.5 * g1 / c2 * u[d] + sgn * 0.5 * nrm[d] / c;
    for (int cI = 0; cI < 3; cI++) for (int rI = 0; rI < 3; rI++) T2(2 + rI, 2 + cI) = tmp1[3 * cI + rI];
    for (int d = 0; d < 3; d++) T2(5, 2 + d) = hh * T2(1, 2 + d);
    T2(1, 5) = 0.5 * g1 / c2;
    for (int d = 0; d < 3; d++) T2(2 + d, 5) = a[d] * T2(1, 5);
    T2(5, 5) = hh * T2(1, 5);
    double lam = fabs(vn + sgn * c);
    for (int q = 0; q < 25; q++) J[q] = J[q] + lam * t2[q];
  }
#undef T2
}

void orc_find_euler_jac(const orc_grid *g, const double *rho, const double *rhou, const double *rhoE,
                        const double *p, double gamm, double *JacC) {
  /* Euler.f90:159-249. JacC(5, 5*(1+maxNF), nC) column-major -> [(c*ncol + col)*5 + row] */
  const int nC = g->nC, k = g->nE, nb = g->maxNF, ncol = 5 * (1 + nb);
  memset(JacC, 0, sizeof(double) * 5 * (size_t)ncol * nC);
  double *u = (double *)malloc(sizeof(double) * 3 * (size_t)k);
  double *H = (double *)malloc(sizeof(double) * (size_t)k);
  for (int i = 0; i < k; i++) {
    for (int d = 0; d < 3; d++) u[3 * i + d] = rhou[3 * i + d] / rho[i];
    H[i] = (rhoE[i] + p[i]) / rho[i];
  }
  double JP[100]; /* JacP(10,10) column-major: JP[10*col + row] */
  double A1[25], A2[25], Aa[25];
  for (int i = 0; i < g->nP; i++) {
    int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    memset(JP, 0, sizeof(JP));
    if (m < k && n < k) {
      double srm = sqrt(rho[m]), srn = sqrt(rho[n]);
      double uAvg[3];
      for (int d = 0; d < 3; d++) uAvg[d] = (srm * u[3 * m + d] + srn * u[3 * n + d]) / (srm + srn);
      double HAvg = (srm * H[m] + srn * H[n]) / (srm + srn);
      double cAvg = sqrt((gamm - 1.0) * (HAvg - 0.5 * dot3(uAvg, uAvg)));
      const double *nr = g->normP + 3 * i;
      orc_euler_nfj(nr, u + 3 * m, H[m], gamm, A1);
      orc_euler_anfj(nr, uAvg, HAvg, cAvg, gamm, Aa);
      for (int cI = 0; cI < 5; cI++)
        for (int rI = 0; rI < 5; rI++) JP[10 * cI + rI] = -0.5 * g->aP[i] * (A1[5 * cI + rI] + Aa[5 * cI + rI]);
      if (n < nC) {
        orc_euler_nfj(nr, u + 3 * n, H[n], gamm, A2);
        for (int cI = 0; cI < 5; cI++)
          for (int rI = 0; rI < 5; rI++) JP[10 * (5 + cI) + rI] = -0.5 * g->aP[i] * (A2[5 * cI + rI] - Aa[5 * cI + rI]);
        for (int cI = 0; cI < 10; cI++)
          for (int rI = 0; rI < 5; rI++) JP[10 * cI + 5 + rI] = -JP[10 * cI + rI];
      }
    }
    /* copy Jacobian to cell (Euler.f90:225-248) */
    if (m < nC)
      for (int cI = 0; cI < 5; cI++) for (int rI = 0; rI < 5; rI++)
        JacC[((size_t)m * ncol + cI) * 5 + rI] += JP[10 * cI + rI];
    if (n < nC)
      for (int cI = 0; cI < 5; cI++) for (int rI = 0; rI < 5; rI++)
        JacC[((size_t)n * ncol + cI) * 5 + rI] += JP[10 * (5 + cI) + 5 + rI];
    if (m < nC && n < nC) {
      for (int j = 0; j < nb; j++)
        if (g->neib[(size_t)nb * m + j] == n) {
          for (int cI = 0; cI < 5; cI++) for (int rI = 0; rI < 5; rI++)
            JacC[((size_t)m * ncol + 5 * (j + 1) + cI) * 5 + rI] += JP[10 * (5 + cI) + rI];
          break;
        }
      for (int j = 0; j < nb; j++)
        if (g->neib[(size_t)nb * n + j] == m) {
          for (int cI = 0; cI < 5; cI++) for (int rI = 0; rI < 5; rI++)
            JacC[((size_t)n * ncol + 5 * (j + 1) + cI) * 5 + rI] += JP[10 * cI + 5 + rI];
          break;
        }
    }
  }
  free(u); free(H);
}

/* WRMS norm with weights 1/(rtol|y_i| + atol + rtol rms(y)): CVODE's 1/(rtol|y_i| + atol) plus a
 * floor at rtol times the state's RMS, so exactly-zero components (zero transverse momentum) do not
 * turn atol = 1e-12 into a 1e12 weight on rounding noise. */
double orc_wrms(int n, const double *v, const double *y, double rtol, double atol) {
  double q = 0.0;
  for (int i = 0; i < n; i++) q += y[i] * y[i];
  double floor_ = atol + rtol * sqrt(q / n);
  double s = 0.0;
  for (int i = 0; i < n; i++) {
    double w = 1.0 / (rtol * fabs(y[i]) + floor_);
    s += (v[i] * w) * (v[i] * w);
  }
  return sqrt(s / n);
}

void orc_dq_jv(const orc_grid *g, const orc_bc *bc, const double *y, const double *fy,
               const double *v, double sigma, double *Jv, double *work, double *scratch) {
  size_t n = 5 * (size_t)g->nC;
  for (size_t i = 0; i < n; i++) work[i] = sigma * v[i] + y[i];
  orc_rhs(g, bc, work, Jv, scratch);
  double si = 1.0 / sigma;
  for (size_t i = 0; i < n; i++) Jv[i] = si * Jv[i] - si * fy[i];
}
