/*
 * CPU oracle for the foeuler hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of the reference algorithms
 * (mianzhi/fosolvers).  Every function cites the reference file:line it
 * follows.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load this library, and only as the checker or the CPU baseline; the
 * product library (fosolvers_amd) never links or calls it.
 *
 * Conventions (0-based, C order):
 *   element index e in [0,nE): cells [0,nC), facets [nC,nC+nF)
 *   iNE[maxNNE*e + j]   node j of element e (-1 padding)
 *   iEP[2*i+0] = m (cell), iEP[2*i+1] = n (cell or facet)
 *   normP[3*i+d], pP[3*i+d], p[3*e+d], pN[3*k+d]
 *   neib[maxNF*c + j]   element through face j of cell c, -1 if none
 *   rhou[3*e+d]         (Fortran rhou(d,e))
 *   x[v*nC + c]         SoA solution vector (foeuler.f90:161-166)
 *
 * Parity status of each function is stated at its definition.  Functions
 * without an independent check say "parity unpinned".
 */
#ifndef FOS_ORACLE_H
#define FOS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* shape codes (polyGrid.f90:10-21) */
enum { ORC_TRI = 10, ORC_QUAD = 11, ORC_TRI6 = 110, ORC_TET = 20, ORC_HEX = 21, ORC_TET10 = 120 };
/* BC codes (foeuler.f90:17-25) */
enum { ORC_BC_WALL = 0, ORC_BC_IN_STATIC = 10, ORC_BC_IN_STATIC_UDF = 110, ORC_BC_IN_TOTAL = 11,
       ORC_BC_IN_TOTAL_UDF = 111, ORC_BC_OUT = 20, ORC_BC_OUT_UDF = 120, ORC_BC_FAR = 30,
       ORC_BC_FAR_UDF = 130 };

typedef struct {
  int nN, nE, nC, nF, maxNNE, maxNF;
  int nP, nPi, nPb;
  double *pN;   /* [3nN] */
  int *sE, *nNE, *iNE, *gid;  /* [nE], [nE], [maxNNE nE], [nE] */
  double *p;    /* [3nE] */
  double *v;    /* [nC] */
  int *iEP;     /* [2nP] */
  double *aP, *normP, *pP; /* [nP], [3nP], [3nP] */
  int *neib;    /* [maxNF nC] */
} orc_grid;

/* grid build: polyX%up + sortPolyGrid + volumes + getNeib + sortPolyFvGrid */
orc_grid *orc_grid_build(int nN, const double *pN, int nE, const int *sE, const int *nNE,
                         int maxNNE, const int *iNE, const int *gid);
void orc_grid_free(orc_grid *g);

/* geometry (geometry.f90) */
double orc_a3p(const double *p1, const double *p2, const double *p3);
void orc_n3p(const double *p1, const double *p2, const double *p3, double *n);
double orc_v4p(const double *p1, const double *p2, const double *p3, const double *p4);

/* condTab + iBC mapping (condition.f90:58-75, foeuler.f90:145-154) */
void orc_map_cond(int nE, const int *gid, int nCond, const int *condGid, int *iBC);

/* euler state: setBC / syncState / findEuler / rhs */
typedef struct {
  int nCond, maxPar;
  const int *t;       /* [nCond] BC type */
  const double *par;  /* [maxPar*nCond] */
  const int *iBC;     /* [nE], -1 = unassigned (wall) */
  const double *udf;  /* [5*nPb] pre-evaluated UDF values of boundary pair k (or NULL) */
  double r, gamm;
} orc_bc;

void orc_set_bc(const orc_grid *g, const orc_bc *bc, const double *x,
                double *rho, double *rhou, double *rhoE);
void orc_sync_state(const orc_grid *g, double r, double gamm, const double *x,
                    double *rho, double *rhou, double *rhoE, double *u, double *p,
                    double *temp, double *c);
void orc_find_euler(const orc_grid *g, const double *rho, const double *rhou, const double *rhoE,
                    const double *p, double gamm, double *dRho, double *dRhou, double *dRhoE);
/* full rhs: dxdt = findEuler(setBC+syncState(x)) / v; uses caller scratch of size 13*nE */
void orc_rhs(const orc_grid *g, const orc_bc *bc, const double *x, double *dxdt, double *scratch);

/* approximate Jacobian (Euler.f90:159-333). JacC[(c*25 + col)*5 + row], col in 0..24 */
void orc_euler_nfj(const double *nrm, const double *u, double H, double gamm, double *J /*5x5 col-major*/);
void orc_euler_anfj(const double *nrm, const double *u, double H, double c, double gamm, double *J);
void orc_find_euler_jac(const orc_grid *g, const double *rho, const double *rhou, const double *rhoE,
                        const double *p, double gamm, double *JacC /* [125*nC] */);

/* DQ Jacobian-vector product, CVODE-style sigma = 1/||v||_WRMS */
double orc_wrms(int n, const double *v, const double *y, double rtol, double atol);
void orc_dq_jv(const orc_grid *g, const orc_bc *bc, const double *y, const double *fy,
               const double *v, double sigma, double *Jv, double *work, double *scratch);

/* SPARSKIT restatements (1-based CSR/MSR arrays as in the reference) */
void orc_coocsr(int nrow, int nnz, const double *a, const int *ir, const int *jc,
                double *ao, int *jao, int *iao);
void orc_amux(int n, const double *x, double *y, const double *a, const int *ja, const int *ia);
int orc_ilu0(int n, const double *a, const int *ja, const int *ia, double *alu, int *jlu, int *ju, int *iw);
int orc_ilut(int n, const double *a, const int *ja, const int *ia, int lfil, double droptol,
             double *alu, int *jlu, int *ju, int iwk, double *w, int *jw);
void orc_lusol(int n, const double *y, double *x, const double *alu, const int *jlu, const int *ju);

/* block ILU(0) on the cell-adjacency 5x5 block pattern */
typedef struct {
  int nC, maxNF;
  const int *neib;   /* [maxNF*nC] element index or -1 */
  double *D;         /* [25 nC]   row-major 5x5: inverse of the U diagonal block */
  double *Off;       /* [25 maxNF nC] row-major 5x5 off-diagonal blocks (L for nb<c, U for nb>c) */
  const int *order;  /* [nC] row eliminated at step s (NULL: natural order); "lower" = earlier */
} orc_bilu;
/* M = I - gamma*JacC/v (block rows); returns 0 or the 1-based cell index of a singular pivot */
int orc_bilu_setup(orc_bilu *b, const double *JacC, const double *v, double gamma);
void orc_bilu_solve(const orc_bilu *b, const double *r, double *z); /* SoA vectors */
void orc_bilu_matvec(const orc_grid *g, const double *JacC, const double *v, double gamma,
                     const double *x, double *y);

/* fixed-step BDF1/BDF2 + Newton + left-preconditioned GMRES (solver.c) */
typedef int (*orc_f)(double t, const double *y, double *ydot, void *ud);
typedef int (*orc_psetup)(double t, const double *y, double gamma, void *ud);
typedef int (*orc_psolve)(const double *r, double *z, double gamma, void *ud);
int orc_bdf_run(int n, orc_f f, orc_psetup pset, orc_psolve psol, void *ud, const double *y0, double t0,
                double dt, int nsteps, double rtol, double atol, double newtonTol, double linTol, int maxl,
                int maxRestarts, int maxNewton, double *states, int *iters);
/* adaptive mode (CVODE's variable-order variable-step BDF): per-run counters */
typedef int (*orc_psetup_j)(double t, const double *y, const double *fy, int jok, int *jcur, double gamma,
                            void *ud);
typedef struct {
  long long nst, nni, nli, npe, netf, ncfn;
} orc_cv_stats;
int orc_cv_run(int n, orc_f f, orc_psetup_j pset, orc_psolve psol, void *ud, const double *y0, double t0,
               double tOut, double tEnd, int maxSteps, double rtol, double atol, int maxl, int maxRestarts,
               int qmax, double h0, double *ts, double *hs, int *qs, double *states, orc_cv_stats *stats);
void orc_cv_coeffs(int q, double h, const double *tau, double *l, double *tq2);
int orc_cv_euler(const orc_grid *g, const orc_bc *bc, const double *y0, double tOut, double tEnd, int maxSteps,
                 double rtol, double atol, int maxl, int maxRestarts, const int *order, double *ts, double *hs,
                 int *qs, double *states, orc_cv_stats *stats);
int orc_bdf_euler(const orc_grid *g, const orc_bc *bc, const double *y0, double dt, int nsteps, double rtol,
                  double atol, double newtonTol, double linTol, int maxl, int maxRestarts, int maxNewton, int useIlu,
                  double *states, int *iters);

#ifdef __cplusplus
}
#endif
#endif
