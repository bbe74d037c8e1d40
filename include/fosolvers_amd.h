/*
 * fosolvers_amd -- MI355X-native (gfx950) hot path of foeuler behind a C ABI.
 *
 * Plain C types only: pointers, sizes, doubles.  Arrays are 0-based; the reference's
 * Fortran arrays are 1-based and column-major, so a Fortran caller passes
 * iEP(2,nP)-1 etc.
 *
 * Return convention (SUNDIALS callbacks, src/libfosolvers/numerics/SUNDIALS.f90):
 *   0 = success, >0 = recoverable failure (integrator retries with a smaller step),
 *   <0 = unrecoverable.  fos_last_error() returns a message for the last failure on
 *   the calling thread.
 *
 * Vectors (y, ydot, v, Jv, r, z) are the foeuler solution layout: 5*nC doubles in SoA
 * variable blocks rho | rhou_x | rhou_y | rhou_z | rhoE (src/foeuler/foeuler.f90:161-166).
 * They are DEVICE pointers unless the context was created with FOS_VECTORS_ON_HOST.
 */
#ifndef FOSOLVERS_AMD_H
#define FOSOLVERS_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------------------
 * Mesh: polyGrid / polyFvGrid  (src/libfosolvers/polyGrid/polyGrid.f90:24-34,
 * polyFvGrid.f90:14-29, polyX.f90:11-25)
 * ------------------------------------------------------------------------------------- */

/* element shape codes (polyGrid.f90:10-21) */
enum { FOS_TRI = 10, FOS_QUAD = 11, FOS_TRI6 = 110, FOS_TET = 20, FOS_HEX = 21, FOS_TET10 = 120 };

/* raw element lists, what readGMSH produces (src/libfosolvers/fileIO/GMSH.f90:4-157) */
typedef struct {
  int nN, nE, maxNNE;
  double *pN;           /* [3 nN] node positions */
  int *sE;              /* [nE] shape */
  int *nNE;             /* [nE] nodes per element */
  int *iNE;             /* [maxNNE nE] node indices, -1 padded */
  int *gid;             /* [nE] geometric group id */
} fos_poly_grid;

/* finite-volume grid after up(): cells [0,nC) first, facets [nC,nE) after
 * (polyGrid.f90:88-136); pairs with boundary pairs last (polyFvGrid.f90:201-227). */
typedef struct {
  int nN, nE, nC, nF, nP, nPi, nPb, maxNF, maxNNE;
  double *pN;           /* [3 nN] */
  int *sE, *nNE, *iNE, *gid;   /* [nE], [nE], [maxNNE nE], [nE] */
  double *p;            /* [3 nE] element centroids */
  double *v;            /* [nC] cell volumes */
  int *iEP;             /* [2 nP] (m, n): m = cell, n = neighbour cell (> m) or facet */
  double *aP;           /* [nP] interface area */
  double *normP;        /* [3 nP] unit normal, outward from m */
  double *pP;           /* [3 nP] interface centre */
  int *neib;            /* [maxNF nC] element through face j of cell c, -1 if none */
} fos_fv_grid;

/* Synthetic box [0,lx]x[0,ly]x[0,lz] of nx*ny*nz hexes, each split into 6 Kuhn tets, plus two
 * boundary triangles per boundary quad.  Interior nodes are jittered by U(-jitter,+jitter) of
 * the spacing per axis, boundary nodes tangentially only.  gids[0..5] tag the sides
 * x-, x+, y-, y+, z-, z+; gids[6] tags the cells.  No reference counterpart (the reference
 * reads meshes from GMSH); it stands in for the mesher. */
int fos_grid_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                 uint64_t seed, const int gids[7], fos_poly_grid *out);
/* GMSH v2.2 / v4.1 ASCII reader with the reference's semantics: gid = elementary entity tag
 * (v2, GMSH.f90:88) or entity tag (v4, :67); unsupported element types are skipped. */
int fos_grid_read_gmsh(const char *path, fos_poly_grid *out);
/* GMSH v2.2 ASCII writer (physical = elementary tag = gid) so the reference reads the same mesh. */
int fos_grid_write_gmsh(const char *path, const fos_poly_grid *g);
/* In-place renumbering of the cells (3-D elements) of g; facets keep their order.
 * FOS_RENUMBER_RCM: reverse Cuthill-McKee on the face-adjacency graph.
 * FOS_RENUMBER_CHUNKS: recursive coordinate bisection into compact chunks of
 * FOS_CHUNK_CELLS cells, the layout the residual kernel stages through LDS; the relative
 * order inside a chunk is kept.  perm (optional, [number of cells]) receives old index of
 * new cell k (cells counted in element order).  Returns -1 (fos_last_error says why) for an
 * unknown method or an element that names a node outside [0, nN); g is then unchanged. */
enum { FOS_RENUMBER_RCM = 1, FOS_RENUMBER_CHUNKS = 2 };
#define FOS_CHUNK_CELLS 256
int fos_grid_renumber(fos_poly_grid *g, int method, int *perm);
/* polyFvGrid%up (polyFvGrid.f90:48-57): sort elements, centroids, volumes, face pairs,
 * neighbour lists, pair order by max(m,n).  Faces are matched by sorted node keys; the set of
 * pairs, their orientation and geometry equal the reference's getNeib (:60-198).  Pairs with the
 * same max(m,n) are ordered by (m, face of m); the reference's own tie order is what
 * fos_grid_sort_pairs_reference gives.  Returns -1 for an element list that fails the node checks
 * (a node id outside [0, nN), a node count that differs from the element's shape). */
int fos_grid_up(const fos_poly_grid *in, fos_fv_grid *out);
/* sortPolyFvGrid's exact pair order (polyFvGrid.f90:201-227), in place on a built grid (iEP, aP,
 * normP, pP permuted): pairs in getNeib's discovery order (cell, face of the cell, neighbour)
 * sorted by max(m, n) with the tie order of the reference's quicksort (misc/sort.f90:30-65).
 * For a host that indexes pairs as the reference does, e.g. the UDF boundary values of
 * fos_euler_set_udf_bc.  Boundary pairs stay last; the residual is the same up to rounding. */
int fos_grid_sort_pairs_reference(fos_fv_grid *g);
/* polyFvGrid%up on the GPU: the same fos_fv_grid as fos_grid_up (host arrays, freed by
 * fos_fv_grid_free) with centroids, volumes, face matching (device radix sorts of face node
 * keys), pair order and pair geometry computed on the device; stream: a hipStream_t or NULL.
 * Refuses meshes whose face records or pair keys exceed 32/64 bits.  No reference counterpart
 * beyond up() itself (its serial O(nE k) search is what this replaces). */
int fos_grid_up_gpu(const fos_poly_grid *in, fos_fv_grid *out, void *stream);
void fos_poly_grid_free(fos_poly_grid *g);
void fos_fv_grid_free(fos_fv_grid *g);

/* ----------------------------------------------------------------------------------------
 * Conditions and UDFs
 * ------------------------------------------------------------------------------------- */

/* BC codes (src/foeuler/foeuler.f90:17-25) */
enum { FOS_BC_WALL = 0, FOS_BC_IN_STATIC = 10, FOS_BC_IN_STATIC_UDF = 110, FOS_BC_IN_TOTAL = 11,
       FOS_BC_IN_TOTAL_UDF = 111, FOS_BC_OUT = 20, FOS_BC_OUT_UDF = 120, FOS_BC_FAR = 30,
       FOS_BC_FAR_UDF = 130 };

/* condTab (src/libfosolvers/condition/condition.f90:8-16): p[maxPar*k + j] = p(j+1,k+1) */
typedef struct {
  int n, maxPar;
  int *gid;
  int *t;
  double *p;
} fos_cond_tab;
/* readCondTab (src/libfosolvers/fileIO/condTab.f90:4-21). Library-allocated arrays. */
int fos_read_cond_tab(const char *path, fos_cond_tab *out);
void fos_cond_tab_free(fos_cond_tab *c);

/* Expression evaluator with libmatheval's C entry points (bound by
 * src/libfosolvers/UDF/UDF.f90:24-46): f(x,y,z,t) with + - * / ^, unary minus, parentheses,
 * numbers, variables, and exp log sqrt sin cos tan asin acos atan sinh cosh tanh abs step
 * delta nandelta erf.  evaluator_create returns NULL on a parse error. */
void *evaluator_create(char *string);
void evaluator_destroy(void *evaluator);
double evaluator_evaluate(void *evaluator, int count, char **names, double *values);
/* readUDFTab (src/libfosolvers/fileIO/UDFTab.f90:4-19): first line n, then one expression per
 * line.  Returns the number of UDFs (<0 on error) and fills handles[0..n-1] (caller array). */
int fos_read_udf_tab(const char *path, void **handles, int maxHandles);

/* ----------------------------------------------------------------------------------------
 * Euler context: device-resident mesh + BC tables, the residual/Jv/preconditioner callbacks
 * ------------------------------------------------------------------------------------- */

typedef struct fos_euler fos_euler;

enum {
  FOS_VECTORS_ON_HOST = 1u << 0,   /* callbacks take host pointers (copied per call) */
  FOS_PREC_FP32 = 1u << 1,         /* Jv (flux) and preconditioner apply in FP32; rhs stays FP64 */
  FOS_LAYOUT_GENERIC = 1u << 2,    /* per-face area and normal even where the tetrahedral layout applies */
  FOS_JAC_EXACT = 1u << 3,         /* psetup / fos_find_euler_jac: the residual's exact Jacobian blocks
                                      (dissipation differentiated, boundary faces through their ghost
                                      states) instead of findEulerJac's frozen-|A~| approximation */
  FOS_ASYNC_CALLBACKS = 1u << 4,   /* the callbacks return without waiting for the GPU, so a non-finite
                                      output is reported by fos_euler_nonfinite, not the return code */
};

/* Upload the polyFvGrid arrays, the condition table and the gas constants (fl file:
 * r, gamma) and build the device layout.  udfBcValues may be NULL when no *_UDF BC is used;
 * otherwise see fos_euler_set_udf_bc.  stream: a hipStream_t or NULL for the default stream. */
fos_euler *fos_euler_create(const fos_fv_grid *grid, const fos_cond_tab *bc, double r,
                            double gamma, unsigned flags, void *stream);
void fos_euler_destroy(fos_euler *ctx);
/* Pre-evaluated UDF BC parameters, 5 doubles per boundary pair in pair order
 * (nPi..nP-1): the reference evaluates them at pP and the module time t of the last
 * accepted step (foeuler.f90:236-299), so they are constant within a step. */
int fos_euler_set_udf_bc(fos_euler *ctx, const double *values /* host [5 nPb] */);

/* CVRhsFn -- foeuler rhs (foeuler.f90:349-375): ydot = findEuler(setBC(y)) / v */
int fos_euler_rhs(double t, const double *y, double *ydot, void *user_data);
/* CVLsJacTimesVecFn -- CVODE's difference-quotient Jv (foeuler sets none, so CVODE's DQ is
 * what the reference runs): Jv = (f(y + sigma v) - fy) / sigma, sigma = 1/||v||_WRMS with
 * weights 1/(rtol|y_i| + atol + rtol rms(y)) -- CVODE's weights plus a floor at rtol times the
 * state's RMS, which keeps exactly-zero components from weighting rounding noise by 1/atol.
 * rtol/atol from fos_euler_set_tolerances (default 1e-6/1e-12, foeuler.f90:14-15).  v = 0 gives
 * Jv = 0.  tmp may be NULL.  One fused pass over the mesh after two small reductions; the weight
 * floor is reused while y is the same array and no fos_euler_rhs call intervened (fy = f(y), as
 * CVODE passes it).  With FOS_PREC_FP32: Jv = (f32(y + sigma v) - f32(y)) / sigma differenced per
 * face, sigma = sqrt(eps32) / (rtol max_i |v_i w_i|). */
int fos_euler_jtimes(const double *v, double *Jv, double t, const double *y, const double *fy,
                     void *user_data, double *tmp);
/* CVLsJacTimesVecFn, the difference quotient of fos_euler_jtimes (bitwise the same result) for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight serving loops: no reference
 * counterpart.  Its reduction partials and sigma live in a file-scope device array of 2*1024 + 8
 * doubles (16.4 KB) per GPU, independent of the mesh size, resolved when the first context is created
 * on the device and held until the process ends; the call allocates nothing.  Device vectors only
 * (-1 with FOS_VECTORS_ON_HOST).  Every context on one GPU shares that array: calls that overlap on
 * one GPU (other streams or threads) race. */
int fos_euler_jtimes_noalloc(const double *v, double *Jv, double t, const double *y, const double *fy,
                             void *user_data, double *tmp);
/* CVLsJacTimesVecFn, analytic: Jv = R'(y) v exactly -- the directional derivative of the residual
 * (setBC ghost states included) by dual-number arithmetic through the same primitive, ghost-state
 * and Roe-flux code, one pass over the mesh, no sigma, no fy (fy and tmp unused).  A kink of
 * |.| in the Roe dissipation takes the mean of its one-sided derivatives.  FP32 arithmetic with
 * FOS_PREC_FP32 (no cancellation, unlike a difference quotient).  No reference counterpart: the
 * reference runs CVODE's difference quotient (fos_euler_jtimes). */
int fos_euler_jtimes_analytic(const double *v, double *Jv, double t, const double *y, const double *fy,
                              void *user_data, double *tmp);
/* The Newton matrix product out = v - gamma J(y) v in one pass over the mesh: the Jv kernels store
 * v - gamma Jv instead of Jv (difference quotient as fos_euler_jtimes, or analytic as
 * fos_euler_jtimes_analytic).  No reference counterpart: CVODE's linear solver forms this product
 * from jtimes in a second vector pass; these fuse the two for the library's integrator
 * (fos_bdf_set_newton_matvec).  Device vectors only; out must not alias v or y. */
typedef int (*fos_newton_matvec_fn)(double gamma, const double *v, double *out, double t, const double *y,
                                    const double *fy, void *user_data);
int fos_euler_newton_matvec(double gamma, const double *v, double *out, double t, const double *y, const double *fy,
                            void *user_data);
int fos_euler_newton_matvec_analytic(double gamma, const double *v, double *out, double t, const double *y,
                                     const double *fy, void *user_data);
/* CVLsPrecSetupFn -- completes foeuler pset (foeuler.f90:378-405): assemble findEulerJac
 * (Euler.f90:159-249), form M = I - gamma*J/v in 5x5 blocks and factor it by block ILU(0)
 * (level scheduled) in the elimination order of fos_euler_ilu_order.  jok != 0 reuses the last
 * Jacobian (*jcur = 0). */
int fos_euler_psetup(double t, const double *y, const double *fy, int jok, int *jcur,
                     double gamma, void *user_data);
/* CVLsPrecSolveFn -- completes foeuler psol (foeuler.f90:408-437): z = (LU)^-1 r */
int fos_euler_psolve(double t, const double *y, const double *fy, const double *r, double *z,
                     double gamma, double delta, int lr, void *user_data);
int fos_euler_set_tolerances(fos_euler *ctx, double rtol, double atol);
/* Non-finite results.  The kernels behind fos_euler_rhs, the jtimes and psolve callbacks and the
 * FP32 variants flag a NaN or Inf among their outputs; the call then returns 1, SUNDIALS' code for a
 * recoverable failure, on which the integrator retries with a smaller step (fopbc's rhs returns 1
 * the same way, src/fopbc/modPbc.f90:641; foeuler's always returns 0, foeuler.f90:373).  A
 * callback waits for its kernels to find out, unless the context has FOS_ASYNC_CALLBACKS: then it
 * returns 0 and fos_euler_nonfinite(ctx, clear) waits for the stream and returns 1 if any call since
 * the last clear produced a non-finite value (clear != 0 resets the record), 0 if none, -1 on error. */
int fos_euler_nonfinite(fos_euler *ctx, int clear);
/* The block ILU's row elimination order: order[s] = cell eliminated at step s (host array of nC
 * ints).  SPARSKIT's ilu0 (ilut.f:1793-1917) eliminates in the natural order; here the 256-cell
 * chunks are coloured so that face-adjacent chunks differ and rows are eliminated colour by colour,
 * chunk by chunk, cells in index order, which lets each sweep run one launch per colour.  Same
 * ILU(0) definition, on the matrix in that order. */
int fos_euler_ilu_order(fos_euler *ctx, int *order);
/* Block Jacobi on the same M = I - gamma JacC/v: z_c = inv(M_cc) r_c.  No reference counterpart; the
 * block-diagonal baseline the learned preconditioner is compared with.  Device vectors only. */
int fos_euler_psetup_bjac(double t, const double *y, const double *fy, int jok, int *jcur, double gamma,
                          void *user_data);
int fos_euler_psolve_bjac(double t, const double *y, const double *fy, const double *r, double *z,
                          double gamma, double delta, int lr, void *user_data);

/* ----------------------------------------------------------------------------------------
 * The same callbacks with SUNDIALS' own signatures (N_Vector arguments), to be handed unchanged to
 * CVodeInit, CVodeSetJacTimes and CVodeSetPreconditioner, as foeuler registers rhs / pset / psol
 * (src/foeuler/foeuler.f90:349-353,378-383,408-412 through SUNDIALS.f90:765-794).  An N_Vector is an
 * opaque pointer here: its data pointer comes from the accessor the host registers once per context,
 * N_VGetArrayPointer for NVECTOR_SERIAL (host data: create the context with FOS_VECTORS_ON_HOST) or a
 * device-array accessor for device-resident vectors -- the reference reaches the data the same way
 * (N_VGetArrayPointer_Serial, SUNDIALS.f90:340-349).  Each forwards to the raw-pointer callback with
 * the context as user_data; return codes as there, and -1 when no accessor is registered or it
 * returns NULL.
 * ------------------------------------------------------------------------------------- */
typedef double *(*fos_nvector_data_fn)(void *nvector);
int fos_euler_set_nvector_access(fos_euler *ctx, fos_nvector_data_fn data);
/* CVRhsFn */
int fos_euler_rhs_nv(double t, void *y, void *ydot, void *user_data);
/* CVLsJacTimesVecFn: the difference quotient (fos_euler_jtimes) and the analytic product */
int fos_euler_jtimes_nv(void *v, void *Jv, double t, void *y, void *fy, void *user_data, void *tmp);
int fos_euler_jtimes_analytic_nv(void *v, void *Jv, double t, void *y, void *fy, void *user_data, void *tmp);
/* CVLsPrecSetupFn (jok and *jcur are SUNDIALS' booleantype, an int) */
int fos_euler_psetup_nv(double t, void *y, void *fy, int jok, int *jcur, double gamma, void *user_data);
/* CVLsPrecSolveFn */
int fos_euler_psolve_nv(double t, void *y, void *fy, void *r, void *z, double gamma, double delta, int lr,
                        void *user_data);

/* findEuler on element arrays (src/libfosolvers/FVM/Euler.f90:25-35): rho[nE], rhou[3 nE]
 * (Fortran rhou(3,nE)), rhoE[nE], p[nE] in; dRho[nC], dRhou[3 nC], dRhoE[nC] out (net flux,
 * not divided by v).  Device pointers. */
int fos_find_euler(fos_euler *ctx, const double *rho, const double *rhou, const double *rhoE,
                   const double *p, double gamma, double *dRho, double *dRhou, double *dRhoE);
/* setBC + syncState (foeuler.f90:186-330) into element arrays of nE: cells from the state y, boundary
 * facets from their pair's ghost state; rho[nE], rhou[3 nE] (Fortran rhou(3,nE)), rhoE[nE],
 * u[3 nE], p[nE], temp[nE], c[nE], any of them NULL to skip.  Facets with no boundary pair keep
 * what the arrays held, as the reference's do.  Device pointers, or host with FOS_VECTORS_ON_HOST
 * (what writeState needs, foeuler.f90:333-346). */
int fos_euler_sync_state(fos_euler *ctx, const double *y, double *rho, double *rhou, double *rhoE, double *u,
                         double *p, double *temp, double *c);
/* writeState (foeuler.f90:333-346): the legacy VTK file writeVTK writes (VTK.f90:83-204) -- the grid's
 * nodes and elements, then density, velocity, pressure and temperature per element -- for state y
 * (device, or host with FOS_VECTORS_ON_HOST) on the context's grid g.  The fields are formed on the
 * GPU (setBC + syncState, facets without a boundary pair 0); numbers are written in their shortest
 * round-trip form. */
int fos_euler_write_vtk(fos_euler *ctx, const fos_fv_grid *g, const double *y, const char *path);
/* findEulerJac's cell blocks JacC(5, 5*(1+maxNF), nC) column-major (Euler.f90:177), device. */
int fos_find_euler_jac(fos_euler *ctx, const double *y, double *JacC);

/* FP32 variants of the hot calls (state, geometry and arithmetic in FP32). */
int fos_euler_rhs_f32(fos_euler *ctx, const float *y, float *ydot);
int fos_euler_jtimes_f32(fos_euler *ctx, const float *v, float *Jv, const float *y,
                         const float *fy, double sigma);

/* Stream the context launches on (hipStream_t); NULL = default stream. */
int fos_euler_set_stream(fos_euler *ctx, void *stream);
/* Counters since creation: [0] rhs calls, [1] Jv calls, [2] psetup calls, [3] psolve calls. */
int fos_euler_counters(const fos_euler *ctx, long long counts[4]);
/* Residual layout in use: 1 = tetrahedral (the default on meshes that qualify: 8-byte face records,
 * area and normal recomputed from node coordinates), 0 = generic (area and normal read per face;
 * any polyhedral mesh, or FOS_LAYOUT_GENERIC).  No reference counterpart. */
int fos_euler_layout(const fos_euler *ctx);

/* ----------------------------------------------------------------------------------------
 * Integrator: BDFNewtonKrylov (src/libfosolvers/numerics/SUNDIALS.f90:65-74, 745-845)
 * Two modes.  Adaptive (the default, what CVodeInit + CV_ONE_STEP give foeuler): CVODE's
 * variable-order (1..5) variable-step BDF in Nordsieck form with its local error test, step and
 * order selection, initial-step estimate and Newton convergence-rate test (at most 3
 * corrections), weights 1/(rtol|y| + atol), GMRES to 0.05 of the Newton tolerance.  Fixed step
 * (after fos_bdf_set_dt): BDF order 1 on the first step then order 2; Newton on G(y) = y - sum
 * a_k y_{n-k} - gamma f(y) formed in FP64 with WRMS weights 1/(rtol|y_n| + atol + rtol rms(y_n))
 * until ||G||_WRMS <= the Newton tolerance (default 1e-8).  In both, (I - gamma J) d = b by
 * left-preconditioned restarted GMRES with FP64 CGS2 orthogonalisation.  The callbacks keep the
 * CVODE signatures (CVRhsFn, CVLsJacTimesVecFn, CVLsPrecSetupFn, CVLsPrecSolveFn) on device
 * vectors; fos_euler_rhs / fos_euler_jtimes / fos_euler_psetup / fos_euler_psolve with the
 * context as user_data are the foeuler set.
 * ------------------------------------------------------------------------------------- */
typedef int (*fos_rhs_fn)(double t, const double *y, double *ydot, void *user_data);
typedef int (*fos_jtimes_fn)(const double *v, double *Jv, double t, const double *y, const double *fy,
                             void *user_data, double *tmp);
typedef int (*fos_psetup_fn)(double t, const double *y, const double *fy, int jok, int *jcur, double gamma,
                             void *user_data);
typedef int (*fos_psolve_fn)(double t, const double *y, const double *fy, const double *r, double *z,
                             double gamma, double delta, int lr, void *user_data);
typedef struct fos_bdf fos_bdf;
typedef struct {
  long long nst, nfe, nni, nli, npe, nps, njv, nncfails, nlinfails, nfallback;
  int lastNewtonIters, lastLinIters;
  double lastResidualNorm;  /* fixed: ||G||_WRMS; adaptive: the last Newton convergence-test ratio */
  long long nreorth;  /* Gram-Schmidt re-orthogonalisation passes (DGKS test) */
  long long netf;     /* local error test failures (adaptive mode) */
  int lastOrder, nextOrder;
  double lastStep, nextStep;
  double lastResidualNormCvode;  /* fixed: the converged ||G||_WRMS under CVODE's weights 1/(rtol|y_n| + atol)
                                    (no floor); adaptive: lastResidualNorm */
} fos_bdf_stats;

/* init(nEq, f, maxl, pSet, pSol) (SUNDIALS.f90:745-795); maxl <= 0 selects 5 like SUNDIALS. */
fos_bdf *fos_bdf_create(int nEq, fos_rhs_fn f, int maxl, fos_psetup_fn pSet, fos_psolve_fn pSol, void *user_data);
void fos_bdf_destroy(fos_bdf *b);
/* jtimes callback (CVODE's internal difference quotient when not set) */
int fos_bdf_set_jtimes(fos_bdf *b, fos_jtimes_fn jtimes);
/* Optional: the Newton matrix product v - gamma J v in one call (fos_euler_newton_matvec[_analytic]),
 * used for GMRES products with device vectors in place of jtimes and a separate combine; it must
 * form the same J v as the jtimes set. */
int fos_bdf_set_newton_matvec(fos_bdf *b, fos_newton_matvec_fn fn);
/* FOS_VECTORS_ON_HOST: f receives host vectors (no jtimes/preconditioner); FOS_PREC_FP32: the GMRES
 * basis is stored in FP32 (products, sums, the Hessenberg matrix, the correction and the Newton
 * residual stay FP64), for use with an FP32 context; stream: hipStream_t.  Returns -1 if the FP32
 * basis cannot be allocated. */
int fos_bdf_set_options(fos_bdf *b, unsigned flags, void *stream);
int fos_bdf_set_tol(fos_bdf *b, double rtol, double atol);                 /* setTol, :691-705 */
int fos_bdf_set_max_steps(fos_bdf *b, int maxSteps);                       /* setMaxSteps, :708-720 */
/* Newton tolerance on ||G||_WRMS, Newton iteration cap, relative GMRES tolerance, GMRES restarts,
 * Jacobian evaluation interval in steps (CVodeSetJacEvalFrequency, default 51; <= 0 keeps the
 * current value) */
int fos_bdf_set_newton(fos_bdf *b, double tol, int maxIter, double linTol, int maxRestarts, int jacInterval);
/* Preconditioner setup control, CVODE's (cvode.c lsetup test; CVodeSetLSetupFrequency and
 * CVodeSetJacEvalFrequency for the first two): psetup is called when msbp steps (default 20) have
 * passed since the last call, when gamma moved by more than dgmax (default 0.3) relative to the
 * gamma of the last call, and on the first step; psetup is asked for a new Jacobian (jok = 0) when
 * msbj steps (default 51) have passed since the last one.  A step whose Newton iteration fails with
 * a preconditioner from an earlier step is retried once after a fresh setup (jok = 0).  Values
 * <= 0 keep the current ones. */
int fos_bdf_set_lsetup_frequency(fos_bdf *b, int msbp, int msbj, double dgmax);
/* Fallback preconditioner: when GMRES with the primary preconditioner has not reached the linear
 * tolerance after gmresCap iterations, pSet is called at the current iterate and the linear solve
 * is repeated with pSol; the fallback stays in use until the step ends.  No reference counterpart
 * (used to put block ILU behind the learned preconditioner). */
int fos_bdf_set_fallback(fos_bdf *b, int gmresCap, fos_psetup_fn pSet, fos_psolve_fn pSol);
/* Fixed steps of size dt from now on (no reference counterpart: the benchmark configurations
 * prescribe the step); called during an adaptive run, fixed steps continue from its current state,
 * restarting with BDF1, and a changed dt restarts with BDF1 too (how a host retries a failed step
 * with a smaller one, as fopbc does, src/fopbc/fopbc.f90:14-45). */
int fos_bdf_set_dt(fos_bdf *b, double dt);
/* Adaptive steps (the default): maximum order (CVodeSetMaxOrd, <= 5; <= 0 keeps the current),
 * initial step (CVodeSetInitStep; 0 = estimated from tOut of the first step), minimum and maximum
 * step (CVodeSetMinStep / CVodeSetMaxStep; 0 = none).  Called during fixed steps, the adaptive
 * history starts afresh from the current state at the next step; during an adaptive run the order
 * limit and the initial step take effect at the next fos_bdf_set_iv, the step bounds at once. */
int fos_bdf_set_adaptive(fos_bdf *b, int maxOrder, double h0, double hmin, double hmax);
int fos_bdf_set_iv(fos_bdf *b, double t, const double *y0);               /* setIV, :672-688 */
/* step(tOut, t, y) (:823-845): one internal step (CV_ONE_STEP; it may pass tOut, which only
 * bounds the first adaptive step); y receives the new state.  Returns 0; >0 when the step failed
 * (fixed: Newton did not converge; adaptive: 10 Newton or 7 error-test failures in one step),
 * with the state unchanged; <0 on a fatal callback error. */
int fos_bdf_step(fos_bdf *b, double tOut, double *t, double *y);
/* getDt (:723-735): the last step taken (CVodeGetLastStep) */
int fos_bdf_get_dt(fos_bdf *b, double *dt);
/* CVodeGetDky for the adaptive mode (no call in the reference; what a host needs to output at an
 * exact time after CV_ONE_STEP passed it): the k-th derivative (0 <= k <= current order) of the
 * interpolating polynomial of the Nordsieck history at t in the last step [t_n - h_u, t_n], into
 * dky (device, or host with FOS_VECTORS_ON_HOST). */
int fos_bdf_get_dky(fos_bdf *b, double t, int k, double *dky);
/* The last accepted step's training triple (device vectors, any may be NULL): the predictor,
 * the Newton residual -G at the predictor, and the converged correction y_new - predictor. */
int fos_bdf_last_triple(fos_bdf *b, double *ypred, double *r, double *delta);
int fos_bdf_get_stats(fos_bdf *b, fos_bdf_stats *out);

/* ----------------------------------------------------------------------------------------
 * Learned preconditioner (new; no reference counterpart -- foeuler's psol is an identity,
 * foeuler.f90:434).  Per cell: 20 features (conservative state / S, log volume, 4 area normals,
 * gamma c / v^(1/3), Mach) -> 3 x 128 SiLU layers with two rounds of one-hop neighbour-mean
 * message passing -> 30 outputs scaled by k = min(1, CFL), the cell's CFL number
 * gamma (c + |u|) / v^(1/3) from features 18 and 19, so the operator tends to the identity as the
 * time step shrinks: 25
 * entries of B^ = I + o and 5 neighbour weights s (the 4 face neighbours in neib slot order, then
 * the cell itself):
 *   t_c = B^_c (r_c/S),  z_c = S ((1 + s_4) t_c + sum_j s_j t_nb(j))   (S: per-variable state RMS)
 * BF16 weights/activations on MFMA with FP32 accumulation; FP32 master weights and Adam.
 * Parameters: fos_learned_nparams() floats, packed W0[20][128] b0 W1[256][128] b1 W2[256][128]
 * b2 W3[128][30] b3.
 * ------------------------------------------------------------------------------------- */
int fos_learned_nparams(void);
/* Glorot-uniform hidden layers, zero output layer (the untrained network is the identity). */
int fos_learned_init(fos_euler *ctx, uint64_t seed);
int fos_learned_set_params(fos_euler *ctx, const float *params);   /* host array */
int fos_learned_get_params(fos_euler *ctx, float *params);         /* host array */
/* Loss |P r - delta|^2 (scaled, mean over the batch) on the cells [batchStart, +batchSize) of a
 * snapshot (device y, r, delta), message passing restricted to the batch; gradient to host. */
int fos_learned_loss_grad(fos_euler *ctx, const double *y, const double *r, const double *delta, double gamma,
                          int batchStart, int batchSize, double *loss, float *grad);
/* The network inputs for state y (device) and gamma: X host [nC][20] (the BF16 values the
 * kernels use), S host [5] (per-variable state scale).  Either may be NULL. */
int fos_learned_inputs(fos_euler *ctx, const double *y, double gamma, float *X, float *S);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on one batch. */
int fos_learned_train_step(fos_euler *ctx, const double *y, const double *r, const double *delta, double gamma,
                           int batchStart, int batchSize, double lr, double *loss);
/* CVLsPrecSetupFn / CVLsPrecSolveFn with the learned operator (psetup runs the network). */
int fos_euler_psetup_learned(double t, const double *y, const double *fy, int jok, int *jcur, double gamma,
                             void *user_data);
int fos_euler_psolve_learned(double t, const double *y, const double *fy, const double *r, double *z,
                             double gamma, double delta, int lr, void *user_data);

/* ----------------------------------------------------------------------------------------
 * ILUT on a general sparse matrix: the reference's ILU type (src/libfosolvers/numerics/sparse.f90:52-63)
 * over SPARSKIT's ilut and lusol (contrib/SPARSKIT2/ITSOL/ilut.f:38-381, 2272-2321), the
 * preconditioner fopbc builds with lfil 50 and drop tolerance 1e-6 (src/fopbc/modPbc.f90:21,647-739).
 * Indices are 0-based.  The factorisation runs on the host (ILUT chooses its fill from the values it
 * produces); the triangular solves and the product run on the GPU, level-scheduled.  Vectors are
 * device pointers unless created with FOS_VECTORS_ON_HOST.  No device memory is touched before the
 * first solve or product.
 * ------------------------------------------------------------------------------------- */
typedef struct fos_ilut fos_ilut;
/* ILU%init(nEq, maxNNz, lFill) (sparse.f90:334-347): the fill per row of L and U is
 * min(lFill, nEq - 1).  stream: a hipStream_t or NULL.  NULL on bad sizes. */
fos_ilut *fos_ilut_create(int nEq, int maxNNz, int lFill, unsigned flags, void *stream);
void fos_ilut_destroy(fos_ilut *p);
/* linEq%setCOO (sparse.f90:249-266, SPARSKIT coocsr): nNz triplets (row, column, value) into CSR,
 * the entries of a row in input order, duplicates kept as separate entries.  -1 for an index
 * outside the matrix or nNz > maxNNz. */
int fos_ilut_set_coo(fos_ilut *p, const int *iA, const int *jA, const double *A, int nNz);
/* linEq%setCSR (sparse.f90:231-247): ia[nEq + 1] (ia[0] = 0), ja, a. */
int fos_ilut_set_csr(fos_ilut *p, const int *ia, const int *ja, const double *a);
/* linEq%mulVec (sparse.f90:269-275, SPARSKIT amux): y = A x. */
int fos_ilut_mul_vec(fos_ilut *p, const double *x, double *y);
/* ILU%fact(eps) (sparse.f90:370-383): ILUT with the fill above and drop tolerance droptol (the
 * reference's default is 0): multipliers of magnitude <= droptol are dropped, U entries of magnitude
 * <= droptol times the row's mean magnitude, then the lfil largest of each part are kept (at most
 * lfil - 1 off-diagonal U entries); a zero pivot is replaced by (1e-4 + droptol) times that mean.
 * The factor is SPARSKIT ilut's, entry for entry.  Unlike the reference's wrapper, storage is not
 * bounded by nEq min(2 lFill + 1, nEq) entries.  -1 for a zero row (ilut's ierr = -5). */
int fos_ilut_fact(fos_ilut *p, double droptol);
/* ILU%solve(rhs, x) (sparse.f90:386-392, SPARSKIT lusol): x = (LU)^-1 rhs. */
int fos_ilut_solve(fos_ilut *p, const double *rhs, double *x);
/* The factor as SPARSKIT's MSR arrays alu, jlu [nEq + 1 + nnz] and ju [nEq], with 1-based values,
 * Fortran position k at C index k - 1 (what ilut hands lusol).  Returns the array length; with all
 * three arrays NULL it only returns the length. */
int fos_ilut_get_msr(const fos_ilut *p, double *alu, int *jlu, int *ju);
/* Number of dependent levels of the forward and backward sweeps (launches per solve). */
int fos_ilut_levels(const fos_ilut *p, int *nForward, int *nBackward);

const char *fos_last_error(void);
const char *fos_version(void);

#ifdef __cplusplus
}
#endif
#endif
