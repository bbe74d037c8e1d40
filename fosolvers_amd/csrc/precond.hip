// Preconditioner for the BDF Newton systems: findEulerJac block assembly and block ILU(0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "context.h"

namespace fos {

struct BlockIlu {};

void destroy_block_ilu(BlockIlu *b) { delete b; }

int save_grid_host(fos_euler *ctx, const fos_fv_grid *g) {
  ctx->hIEP.assign(g->iEP, g->iEP + 2 * (size_t)g->nP);
  ctx->hAP.assign(g->aP, g->aP + g->nP);
  ctx->hNormP.assign(g->normP, g->normP + 3 * (size_t)g->nP);
  ctx->hV.assign(g->v, g->v + g->nC);
  ctx->hNeib.assign(g->neib, g->neib + (size_t)g->maxNF * g->nC);
  return 0;
}

}  // namespace fos

using namespace fos;

extern "C" {

int fos_euler_psetup(double t, const double *y, const double *fy, int jok, int *jcur, double gamma, void *user_data) {
  (void)t; (void)y; (void)fy; (void)jok; (void)jcur; (void)gamma; (void)user_data;
  set_error("fos_euler_psetup: not available in this build");
  return -1;
}

int fos_euler_psolve(double t, const double *y, const double *fy, const double *r, double *z, double gamma,
                     double delta, int lr, void *user_data) {
  (void)t; (void)y; (void)fy; (void)r; (void)z; (void)gamma; (void)delta; (void)lr; (void)user_data;
  set_error("fos_euler_psolve: not available in this build");
  return -1;
}

int fos_find_euler_jac(fos_euler *ctx, const double *y, double *JacC) {
  (void)ctx; (void)y; (void)JacC;
  set_error("fos_find_euler_jac: not available in this build");
  return -1;
}

}  // extern "C"
