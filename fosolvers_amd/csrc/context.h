// The fos_euler context (opaque in the public header) and internal entry points shared by
// the device sources.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/fosolvers_amd.h"
#include "euler_kernels.h"
#include "flux.h"
#include "hip_util.h"

namespace fos {
struct BlockIlu;  // precond.hip
void destroy_block_ilu(BlockIlu *);
struct Learned;   // learned.hip
void destroy_learned(Learned *);
}  // namespace fos

struct fos_euler {
  int nC = 0, nE = 0, nP = 0, nPi = 0, nPb = 0, maxNF = 0, nChunks = 0;
  size_t nChunkFaces = 0;
  unsigned flags = 0;
  int device = 0;  // the HIP device current at creation
  hipStream_t stream = nullptr;
  double r = 1.0, gamm = 1.4, rtol = 1e-6, atol = 1e-12;
  long long counters[4] = {0, 0, 0, 0};

  // boundary pairs in pair order: BC code and 5 parameters
  std::vector<int> bndType;
  std::vector<double> bndPar;
  std::vector<char> bndIsUdf;
  bool needUdf = false, haveUdf = false;

  // host copy of the polyFvGrid arrays the preconditioner and findEuler paths need
  std::vector<int> hIEP, hNeib;
  std::vector<double> hAP, hNormP, hV;

  // residual layout (chunked face lists) in both precisions
  fos::FluxLayout<double> L64;
  fos::FluxLayout<float> L32;
  fos::DBuf<int> dChunkBeg;
  fos::DBuf<int2> dLR;
  fos::DBuf<fos::AreaNormal<double>> dGeo64;
  fos::DBuf<fos::AreaNormal<float>> dGeo32;
  fos::DBuf<short> dSlots;
  fos::DBuf<double> dInvV64;
  fos::DBuf<float> dInvV32;
  fos::DBuf<int> dBndType;
  fos::DBuf<double> dBndPar64;
  fos::DBuf<float> dBndPar32;

  // tetrahedral layout (node-based face geometry); used when the mesh allows it
  bool useTet = false;
  fos::TetLayout<double> T64;
  fos::TetLayout<float> T32;
  fos::DBuf<uint2> dTetFace;
  fos::DBuf<int> dTetNodeBeg;
  fos::DBuf<double> dTetNode64;
  fos::DBuf<float> dTetNode32;
  fos::DBuf<int> dBndCell, dBndSlotPair;
  fos::DBuf<double> dBndNorm64, dGhost64;
  fos::DBuf<float> dBndNorm32, dGhost32;

  // reductions: dSigma = [1/||v||_WRMS, weight floor, max-norm increment]; the floor is reused
  // for the same y between rhs calls
  fos::DBuf<double> redPart, dSigma;
  const double *floorY = nullptr;
  long long rhsEpoch = 0, floorEpoch = -1;

  // host-vector mode staging
  fos::DBuf<double> hY, hVec, hFy, hOut, hR, hZ;

  // findEuler on element arrays
  fos::DBuf<int> dPairIEP, dCellPairPtr, dCellPairIdx;
  fos::DBuf<double> dPairA, dPairN, dPairF;

  // preconditioner state (precond.hip)
  fos::BlockIlu *ilu = nullptr;
  fos::DBuf<int> dStatus;
  fos::Learned *learned = nullptr;

  // non-finite results: the kernels that write a callback's output set this device word; a
  // callback reads it back (pinned word) and clears it before returning, unless
  // FOS_ASYNC_CALLBACKS, where it accumulates until fos_euler_nonfinite
  fos::DBuf<unsigned> dNonFinite;
  unsigned *hNonFinite = nullptr;

  // the host's N_Vector data accessor for the *_nv callbacks (N_VGetArrayPointer or its device form)
  fos_nvector_data_fn nvData = nullptr;

  ~fos_euler() {
    fos::destroy_block_ilu(ilu);
    fos::destroy_learned(learned);
    if (hNonFinite) (void)hipHostFree(hNonFinite);
  }
};

namespace fos {
int build_layout(fos_euler *ctx, const fos_fv_grid *g, const fos_cond_tab *bc);
int save_grid_host(fos_euler *ctx, const fos_fv_grid *g);
int wrms_sigma(fos_euler *ctx, const double *v, const double *y, double *sigmaDev);
int residual_dev(fos_euler *ctx, const double *y, double *ydot);
int jv_dev(fos_euler *ctx, const double *v, double *Jv, const double *y, const double *fy, int newton = 0,
           double ngam = 0.0);
int ensure_vec(DBuf<double> &b, size_t n);
int stage_in(fos_euler *ctx, DBuf<double> &b, const double *h);
int stage_out(fos_euler *ctx, const DBuf<double> &b, double *h);
int jacobian_dev(fos_euler *ctx, const double *y, double *jac);
int psetup_dev(fos_euler *ctx, const double *y, int jok, int *jcur, double gamma);
int ilu_solve(fos_euler *ctx, const double *r, double *z);
int ensure_pair_arrays(fos_euler *ctx);
// End of a callback: unless FOS_ASYNC_CALLBACKS, wait for the stream and return 1 (recoverable, the
// SUNDIALS convention) when a kernel flagged a non-finite output since the last check, 0 otherwise
// (the word is cleared); -1 on a HIP error.
int finish_call(fos_euler *ctx, const char *who);
// Resolves, once per device, the address of the file-scope scratch array behind
// fos_euler_jtimes_noalloc (with that device current).  Called by fos_euler_create.
int resolve_noalloc_scratch(int device);
}  // namespace fos
