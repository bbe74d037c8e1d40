// Preconditioner for the BDF Newton systems (completes foeuler's pset/psol,
// src/foeuler/foeuler.f90:378-437):
//   psetup: JacC = findEulerJac(y) (src/libfosolvers/FVM/Euler.f90:159-249), M = I - gamma JacC/v
//           in 5x5 blocks on the cell-adjacency pattern, block ILU(0) factor by level schedule;
//   psolve: z = U^-1 L^-1 r by level-scheduled forward and backward sweeps.
// Block ILU(0) on dense 5x5 blocks keeps the same updates as SPARSKIT's scalar ilu0
// (contrib/SPARSKIT2/ITSOL/ilut.f:1793-1917) on the expanded pattern; lusol's convention of
// storing the inverted diagonal (:2272-2321) is kept per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <utility>
#include <vector>

#include "context.h"

namespace fos {

constexpr int NB = 4;  // neighbour slots per cell (tetrahedra)

// A fixed chain of dependent kernel launches recorded once as a graph of kernel nodes (built
// explicitly, no stream capture, so it works on any stream including the default one) and replayed
// with one hipGraphLaunch.  The level-scheduled sweeps and the factorisation are hundreds of short
// dependent kernels whose launch gaps would otherwise exceed their run time.
struct KernelChain {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  hipGraphNode_t last = nullptr;
  KernelChain() = default;
  KernelChain(const KernelChain &) = delete;
  KernelChain &operator=(const KernelChain &) = delete;
  ~KernelChain() { reset(); }
  void reset() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
    e = nullptr;
    g = nullptr;
    last = nullptr;
  }
  bool ready() const { return e != nullptr; }
  // one kernel node depending on the previous one; the argument values are copied into the node
  template <class... P, class... A>
  int add(void (*k)(P...), dim3 grid, dim3 block, A... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<P...> vals(static_cast<P>(a)...);
    return add_node(reinterpret_cast<void *>(k), grid, block, vals, std::index_sequence_for<P...>{});
  }
  template <class T, size_t... I>
  int add_node(void *f, dim3 grid, dim3 block, T &vals, std::index_sequence<I...>) {
    void *args[] = {reinterpret_cast<void *>(&std::get<I>(vals))...};
    hipKernelNodeParams p{};
    p.func = f;
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    if (!g && (err = hipGraphCreate(&g, 0)) != hipSuccess) return -1;
    hipGraphNode_t n = nullptr;
    if ((err = hipGraphAddKernelNode(&n, g, last ? &last : nullptr, last ? 1 : 0, &p)) != hipSuccess) return -1;
    last = n;
    return 0;
  }
  int instantiate() {
    if (!g) return -1;
    err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err != hipSuccess) e = nullptr;
    return err == hipSuccess ? 0 : -1;
  }
  int launch(hipStream_t s) { return (err = hipGraphLaunch(e, s)) == hipSuccess ? 0 : -1; }
  // the error of the failed call; the chain is dropped so that the next call records it afresh
  const char *fail() {
    const char *why = hipGetErrorString(err);
    reset();
    return why;
  }
  hipError_t err = hipSuccess;
};

struct BlockIlu {
  int nC = 0, nLev = 0, nULev = 0;
  bool haveJac = false, bjac = false;  // factors hold block ILU(0), or block Jacobi inverses
  std::vector<int> levPtr, uLevPtr;
  DBuf<int> dNbr;       // [NB nC] neighbour cell per slot (neib order), -1 = none / facet
  DBuf<int> dLowOrd;    // [NB nC] slots of lower neighbours in ascending neighbour order, -1 padded
  DBuf<int> dLevRows;   // rows grouped by forward level (forward position p -> row)
  DBuf<int> dULevRows;  // rows grouped by backward level (backward position p -> row)
  DBuf<int> dJSlot;     // [nSlots nC] neib slot j of residual-layout slot s (-1 boundary / none)
  DBuf<double> dJac;    // [125 nC] JacC(5, 25, nC) column-major
  DBuf<double> dD, dOff;  // factor: inverted U diagonal blocks [25 nC], off-diagonal blocks [25 NB nC]
  DBuf<float> dD32;       // FP32 copy of dD (block Jacobi apply)
  // Level-ordered copies of the factors for the sweeps (built after each factorisation in the
  // context's apply precision): the forward sweep reads position p's lower neighbours as forward
  // positions (slot order, -1 padded) and its L blocks packed from lPtr[p]; the backward sweep
  // likewise its upper neighbours as backward positions, its U blocks from uPtr[p] and the inverted
  // diagonal block at p, and yPos[p] = the forward position of the same row.
  DBuf<int4> dLNb, dUNb;
  DBuf<int> dLPtr, dUPtr, dYPos, dFPos, dBPos;  // dFPos/dBPos: row -> forward/backward position
  DBuf<int> dLSrc, dUSrc;                       // packed block -> source block NB*row + slot in dOff
  long long nLBlk = 0, nUBlk = 0;
  DBuf<double> dLf, dUf, dDb;
  DBuf<float> dLf32, dUf32, dDb32;
  DBuf<double> dY, dW;  // forward sweep result (forward positions), backward result (backward positions), [nC][5]
  // level sequences replayed as graphs (FOS_ILU_EAGER=1 in the environment launches them one by one)
  bool eager = false;
  KernelChain factorChain, solveChain;
};

void destroy_block_ilu(BlockIlu *b) { delete b; }

int save_grid_host(fos_euler *ctx, const fos_fv_grid *g) {
  ctx->hIEP.assign(g->iEP, g->iEP + 2 * (size_t)g->nP);
  ctx->hAP.assign(g->aP, g->aP + g->nP);
  ctx->hNormP.assign(g->normP, g->normP + 3 * (size_t)g->nP);
  ctx->hV.assign(g->v, g->v + g->nC);
  ctx->hNeib.assign(g->neib, g->neib + (size_t)g->maxNF * g->nC);
  return 0;
}

namespace {

// ---------------------------------------------------------------- flux Jacobians (Euler.f90:252-333)

// A(n, u, H): exact normal-flux Jacobian, J[col*5 + row] (column-major like the Fortran)
__device__ inline void nfj(double nx, double ny, double nz, double ux, double uy, double uz, double H, double g,
                           double *J) {
  const double n[3] = {nx, ny, nz}, u[3] = {ux, uy, uz};
  const double v2 = ux * ux + uy * uy + uz * uz, vn = ux * nx + uy * ny + uz * nz, g1 = g - 1.0;
  J[0] = 0.0;
  for (int d = 0; d < 3; d++) J[1 + d] = 0.5 * g1 * v2 * n[d] - u[d] * vn;
  J[4] = (0.5 * g1 * v2 - H) * vn;
  for (int c = 0; c < 3; c++) {
    J[5 * (1 + c)] = n[c];
    for (int r = 0; r < 3; r++) J[5 * (1 + c) + 1 + r] = u[r] * n[c] + (1.0 - g) * n[r] * u[c] + (r == c ? vn : 0.0);
    J[5 * (1 + c) + 4] = H * n[c] - g1 * u[c] * vn;
  }
  J[20] = 0.0;
  for (int d = 0; d < 3; d++) J[21 + d] = g1 * n[d];
  J[24] = g * vn;
}

// |A~| at the Roe state, the reference's eigen-decomposition sum (Euler.f90:279-333)
__device__ inline void anfj(double nx, double ny, double nz, double ux, double uy, double uz, double H, double c,
                            double g, double *J) {
  const double n[3] = {nx, ny, nz}, u[3] = {ux, uy, uz};
  const double v2 = ux * ux + uy * uy + uz * uz, vn = ux * nx + uy * ny + uz * nz;
  const double g1 = g - 1.0, c2 = c * c, M2 = v2 / c2, Mn = vn / c;
  const double av = fabs(vn);
  // entropy + shear part, weight |vn|
  J[0] = av * (1.0 - 0.5 * g1 * M2);
  for (int d = 0; d < 3; d++) J[1 + d] = av * (-(1.0 + 0.5 * g1 * M2) * u[d] + vn * n[d]);
  J[4] = av * (vn * vn - 0.5 * v2 * (1.0 + 0.5 * g1 * M2));
  for (int cc = 0; cc < 3; cc++) {
    J[5 * (1 + cc)] = av * (g1 / c2 * u[cc]);
    for (int r = 0; r < 3; r++)
      J[5 * (1 + cc) + 1 + r] = av * (g1 / c2 * u[r] * u[cc] - n[r] * n[cc] + (r == cc ? 1.0 : 0.0));
    J[5 * (1 + cc) + 4] = av * ((1.0 + 0.5 * g1 * M2) * u[cc] - vn * n[cc]);
  }
  J[20] = av * (-g1 / c2);
  for (int d = 0; d < 3; d++) J[21 + d] = av * (-g1 / c2 * u[d]);
  J[24] = av * (-0.5 * g1 * M2);
  // acoustic parts, weights |vn -+ c|
  for (int s = -1; s <= 1; s += 2) {
    double a[3], b[3];
    for (int d = 0; d < 3; d++) {
      a[d] = u[d] + s * c * n[d];
      b[d] = -0.5 * g1 / c2 * u[d] + s * 0.5 * n[d] / c;
    }
    const double hh = H + s * vn * c, lam = fabs(vn + s * c);
    const double t11 = 0.25 * g1 * M2 - s * 0.5 * Mn, t15 = 0.5 * g1 / c2;
    J[0] += lam * t11;
    for (int d = 0; d < 3; d++) J[1 + d] += lam * (a[d] * t11);
    J[4] += lam * (hh * t11);
    for (int cc = 0; cc < 3; cc++) {
      J[5 * (1 + cc)] += lam * b[cc];
      for (int r = 0; r < 3; r++) J[5 * (1 + cc) + 1 + r] += lam * (a[r] * b[cc]);
      J[5 * (1 + cc) + 4] += lam * (hh * b[cc]);
    }
    J[20] += lam * t15;
    for (int d = 0; d < 3; d++) J[21 + d] += lam * (a[d] * t15);
    J[24] += lam * (hh * t15);
  }
}

__device__ inline void load_state(const double *y, int nC, int c, double g1, Prim<double> &q) {
  make_prim(y[c], y[(size_t)nC + c], y[2 * (size_t)nC + c], y[3 * (size_t)nC + c], y[4 * (size_t)nC + c], g1, q);
}

// One thread per (cell, face slot): NS consecutive lanes hold one cell.  Each lane forms its face's
// neighbour block of JacC (written directly: a neighbour shares one face with the cell) and its
// contribution to the diagonal block, which the cell's lanes sum by butterfly shuffles; the lanes
// then zero the neighbour blocks no face wrote (facets).
template <int NS>
__global__ __launch_bounds__(256) void jac_kernel(FluxLayout<double> L, const int *__restrict__ jslot,
                                                  const double *__restrict__ y, double *__restrict__ jac) {
  static_assert(NS >= NB && (NS & (NS - 1)) == 0 && 64 % NS == 0, "slot lanes");
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = t / NS, s = t % NS;
  const bool live = c < L.nC;  // every lane takes part in the shuffles
  const int nC = L.nC;
  const double g = L.gamma, g1 = g - 1.0;
  double D[25];
#pragma unroll
  for (int k = 0; k < 25; k++) D[k] = 0.0;
  int jj = -1;
  double *out = jac + (size_t)125 * c;
  const int code = live ? L.slots[(size_t)NS * c + s] : -1;
  if (code >= 0) {
    const int f = L.chunkBeg[c / CHUNK] + (code >> 1);
    const bool isR = code & 1;  // this cell is the pair's n
    const int2 lr = L.lr[f];
    const AreaNormal<double> an = L.geo[f];
    double4 gm;
    face_geom(an.x, an.y, an.z, gm.x, gm.y, gm.z, gm.w);
    Prim<double> Q, Pm, Pn;
    load_state(y, nC, c, g1, Q);
    load_state(y, nC, lr.x, g1, Pm);
    if (lr.y >= 0) {
      load_state(y, nC, lr.y, g1, Pn);
    } else {
      const int bp = -lr.y - 1;
      double gr, gx, gy, gz, gE;
      ghost_state<double>(L.bndType[bp], L.bndPar + 5 * (size_t)bp, gm.x, gm.y, gm.z, Pm, g, L.r, gr, gx, gy, gz, gE);
      make_prim(gr, gx, gy, gz, gE, g1, Pn);
    }
    // Roe averages (Euler.f90:207-209)
    const double srm = Pm.sr, srn = Pn.sr, inv = 1.0 / (srm + srn);
    const double um[3] = {Pm.rux * Pm.ir, Pm.ruy * Pm.ir, Pm.ruz * Pm.ir};
    const double un[3] = {Pn.rux * Pn.ir, Pn.ruy * Pn.ir, Pn.ruz * Pn.ir};
    const double Hm = (Pm.rE + Pm.p) * Pm.ir, Hn = (Pn.rE + Pn.p) * Pn.ir;
    double ua[3];
    for (int d = 0; d < 3; d++) ua[d] = (srm * um[d] + srn * un[d]) * inv;
    const double Ha = (srm * Hm + srn * Hn) * inv;
    const double ca = sqrt(g1 * (Ha - 0.5 * (ua[0] * ua[0] + ua[1] * ua[1] + ua[2] * ua[2])));
    double Aa[25], B[25];
    anfj(gm.x, gm.y, gm.z, ua[0], ua[1], ua[2], Ha, ca, g, Aa);
    const double h = 0.5 * gm.w;
    // neighbour block first (B and |A| live), then the diagonal part (A and |A| live)
    if (!isR) {
      if (lr.y >= 0) {
        // JacP(1:5,6:10) = -a/2 (A_n - |A|): d(dU_m)/dU_n
        nfj(gm.x, gm.y, gm.z, un[0], un[1], un[2], Hn, g, B);
        jj = jslot[(size_t)NS * c + s];
        if (jj >= 0)
          for (int k = 0; k < 25; k++) out[25 * (1 + jj) + k] = -h * (B[k] - Aa[k]);
      }
    } else {
      // JacP(6:10,1:5) = a/2 (A_m + |A|): d(dU_n)/dU_m
      nfj(gm.x, gm.y, gm.z, um[0], um[1], um[2], Hm, g, B);
      jj = jslot[(size_t)NS * c + s];
      if (jj >= 0)
        for (int k = 0; k < 25; k++) out[25 * (1 + jj) + k] = h * (B[k] + Aa[k]);
    }
    const double uc[3] = {Q.rux * Q.ir, Q.ruy * Q.ir, Q.ruz * Q.ir};
    const double Hc = (Q.rE + Q.p) * Q.ir;
    nfj(gm.x, gm.y, gm.z, uc[0], uc[1], uc[2], Hc, g, B);  // A(u_c): this cell's own state
    if (!isR) {
      // JacP(1:5,1:5) = -a/2 (A_m + |A|): d(dU_m)/dU_m
      for (int k = 0; k < 25; k++) D[k] = -h * (B[k] + Aa[k]);
    } else {
      // JacP(6:10,6:10) = a/2 (A_n - |A|): d(dU_n)/dU_n
      for (int k = 0; k < 25; k++) D[k] = h * (B[k] - Aa[k]);
    }
  }
  // diagonal block: sum over the cell's lanes; neighbour blocks written: OR of the lanes' bits
  int written = jj >= 0 ? 1 << jj : 0;
#pragma unroll
  for (int m = 1; m < NS; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 25; k++) D[k] += __shfl_xor(D[k], m, NS);
    written |= __shfl_xor(written, m, NS);
  }
  if (!live) return;
#pragma unroll
  for (int k = 0; k < 25; k++)
    if (k % NS == s) out[k] = D[k];
  if (s < NB && !((written >> s) & 1))
    for (int k = 0; k < 25; k++) out[25 * (1 + s) + k] = 0.0;
}

// ---------------------------------------------------------------- block ILU(0)

__device__ inline void mm5(const double *A, const double *B, double *C) {  // row-major C = A B
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      double s = 0.0;
      for (int k = 0; k < 5; k++) s += A[5 * i + k] * B[5 * k + j];
      C[5 * i + j] = s;
    }
}

__device__ inline bool inv5(const double *A, double *Ai) {  // Gauss-Jordan, partial pivoting, row-major
  // every index is a compile-time constant once the loops unroll (the pivot row is swapped in by
  // selects), so M stays in registers
  double M[5][10];
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 10; j++) M[i][j] = j < 5 ? A[5 * i + j] : (j - 5 == i ? 1.0 : 0.0);
#pragma unroll
  for (int c = 0; c < 5; c++) {
    int piv = c;
    double best = fabs(M[c][c]);
#pragma unroll
    for (int r = c + 1; r < 5; r++)
      if (fabs(M[r][c]) > best) { best = fabs(M[r][c]); piv = r; }
    if (best == 0.0) return false;
#pragma unroll
    for (int r = c + 1; r < 5; r++)
      if (r == piv)
#pragma unroll
        for (int j = 0; j < 10; j++) { double t = M[c][j]; M[c][j] = M[r][j]; M[r][j] = t; }
    const double d = 1.0 / M[c][c];
#pragma unroll
    for (int j = 0; j < 10; j++) M[c][j] *= d;
#pragma unroll
    for (int r = 0; r < 5; r++) {
      if (r == c) continue;
      const double f = M[r][c];
      if (f != 0.0)
#pragma unroll
        for (int j = 0; j < 10; j++) M[r][j] -= f * M[c][j];
    }
  }
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) Ai[5 * i + j] = M[i][j + 5];
  return true;
}

// M = I - gamma JacC / v, blocks row-major: D <- diagonal block (not yet inverted), Off <- neighbours
__global__ __launch_bounds__(256) void form_m_kernel(int nC, const double *__restrict__ jac,
                                                     const double *__restrict__ invV, const int *__restrict__ nbr,
                                                     double gamma, double *__restrict__ D, double *__restrict__ Off) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nC * 25 * (1 + NB)) return;
  const int c = (int)(t / (25 * (1 + NB)));
  const int e = (int)(t % (25 * (1 + NB)));
  const int blk = e / 25, k = e % 25, r = k / 5, q = k % 5;
  const double s = -gamma * invV[c];
  const double jv = jac[(size_t)125 * c + 25 * blk + 5 * q + r];  // JacC column-major: (row r, col q)
  if (blk == 0) {
    D[(size_t)25 * c + k] = (r == q ? 1.0 : 0.0) + s * jv;
  } else {
    const int j = blk - 1;
    const int nb = nbr[(size_t)NB * c + j];
    Off[((size_t)NB * c + j) * 25 + k] = nb >= 0 ? s * jv : 0.0;
  }
}

__device__ inline int sel4(int4 v, int a) { return a == 0 ? v.x : a == 1 ? v.y : a == 2 ? v.z : v.w; }

// C -= A B (row-major 5x5), B read from memory
__device__ inline void mms5(const double *A, const double *__restrict__ B, double *C) {
  double b[25];
#pragma unroll
  for (int q = 0; q < 25; q++) b[q] = B[q];
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) {
      double s = A[5 * i] * b[j];
#pragma unroll
      for (int k = 1; k < 5; k++) s += A[5 * i + k] * b[5 * k + j];
      C[5 * i + j] -= s;
    }
}

// One level of the IKJ block factorisation: one thread per block row.  The diagonal block and the
// current L_ik stay in registers (static indices only); the rare update of another off-diagonal
// block of row i (a neighbour of both i and k) goes to memory in place.
__global__ __launch_bounds__(128) void factor_level_kernel(const int *__restrict__ rows, int n,
                                                           const int *__restrict__ nbr, const int *__restrict__ lowOrd,
                                                           double *__restrict__ D, double *__restrict__ Off,
                                                           int *__restrict__ status) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = rows[t];
  const int4 nbi = reinterpret_cast<const int4 *>(nbr)[i];
  const int4 lo = reinterpret_cast<const int4 *>(lowOrd)[i];
  double Aii[25];
#pragma unroll
  for (int q = 0; q < 25; q++) Aii[q] = D[(size_t)25 * i + q];
#pragma unroll
  for (int a = 0; a < NB; a++) {
    const int js = sel4(lo, a);
    if (js < 0) break;
    const int k = sel4(nbi, js);
    const int4 nbk = reinterpret_cast<const int4 *>(nbr)[k];
    double *Lik = Off + ((size_t)NB * i + js) * 25;
    double A[25], Lk[25];
#pragma unroll
    for (int q = 0; q < 25; q++) { A[q] = Lik[q]; Lk[q] = 0.0; }
    mms5(A, D + (size_t)25 * k, Lk);  // Lk = -A_ik inv(U_kk)
#pragma unroll
    for (int q = 0; q < 25; q++) { Lk[q] = -Lk[q]; Lik[q] = Lk[q]; }
#pragma unroll
    for (int jk = 0; jk < NB; jk++) {
      const int j = sel4(nbk, jk);
      if (j <= k) continue;  // U part of row k (also skips -1)
      const double *U = Off + ((size_t)NB * k + jk) * 25;
      if (j == i) {
        mms5(Lk, U, Aii);
      } else {
        const int ji = nbi.x == j ? 0 : nbi.y == j ? 1 : nbi.z == j ? 2 : nbi.w == j ? 3 : -1;
        if (ji < 0) continue;
        double *dst = Off + ((size_t)NB * i + ji) * 25;
        double C[25];
#pragma unroll
        for (int q = 0; q < 25; q++) C[q] = dst[q];
        mms5(Lk, U, C);
#pragma unroll
        for (int q = 0; q < 25; q++) dst[q] = C[q];
      }
    }
  }
  double Di[25];
  if (!inv5(Aii, Di)) {
    atomicMin(status, i);
#pragma unroll
    for (int q = 0; q < 25; q++) Di[q] = 0.0;
  }
#pragma unroll
  for (int q = 0; q < 25; q++) D[(size_t)25 * i + q] = Di[q];
}

// Factor blocks into level order: to[m] = (TF) from[src[m]], 25 values per block.
template <class TF>
__global__ __launch_bounds__(256) void gather_blocks_kernel(long long nBlk, const int *__restrict__ src,
                                                            const double *__restrict__ from, TF *__restrict__ to) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nBlk * 25) return;
  const size_t m = t / 25;
  to[t] = (TF)from[(size_t)src[m] * 25 + (t - m * 25)];
}

// y[fpos[i]][q] = r[q][i]: the right-hand side into forward-position order, cell-major.
__global__ __launch_bounds__(256) void ilu_permute_in_kernel(int nC, const int *__restrict__ fpos,
                                                             const double *__restrict__ r, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nC) return;
  double *yp = y + (size_t)5 * fpos[i];
  for (int q = 0; q < 5; q++) yp[q] = r[(size_t)q * nC + i];
}

// z[q][i] = w[bpos[i]][q]: the result back to the SoA layout.
__global__ __launch_bounds__(256) void ilu_permute_out_kernel(int nC, const int *__restrict__ bpos,
                                                              const double *__restrict__ w, double *__restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nC) return;
  const double *wp = w + (size_t)5 * bpos[i];
  for (int q = 0; q < 5; q++) z[(size_t)q * nC + i] = wp[q];
}

template <class TF>
__device__ inline double dot5(const TF *__restrict__ a, const double *__restrict__ b) {
  double s = (double)a[0] * b[0];
  for (int p = 1; p < 5; p++) s += (double)a[p] * b[p];
  return s;
}

// Forward sweep over the level [p0, p0+n): y_p -= sum_k L_pk y_k, one thread per scalar row (5 per
// block row), in place over the permuted right-hand side.  Every load but the neighbours' y is
// indexed by the position alone, and consecutive positions hold consecutive L blocks.
template <class TF>
__global__ __launch_bounds__(256) void fwd_level_kernel(int p0, int n, const int4 *__restrict__ lnb,
                                                        const int *__restrict__ lptr, const TF *__restrict__ Lf,
                                                        double *__restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 5 * n) return;
  const int p = p0 + t / 5, q = t % 5;
  const int4 nb = lnb[p];
  const TF *L = Lf + (size_t)lptr[p] * 25 + 5 * q;
  double s = y[(size_t)5 * p + q];
  if (nb.x >= 0) {
    s -= dot5(L, y + (size_t)5 * nb.x);
    if (nb.y >= 0) {
      s -= dot5(L + 25, y + (size_t)5 * nb.y);
      if (nb.z >= 0) {
        s -= dot5(L + 50, y + (size_t)5 * nb.z);
        if (nb.w >= 0) s -= dot5(L + 75, y + (size_t)5 * nb.w);
      }
    }
  }
  y[(size_t)5 * p + q] = s;
}

template <class TF>
__device__ inline void sub_block(const TF *__restrict__ U, const double *__restrict__ v, double *s) {
  double x[5];
  for (int c = 0; c < 5; c++) x[c] = v[c];
  for (int a = 0; a < 5; a++) s[a] -= dot5(U + 5 * a, x);
}

// Backward sweep over the level [p0, p0+n): w_p = inv(U_pp) (y_p - sum_k U_pk w_k), one thread per
// scalar row; each thread forms the row's full 5-vector (the 5x5 product needs all of it).
template <class TF>
__global__ __launch_bounds__(256) void bwd_level_kernel(int p0, int n, const int *__restrict__ ypos,
                                                        const int4 *__restrict__ unb, const int *__restrict__ uptr,
                                                        const TF *__restrict__ Uf, const TF *__restrict__ Db,
                                                        const double *__restrict__ y, double *__restrict__ w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 5 * n) return;
  const int p = p0 + t / 5, q = t % 5;
  const int4 nb = unb[p];
  const TF *U = Uf + (size_t)uptr[p] * 25;
  const double *yp = y + (size_t)5 * ypos[p];
  double s[5];
  for (int a = 0; a < 5; a++) s[a] = yp[a];
  if (nb.x >= 0) {
    sub_block(U, w + (size_t)5 * nb.x, s);
    if (nb.y >= 0) {
      sub_block(U + 25, w + (size_t)5 * nb.y, s);
      if (nb.z >= 0) {
        sub_block(U + 50, w + (size_t)5 * nb.z, s);
        if (nb.w >= 0) sub_block(U + 75, w + (size_t)5 * nb.w, s);
      }
    }
  }
  const TF *Di = Db + (size_t)25 * p + 5 * q;
  double acc = (double)Di[0] * s[0];
  for (int a = 1; a < 5; a++) acc += (double)Di[a] * s[a];
  w[(size_t)5 * p + q] = acc;
}

// Block Jacobi: D <- inv(D) per cell (also the FP32 copy); status = lowest singular cell
__global__ __launch_bounds__(128) void bjac_invert_kernel(int nC, double *__restrict__ D, float *__restrict__ D32,
                                                          int *__restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nC) return;
  double A[25], Ai[25];
  for (int q = 0; q < 25; q++) A[q] = D[(size_t)25 * i + q];
  if (!inv5(A, Ai)) {
    atomicMin(status, i);
    for (int q = 0; q < 25; q++) Ai[q] = 0.0;
  }
  for (int q = 0; q < 25; q++) {
    D[(size_t)25 * i + q] = Ai[q];
    D32[(size_t)25 * i + q] = (float)Ai[q];
  }
}

// z_c = inv(D_c) r_c: one thread per cell
template <class TF>
__global__ __launch_bounds__(256) void bjac_apply_kernel(int nC, const TF *__restrict__ Di,
                                                         const double *__restrict__ r, double *__restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nC) return;
  double w[5];
  for (int p = 0; p < 5; p++) w[p] = r[(size_t)p * nC + i];
  const TF *M = Di + (size_t)25 * i;
  for (int a = 0; a < 5; a++) {
    double s = 0.0;
    for (int p = 0; p < 5; p++) s += (double)M[5 * a + p] * w[p];
    z[(size_t)a * nC + i] = s;
  }
}

int build_ilu(fos_euler *ctx) {
  if (ctx->maxNF > NB) {
    set_error("psetup: cells with %d faces; the block ILU supports tetrahedral meshes (%d)", ctx->maxNF, NB);
    return -1;
  }
  BlockIlu *b = new BlockIlu();
  ctx->ilu = b;
  const int nC = ctx->nC, mf = ctx->maxNF;
  b->nC = nC;
  const char *ev = getenv("FOS_ILU_EAGER");
  b->eager = ev && ev[0] == '1';
  std::vector<int> nbr((size_t)NB * nC, -1), low((size_t)NB * nC, -1);
  for (int c = 0; c < nC; c++)
    for (int j = 0; j < mf; j++) {
      int e = ctx->hNeib[(size_t)mf * c + j];
      nbr[(size_t)NB * c + j] = (e >= 0 && e < nC) ? e : -1;
    }
  // levels: forward by lower neighbours, backward by upper neighbours
  std::vector<int> lev(nC, 0), ulev(nC, 0);
  int nl = 0, nu = 0;
  for (int c = 0; c < nC; c++) {
    int m = 0;
    std::vector<std::pair<int, int>> lo;
    for (int j = 0; j < NB; j++) {
      int k = nbr[(size_t)NB * c + j];
      if (k >= 0 && k < c) { m = std::max(m, lev[k] + 1); lo.push_back({k, j}); }
    }
    std::sort(lo.begin(), lo.end());
    for (size_t a = 0; a < lo.size(); a++) low[(size_t)NB * c + a] = lo[a].second;
    lev[c] = m;
    nl = std::max(nl, m + 1);
  }
  for (int c = nC - 1; c >= 0; c--) {
    int m = 0;
    for (int j = 0; j < NB; j++) {
      int k = nbr[(size_t)NB * c + j];
      if (k > c) m = std::max(m, ulev[k] + 1);
    }
    ulev[c] = m;
    nu = std::max(nu, m + 1);
  }
  auto group = [&](const std::vector<int> &L, int n, std::vector<int> &ptr, std::vector<int> &rows) {
    ptr.assign(n + 1, 0);
    for (int c = 0; c < nC; c++) ptr[L[c] + 1]++;
    for (int l = 0; l < n; l++) ptr[l + 1] += ptr[l];
    rows.resize(nC);
    std::vector<int> f(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < nC; c++) rows[f[L[c]]++] = c;
  };
  std::vector<int> rows, urows;
  group(lev, nl, b->levPtr, rows);
  group(ulev, nu, b->uLevPtr, urows);
  b->nLev = nl;
  b->nULev = nu;
  // residual-layout slot -> neib slot of the same neighbour
  const int NS = ctx->L64.nSlots;
  std::vector<short> slots((size_t)NS * nC);
  std::vector<int2> lr(ctx->nChunkFaces);
  std::vector<int> cb(ctx->nChunks + 1);
  hipStream_t s = ctx->stream;
  FOS_HIP_TRY(hipMemcpyAsync(slots.data(), ctx->dSlots.p, sizeof(short) * slots.size(), hipMemcpyDeviceToHost, s));
  FOS_HIP_TRY(hipMemcpyAsync(lr.data(), ctx->dLR.p, sizeof(int2) * lr.size(), hipMemcpyDeviceToHost, s));
  FOS_HIP_TRY(hipMemcpyAsync(cb.data(), ctx->dChunkBeg.p, sizeof(int) * cb.size(), hipMemcpyDeviceToHost, s));
  FOS_HIP_TRY(hipStreamSynchronize(s));
  std::vector<int> jslot((size_t)NS * nC, -1);
  for (int c = 0; c < nC; c++)
    for (int q = 0; q < NS; q++) {
      int code = slots[(size_t)NS * c + q];
      if (code < 0) continue;
      int2 p = lr[cb[c / CHUNK] + (code >> 1)];
      int other = (code & 1) ? p.x : p.y;
      if (other < 0) continue;
      for (int j = 0; j < mf; j++)
        if (ctx->hNeib[(size_t)mf * c + j] == other) { jslot[(size_t)NS * c + q] = j; break; }
    }
  // level order of the sweeps: positions, neighbour positions and the packed block sources
  std::vector<int> fpos(nC), bpos(nC), lptr(nC + 1), uptr(nC + 1), ypos(nC), lsrc, usrc;
  std::vector<int4> lnb(nC), unb(nC);
  for (int p = 0; p < nC; p++) {
    fpos[rows[p]] = p;
    bpos[urows[p]] = p;
  }
  auto pack = [&](const std::vector<int> &order, const std::vector<int> &pos, bool lower, std::vector<int4> &nbp,
                  std::vector<int> &ptr, std::vector<int> &src) {
    for (int p = 0; p < nC; p++) {
      const int i = order[p];
      int e[NB] = {-1, -1, -1, -1}, m = 0;
      ptr[p] = (int)src.size();
      for (int j = 0; j < NB; j++) {
        const int k = nbr[(size_t)NB * i + j];
        if (k < 0 || (lower ? k >= i : k <= i)) continue;
        e[m++] = pos[k];
        src.push_back(NB * i + j);
      }
      nbp[p] = make_int4(e[0], e[1], e[2], e[3]);
    }
    ptr[nC] = (int)src.size();
  };
  pack(rows, fpos, true, lnb, lptr, lsrc);
  pack(urows, bpos, false, unb, uptr, usrc);
  for (int p = 0; p < nC; p++) ypos[p] = fpos[urows[p]];
  b->nLBlk = (long long)lsrc.size();
  b->nUBlk = (long long)usrc.size();
  if (b->dNbr.upload(nbr.data(), nbr.size(), s) || b->dLowOrd.upload(low.data(), low.size(), s) ||
      b->dLevRows.upload(rows.data(), rows.size(), s) || b->dULevRows.upload(urows.data(), urows.size(), s) ||
      b->dJSlot.upload(jslot.data(), jslot.size(), s) || b->dLNb.upload(lnb.data(), lnb.size(), s) ||
      b->dUNb.upload(unb.data(), unb.size(), s) || b->dLPtr.upload(lptr.data(), lptr.size(), s) ||
      b->dUPtr.upload(uptr.data(), uptr.size(), s) || b->dYPos.upload(ypos.data(), ypos.size(), s) ||
      b->dFPos.upload(fpos.data(), fpos.size(), s) || b->dBPos.upload(bpos.data(), bpos.size(), s) ||
      b->dLSrc.upload(lsrc.data(), lsrc.size(), s) || b->dUSrc.upload(usrc.data(), usrc.size(), s) ||
      b->dJac.alloc(125 * (size_t)nC) || b->dD.alloc(25 * (size_t)nC) || b->dOff.alloc(25 * (size_t)NB * nC) ||
      b->dD32.alloc(25 * (size_t)nC) || b->dY.alloc(5 * (size_t)nC) || b->dW.alloc(5 * (size_t)nC) ||
      ctx->dStatus.alloc(1))
    return -1;
  return 0;
}

}  // namespace

int jacobian_dev(fos_euler *ctx, const double *y, double *jac) {
  if (!ctx->ilu && build_ilu(ctx)) return -1;
  const int nC = ctx->nC;
  if (nC == 0) return 0;
  const int NS = ctx->L64.nSlots;
  const dim3 blk(256);
  if (NS == 4)
    hipLaunchKernelGGL(jac_kernel<4>, dim3((unsigned)(((size_t)nC * 4 + 255) / 256)), blk, 0, ctx->stream, ctx->L64,
                       ctx->ilu->dJSlot.p, y, jac);
  else
    hipLaunchKernelGGL(jac_kernel<8>, dim3((unsigned)(((size_t)nC * 8 + 255) / 256)), blk, 0, ctx->stream, ctx->L64,
                       ctx->ilu->dJSlot.p, y, jac);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

// Level-ordered copies of the factors in the apply precision TF.
template <class TF>
static int gather_factors(fos_euler *ctx, DBuf<TF> &Lf, DBuf<TF> &Uf, DBuf<TF> &Db) {
  BlockIlu *b = ctx->ilu;
  const int nC = ctx->nC;
  if ((!Lf.p && Lf.alloc(25 * (size_t)b->nLBlk)) || (!Uf.p && Uf.alloc(25 * (size_t)b->nUBlk)) ||
      (!Db.p && Db.alloc(25 * (size_t)nC)))
    return -1;
  auto go = [&](long long nBlk, const int *src, const double *from, TF *to) {
    if (nBlk == 0) return;
    const size_t nt = (size_t)nBlk * 25;
    hipLaunchKernelGGL(gather_blocks_kernel<TF>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream, nBlk,
                       src, from, to);
  };
  go(b->nLBlk, b->dLSrc.p, b->dOff.p, Lf.p);
  go(b->nUBlk, b->dUSrc.p, b->dOff.p, Uf.p);
  go(nC, b->dULevRows.p, b->dD.p, Db.p);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

int ilu_factor(fos_euler *ctx, double gamma) {
  BlockIlu *b = ctx->ilu;
  const int nC = ctx->nC;
  const size_t nt = (size_t)nC * 25 * (1 + NB);
  hipLaunchKernelGGL(form_m_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream, nC, b->dJac.p,
                     ctx->dInvV64.p, b->dNbr.p, gamma, b->dD.p, b->dOff.p);
  const int big = 0x7fffffff;
  FOS_HIP_TRY(hipMemcpyAsync(ctx->dStatus.p, &big, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  auto level = [&](int l, auto &&launch) {
    const int n = b->levPtr[l + 1] - b->levPtr[l];
    return launch(factor_level_kernel, dim3((n + 127) / 128), dim3(128), b->dLevRows.p + b->levPtr[l], n, b->dNbr.p,
                  b->dLowOrd.p, b->dD.p, b->dOff.p, ctx->dStatus.p);
  };
  if (!b->eager && !b->factorChain.ready()) {
    int bad = 0;
    for (int l = 0; l < b->nLev && !bad; l++)
      bad = level(l, [&](auto k, dim3 g, dim3 blk, auto... a) { return b->factorChain.add(k, g, blk, a...); });
    if (bad || b->factorChain.instantiate()) {
      set_error("psetup: graph of the factorisation levels could not be built (%s)", b->factorChain.fail());
      return -1;
    }
  }
  if (b->eager) {
    for (int l = 0; l < b->nLev; l++)
      level(l, [&](auto k, dim3 g, dim3 blk, auto... a) {
        hipLaunchKernelGGL(k, g, blk, 0, ctx->stream, a...);
        return 0;
      });
  } else if (b->factorChain.launch(ctx->stream)) {
    set_error("psetup: graph launch failed (%s)", hipGetErrorString(b->factorChain.err));
    return -1;
  }
  FOS_HIP_TRY(hipGetLastError());
  if ((ctx->flags & FOS_PREC_FP32) ? gather_factors(ctx, b->dLf32, b->dUf32, b->dDb32)
                                   : gather_factors(ctx, b->dLf, b->dUf, b->dDb))
    return -1;
  int st = big;
  FOS_HIP_TRY(hipMemcpyAsync(&st, ctx->dStatus.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (st != big) {
    set_error("psetup: singular pivot block at cell %d", st);
    return 1;  // recoverable: the integrator may retry with a smaller step
  }
  return 0;
}

// Both sweeps in level order over dY (forward positions) and dW (backward positions); launch(kernel,
// grid, block, args...) either records or launches each level.
template <class TF, class Launch>
static int ilu_levels(BlockIlu *b, const TF *Lf, const TF *Uf, const TF *Db, Launch &&launch) {
  const dim3 blk(256);
  for (int l = 0; l < b->nLev; l++) {
    const int p0 = b->levPtr[l], n = b->levPtr[l + 1] - p0;
    if (launch(fwd_level_kernel<TF>, dim3((5 * n + 255) / 256), blk, p0, n, b->dLNb.p, b->dLPtr.p, Lf, b->dY.p))
      return -1;
  }
  for (int l = 0; l < b->nULev; l++) {
    const int p0 = b->uLevPtr[l], n = b->uLevPtr[l + 1] - p0;
    if (launch(bwd_level_kernel<TF>, dim3((5 * n + 255) / 256), blk, p0, n, b->dYPos.p, b->dUNb.p, b->dUPtr.p, Uf, Db,
               b->dY.p, b->dW.p))
      return -1;
  }
  return 0;
}

template <class TF>
static int ilu_sweeps(fos_euler *ctx, const TF *Lf, const TF *Uf, const TF *Db, const double *r, double *z) {
  BlockIlu *b = ctx->ilu;
  const int nC = ctx->nC;
  const dim3 gc((nC + 255) / 256), blk(256);
  if (!b->eager && !b->solveChain.ready()) {
    if (ilu_levels(b, Lf, Uf, Db,
                   [&](auto k, dim3 g, dim3 bl, auto... a) { return b->solveChain.add(k, g, bl, a...); }) ||
        b->solveChain.instantiate()) {
      set_error("psolve: graph of the sweep levels could not be built (%s)", b->solveChain.fail());
      return -1;
    }
  }
  hipLaunchKernelGGL(ilu_permute_in_kernel, gc, blk, 0, ctx->stream, nC, b->dFPos.p, r, b->dY.p);
  if (b->eager) {
    ilu_levels(b, Lf, Uf, Db, [&](auto k, dim3 g, dim3 bl, auto... a) {
      hipLaunchKernelGGL(k, g, bl, 0, ctx->stream, a...);
      return 0;
    });
  } else if (b->solveChain.launch(ctx->stream)) {
    set_error("psolve: graph launch failed (%s)", hipGetErrorString(b->solveChain.err));
    return -1;
  }
  hipLaunchKernelGGL(ilu_permute_out_kernel, gc, blk, 0, ctx->stream, nC, b->dBPos.p, b->dW.p, z);
  return 0;
}

int ilu_solve(fos_euler *ctx, const double *r, double *z) {
  BlockIlu *b = ctx->ilu;
  if (ctx->nC == 0) return 0;
  if ((ctx->flags & FOS_PREC_FP32) ? ilu_sweeps(ctx, b->dLf32.p, b->dUf32.p, b->dDb32.p, r, z)
                                   : ilu_sweeps(ctx, b->dLf.p, b->dUf.p, b->dDb.p, r, z))
    return -1;
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

int bjac_factor(fos_euler *ctx, double gamma) {
  BlockIlu *b = ctx->ilu;
  const int nC = ctx->nC;
  const size_t nt = (size_t)nC * 25 * (1 + NB);
  hipLaunchKernelGGL(form_m_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream, nC, b->dJac.p,
                     ctx->dInvV64.p, b->dNbr.p, gamma, b->dD.p, b->dOff.p);
  const int big = 0x7fffffff;
  FOS_HIP_TRY(hipMemcpyAsync(ctx->dStatus.p, &big, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(bjac_invert_kernel, dim3((nC + 127) / 128), dim3(128), 0, ctx->stream, nC, b->dD.p, b->dD32.p,
                     ctx->dStatus.p);
  FOS_HIP_TRY(hipGetLastError());
  int st = big;
  FOS_HIP_TRY(hipMemcpyAsync(&st, ctx->dStatus.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (st != big) {
    set_error("psetup_bjac: singular diagonal block at cell %d", st);
    return 1;
  }
  b->bjac = true;
  return 0;
}

int bjac_solve(fos_euler *ctx, const double *r, double *z) {
  BlockIlu *b = ctx->ilu;
  const int nC = ctx->nC;
  if (ctx->flags & FOS_PREC_FP32)
    hipLaunchKernelGGL(bjac_apply_kernel<float>, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, b->dD32.p, r, z);
  else
    hipLaunchKernelGGL(bjac_apply_kernel<double>, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, b->dD.p, r, z);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

int psetup_dev(fos_euler *ctx, const double *y, int jok, int *jcur, double gamma) {
  if (!ctx->ilu && build_ilu(ctx)) return -1;
  BlockIlu *b = ctx->ilu;
  if (jok && b->haveJac) {
    *jcur = 0;
  } else {
    if (jacobian_dev(ctx, y, b->dJac.p)) return -1;
    b->haveJac = true;
    *jcur = 1;
  }
  b->bjac = false;
  return ilu_factor(ctx, gamma);
}

int psetup_bjac_dev(fos_euler *ctx, const double *y, int jok, int *jcur, double gamma) {
  if (!ctx->ilu && build_ilu(ctx)) return -1;
  BlockIlu *b = ctx->ilu;
  if (jok && b->haveJac) {
    *jcur = 0;
  } else {
    if (jacobian_dev(ctx, y, b->dJac.p)) return -1;
    b->haveJac = true;
    *jcur = 1;
  }
  return bjac_factor(ctx, gamma);
}

}  // namespace fos

using namespace fos;

extern "C" {

int fos_euler_psetup(double t, const double *y, const double *fy, int jok, int *jcur, double gamma, void *user_data) {
  (void)t;
  (void)fy;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx || !jcur) return -1;
  ctx->counters[2]++;
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("psetup: a *_UDF boundary condition is used but fos_euler_set_udf_bc was not called");
    return -1;
  }
  if (ctx->flags & FOS_VECTORS_ON_HOST) {
    if (stage_in(ctx, ctx->hY, y)) return -1;
    return psetup_dev(ctx, ctx->hY.p, jok, jcur, gamma);
  }
  return psetup_dev(ctx, y, jok, jcur, gamma);
}

int fos_euler_psolve(double t, const double *y, const double *fy, const double *r, double *z, double gamma,
                     double delta, int lr, void *user_data) {
  (void)t; (void)y; (void)fy; (void)gamma; (void)delta; (void)lr;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) return -1;
  ctx->counters[3]++;
  if (!ctx->ilu || !ctx->ilu->haveJac || ctx->ilu->bjac) {
    set_error("psolve: called before psetup");
    return -1;
  }
  if (ctx->flags & FOS_VECTORS_ON_HOST) {
    if (stage_in(ctx, ctx->hR, r) || ensure_vec(ctx->hZ, 5 * (size_t)ctx->nC)) return -1;
    if (ilu_solve(ctx, ctx->hR.p, ctx->hZ.p)) return -1;
    return stage_out(ctx, ctx->hZ, z);
  }
  return ilu_solve(ctx, r, z);
}

int fos_euler_psetup_bjac(double t, const double *y, const double *fy, int jok, int *jcur, double gamma,
                          void *user_data) {
  (void)t;
  (void)fy;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx || !jcur || (ctx->flags & FOS_VECTORS_ON_HOST)) return -1;
  ctx->counters[2]++;
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("psetup_bjac: a *_UDF boundary condition is used but fos_euler_set_udf_bc was not called");
    return -1;
  }
  return psetup_bjac_dev(ctx, y, jok, jcur, gamma);
}

int fos_euler_psolve_bjac(double t, const double *y, const double *fy, const double *r, double *z, double gamma,
                          double delta, int lr, void *user_data) {
  (void)t; (void)y; (void)fy; (void)gamma; (void)delta; (void)lr;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx || (ctx->flags & FOS_VECTORS_ON_HOST)) return -1;
  ctx->counters[3]++;
  if (!ctx->ilu || !ctx->ilu->bjac) {
    set_error("psolve_bjac: called before psetup_bjac");
    return -1;
  }
  return bjac_solve(ctx, r, z);
}

int fos_find_euler_jac(fos_euler *ctx, const double *y, double *JacC) {
  if (!ctx) return -1;
  if (ctx->maxNF > NB) {
    set_error("findEulerJac: cells with %d faces are not supported", ctx->maxNF);
    return -1;
  }
  return jacobian_dev(ctx, y, JacC);
}

}  // extern "C"
