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
#include <type_traits>
#include <utility>
#include <vector>

#include "context.h"

namespace fos {

constexpr int NB = 4;  // neighbour slots per cell (tetrahedra)

struct BlockIlu {
  int nC = 0, nLev = 0, nCol = 0;
  bool haveJac = false, bjac = false;  // factors hold block ILU(0), or block Jacobi inverses
  std::vector<int> levPtr, colPtr, hOrder;
  DBuf<int> dNbr;       // [NB nC] neighbour cell per slot (neib order), -1 = none / facet
  DBuf<int> dRank;      // [nC] elimination step of each row (its position in the sweeps' vectors)
  DBuf<int> dOrder;     // [nC] row eliminated at each step
  DBuf<int> dLowOrd;    // [NB nC] slots of lower neighbours in elimination order, -1 padded
  DBuf<int> dLevRows;   // rows grouped by factorisation level
  DBuf<int> dJSlot;     // [nSlots nC] neib slot j of residual-layout slot s (-1 boundary / none)
  DBuf<double> dJac;    // [125 nC] JacC(5, 25, nC) column-major
  DBuf<double> dD, dOff;  // factor: inverted U diagonal blocks [25 nC], off-diagonal blocks [25 NB nC]
  DBuf<float> dD32;       // FP32 copy of dD (block Jacobi apply)
  // Copies of the factors for the sweeps (built after each factorisation in the context's apply
  // precision): per chunk its rows' L blocks packed in the forward sweep's step order and its U
  // blocks in the backward sweep's, each row's blocks consecutive; the inverted diagonal block of
  // elimination position p at p.
  DBuf<int> dLSrc, dUSrc;  // packed block -> source block NB*row + slot in dOff
  long long nLBlk = 0, nUBlk = 0;
  DBuf<double> dLf, dUf, dDb;
  DBuf<float> dLf32, dUf32, dDb32;
  // sweep schedule: the chunks of each colour (colPtr) in launch order, per launch slot the chunk's
  // metadata (two int4, forward and backward), per chunk its steps (forward by the lower, backward
  // by the upper neighbours inside the chunk) as step records of chunk-local codes, and its
  // neighbours outside the chunk (cells; the records code them REC_OUT0 + index)
  DBuf<int4> dFMeta, dBMeta;
  DBuf<uint2> dFRec, dBRec;
  DBuf<int> dFOutCell, dBOutCell;
  int fMaxOut = 0, bMaxOut = 0;
  DBuf<const double *> dIo;  // [2] the current apply's right-hand side r and result z (SoA, by cell)
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

// Column q of d flow / d U_side for one face: the Roe flux in dual numbers with component q of one
// side's conservative state seeded (side 0: the pair's m, 1: its n); with ghost, the n side is
// setBC's ghost of m, so the m column carries the ghost's dependence on m as well.
__device__ inline void flux_col(const double (&Um)[5], const double (&Un)[5], bool ghost, int bt, const double *bp,
                                double nx, double ny, double nz, double area, double g, double r, int side, int q,
                                double col[5]) {
  using DT = Dual<double>;
  const DT g1(g - 1.0);
  DT a[5];
#pragma unroll
  for (int k = 0; k < 5; k++) a[k] = DT(Um[k], side == 0 && k == q ? 1.0 : 0.0);
  Prim<DT> A, B;
  make_prim(a[0], a[1], a[2], a[3], a[4], g1, A);
  if (ghost) {
    DT gr, gx, gy, gz, gE;
    ghost_state<DT>(bt, bp, DT(nx), DT(ny), DT(nz), A, DT(g), DT(r), gr, gx, gy, gz, gE);
    make_prim(gr, gx, gy, gz, gE, g1, B);
  } else {
    DT b[5];
#pragma unroll
    for (int k = 0; k < 5; k++) b[k] = DT(Un[k], side == 1 && k == q ? 1.0 : 0.0);
    make_prim(b[0], b[1], b[2], b[3], b[4], g1, B);
  }
  DT F[5];
  roe_flux<DT>(A, B, DT(nx), DT(ny), DT(nz), DT(area), g1, F);
#pragma unroll
  for (int k = 0; k < 5; k++) col[k] = F[k].d;
}

// The exact, boundary-consistent JacC (FOS_JAC_EXACT): the blocks of the residual's true
// derivative (not divided by v), column by column from flux_col -- the Roe dissipation
// differentiated with its averages, boundary faces through their ghost states -- where
// findEulerJac freezes |A~| and drops the ghost dependence (Euler.f90:210-219).  Same lanes,
// layout and zeroing as jac_kernel.
template <int NS>
__global__ __launch_bounds__(256) void jac_exact_kernel(FluxLayout<double> L, const int *__restrict__ jslot,
                                                        const double *__restrict__ y, double *__restrict__ jac) {
  static_assert(NS >= NB && (NS & (NS - 1)) == 0 && 64 % NS == 0, "slot lanes");
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = t / NS, s = t % NS;
  const bool live = c < L.nC;
  const int nC = L.nC;
  double D[25];
#pragma unroll
  for (int k = 0; k < 25; k++) D[k] = 0.0;
  int jj = -1;
  double *out = jac + (size_t)125 * c;
  const int code = live ? L.slots[(size_t)NS * c + s] : -1;
  if (code >= 0) {
    const int f = L.chunkBeg[c / CHUNK] + (code >> 1);
    const bool isR = code & 1;  // this cell is the pair's n: dU_c = +flow, else dU_c = -flow
    const int2 lr = L.lr[f];
    const AreaNormal<double> an = L.geo[f];
    double nx, ny, nz, area;
    face_geom(an.x, an.y, an.z, nx, ny, nz, area);
    double Um[5], Un[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 5; k++) Um[k] = y[(size_t)k * nC + lr.x];
    const bool ghost = lr.y < 0;
    if (!ghost)
#pragma unroll
      for (int k = 0; k < 5; k++) Un[k] = y[(size_t)k * nC + lr.y];
    const int bnd = ghost ? -lr.y - 1 : 0;
    const int bt = ghost ? L.bndType[bnd] : 0;
    const double *bp = L.bndPar + 5 * (size_t)bnd;
    const double sg = isR ? 1.0 : -1.0;
    double col[5];
    if (!ghost) {  // the block of the other cell
      jj = jslot[(size_t)NS * c + s];
      if (jj >= 0) {
#pragma unroll 1
        for (int q = 0; q < 5; q++) {
          flux_col(Um, Un, false, 0, bp, nx, ny, nz, area, L.gamma, L.r, isR ? 0 : 1, q, col);
          for (int k = 0; k < 5; k++) out[25 * (1 + jj) + 5 * q + k] = sg * col[k];
        }
      }
    }
#pragma unroll 1
    for (int q = 0; q < 5; q++) {  // this cell's own column of the face's flux
      flux_col(Um, Un, ghost, bt, bp, nx, ny, nz, area, L.gamma, L.r, isR ? 1 : 0, q, col);
#pragma unroll
      for (int k = 0; k < 5; k++) {
#pragma unroll
        for (int qq = 0; qq < 5; qq++)
          if (qq == q) D[5 * qq + k] = sg * col[k];
      }
    }
  }
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

// One level of the IKJ block factorisation: one thread per block row, "lower" meaning eliminated
// earlier (rank).  The diagonal block and the current L_ik stay in registers (static indices only);
// the rare update of another off-diagonal block of row i (a neighbour of both i and k) goes to
// memory in place.
__global__ __launch_bounds__(128) void factor_level_kernel(const int *__restrict__ rows, int n,
                                                           const int *__restrict__ nbr, const int *__restrict__ lowOrd,
                                                           const int *__restrict__ rank, double *__restrict__ D,
                                                           double *__restrict__ Off, int *__restrict__ status) {
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
    const int rk = rank[k];
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
      if (j < 0 || rank[j] <= rk) continue;  // U part of row k: eliminated after k
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

// y[rank[i]][q] = r[q][i]: the right-hand side into elimination order, cell-major.

template <class TF>
__device__ inline double dot5(const TF *__restrict__ a, const double *__restrict__ b) {
  double s = (double)a[0] * b[0];
  for (int p = 1; p < 5; p++) s += (double)a[p] * b[p];
  return s;
}

// Cells per chunk of the block ILU's elimination order and sweeps: consecutive cells, so with the
// residual layout's compact 256-cell chunks (coordinate bisection) they are compact sub-blocks.
#ifndef FOS_ILU_CHUNK
#define FOS_ILU_CHUNK 64
#endif
constexpr int ICH = FOS_ILU_CHUNK;
static_assert(ICH > 0 && CHUNK % ICH == 0, "ILU chunks nest in residual chunks");

// Rows per step of the colour sweeps: 12 rows x 5 components on one wave.
constexpr int SWEEP_ROWS = 12;

// Step records: one uint2 per row slot, REC_SLOTS per step (the last slot, read by lanes 60-63, is
// always empty).  The row's four neighbours on the swept side have 9-bit codes (0..63: a row of the
// chunk, 64 + k: entry k of the chunk's outside list, 511: none, after the row's last neighbour),
// low 8 bits in the bytes of x, bit 8 in bits 16-19 of y; y bits 0-7: the row's local index in the
// chunk, bits 8-15: the offset of its first packed block from the chunk's first block, bit 20: the
// step starts a new local level (its rows read rows of the previous step), bit 21: empty slot (row
// 0, no neighbours), bits 22-27: the step's blocks (the chunk's blocks are packed in step order, a
// row's consecutively, so a step's blocks are one run).  Slot 0 of a step always holds a row.  A chunk has at most ICH steps, followed
// by two empty steps (so a lane reads its records two steps ahead without a bounds test), and at
// most 4 ICH outside neighbours.
constexpr int REC_SLOTS = SWEEP_ROWS + 1, REC_PAD_STEPS = 2;
constexpr unsigned REC_NB_NONE = 511u, REC_NEW = 1u << 20, REC_EMPTY = 1u << 21, REC_BLK_SHIFT = 22;
constexpr int REC_OUT0 = 64, MAX_OUT = 4 * ICH;
static_assert(SWEEP_ROWS * 5 <= 64 && REC_SLOTS * 5 > 64, "lanes past the last row slot read the empty slot");
static_assert(ICH <= REC_OUT0 && 4 * (ICH - 1) < 256 && REC_OUT0 + MAX_OUT <= (int)REC_NB_NONE && 4 * SWEEP_ROWS < 64,
              "chunk-local record fields fit their bits");
__host__ __device__ inline unsigned rec_nb(uint2 rc, int j) {
  return ((rc.x >> (8 * j)) & 255u) | ((rc.y >> (16 + j)) & 1u) << 8;
}

// Per launch slot of a colour launch (chunks in launch order), two int4: {first cell c0, first step
// record s0, steps ns, first outside entry o0}, {outside entries nOut, first packed block, first
// elimination position pb, 0}.
//
// dynamic LDS of a colour-sweep launch: 5 ICH doubles of row values, 5 SWEEP_ROWS doubles of
// exchange, 5 maxOut outside values
// The sweeps stage each step's blocks through LDS (one run of at most 4 SWEEP_ROWS blocks, read as
// stage_pieces 16-byte pieces per lane from its 16-byte aligned start) instead of gathering each
// lane's block rows from memory: on for FP32 factors, a build switch for FP64.
#ifndef FOS_SWEEP_STAGED32
#define FOS_SWEEP_STAGED32 1
#endif
#ifndef FOS_SWEEP_STAGED64
#define FOS_SWEEP_STAGED64 0
#endif
template <class TF>
constexpr bool sweep_staged() { return sizeof(TF) == 4 ? FOS_SWEEP_STAGED32 : FOS_SWEEP_STAGED64; }
template <class TF>
constexpr int stage_pieces() {  // 16-byte pieces per lane: a step's run plus the alignment slack
  return (4 * SWEEP_ROWS * 25 + 2 * (int)(16 / sizeof(TF)) - 1) / (int)(16 / sizeof(TF) * 64) + 1;
}
template <class TF>
constexpr size_t stage_bytes() { return sweep_staged<TF>() ? (size_t)stage_pieces<TF>() * 64 * 16 : 0; }
template <class TF>
constexpr size_t sweep_lds(int maxOut) {
  return stage_bytes<TF>() + sizeof(double) * 5 * (ICH + SWEEP_ROWS + (size_t)maxOut);
}
static_assert(sweep_lds<float>(MAX_OUT) <= 65536 && sweep_lds<double>(MAX_OUT) <= 65536, "a chunk's sweep fits LDS");
template <class TF>
struct SweepLds {
  TF *blk;  // staged step blocks (16-byte aligned)
  double *val, *ex, *out;
  __device__ explicit SweepLds(unsigned char *raw) {
    blk = reinterpret_cast<TF *>(raw);
    val = reinterpret_cast<double *>(raw + stage_bytes<TF>());
    ex = val + 5 * ICH;
    out = ex + 5 * SWEEP_ROWS;
  }
};

// The vectors come through a device record of pointers, so their address space is stated here:
// plain pointer accesses would compile to flat loads and stores.
typedef __attribute__((address_space(1))) double gdouble;
__device__ inline double ld_g(const double *p) { return *(const gdouble *)p; }
__device__ inline void st_g(double *p, double v) { *(gdouble *)p = v; }

// A chunk's sweep is a sequence of steps of up to SWEEP_ROWS rows of one local level each (a level
// with more rows takes several steps); the step records, in schedule order, give per row slot the
// row, its neighbours on the swept side and the offset of its packed blocks.  A neighbour inside
// the chunk is read from LDS (written at an earlier level of the launch); one outside the chunk,
// finished by an earlier launch, is an entry of the chunk's outside list (cell indices), whose
// values the kernel stages in LDS before the first step.  Each lane reads its slot's records from
// memory two steps ahead and the blocks one step ahead of the arithmetic, so a step costs one
// memory round trip; before the first step a chunk makes three: its metadata, then its records of
// steps 0 and 1 with the outside cell indices, then step 0's blocks with the outside values.  Vectors stay in the
// callers' layout (component-major, by cell): a chunk's rows are consecutive cells, so local row i
// is cell c0 + i (and elimination position pb + i).  The forward sweep reads r and writes y into z;
// the backward sweep runs in place on z.

// the chunk's outside cell indices (up to MAX_OUT, 4 per lane), then their values into LDS; the
// index loads are unconditional, clamped into the list (which ends with one padding entry), so that
// they issue together
struct OutStage {
  int cell[MAX_OUT / 64];
  double v[MAX_OUT / 64][5];
};
__device__ inline void out_cells(OutStage &g, const int *__restrict__ outCell, int o0, int nOut) {
#pragma unroll
  for (int j = 0; j < MAX_OUT / 64; j++) g.cell[j] = outCell[o0 + max(min((int)threadIdx.x + 64 * j, nOut - 1), 0)];
}
__device__ inline void out_values(OutStage &g, const double *vals, int nC, int nOut) {
#pragma unroll
  for (int j = 0; j < MAX_OUT / 64; j++)
    if ((int)threadIdx.x + 64 * j < nOut)
#pragma unroll
      for (int q = 0; q < 5; q++) g.v[j][q] = ld_g(vals + (size_t)q * nC + g.cell[j]);
}
__device__ inline void out_store(double *out, const OutStage &g, int nOut) {
#pragma unroll
  for (int j = 0; j < MAX_OUT / 64; j++) {
    const int k = threadIdx.x + 64 * j;
    if (k < nOut)
#pragma unroll
      for (int q = 0; q < 5; q++) out[5 * k + q] = g.v[j][q];
  }
}

// sum_j B_j[q][:] . x_j over the row's neighbours (B row q per neighbour in Bq, x from LDS; rc the
// row's step record)
template <class TF>
__device__ inline double nb_dot(uint2 rc, const TF (&Bq)[4][5], const double *sv, const double *so) {
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned k = rec_nb(rc, j);
    if (k == REC_NB_NONE) break;
    const double *x = k < (unsigned)REC_OUT0 ? sv + 5 * k : so + 5 * (k - REC_OUT0);
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 5; c++) t += (double)Bq[j][c] * x[c];
    acc += t;
  }
  return acc;
}

// row q of each of the row's blocks from its record (none for an empty slot)
template <class TF>
__device__ inline void load_blocks(uint2 rc, const TF *__restrict__ B, int blk0, int q, TF (&Bq)[4][5]) {
  const TF *Bp = B + (size_t)(blk0 + (int)((rc.y >> 8) & 255u)) * 25 + 5 * q;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (rec_nb(rc, j) == REC_NB_NONE) break;
#pragma unroll
    for (int c = 0; c < 5; c++) Bq[j][c] = Bp[25 * j + c];
  }
}

// A step's block run (staged sweeps): its bounds come from slot 0's record (the same in every lane);
// every lane loads stage_pieces 16-byte pieces (clamped into the run, so the loads are unconditional
// and issue together), later written to LDS.  The packed arrays end with a spare block, so the
// aligned pieces stay inside them.
typedef float pf32 __attribute__((ext_vector_type(4)));
typedef double pf64 __attribute__((ext_vector_type(2)));
template <class TF>
struct StepRun {
  using P = typename std::conditional<sizeof(TF) == 4, pf32, pf64>::type;  // one 16-byte piece
  static constexpr int PV = 16 / sizeof(TF);                               // values per piece
  P pc[stage_pieces<TF>()];
  int off, npc, start;  // first value of the run in the staged data, pieces, first block (chunk-local)
};
template <class TF>
__device__ inline void run_load(StepRun<TF> &R, uint2 rc, const TF *__restrict__ B, int blk0) {
  using P = typename StepRun<TF>::P;
  constexpr int PV = StepRun<TF>::PV;
  const unsigned y = __builtin_amdgcn_readfirstlane(rc.y);
  R.start = (int)((y >> 8) & 255u);
  const size_t b0 = (size_t)(blk0 + R.start) * 25, a0 = b0 & ~(size_t)(PV - 1);
  R.off = (int)(b0 - a0);
  R.npc = (R.off + (int)((y >> REC_BLK_SHIFT) & 63u) * 25 + PV - 1) / PV;
  const P *src = reinterpret_cast<const P *>(B + a0);
#pragma unroll
  for (int k = 0; k < stage_pieces<TF>(); k++) R.pc[k] = src[max(min((int)threadIdx.x + 64 * k, R.npc - 1), 0)];
}
template <class TF>
__device__ inline void run_store(TF *blk, const StepRun<TF> &R) {
#pragma unroll
  for (int k = 0; k < stage_pieces<TF>(); k++) {
    const int pi = threadIdx.x + 64 * k;
    if (pi < R.npc) reinterpret_cast<typename StepRun<TF>::P *>(blk)[pi] = R.pc[k];
  }
}
// nb_dot with row q of the row's blocks read from the staged run (bl: its first block's row q)
template <class TF>
__device__ inline double nb_dot_staged(uint2 rc, const TF *bl, const double *sv, const double *so) {
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned k = rec_nb(rc, j);
    if (k == REC_NB_NONE) break;
    const double *x = k < (unsigned)REC_OUT0 ? sv + 5 * k : so + 5 * (k - REC_OUT0);
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 5; c++) t += (double)bl[25 * j + c] * x[c];
    acc += t;
  }
  return acc;
}

// The chunk's results, which the steps keep in LDS, to z once the chunk is done: one coalesced
// 512-byte run per component instead of the steps' scattered 8-byte stores (no chunk reads another
// chunk's rows of z before their launch ends).
__device__ inline void chunk_store(double *z, const double *val, int nC, int c0) {
  __syncthreads();  // the last step's rows are in LDS
  const int t = threadIdx.x;
  if (t < min(ICH, nC - c0))
#pragma unroll
    for (int q = 0; q < 5; q++) st_g(z + (size_t)q * nC + c0 + t, val[5 * t + q]);
}

// The chunk's own rows of src (r for the forward sweep, y for the backward one) into LDS, one
// coalesced run per component; each step reads its rows there and overwrites them with its results.
__device__ inline void chunk_load(double *val, const double *src, int nC, int c0) {
  const int t = threadIdx.x;
  if (t < min(ICH, nC - c0))
#pragma unroll
    for (int q = 0; q < 5; q++) val[5 * t + q] = ld_g(src + (size_t)q * nC + c0 + t);
}

// Forward sweep over the chunks of one colour, one wave per chunk (meta: the colour's launch
// slots): y_p = r_p - sum_k L_pk y_k step by step, from r = io[0] into z = io[1].
template <class TF>
__global__ __launch_bounds__(64) void fwd_colour_kernel(int nC, const int4 *__restrict__ meta,
                                                        const uint2 *__restrict__ rec, const int *__restrict__ outCell,
                                                        const TF *__restrict__ Lf, const double *const *__restrict__ io) {
  constexpr bool STG = sweep_staged<TF>();
  extern __shared__ unsigned char sweep_raw[];
  const SweepLds<TF> S(sweep_raw);
  const double *r = io[0];
  double *z = const_cast<double *>(io[1]);
  const int4 m0 = meta[2 * blockIdx.x], m1 = meta[2 * blockIdx.x + 1];
  const int c0 = m0.x, s0 = m0.y, ns = m0.z, o0 = m0.w, nOut = m1.x, blk0 = m1.y;  // ns <= ICH, nOut <= MAX_OUT
  const int slot = threadIdx.x / 5, q = threadIdx.x - 5 * slot;
  // this lane's records (lanes 60-63 read the empty slot); every load below is unconditional or
  // writes registers no other path writes, so no wait is forced before the next step's loads issue
  const uint2 *rl = rec + (size_t)s0 * REC_SLOTS + slot;
  if constexpr (STG) {
    OutStage g;
    uint2 rc = rl[0], rcN = rl[REC_SLOTS];
    out_cells(g, outCell, o0, nOut);
    StepRun<TF> R;
    run_load(R, rc, Lf, blk0);
    out_values(g, z, nC, nOut);
    chunk_load(S.val, r, nC, c0);
    out_store(S.out, g, nOut);
    run_store(S.blk, R);
    int off = R.off, start = R.start;
    __syncthreads();
    for (int st = 0; st < ns; st++) {
      const uint2 rcNN = rl[(size_t)(st + 2) * REC_SLOTS];
      run_load(R, rcN, Lf, blk0);  // step st + 1's run, in flight during this step
      if (st > 0 && (__builtin_amdgcn_readfirstlane(rc.y) & REC_NEW)) __syncthreads();  // previous level in LDS
      const unsigned i = rc.y & 255u;
      if (!(rc.y & REC_EMPTY)) {
        const TF *bl = S.blk + off + ((int)((rc.y >> 8) & 255u) - start) * 25 + 5 * q;
        const double v = S.val[5 * i + q] - nb_dot_staged(rc, bl, S.val, S.out);
        S.val[5 * i + q] = v;
      }
      run_store(S.blk, R);  // after this step's reads of the staged run
      off = R.off;
      start = R.start;
      rc = rcN;
      rcN = rcNN;
    }
    chunk_store(z, S.val, nC, c0);
    return;
  }
  // the loads of step st: row q of each L block
  auto fetch = [&](uint2 rc, TF (&L)[4][5]) { load_blocks(rc, Lf, blk0, q, L); };
  OutStage g;
  uint2 rc = rl[0], rcN = rl[REC_SLOTS];
  out_cells(g, outCell, o0, nOut);
  TF Lc[4][5], Ln[4][5];
  fetch(rc, Lc);
  out_values(g, z, nC, nOut);
  chunk_load(S.val, r, nC, c0);
  out_store(S.out, g, nOut);
  __syncthreads();
  for (int st = 0; st < ns; st++) {
    const uint2 rcNN = rl[(size_t)(st + 2) * REC_SLOTS];
    fetch(rcN, Ln);
    if (st > 0 && (__builtin_amdgcn_readfirstlane(rc.y) & REC_NEW)) __syncthreads();  // previous level in LDS
    const unsigned i = rc.y & 255u;
    if (!(rc.y & REC_EMPTY)) {
      const double v = S.val[5 * i + q] - nb_dot(rc, Lc, S.val, S.out);
      S.val[5 * i + q] = v;
    }
    rc = rcN;
    rcN = rcNN;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 5; c++) Lc[j][c] = Ln[j][c];
  }
  chunk_store(z, S.val, nC, c0);
}

// Backward sweep over the chunks of one colour (colours in descending order), in place on z = io[1]:
// w_p = inv(U_pp) (y_p - sum_k U_pk w_k).  Lane (row, q) forms component q of y_p - sum_k U_pk w_k
// from row q of each U block, the row's 5 components meet in LDS, and lane q applies row q of
// inv(U_pp) (stored by elimination position).  A non-finite result sets *nf.
template <class TF>
__global__ __launch_bounds__(64) void bwd_colour_kernel(int nC, const int4 *__restrict__ meta,
                                                        const uint2 *__restrict__ rec, const int *__restrict__ outCell,
                                                        const TF *__restrict__ Uf, const TF *__restrict__ Db,
                                                        const double *const *__restrict__ io, unsigned *__restrict__ nf) {
  constexpr bool STG = sweep_staged<TF>();
  extern __shared__ unsigned char sweep_raw[];
  const SweepLds<TF> S(sweep_raw);
  double *sT = S.ex;
  double *z = const_cast<double *>(io[1]);
  const int4 m0 = meta[2 * blockIdx.x], m1 = meta[2 * blockIdx.x + 1];
  const int c0 = m0.x, s0 = m0.y, ns = m0.z, o0 = m0.w, nOut = m1.x, blk0 = m1.y, pb = m1.z;
  const int slot = threadIdx.x / 5, q = threadIdx.x - 5 * slot;
  const uint2 *rl = rec + (size_t)s0 * REC_SLOTS + slot;  // as in the forward sweep
  if constexpr (STG) {
    // as the forward sweep: the U blocks staged per step; row q of inv(U_pp) per lane
    auto fetch_row = [&](uint2 rc, TF (&Dq)[5]) {
      const unsigned i = rc.y & 255u;
#pragma unroll
      for (int c = 0; c < 5; c++) Dq[c] = Db[(size_t)25 * (pb + (int)i) + 5 * q + c];
    };
    OutStage g;
    uint2 rc = rl[0], rcN = rl[REC_SLOTS];
    out_cells(g, outCell, o0, nOut);
    TF Dc[5], Dn[5];
    bool bad = false;
    fetch_row(rc, Dc);
    StepRun<TF> R;
    run_load(R, rc, Uf, blk0);
    out_values(g, z, nC, nOut);
    chunk_load(S.val, z, nC, c0);
    out_store(S.out, g, nOut);
    run_store(S.blk, R);
    int off = R.off, start = R.start;
    __syncthreads();
    for (int st = 0; st < ns; st++) {
      const uint2 rcNN = rl[(size_t)(st + 2) * REC_SLOTS];
      fetch_row(rcN, Dn);
      run_load(R, rcN, Uf, blk0);
      if (st > 0 && (__builtin_amdgcn_readfirstlane(rc.y) & REC_NEW)) __syncthreads();  // previous level in LDS
      const unsigned i = rc.y & 255u;
      const bool full = !(rc.y & REC_EMPTY);  // slot < SWEEP_ROWS
      if (full) {
        const TF *bl = S.blk + off + ((int)((rc.y >> 8) & 255u) - start) * 25 + 5 * q;
        sT[5 * slot + q] = S.val[5 * i + q] - nb_dot_staged(rc, bl, S.val, S.out);
      }
      __syncthreads();  // the row's 5 components of y_p - sum U w
      if (full) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 5; c++) acc += (double)Dc[c] * sT[5 * slot + c];
        S.val[5 * i + q] = acc;
        bad |= !isfinite(acc);
      }
      __syncthreads();  // sT is reused by the next step
      run_store(S.blk, R);
      off = R.off;
      start = R.start;
      rc = rcN;
      rcN = rcNN;
#pragma unroll
      for (int c = 0; c < 5; c++) Dc[c] = Dn[c];
    }
    chunk_store(z, S.val, nC, c0);
    if (bad) atomicOr(nf, 1u);
    return;
  }
  // the loads of step st: row q of each U block and of inv(U_pp) (y_p comes from LDS)
  auto fetch = [&](uint2 rc, TF (&U)[4][5], TF (&Dq)[5]) {
    const unsigned i = rc.y & 255u;
#pragma unroll
    for (int c = 0; c < 5; c++) Dq[c] = Db[(size_t)25 * (pb + (int)i) + 5 * q + c];
    load_blocks(rc, Uf, blk0, q, U);
  };
  OutStage g;
  uint2 rc = rl[0], rcN = rl[REC_SLOTS];
  out_cells(g, outCell, o0, nOut);
  TF Uc[4][5], Un[4][5], Dc[5], Dn[5];
  bool bad = false;
  fetch(rc, Uc, Dc);
  out_values(g, z, nC, nOut);
  chunk_load(S.val, z, nC, c0);
  out_store(S.out, g, nOut);
  __syncthreads();
  for (int st = 0; st < ns; st++) {
    const uint2 rcNN = rl[(size_t)(st + 2) * REC_SLOTS];
    fetch(rcN, Un, Dn);
    if (st > 0 && (__builtin_amdgcn_readfirstlane(rc.y) & REC_NEW)) __syncthreads();  // previous level in LDS
    const unsigned i = rc.y & 255u;
    const bool full = !(rc.y & REC_EMPTY);  // slot < SWEEP_ROWS
    if (full) sT[5 * slot + q] = S.val[5 * i + q] - nb_dot(rc, Uc, S.val, S.out);
    __syncthreads();  // the row's 5 components of y_p - sum U w
    if (full) {
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < 5; c++) acc += (double)Dc[c] * sT[5 * slot + c];
      S.val[5 * i + q] = acc;
      bad |= !isfinite(acc);
    }
    __syncthreads();  // sT is reused by the next step
    rc = rcN;
    rcN = rcNN;
#pragma unroll
    for (int c = 0; c < 5; c++) Dc[c] = Dn[c];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 5; c++) Uc[j][c] = Un[j][c];
  }
  chunk_store(z, S.val, nC, c0);
  if (bad) atomicOr(nf, 1u);
}

// the current apply's vectors for the recorded sweeps
__global__ void ilu_io_kernel(const double **io, const double *r, double *z) {
  if (threadIdx.x == 0) {
    io[0] = r;
    io[1] = z;
  }
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
                                                         const double *__restrict__ r, double *__restrict__ z,
                                                         unsigned *__restrict__ nf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nC) return;
  double w[5];
  for (int p = 0; p < 5; p++) w[p] = r[(size_t)p * nC + i];
  const TF *M = Di + (size_t)25 * i;
  bool bad = false;
  for (int a = 0; a < 5; a++) {
    double s = 0.0;
    for (int p = 0; p < 5; p++) s += (double)M[5 * a + p] * w[p];
    z[(size_t)a * nC + i] = s;
    bad |= !isfinite(s);
  }
  if (bad) atomicOr(nf, 1u);
}

int build_ilu_into(fos_euler *ctx, BlockIlu *b);

// The ILU structure is attached to the context only once it is complete, so a failed build leaves
// no half-built state behind for the next call.
int build_ilu(fos_euler *ctx) {
  if (ctx->maxNF > NB) {
    set_error("psetup: cells with %d faces; the block ILU supports tetrahedral meshes (%d)", ctx->maxNF, NB);
    return -1;
  }
  BlockIlu *b = new BlockIlu();
  if (build_ilu_into(ctx, b)) {
    delete b;
    return -1;
  }
  ctx->ilu = b;
  return 0;
}

int build_ilu_into(fos_euler *ctx, BlockIlu *b) {
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
  // Elimination order: the ICH-cell chunks are coloured greedily (in chunk
  // order, the smallest colour no face-adjacent chunk has); rows are eliminated colour by colour,
  // chunk by chunk, cells in index order.  Inside a chunk this is the natural order; between chunks
  // "lower" means "of a lower colour", so every chunk of one colour can be swept at once.
  const int nCh = (nC + ICH - 1) / ICH;
  std::vector<int> colour(nCh, -1);
  int nCol = 0;
  for (int ch = 0; ch < nCh; ch++) {
    unsigned long long used = 0;
    const int c1 = std::min(nC, (ch + 1) * ICH);
    for (int c = ch * ICH; c < c1; c++)
      for (int j = 0; j < NB; j++) {
        const int k = nbr[(size_t)NB * c + j];
        if (k >= 0 && k / ICH != ch && colour[k / ICH] >= 0 && colour[k / ICH] < 64)
          used |= 1ull << colour[k / ICH];
      }
    int col = 0;
    while (col < 64 && ((used >> col) & 1ull)) col++;
    if (col == 64) {
      set_error("psetup: chunk adjacency needs more than 64 colours");
      return -1;
    }
    colour[ch] = col;
    nCol = std::max(nCol, col + 1);
  }
  std::vector<int> colPtr(nCol + 1, 0), colChunks(nCh), chunkStart(nCh), order(nC), rank(nC);
  for (int ch = 0; ch < nCh; ch++) colPtr[colour[ch] + 1]++;
  for (int c = 0; c < nCol; c++) colPtr[c + 1] += colPtr[c];
  {
    std::vector<int> f(colPtr.begin(), colPtr.end() - 1);
    for (int ch = 0; ch < nCh; ch++) colChunks[f[colour[ch]]++] = ch;
    int st = 0;
    for (int q = 0; q < nCh; q++) {
      const int ch = colChunks[q];
      chunkStart[ch] = st;
      for (int c = ch * ICH; c < std::min(nC, (ch + 1) * ICH); c++) {
        rank[c] = st;
        order[st++] = c;
      }
    }
  }
  // factorisation levels over the elimination DAG; lower neighbours in elimination order
  std::vector<int> lev(nC, 0);
  int nl = 0;
  // chunks of one colour share no face, so their rows' levels depend only on lower colours and on
  // earlier rows of the same chunk: host threads over each colour's chunks
  for (int col = 0; col < nCol; col++) {
    const int q0 = colPtr[col], nq = colPtr[col + 1] - q0;
    const int lparts = host_parts((size_t)nq, 512);
    std::vector<int> partMax(lparts, 0);
    host_parallel((size_t)nq, lparts, [&](size_t a0, size_t a1, int pi) {
      for (int q = q0 + (int)a0; q < q0 + (int)a1; q++) {
        const int ch = colChunks[q];
        for (int c = ch * ICH; c < std::min(nC, (ch + 1) * ICH); c++) {
          const int st = rank[c];
          int m = 0, nlo = 0;
          int loR[NB], loJ[NB];  // lower neighbours by rank (insertion sort of at most NB)
          for (int j = 0; j < NB; j++) {
            const int k = nbr[(size_t)NB * c + j];
            if (k >= 0 && rank[k] < st) {
              m = std::max(m, lev[k] + 1);
              int a = nlo++;
              for (; a > 0 && loR[a - 1] > rank[k]; a--) { loR[a] = loR[a - 1]; loJ[a] = loJ[a - 1]; }
              loR[a] = rank[k];
              loJ[a] = j;
            }
          }
          for (int a = 0; a < nlo; a++) low[(size_t)NB * c + a] = loJ[a];
          lev[c] = m;
          partMax[pi] = std::max(partMax[pi], m + 1);
        }
      }
    });
    for (int v : partMax) nl = std::max(nl, v);
  }
  std::vector<int> rows(nC);
  b->levPtr.assign(nl + 1, 0);
  for (int c = 0; c < nC; c++) b->levPtr[lev[c] + 1]++;
  for (int l = 0; l < nl; l++) b->levPtr[l + 1] += b->levPtr[l];
  {
    std::vector<int> f(b->levPtr.begin(), b->levPtr.end() - 1);
    for (int st = 0; st < nC; st++) rows[f[lev[order[st]]]++] = order[st];
  }
  b->nLev = nl;
  b->nCol = nCol;
  b->colPtr = colPtr;
  b->hOrder = order;
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
  // the loops below are independent per row (per chunk for the schedules): host threads over
  // contiguous ranges, combined in range order, so the arrays equal a sequential build's
  const int parts = host_parts((size_t)nC, 1 << 16);
  host_parallel((size_t)nC, parts, [&](size_t c0_, size_t c1_, int) {
    for (int c = (int)c0_; c < (int)c1_; c++)
      for (int q = 0; q < NS; q++) {
        int code = slots[(size_t)NS * c + q];
        if (code < 0) continue;
        int2 p = lr[cb[c / CHUNK] + (code >> 1)];
        int other = (code & 1) ? p.x : p.y;
        if (other < 0) continue;
        for (int j = 0; j < mf; j++)
          if (ctx->hNeib[(size_t)mf * c + j] == other) { jslot[(size_t)NS * c + q] = j; break; }
      }
  });
  // neighbours as steps and the packed block sources, in elimination order
  std::vector<int> lptr(nC + 1), uptr(nC + 1), lsrc, usrc;
  std::vector<int4> lnb(nC), unb(nC);
  // per row in elimination order: its lower (upper) neighbours' positions and, from ptr[st], the
  // sources of its blocks; counted, prefix-summed, then filled
  auto pack = [&](bool lower, std::vector<int4> &nbp, std::vector<int> &ptr, std::vector<int> &src) {
    host_parallel((size_t)nC, parts, [&](size_t s0_, size_t s1_, int) {
      for (int st = (int)s0_; st < (int)s1_; st++) {
        const int i = order[st];
        int e[NB] = {-1, -1, -1, -1}, m = 0;
        for (int j = 0; j < NB; j++) {
          const int k = nbr[(size_t)NB * i + j];
          if (k < 0 || (lower ? rank[k] >= st : rank[k] <= st)) continue;
          e[m++] = rank[k];
        }
        nbp[st] = make_int4(e[0], e[1], e[2], e[3]);
        ptr[st + 1] = m;
      }
    });
    ptr[0] = 0;
    for (int st = 0; st < nC; st++) ptr[st + 1] += ptr[st];
    src.resize(ptr[nC]);
    host_parallel((size_t)nC, parts, [&](size_t s0_, size_t s1_, int) {
      for (int st = (int)s0_; st < (int)s1_; st++) {
        const int i = order[st];
        int m = ptr[st];
        for (int j = 0; j < NB; j++) {
          const int k = nbr[(size_t)NB * i + j];
          if (k < 0 || (lower ? rank[k] >= st : rank[k] <= st)) continue;
          src[m++] = NB * i + j;
        }
      }
    });
  };
  pack(true, lnb, lptr, lsrc);
  pack(false, unb, uptr, usrc);
  // sweep steps per chunk: its rows by local level (forward over the lower neighbours inside the
  // chunk, backward over the upper ones), SWEEP_ROWS rows per step
  std::vector<int> fStepOff(nCh + 1), bStepOff(nCh + 1);
  std::vector<uint2> fRec, bRec;
  std::vector<int> fOutOff(nCh + 1), fOutCell, bOutOff(nCh + 1), bOutCell, fChunkBlk(nCh), bChunkBlk(nCh);
  for (int ch = 0; ch < nCh; ch++) {
    fChunkBlk[ch] = lptr[chunkStart[ch]];
    bChunkBlk[ch] = uptr[chunkStart[ch]];
  }
  struct Sched {  // one range of chunks' steps, records and outside lists
    std::vector<int> outPos, stepRel, outRel;
    std::vector<uint2> rec;
    int nStep = 0, maxOut = 0;
  };
  // packed block sources in step order (each chunk keeps its range; rows in the order of its steps)
  std::vector<int> lsrcStep(lsrc.size()), usrcStep(usrc.size());
  auto schedule = [&](bool fwd, std::vector<int> &stepOff, std::vector<uint2> &rec, std::vector<int> &outOff,
                      std::vector<int> &outPos, int &maxOut) {
    const int cparts = host_parts((size_t)nCh, 1 << 10);
    std::vector<Sched> part(cparts);
    host_parallel((size_t)nCh, cparts, [&](size_t k0_, size_t k1_, int pi) {
      Sched &S = part[pi];
      std::vector<int> outs, loc(ICH), byLev(ICH), pend(ICH);
      const size_t cap = (size_t)(k1_ - k0_) * (ICH + SWEEP_ROWS * 4);  // rows plus padding of a few levels
      S.rec.reserve(cap);
      for (int ch = (int)k0_; ch < (int)k1_; ch++) {
        const int c0 = ch * ICH, n = std::min(ICH, nC - c0);
        // the chunk's rows as a DAG over its in-chunk neighbours on the swept side (forward: lower
        // index; backward: higher); height = the longest chain of rows waiting on a row
        auto onSide = [&](int c, int k) { return k >= c0 && k < c0 + n && (fwd ? k < c : k > c); };
        for (int t = n - 1; t >= 0; t--) {  // dependants before the rows they wait on
          const int i = fwd ? t : n - 1 - t, c = c0 + i;
          int h = 0;
          for (int j = 0; j < NB; j++) {
            const int k = nbr[(size_t)NB * c + j];
            if (k >= c0 && k < c0 + n && onSide(k, c)) h = std::max(h, loc[k - c0] + 1);
          }
          loc[i] = h;  // height
        }
        for (int i = 0; i < n; i++) {
          int d = 0;
          for (int j = 0; j < NB; j++) d += onSide(c0 + i, nbr[(size_t)NB * (c0 + i) + j]) ? 1 : 0;
          pend[i] = d;  // unfinished rows it waits on
        }
        S.stepRel.push_back(S.nStep);
        // the chunk's neighbours outside it, and the records' code for each neighbour
        const int pb = rank[c0];
        outs.clear();
        for (int i = 0; i < n; i++) {
          const int4 nb = fwd ? lnb[rank[c0 + i]] : unb[rank[c0 + i]];
          for (int k : {nb.x, nb.y, nb.z, nb.w})
            if (k >= 0 && (k < pb || k >= pb + n)) outs.push_back(k);
        }
        std::sort(outs.begin(), outs.end());
        outs.erase(std::unique(outs.begin(), outs.end()), outs.end());
        S.outRel.push_back((int)S.outPos.size());
        for (int k : outs) S.outPos.push_back(order[k]);  // the kernels read outside values by cell
        S.maxOut = std::max(S.maxOut, (int)outs.size());
        auto code = [&](int k) -> unsigned {
          if (k < 0) return REC_NB_NONE;
          if (k >= pb && k < pb + n) return (unsigned)(k - pb);
          return (unsigned)(REC_OUT0 + (std::lower_bound(outs.begin(), outs.end(), k) - outs.begin()));
        };
        const int *bptr = fwd ? lptr.data() : uptr.data();
        const std::vector<int> &srcOld = fwd ? lsrc : usrc;
        std::vector<int> &srcNew = fwd ? lsrcStep : usrcStep;
        int acc = 0;  // blocks of the chunk placed so far, in step order
        // list schedule: each step takes up to SWEEP_ROWS rows whose swept-side neighbours in the
        // chunk are done, greatest height first (index order among equals); every step after the
        // first waits for the previous one's rows in LDS
        const int s0 = S.nStep;
        const uint2 empty = make_uint2(~0u, 0xF0000u | REC_EMPTY);  // row 0, no neighbours
        int left = n;
        while (left > 0) {
          int nr = 0;
          for (int i = 0; i < n; i++)
            if (pend[i] == 0) byLev[nr++] = i;
          std::sort(byLev.begin(), byLev.begin() + nr, [&](int a, int b) { return loc[a] != loc[b] ? loc[a] > loc[b] : a < b; });
          nr = std::min(nr, SWEEP_ROWS);
          const unsigned fresh = S.nStep > s0 ? REC_NEW : 0u;  // every step after the first
          S.nStep++;
          const size_t r0 = S.rec.size();
          const int acc0 = acc;
          for (int a = 0; a < nr; a++) {
            const int i = byLev[a];
            pend[i] = -1;  // scheduled
            const int pos = rank[c0 + i];
            const int4 nb = fwd ? lnb[pos] : unb[pos];
            uint2 rc = make_uint2(0u, (unsigned)i | (unsigned)acc << 8 | fresh);
            for (int m = bptr[pos]; m < bptr[pos + 1]; m++) srcNew[bptr[pb] + acc++] = srcOld[m];
            const int nbs[4] = {nb.x, nb.y, nb.z, nb.w};
            for (int j = 0; j < 4; j++) {
              const unsigned k = code(nbs[j]);
              rc.x |= (k & 255u) << (8 * j);
              rc.y |= (k >> 8) << (16 + j);
            }
            S.rec.push_back(rc);
          }
          for (int a = nr; a < REC_SLOTS; a++) S.rec.push_back(empty);  // pad the step
          for (size_t m = r0; m < S.rec.size(); m++) S.rec[m].y |= (unsigned)(acc - acc0) << REC_BLK_SHIFT;
          for (int a = 0; a < nr; a++) {  // the rows waiting on this step's rows move up
            const int c = c0 + byLev[a];
            for (int j = 0; j < NB; j++) {
              const int k = nbr[(size_t)NB * c + j];
              if (k >= c0 && k < c0 + n && onSide(k, c)) pend[k - c0]--;
            }
          }
          left -= nr;
        }
        for (int a = 0; a < REC_PAD_STEPS * REC_SLOTS; a++) S.rec.push_back(empty);
        S.nStep += REC_PAD_STEPS;
      }
    });
    // the ranges in order
    size_t nStep = 0, nOutAll = 0;
    for (const Sched &S : part) {
      nStep += S.nStep;
      nOutAll += S.outPos.size();
    }
    rec.clear();
    outPos.clear();
    rec.reserve(nStep * REC_SLOTS);
    outPos.reserve(nOutAll);
    maxOut = 0;
    int ch = 0, sb = 0;
    for (const Sched &S : part) {
      const int ob = (int)outPos.size();
      for (size_t q = 0; q < S.stepRel.size(); q++, ch++) {
        stepOff[ch] = sb + S.stepRel[q];
        outOff[ch] = ob + S.outRel[q];
      }
      sb += S.nStep;
      rec.insert(rec.end(), S.rec.begin(), S.rec.end());
      outPos.insert(outPos.end(), S.outPos.begin(), S.outPos.end());
      maxOut = std::max(maxOut, S.maxOut);
    }
    stepOff[nCh] = sb;
    outOff[nCh] = (int)outPos.size();
    outPos.push_back(0);  // padding: the kernels' clamped index loads stay inside the list
  };
  schedule(true, fStepOff, fRec, fOutOff, fOutCell, b->fMaxOut);
  schedule(false, bStepOff, bRec, bOutOff, bOutCell, b->bMaxOut);
  lsrc.swap(lsrcStep);
  usrc.swap(usrcStep);
  // launch order inside each colour: chunks with the most sweep steps first, so the short ones fill
  // the end of each launch (the chunks of one colour are independent; results do not depend on it)
  for (int col = 0; col < nCol; col++)
    std::stable_sort(colChunks.begin() + colPtr[col], colChunks.begin() + colPtr[col + 1], [&](int a, int c) {
      return fStepOff[a + 1] - fStepOff[a] + bStepOff[a + 1] - bStepOff[a] >
             fStepOff[c + 1] - fStepOff[c] + bStepOff[c + 1] - bStepOff[c];
    });
  // per launch slot the chunk's metadata, forward and backward
  std::vector<int4> fMeta(2 * (size_t)nCh), bMeta(2 * (size_t)nCh);
  for (int p = 0; p < nCh; p++) {
    const int ch = colChunks[p];
    fMeta[2 * p] = make_int4(ch * ICH, fStepOff[ch], fStepOff[ch + 1] - fStepOff[ch] - REC_PAD_STEPS, fOutOff[ch]);
    fMeta[2 * p + 1] = make_int4(fOutOff[ch + 1] - fOutOff[ch], fChunkBlk[ch], chunkStart[ch], 0);
    bMeta[2 * p] = make_int4(ch * ICH, bStepOff[ch], bStepOff[ch + 1] - bStepOff[ch] - REC_PAD_STEPS, bOutOff[ch]);
    bMeta[2 * p + 1] = make_int4(bOutOff[ch + 1] - bOutOff[ch], bChunkBlk[ch], chunkStart[ch], 0);
  }
  b->nLBlk = (long long)lsrc.size();
  b->nUBlk = (long long)usrc.size();
  if (b->dNbr.upload(nbr.data(), nbr.size(), s) || b->dLowOrd.upload(low.data(), low.size(), s) ||
      b->dRank.upload(rank.data(), rank.size(), s) || b->dOrder.upload(order.data(), order.size(), s) ||
      b->dLevRows.upload(rows.data(), rows.size(), s) || b->dJSlot.upload(jslot.data(), jslot.size(), s) ||
      b->dLSrc.upload(lsrc.data(), lsrc.size(), s) || b->dUSrc.upload(usrc.data(), usrc.size(), s) ||
      b->dFMeta.upload(fMeta.data(), fMeta.size(), s) || b->dBMeta.upload(bMeta.data(), bMeta.size(), s) ||
      b->dFRec.upload(fRec.data(), fRec.size(), s) || b->dBRec.upload(bRec.data(), bRec.size(), s) ||
      b->dFOutCell.upload(fOutCell.data(), fOutCell.size(), s) || b->dBOutCell.upload(bOutCell.data(), bOutCell.size(), s) ||
      b->dJac.alloc(125 * (size_t)nC) || b->dD.alloc(25 * (size_t)nC) || b->dOff.alloc(25 * (size_t)NB * nC) ||
      b->dD32.alloc(25 * (size_t)nC) || b->dIo.alloc(2) ||
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
  const bool exact = ctx->flags & FOS_JAC_EXACT;
  if (NS == 4)
    hipLaunchKernelGGL(exact ? jac_exact_kernel<4> : jac_kernel<4>, dim3((unsigned)(((size_t)nC * 4 + 255) / 256)), blk,
                       0, ctx->stream, ctx->L64, ctx->ilu->dJSlot.p, y, jac);
  else
    hipLaunchKernelGGL(exact ? jac_exact_kernel<8> : jac_kernel<8>, dim3((unsigned)(((size_t)nC * 8 + 255) / 256)), blk,
                       0, ctx->stream, ctx->L64, ctx->ilu->dJSlot.p, y, jac);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

// Level-ordered copies of the factors in the apply precision TF.
template <class TF>
static int gather_factors(fos_euler *ctx, DBuf<TF> &Lf, DBuf<TF> &Uf, DBuf<TF> &Db) {
  BlockIlu *b = ctx->ilu;
  const int nC = ctx->nC;
  // one spare block each: the FP32 sweeps' 16-byte aligned pieces may reach past the last run
  if ((!Lf.p && Lf.alloc(25 * (size_t)(b->nLBlk + 1))) || (!Uf.p && Uf.alloc(25 * (size_t)(b->nUBlk + 1))) ||
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
  go(nC, b->dOrder.p, b->dD.p, Db.p);
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
    return launch(factor_level_kernel, dim3((n + 127) / 128), dim3(128), (size_t)0, b->dLevRows.p + b->levPtr[l], n,
                  b->dNbr.p, b->dLowOrd.p, b->dRank.p, b->dD.p, b->dOff.p, ctx->dStatus.p);
  };
  if (!b->eager && !b->factorChain.ready()) {
    int bad = 0;
    for (int l = 0; l < b->nLev && !bad; l++)
      bad = level(l, [&](auto k, dim3 g, dim3 blk, size_t lds, auto... a) {
        return b->factorChain.add(k, g, blk, lds, a...);
      });
    if (bad || b->factorChain.instantiate()) {
      set_error("psetup: graph of the factorisation levels could not be built (%s)", b->factorChain.fail());
      return -1;
    }
  }
  if (b->eager) {
    for (int l = 0; l < b->nLev; l++)
      level(l, [&](auto k, dim3 g, dim3 blk, size_t lds, auto... a) {
        hipLaunchKernelGGL(k, g, blk, lds, ctx->stream, a...);
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

// Both sweeps, from b->dIo[0] into b->dIo[1]: the forward sweep one launch per colour in ascending
// order, the backward sweep one per colour in descending order; launch(kernel, grid, block, args...)
// either records or launches each.
template <class TF, class Launch>
static int ilu_levels(BlockIlu *b, const TF *Lf, const TF *Uf, const TF *Db, unsigned *nf, Launch &&launch) {
  const dim3 blk(64);
  const int nC = b->nC;
  const size_t fl = sweep_lds<TF>(b->fMaxOut), bl = sweep_lds<TF>(b->bMaxOut);
  for (int c = 0; c < b->nCol; c++) {
    const int n = b->colPtr[c + 1] - b->colPtr[c];
    if (n && launch(fwd_colour_kernel<TF>, dim3(n), blk, fl, nC, b->dFMeta.p + 2 * (size_t)b->colPtr[c],
                    b->dFRec.p, b->dFOutCell.p, Lf, (const double *const *)b->dIo.p))
      return -1;
  }
  for (int c = b->nCol - 1; c >= 0; c--) {
    const int n = b->colPtr[c + 1] - b->colPtr[c];
    if (n && launch(bwd_colour_kernel<TF>, dim3(n), blk, bl, nC, b->dBMeta.p + 2 * (size_t)b->colPtr[c],
                    b->dBRec.p, b->dBOutCell.p, Uf, Db, (const double *const *)b->dIo.p, nf))
      return -1;
  }
  return 0;
}

template <class TF>
static int ilu_sweeps(fos_euler *ctx, const TF *Lf, const TF *Uf, const TF *Db, const double *r, double *z) {
  BlockIlu *b = ctx->ilu;
  if (!b->eager && !b->solveChain.ready()) {
    if (ilu_levels(b, Lf, Uf, Db, ctx->dNonFinite.p,
                   [&](auto k, dim3 g, dim3 bl, size_t lds, auto... a) {
                     return b->solveChain.add(k, g, bl, lds, a...);
                   }) ||
        b->solveChain.instantiate()) {
      set_error("psolve: graph of the sweeps could not be built (%s)", b->solveChain.fail());
      return -1;
    }
  }
  hipLaunchKernelGGL(ilu_io_kernel, dim3(1), dim3(64), 0, ctx->stream, b->dIo.p, r, z);
  if (b->eager) {
    ilu_levels(b, Lf, Uf, Db, ctx->dNonFinite.p, [&](auto k, dim3 g, dim3 bl, size_t lds, auto... a) {
      hipLaunchKernelGGL(k, g, bl, lds, ctx->stream, a...);
      return 0;
    });
  } else if (b->solveChain.launch(ctx->stream)) {
    set_error("psolve: graph launch failed (%s)", hipGetErrorString(b->solveChain.err));
    return -1;
  }
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
    hipLaunchKernelGGL(bjac_apply_kernel<float>, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, b->dD32.p, r, z,
                       ctx->dNonFinite.p);
  else
    hipLaunchKernelGGL(bjac_apply_kernel<double>, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, b->dD.p, r, z,
                       ctx->dNonFinite.p);
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
    if (ilu_solve(ctx, ctx->hR.p, ctx->hZ.p) || stage_out(ctx, ctx->hZ, z)) return -1;
    return finish_call(ctx, "psolve");
  }
  if (ilu_solve(ctx, r, z)) return -1;
  return finish_call(ctx, "psolve");
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
  if (bjac_solve(ctx, r, z)) return -1;
  return finish_call(ctx, "psolve_bjac");
}

int fos_euler_ilu_order(fos_euler *ctx, int *order) {
  if (!ctx || !order) return -1;
  if (!ctx->ilu && build_ilu(ctx)) return -1;
  std::copy(ctx->ilu->hOrder.begin(), ctx->ilu->hOrder.end(), order);
  return 0;
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
