// First-order Roe residual and fused DQ Jacobian-vector product on gfx950.
//
// Reference: src/libfosolvers/FVM/Euler.f90:25-156 (findEuler) with the ghost states of
// src/foeuler/foeuler.f90:209-330 (setBC) and the division by the cell volume of
// src/foeuler/foeuler.f90:366-372 (rhs).
//
// Design: cells come in chunks of CHUNK consecutive cells (the mesh is renumbered so a chunk
// is spatially compact).  One workgroup owns a chunk:
//   A. each thread loads one cell's conservative state (SoA, coalesced), computes p, sqrt(rho)
//      and 1/rho once, and stages them in LDS;
//   B. threads sweep the chunk's face list (every face touching the chunk, stored contiguously:
//      L/R indices + (n, area)), take local states from LDS and halo/ghost states from HBM or
//      the BC formulas, compute the Roe flux once per face and park it in LDS;
//   C. each thread sums its cell's face fluxes from LDS in a fixed slot order, divides by v
//      and writes the SoA result (coalesced).
// No atomics: the result is bitwise reproducible run to run.  A face whose two cells lie in
// different chunks is evaluated by both chunks (each keeps its own side).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "euler_kernels.h"
#include "flux.h"

namespace fos {
namespace {
template <class TS, class TC, bool PERTURB>
__device__ inline void load_cell(const TS *__restrict__ y, const TS *__restrict__ v, TS sigma, int nC, int c,
                                 TC g1, Prim<TC> &q) {
  TC u[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    size_t i = (size_t)k * nC + c;
    if (PERTURB) u[k] = (TC)(y[i] + sigma * v[i]);
    else u[k] = (TC)y[i];
  }
  make_prim(u[0], u[1], u[2], u[3], u[4], g1, q);
}

// modes of the residual kernel
enum { RES = 0, DQ_FY = 1, DQ_DIFF = 2 };

template <class TC>
__device__ inline void store_prim(TC *sQ, int lc, const Prim<TC> &q) {
  sQ[0 * CHUNK + lc] = q.rho;
  sQ[1 * CHUNK + lc] = q.rux;
  sQ[2 * CHUNK + lc] = q.ruy;
  sQ[3 * CHUNK + lc] = q.ruz;
  sQ[4 * CHUNK + lc] = q.rE;
  sQ[5 * CHUNK + lc] = q.p;
  sQ[6 * CHUNK + lc] = q.sr;
  sQ[7 * CHUNK + lc] = q.ir;
}

template <class TC>
__device__ inline void read_prim(const TC *sQ, int lc, Prim<TC> &q) {
  q.rho = sQ[0 * CHUNK + lc];
  q.rux = sQ[1 * CHUNK + lc];
  q.ruy = sQ[2 * CHUNK + lc];
  q.ruz = sQ[3 * CHUNK + lc];
  q.rE = sQ[4 * CHUNK + lc];
  q.p = sQ[5 * CHUNK + lc];
  q.sr = sQ[6 * CHUNK + lc];
  q.ir = sQ[7 * CHUNK + lc];
}

template <class TS, class TC, int MODE, int NS>
__global__ __launch_bounds__(THREADS) void residual_kernel(FluxLayout<TC> L, const TS *__restrict__ y,
                                                           const TS *__restrict__ v, TS sigmaVal,
                                                           const double *__restrict__ sigmaPtr,
                                                           const TS *__restrict__ fy, TS *__restrict__ out) {
  constexpr bool JV = MODE != RES;
  constexpr int NQ = MODE == DQ_DIFF ? 16 : 8;  // staged values per cell: base (+ perturbed)
  const TS sigma = JV && sigmaPtr ? (TS)*sigmaPtr : sigmaVal;
  extern __shared__ unsigned char smem_raw[];
  TC *sQ = reinterpret_cast<TC *>(smem_raw);  // [NQ][CHUNK]
  TC *sF = sQ + NQ * CHUNK;                   // [5][maxChunkFaces]
  const int mcf = L.maxChunkFaces;
  const int chunk = xcd_remap(blockIdx.x, gridDim.x);
  const int nC = L.nC;
  const int c0 = chunk * CHUNK;
  const int ncell = min(CHUNK, nC - c0);
  const TC g1 = L.gamma - TC(1);
  const int t = threadIdx.x;
  // A: stage the chunk's cells (DQ_FY: perturbed state only; DQ_DIFF: base and perturbed)
  if (t < ncell) {
    Prim<TC> q;
    if (MODE == RES) {
      load_cell<TS, TC, false>(y, v, sigma, nC, c0 + t, g1, q);
      store_prim(sQ, t, q);
    } else if (MODE == DQ_FY) {
      load_cell<TS, TC, true>(y, v, sigma, nC, c0 + t, g1, q);
      store_prim(sQ, t, q);
    } else {
      load_cell<TS, TC, false>(y, v, sigma, nC, c0 + t, g1, q);
      store_prim(sQ, t, q);
      load_cell<TS, TC, true>(y, v, sigma, nC, c0 + t, g1, q);
      store_prim(sQ + 8 * CHUNK, t, q);
    }
  }
  __syncthreads();
  // B: one Roe flux (or flux difference) per chunk face
  const int fb = L.chunkBeg[chunk], fe = L.chunkBeg[chunk + 1];
  for (int f = fb + t; f < fe; f += THREADS) {
    const int2 lr = L.lr[f];
    const auto gm = L.geo[f];
    TC F[5];
#pragma unroll
    for (int k = 0; k < 5; k++) F[k] = TC(0);
#pragma unroll
    for (int pass = 0; pass < (MODE == DQ_DIFF ? 2 : 1); pass++) {
      // pass 0: the staged state of RES / DQ_FY, or the base state of DQ_DIFF; pass 1: perturbed
      const bool pert = (MODE == DQ_FY) || (MODE == DQ_DIFF && pass == 1);
      const TC *sq = sQ + (MODE == DQ_DIFF && pass == 1 ? 8 * CHUNK : 0);
      Prim<TC> A, B;
      auto fetch = [&](int c, Prim<TC> &q) {
        const int lc = c - c0;
        if ((unsigned)lc < (unsigned)ncell) read_prim(sq, lc, q);
        else if (pert) load_cell<TS, TC, true>(y, v, sigma, nC, c, g1, q);
        else load_cell<TS, TC, false>(y, v, sigma, nC, c, g1, q);
      };
      fetch(lr.x, A);
      if (lr.y >= 0) {
        fetch(lr.y, B);
      } else {
        const int b = -lr.y - 1;
        TC gr, gx, gy, gz, gE;
        ghost_state<TC>(L.bndType[b], L.bndPar + 5 * (size_t)b, gm.x, gm.y, gm.z, A, L.gamma, L.r, gr, gx, gy, gz,
                        gE);
        make_prim(gr, gx, gy, gz, gE, g1, B);
      }
      TC Fp[5];
      roe_flux<TC>(A, B, gm.x, gm.y, gm.z, gm.w, g1, Fp);
#pragma unroll
      for (int k = 0; k < 5; k++) F[k] = (MODE == DQ_DIFF && pass == 0) ? -Fp[k] : F[k] + Fp[k];
    }
    const int lf = f - fb;
#pragma unroll
    for (int k = 0; k < 5; k++) sF[k * mcf + lf] = F[k];
  }
  __syncthreads();
  // C: gather the cell's faces in slot order, divide by v
  if (t < ncell) {
    const int c = c0 + t;
    TC acc[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
    const short *sl = L.slots + (size_t)NS * c;
#pragma unroll
    for (int j = 0; j < NS; j++) {
      const int s = sl[j];
      if (s >= 0) {
        const int lf = s >> 1;
        if (s & 1) {
#pragma unroll
          for (int k = 0; k < 5; k++) acc[k] += sF[k * mcf + lf];
        } else {
#pragma unroll
          for (int k = 0; k < 5; k++) acc[k] -= sF[k * mcf + lf];
        }
      }
    }
    const TC iv = L.invV[c];
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const size_t i = (size_t)k * nC + c;
      if (MODE == DQ_FY) {
        const TS isig = sigma != TS(0) ? TS(1) / sigma : TS(0);  // sigma = 0: v = 0, J v = 0
        out[i] = sigma != TS(0) ? ((TS)(acc[k] * iv) - fy[i]) * isig : TS(0);
      } else if (MODE == DQ_DIFF) {
        const TS isig = sigma != TS(0) ? TS(1) / sigma : TS(0);
        out[i] = (TS)(acc[k] * iv) * isig;
      } else {
        out[i] = (TS)(acc[k] * iv);
      }
    }
  }
}

template <class TS, class TC, int MODE>
int launch(const FluxLayout<TC> &L, const TS *y, const TS *v, TS sigma, const double *sigmaPtr, const TS *fy,
           TS *out, hipStream_t s) {
  if (L.nC == 0) return 0;
  const size_t smem = sizeof(TC) * ((MODE == DQ_DIFF ? 16 : 8) * (size_t)CHUNK + 5 * (size_t)L.maxChunkFaces);
  if (smem > 65536) {
    set_error("residual: chunk face list (%d faces) exceeds LDS", L.maxChunkFaces);
    return -1;
  }
  dim3 grid(L.nChunks), block(THREADS);
  if (L.nSlots <= 4)
    hipLaunchKernelGGL((residual_kernel<TS, TC, MODE, 4>), grid, block, smem, s, L, y, v, sigma, sigmaPtr, fy, out);
  else if (L.nSlots <= 8)
    hipLaunchKernelGGL((residual_kernel<TS, TC, MODE, 8>), grid, block, smem, s, L, y, v, sigma, sigmaPtr, fy, out);
  else {
    set_error("residual: %d face slots per cell not supported", L.nSlots);
    return -1;
  }
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

template <class TS, class TC>
int launch_residual(const FluxLayout<TC> &L, const TS *y, TS *ydot, hipStream_t s) {
  return launch<TS, TC, RES>(L, y, nullptr, TS(0), nullptr, nullptr, ydot, s);
}

template <class TS, class TC>
int launch_dq_jv(const FluxLayout<TC> &L, const TS *y, const TS *v, const TS *fy, double sigma,
                 const double *sigmaDev, TS *Jv, hipStream_t s) {
  return launch<TS, TC, DQ_FY>(L, y, v, (TS)sigma, sigmaDev, fy, Jv, s);
}

template <class TS, class TC>
int launch_dq_jv_diff(const FluxLayout<TC> &L, const TS *y, const TS *v, double sigma, const double *sigmaDev, TS *Jv,
                      hipStream_t s) {
  return launch<TS, TC, DQ_DIFF>(L, y, v, (TS)sigma, sigmaDev, nullptr, Jv, s);
}

template int launch_residual<double, double>(const FluxLayout<double> &, const double *, double *, hipStream_t);
template int launch_residual<float, float>(const FluxLayout<float> &, const float *, float *, hipStream_t);
template int launch_residual<double, float>(const FluxLayout<float> &, const double *, double *, hipStream_t);
template int launch_dq_jv<double, double>(const FluxLayout<double> &, const double *, const double *, const double *,
                                          double, const double *, double *, hipStream_t);
template int launch_dq_jv<float, float>(const FluxLayout<float> &, const float *, const float *, const float *,
                                        double, const double *, float *, hipStream_t);
template int launch_dq_jv<double, float>(const FluxLayout<float> &, const double *, const double *, const double *,
                                         double, const double *, double *, hipStream_t);
template int launch_dq_jv_diff<double, float>(const FluxLayout<float> &, const double *, const double *, double,
                                              const double *, double *, hipStream_t);
template int launch_dq_jv_diff<double, double>(const FluxLayout<double> &, const double *, const double *, double,
                                               const double *, double *, hipStream_t);

}  // namespace fos
