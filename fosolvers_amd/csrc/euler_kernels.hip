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
//   B. threads sweep the chunk's face list (every face touching the chunk, stored contiguously),
//      take local states from LDS and halo states from HBM, compute the Roe flux once per face;
//   C. each thread divides its cell's flux sum by v and writes the SoA result (coalesced).
// A face whose two cells lie in different chunks is evaluated by both chunks (each keeps its own
// side).  Two kernels share this plan:
//  * residual_kernel (generic layout, any polyhedral mesh): the fluxes are parked in LDS and each
//    cell sums its faces in a fixed slot order, so the result is bitwise reproducible; the ghost
//    states of boundary faces are formed in the face loop.
//  * tet_kernel (tetrahedral layout, the default): the ghost states of the boundary pairs are formed
//    by a small kernel ahead of the face pass (ghost_kernel), which keeps the boundary-condition
//    code and its registers out of the face loop, and each face adds its flux to its chunk-local
//    cells' sums in LDS with atomic adds (as the reference's OpenMP atomics do, Euler.f90:127-151),
//    which needs 5 values per cell instead of 5 per face.  The sums are FP64 for both arithmetic
//    types: gfx950's FP32 LDS atomic add (ds_add_f32) ran the FP32 Jv kernels 2.5x slower than
//    ds_add_f64 does.  The order of the adds varies from run to run, so results agree to rounding
//    (~1e-16 relative per cell sum), not bitwise.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "euler_kernels.h"
#include "flux.h"

// Minimum waves per SIMD requested for the generic residual kernel (1: the compiler's choice).
#ifndef FOS_RESIDUAL_MIN_WAVES
#define FOS_RESIDUAL_MIN_WAVES 1
#endif
// ... for the tetrahedral-layout kernel in its residual and difference-quotient modes: with the ghost
// states formed ahead of it and the fluxes summed in LDS, 96 registers and 5 workgroups per CU fit
#ifndef FOS_TET_MIN_WAVES
#define FOS_TET_MIN_WAVES 5
#endif
// ... and for its analytic-Jv mode, whose dual arithmetic needs ~160 registers (FP64) at the
// compiler's choice
#ifndef FOS_JVP_MIN_WAVES
#define FOS_JVP_MIN_WAVES 1
#endif

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

// modes of the residual kernel: the residual, the difference-quotient Jv (perturbed state only,
// or base and perturbed), the analytic Jv (state and direction as dual numbers)
enum { RES = 0, DQ_FY = 1, DQ_DIFF = 2, JVP = 3 };

// Staged cell values: 8 arrays of `stride` entries (SoA, so lanes reading one field of different
// cells spread over the LDS banks)
template <class TC>
__device__ inline void store_prim(TC *sQ, int stride, int lc, const Prim<TC> &q) {
  sQ[0 * stride + lc] = q.rho;
  sQ[1 * stride + lc] = q.rux;
  sQ[2 * stride + lc] = q.ruy;
  sQ[3 * stride + lc] = q.ruz;
  sQ[4 * stride + lc] = q.rE;
  sQ[5 * stride + lc] = q.p;
  sQ[6 * stride + lc] = q.sr;
  sQ[7 * stride + lc] = q.ir;
}

template <class TC>
__device__ inline void read_prim(const TC *sQ, int stride, int lc, Prim<TC> &q) {
  q.rho = sQ[0 * stride + lc];
  q.rux = sQ[1 * stride + lc];
  q.ruy = sQ[2 * stride + lc];
  q.ruz = sQ[3 * stride + lc];
  q.rE = sQ[4 * stride + lc];
  q.p = sQ[5 * stride + lc];
  q.sr = sQ[6 * stride + lc];
  q.ir = sQ[7 * stride + lc];
}

// staged values per cell: base (+ perturbed, or + tangent)
constexpr int nq_of(int mode) { return mode == DQ_DIFF || mode == JVP ? 16 : 8; }

// cell c as dual primitives: the state y and the direction v (FP64 loads, TC arithmetic)
template <class TS, class TC>
__device__ inline void load_cell_dual(const TS *__restrict__ y, const TS *__restrict__ v, int nC, int c, TC g1,
                                      Prim<Dual<TC>> &q) {
  Dual<TC> u[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const size_t i = (size_t)k * nC + c;
    u[k] = Dual<TC>((TC)y[i], (TC)v[i]);
  }
  make_prim(u[0], u[1], u[2], u[3], u[4], Dual<TC>(g1), q);
}

// dual primitives split over two staged sets of 8 rows: values, then tangents
template <class TC>
__device__ inline void store_prim_dual(TC *sQ, int stride, int lc, const Prim<Dual<TC>> &q) {
  store_prim(sQ, stride, lc, Prim<TC>{q.rho.v, q.rux.v, q.ruy.v, q.ruz.v, q.rE.v, q.p.v, q.sr.v, q.ir.v});
  store_prim(sQ + 8 * stride, stride, lc, Prim<TC>{q.rho.d, q.rux.d, q.ruy.d, q.ruz.d, q.rE.d, q.p.d, q.sr.d, q.ir.d});
}
template <class TC>
__device__ inline void read_prim_dual(const TC *sQ, int stride, int lc, Prim<Dual<TC>> &q) {
  Prim<TC> a, b;
  read_prim(sQ, stride, lc, a);
  read_prim(sQ + 8 * stride, stride, lc, b);
  q.rho = {a.rho, b.rho}; q.rux = {a.rux, b.rux}; q.ruy = {a.ruy, b.ruy}; q.ruz = {a.ruz, b.ruz};
  q.rE = {a.rE, b.rE}; q.p = {a.p, b.p}; q.sr = {a.sr, b.sr}; q.ir = {a.ir, b.ir};
}

// The analytic Jv's face term: the directional derivative of the Roe flux across a face whose
// left state A is given, right state B from neighbour or ghost (ghost = -1 - boundary pair index,
// its derivative through setBC included).
template <class TC, class TP>
__device__ inline void jvp_face(const Prim<Dual<TC>> &A, const Prim<Dual<TC>> *Bn, int bnd, const int *bndType,
                                const TP *bndPar, TC gamma, TC r, TC g1, TC nx, TC ny, TC nz, TC area, TC F[5]) {
  using DT = Dual<TC>;
  Prim<DT> B;
  if (Bn) {
    B = *Bn;
  } else {
    DT gr, gx, gy, gz, gE;
    ghost_state<DT>(bndType[bnd], bndPar + 5 * (size_t)bnd, DT(nx), DT(ny), DT(nz), A, DT(gamma), DT(r), gr, gx, gy,
                    gz, gE);
    make_prim(gr, gx, gy, gz, gE, DT(g1), B);
  }
  DT Fd[5];
  roe_flux<DT>(A, B, DT(nx), DT(ny), DT(nz), DT(area), DT(g1), Fd);
#pragma unroll
  for (int k = 0; k < 5; k++) F[k] = Fd[k].d;
}

// Cell c's staged values for a residual mode: RES the state, DQ_FY the perturbed state, DQ_DIFF
// both (the perturbed set NQ/2 rows further on).
template <class TS, class TC, int MODE>
__device__ inline void stage_cell(const TS *__restrict__ y, const TS *__restrict__ v, TS sigma, int nC, int c,
                                  TC g1, TC *sQ, int stride, int lc) {
  Prim<TC> q;
  if (MODE == JVP) {
    Prim<Dual<TC>> qd;
    load_cell_dual<TS, TC>(y, v, nC, c, g1, qd);
    store_prim_dual(sQ, stride, lc, qd);
  } else if (MODE == RES) {
    load_cell<TS, TC, false>(y, v, sigma, nC, c, g1, q);
    store_prim(sQ, stride, lc, q);
  } else if (MODE == DQ_FY) {
    load_cell<TS, TC, true>(y, v, sigma, nC, c, g1, q);
    store_prim(sQ, stride, lc, q);
  } else {
    load_cell<TS, TC, false>(y, v, sigma, nC, c, g1, q);
    store_prim(sQ, stride, lc, q);
    load_cell<TS, TC, true>(y, v, sigma, nC, c, g1, q);
    store_prim(sQ + 8 * stride, stride, lc, q);
  }
}

// Final per-cell combine: the residual, or the Jv difference quotient; with `newton` set (Jv modes)
// the Newton matrix product v - ngam J v instead of J v.  A non-finite result sets *nf (when given):
// the callbacks report it as a recoverable failure.
template <class TS, class TC, int MODE, bool NT = false>
__device__ inline void write_cell(const TC acc[5], TC iv, TS sigma, const TS *__restrict__ fy, TS *__restrict__ out,
                                  int nC, int c, unsigned *__restrict__ nf, const TS *__restrict__ v, int newton,
                                  TS ngam) {
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const size_t i = (size_t)k * nC + c;
    TS o;
    if (MODE == DQ_FY) {
      const TS isig = sigma != TS(0) ? TS(1) / sigma : TS(0);  // sigma = 0: v = 0, J v = 0
      o = sigma != TS(0) ? ((TS)(acc[k] * iv) - fy[i]) * isig : TS(0);
    } else if (MODE == DQ_DIFF) {
      const TS isig = sigma != TS(0) ? TS(1) / sigma : TS(0);
      o = (TS)(acc[k] * iv) * isig;
    } else {
      o = (TS)(acc[k] * iv);
    }
    if (MODE != RES && newton) o = v[i] - ngam * o;
    if (NT) __builtin_nontemporal_store(o, &out[i]);  // written once, read by a later launch
    else out[i] = o;
    bad |= !isfinite(o);
  }
  if (bad && nf) atomicOr(nf, 1u);
}

template <class TS, class TC, int MODE, int NS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(FOS_RESIDUAL_MIN_WAVES))) void residual_kernel(FluxLayout<TC> L, const TS *__restrict__ y,
                                                           const TS *__restrict__ v, TS sigmaVal,
                                                           const double *__restrict__ sigmaPtr,
                                                           const TS *__restrict__ fy, TS *__restrict__ out,
                                                           unsigned *__restrict__ nf, int newton, TS ngam) {
  constexpr bool JV = MODE != RES;
  constexpr int NQ = nq_of(MODE);  // staged values per cell: base (+ perturbed or tangent)
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
  // A: stage the chunk's cells (DQ_FY: perturbed state only; DQ_DIFF: base and perturbed); the
  // cell's face slots and 1/v are loaded here too, ahead of their use in C
  short slv[NS];
  TC iv = TC(0);
  if (t < ncell) {
#pragma unroll
    for (int j = 0; j < NS; j++) slv[j] = L.slots[(size_t)NS * (c0 + t) + j];
    iv = L.invV[c0 + t];
  }
  // the first face record of B is loaded before the cells are staged, so both latencies overlap
  const int fb = L.chunkBeg[chunk], fe = L.chunkBeg[chunk + 1];
  int2 lrN = make_int2(0, 0);
  AreaNormal<TC> gmN{};
  if (fb + t < fe) {
    lrN = L.lr[fb + t];
    gmN = L.geo[fb + t];
  }
  if (t < ncell) stage_cell<TS, TC, MODE>(y, v, sigma, nC, c0 + t, g1, sQ, CHUNK, t);
  __syncthreads();
  // B: one Roe flux (or flux difference) per chunk face; the next face's record is loaded while
  // this one is computed
  for (int f = fb + t; f < fe; f += THREADS) {
    const int2 lr = lrN;
    TC nx, ny, nz, area;
    face_geom(gmN.x, gmN.y, gmN.z, nx, ny, nz, area);
    if (f + THREADS < fe) {
      lrN = L.lr[f + THREADS];
      gmN = L.geo[f + THREADS];
    }
    TC F[5];
#pragma unroll
    for (int k = 0; k < 5; k++) F[k] = TC(0);
    if constexpr (MODE == JVP) {
      Prim<Dual<TC>> A, B;
      auto fetch = [&](int c, Prim<Dual<TC>> &q) {
        const int lc = c - c0;
        if ((unsigned)lc < (unsigned)ncell) read_prim_dual(sQ, CHUNK, lc, q);
        else load_cell_dual<TS, TC>(y, v, nC, c, g1, q);
      };
      fetch(lr.x, A);
      if (lr.y >= 0) fetch(lr.y, B);
      jvp_face(A, lr.y >= 0 ? &B : nullptr, -lr.y - 1, L.bndType, L.bndPar, L.gamma, L.r, g1, nx, ny, nz, area, F);
    } else
#pragma unroll
    for (int pass = 0; pass < (MODE == DQ_DIFF ? 2 : 1); pass++) {
      // pass 0: the staged state of RES / DQ_FY, or the base state of DQ_DIFF; pass 1: perturbed
      const bool pert = (MODE == DQ_FY) || (MODE == DQ_DIFF && pass == 1);
      const TC *sq = sQ + (MODE == DQ_DIFF && pass == 1 ? 8 * CHUNK : 0);
      Prim<TC> A, B;
      auto fetch = [&](int c, Prim<TC> &q) {
        const int lc = c - c0;
        if ((unsigned)lc < (unsigned)ncell) read_prim(sq, CHUNK, lc, q);
        else if (pert) load_cell<TS, TC, true>(y, v, sigma, nC, c, g1, q);
        else load_cell<TS, TC, false>(y, v, sigma, nC, c, g1, q);
      };
      fetch(lr.x, A);
      if (lr.y >= 0) {
        fetch(lr.y, B);
      } else {
        const int b = -lr.y - 1;
        TC gr, gx, gy, gz, gE;
        ghost_state<TC>(L.bndType[b], L.bndPar + 5 * (size_t)b, nx, ny, nz, A, L.gamma, L.r, gr, gx, gy, gz, gE);
        make_prim(gr, gx, gy, gz, gE, g1, B);
      }
      TC Fp[5];
      roe_flux<TC>(A, B, nx, ny, nz, area, g1, Fp);
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
#pragma unroll
    for (int j = 0; j < NS; j++) {
      const int s = slv[j];
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
    write_cell<TS, TC, MODE>(acc, iv, sigma, fy, out, nC, c, nf, v, newton, ngam);
  }
}

template <class TS, class TC, int MODE>
int launch(const FluxLayout<TC> &L, const TS *y, const TS *v, TS sigma, const double *sigmaPtr, const TS *fy,
           TS *out, unsigned *nf, hipStream_t s, int newton = 0, TS ngam = TS(0)) {
  if (L.nC == 0) return 0;
  const size_t smem = sizeof(TC) * (nq_of(MODE) * (size_t)CHUNK + 5 * (size_t)L.maxChunkFaces);
  if (smem > 65536) {
    set_error("residual: chunk face list (%d faces) exceeds LDS", L.maxChunkFaces);
    return -1;
  }
  dim3 grid(L.nChunks), block(THREADS);
  if (L.nSlots <= 4)
    hipLaunchKernelGGL((residual_kernel<TS, TC, MODE, 4>), grid, block, smem, s, L, y, v, sigma, sigmaPtr, fy, out,
                       nf, newton, ngam);
  else if (L.nSlots <= 8)
    hipLaunchKernelGGL((residual_kernel<TS, TC, MODE, 8>), grid, block, smem, s, L, y, v, sigma, sigmaPtr, fy, out,
                       nf, newton, ngam);
  else {
    set_error("residual: %d face slots per cell not supported", L.nSlots);
    return -1;
  }
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- tetrahedral layout
//
// The generic kernel's three phases with 8 bytes per face instead of 32: the face record names the
// cells (chunk-local, or global for a halo cell, or the boundary pair) and the face's three nodes;
// phase A stages the chunk's node coordinates next to its cells and zeroes the cells' flux sums, and
// phase B forms the triangle's area and unit normal from them (c = (p1 - p0) x (p2 - p0),
// area = |c|/2, n = +-c/|c|: geometry.f90:17-49 up to rounding; one refined rsqrt per face), reads
// a boundary face's ghost primitives from ghost_kernel's output, and adds the face flux into the
// chunk-local cells' sums.  Both chunks of a chunk-crossing face evaluate it with the same operand
// order and the same FP64 coordinates, so its two copies agree bitwise.
template <class TS, class TC, int MODE, int NS, bool NT>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(
    MODE == 3 ? FOS_JVP_MIN_WAVES : FOS_TET_MIN_WAVES))) void tet_kernel(
    TetLayout<TC> L, const TS *__restrict__ y, const TS *__restrict__ v, TS sigmaVal,
    const double *__restrict__ sigmaPtr, const TS *__restrict__ fy, TS *__restrict__ out, unsigned *__restrict__ nf,
    int newton, TS ngam) {
  constexpr bool JV = MODE != RES;
  constexpr int NQ = nq_of(MODE);
  const TS sigma = JV && sigmaPtr ? (TS)*sigmaPtr : sigmaVal;
  extern __shared__ unsigned char smem_raw[];
  const int mn = L.maxNodes;
  TC *sQ = reinterpret_cast<TC *>(smem_raw);                 // [NQ][CHUNK]
  double *sA = reinterpret_cast<double *>(sQ + NQ * CHUNK);  // [5][CHUNK] per-cell flux sums (FP64)
  TC *sN = reinterpret_cast<TC *>(sA + 5 * CHUNK);           // [3][mn] node coordinates
  const int chunk = xcd_remap(blockIdx.x, gridDim.x);
  const int nC = L.nC, c0 = chunk * CHUNK, ncell = min(CHUNK, nC - c0);
  const TC g1 = L.gamma - TC(1);
  const int t = threadIdx.x;
  // A: the first face record and the nodes are loaded before the cells are staged
#pragma unroll
  for (int k = 0; k < 5; k++) sA[k * CHUNK + t] = 0.0;
  const int fb = L.chunkBeg[chunk], fe = L.chunkBeg[chunk + 1];
  uint2 recN = make_uint2(0u, 0u);
  // NT (working sets beyond the 256 MB last-level cache): the face records, node coordinates and
  // 1/v, read once per launch, and the output, written once, go through non-temporal loads and
  // stores, so they do not evict the cell states that neighbouring chunks read again as halo (6M
  // tets: residual 240 -> 226 us).  A working set that fits the cache is faster without (1.5M tets:
  // 59 -> 63 us with).
  auto ld = [](const auto *p) { return NT ? __builtin_nontemporal_load(p) : *p; };
  auto ldrec = [&](int f) {
    const unsigned long long r = ld(reinterpret_cast<const unsigned long long *>(L.face) + f);
    return make_uint2((unsigned)r, (unsigned)(r >> 32));
  };
  if (fb + t < fe) recN = ldrec(fb + t);
  const int nb = L.nodeBeg[chunk], nn = L.nodeBeg[chunk + 1] - nb;
  if (t < nn) {  // nn <= TET_MAX_NODES = THREADS
    const TC *p = L.node + 3 * (size_t)(nb + t);
    sN[t] = ld(p);
    sN[mn + t] = ld(p + 1);
    sN[2 * mn + t] = ld(p + 2);
  }
  if (t < ncell) stage_cell<TS, TC, MODE>(y, v, sigma, nC, c0 + t, g1, sQ, CHUNK, t);
  __syncthreads();
  // B: one Roe flux (or flux difference) per chunk face
  const TC *sy = sN + mn, *sz = sN + 2 * mn;
  for (int f = fb + t; f < fe; f += THREADS) {
    const uint2 rec = recN;
    if (f + THREADS < fe) recN = ldrec(f + THREADS);
    const int payload = (int)(rec.x & (TET_MAX_PAYLOAD - 1)), kind = (rec.x >> 29) & 3u;
    const bool flip = rec.x >> 31;
    const int lc = rec.y & 255u, q0 = (rec.y >> 8) & 255u, q1 = (rec.y >> 16) & 255u, q2 = rec.y >> 24;
    const TC ax = sN[q1] - sN[q0], ay = sy[q1] - sy[q0], az = sz[q1] - sz[q0];
    const TC bx = sN[q2] - sN[q0], by = sy[q2] - sy[q0], bz = sz[q2] - sz[q0];
    const TC cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    const TC c2 = cx * cx + cy * cy + cz * cz;
    const TC yc = frsqrt(c2);  // 1/|c|
    const TC inv = flip ? -yc : yc;
    const TC nx = cx * inv, ny = cy * inv, nz = cz * inv, area = TC(0.5) * (c2 * yc);
    TC F[5];
#pragma unroll
    for (int k = 0; k < 5; k++) F[k] = TC(0);
    if constexpr (MODE == JVP) {
      Prim<Dual<TC>> A, B;
      if (kind == 1) load_cell_dual<TS, TC>(y, v, nC, payload, g1, A);
      else read_prim_dual(sQ, CHUNK, lc, A);
      if (kind == 0) read_prim_dual(sQ, CHUNK, payload, B);
      else if (kind == 1) read_prim_dual(sQ, CHUNK, lc, B);
      else if (kind == 2) load_cell_dual<TS, TC>(y, v, nC, payload, g1, B);
      else read_prim_dual(L.ghost, L.nPb, payload, B);
      Dual<TC> Fd[5];
      roe_flux<Dual<TC>>(A, B, Dual<TC>(nx), Dual<TC>(ny), Dual<TC>(nz), Dual<TC>(area), Dual<TC>(g1), Fd);
#pragma unroll
      for (int k = 0; k < 5; k++) F[k] = Fd[k].d;
    } else
#pragma unroll
    for (int pass = 0; pass < (MODE == DQ_DIFF ? 2 : 1); pass++) {
      const bool pert = (MODE == DQ_FY) || (MODE == DQ_DIFF && pass == 1);
      const TC *sq = sQ + (MODE == DQ_DIFF && pass == 1 ? 8 * CHUNK : 0);
      Prim<TC> A, B;
      auto halo = [&](int c, Prim<TC> &q) {
        if (pert) load_cell<TS, TC, true>(y, v, sigma, nC, c, g1, q);
        else load_cell<TS, TC, false>(y, v, sigma, nC, c, g1, q);
      };
      if (kind == 1) halo(payload, A);
      else read_prim(sq, CHUNK, lc, A);
      if (kind == 0) {
        read_prim(sq, CHUNK, payload, B);
      } else if (kind == 1) {
        read_prim(sq, CHUNK, lc, B);
      } else if (kind == 2) {
        halo(payload, B);
      } else {
        read_prim(L.ghost + (MODE == DQ_DIFF && pass == 1 ? 8 * (size_t)L.nPb : 0), L.nPb, payload, B);
      }
      TC Fp[5];
      roe_flux<TC>(A, B, nx, ny, nz, area, g1, Fp);
#pragma unroll
      for (int k = 0; k < 5; k++) F[k] = (MODE == DQ_DIFF && pass == 0) ? -Fp[k] : F[k] + Fp[k];
    }
    // the flux leaves L and enters R: each chunk-local side accumulates it (LDS atomics)
    if (kind == 1) {
#pragma unroll
      for (int k = 0; k < 5; k++) atomicAdd(&sA[k * CHUNK + lc], (double)F[k]);
    } else {
#pragma unroll
      for (int k = 0; k < 5; k++) atomicAdd(&sA[k * CHUNK + lc], -(double)F[k]);
      if (kind == 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) atomicAdd(&sA[k * CHUNK + payload], (double)F[k]);
      }
    }
  }
  __syncthreads();
  // C: the cell's sum, divided by v
  if (t < ncell) {
    TC acc[5];
#pragma unroll
    for (int k = 0; k < 5; k++) acc[k] = (TC)sA[k * CHUNK + t];
    write_cell<TS, TC, MODE, NT>(acc, ld(&L.invV[c0 + t]), sigma, fy, out, nC, c0 + t, nf, v, newton, ngam);
  }
}

// Ghost states of the boundary faces (setBC, foeuler.f90:209-330) for a tet-layout launch: per ghost
// slot (chunk order) the cell's values as the face pass forms them (state, perturbed state, or dual
// state), the ghost state from the boundary condition of the slot's pair with the pair's normal, and
// its primitives, into L.ghost (the same rows as the staged cells, stride nPb).
template <class TS, class TC, int MODE>
__global__ __launch_bounds__(256) void ghost_kernel(TetLayout<TC> L, const TS *__restrict__ y,
                                                    const TS *__restrict__ v, TS sigmaVal,
                                                    const double *__restrict__ sigmaPtr) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;  // ghost slot
  if (g >= L.nPb) return;
  const TS sigma = MODE != RES && sigmaPtr ? (TS)*sigmaPtr : sigmaVal;
  const int c = L.bndCell[g], b = L.bndPair[g], nC = L.nC, P = L.nPb;
  const TC g1 = L.gamma - TC(1);
  const TC nx = L.bndNorm[g], ny = L.bndNorm[P + g], nz = L.bndNorm[2 * P + g];
  if constexpr (MODE == JVP) {
    using DT = Dual<TC>;
    Prim<DT> A, B;
    load_cell_dual<TS, TC>(y, v, nC, c, g1, A);
    DT gr, gx, gy, gz, gE;
    ghost_state<DT>(L.bndType[b], L.bndPar + 5 * (size_t)b, DT(nx), DT(ny), DT(nz), A, DT(L.gamma), DT(L.r), gr, gx,
                    gy, gz, gE);
    make_prim(gr, gx, gy, gz, gE, DT(g1), B);
    store_prim_dual(L.ghost, P, g, B);
  } else {
#pragma unroll
    for (int pass = 0; pass < (MODE == DQ_DIFF ? 2 : 1); pass++) {
      const bool pert = (MODE == DQ_FY) || (MODE == DQ_DIFF && pass == 1);
      Prim<TC> A, B;
      if (pert) load_cell<TS, TC, true>(y, v, sigma, nC, c, g1, A);
      else load_cell<TS, TC, false>(y, v, sigma, nC, c, g1, A);
      TC gr, gx, gy, gz, gE;
      ghost_state<TC>(L.bndType[b], L.bndPar + 5 * (size_t)b, nx, ny, nz, A, L.gamma, L.r, gr, gx, gy, gz, gE);
      make_prim(gr, gx, gy, gz, gE, g1, B);
      store_prim(L.ghost + (size_t)pass * 8 * P, P, g, B);
    }
  }
}

constexpr size_t TET_LDS_LIMIT = 65536;

template <class TC>
size_t tet_lds(const TetLayout<TC> &L, int mode) {
  const size_t nq = nq_of(mode);
  const size_t b = sizeof(double) * 5 * (size_t)CHUNK + sizeof(TC) * (nq * (size_t)CHUNK + 3 * (size_t)L.maxNodes);
  return b <= TET_LDS_LIMIT ? b : 0;
}

template <class TS, class TC, int MODE>
int launch_tet(const TetLayout<TC> &L, const TS *y, const TS *v, TS sigma, const double *sigmaPtr, const TS *fy,
               TS *out, unsigned *nf, hipStream_t s, int newton = 0, TS ngam = TS(0)) {
  if (L.nC == 0) return 0;
  const size_t smem = tet_lds(L, MODE);
  if (!smem) {
    set_error("residual: tetrahedral chunk data exceeds LDS");
    return -1;
  }
  if (L.nPb > 0)
    hipLaunchKernelGGL((ghost_kernel<TS, TC, MODE>), dim3((L.nPb + 255) / 256), dim3(256), 0, s, L, y, v, sigma,
                       sigmaPtr);
  if (L.nSlots <= 4 && L.streaming)
    hipLaunchKernelGGL((tet_kernel<TS, TC, MODE, 4, true>), dim3(L.nChunks), dim3(THREADS), smem, s, L, y, v, sigma,
                       sigmaPtr, fy, out, nf, newton, ngam);
  else if (L.nSlots <= 4)
    hipLaunchKernelGGL((tet_kernel<TS, TC, MODE, 4, false>), dim3(L.nChunks), dim3(THREADS), smem, s, L, y, v, sigma,
                       sigmaPtr, fy, out, nf, newton, ngam);
  else {
    set_error("residual: tetrahedral layout with %d face slots per cell", L.nSlots);
    return -1;
  }
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

template <class TS, class TC>
int launch_residual(const FluxLayout<TC> &L, const TS *y, TS *ydot, hipStream_t s, unsigned *nf) {
  return launch<TS, TC, RES>(L, y, nullptr, TS(0), nullptr, nullptr, ydot, nf, s);
}

template <class TS, class TC>
int launch_dq_jv(const FluxLayout<TC> &L, const TS *y, const TS *v, const TS *fy, double sigma,
                 const double *sigmaDev, TS *Jv, hipStream_t s, unsigned *nf, int newton, double ngam) {
  return launch<TS, TC, DQ_FY>(L, y, v, (TS)sigma, sigmaDev, fy, Jv, nf, s, newton, (TS)ngam);
}

template <class TS, class TC>
int launch_dq_jv_diff(const FluxLayout<TC> &L, const TS *y, const TS *v, double sigma, const double *sigmaDev, TS *Jv,
                      hipStream_t s, unsigned *nf, int newton, double ngam) {
  return launch<TS, TC, DQ_DIFF>(L, y, v, (TS)sigma, sigmaDev, nullptr, Jv, nf, s, newton, (TS)ngam);
}

template <class TS, class TC>
int launch_jvp(const FluxLayout<TC> &L, const TS *y, const TS *v, TS *Jv, hipStream_t s, unsigned *nf, int newton,
               double ngam) {
  return launch<TS, TC, JVP>(L, y, v, TS(0), nullptr, nullptr, Jv, nf, s, newton, (TS)ngam);
}

template int launch_jvp<double, double>(const FluxLayout<double> &, const double *, const double *, double *,
                                        hipStream_t, unsigned *, int, double);
template int launch_jvp<double, float>(const FluxLayout<float> &, const double *, const double *, double *,
                                       hipStream_t, unsigned *, int, double);
template int launch_residual<double, double>(const FluxLayout<double> &, const double *, double *, hipStream_t, unsigned *);
template int launch_residual<float, float>(const FluxLayout<float> &, const float *, float *, hipStream_t, unsigned *);
template int launch_residual<double, float>(const FluxLayout<float> &, const double *, double *, hipStream_t, unsigned *);
template int launch_dq_jv<double, double>(const FluxLayout<double> &, const double *, const double *, const double *,
                                          double, const double *, double *, hipStream_t, unsigned *, int, double);
template int launch_dq_jv<float, float>(const FluxLayout<float> &, const float *, const float *, const float *,
                                        double, const double *, float *, hipStream_t, unsigned *, int, double);
template int launch_dq_jv<double, float>(const FluxLayout<float> &, const double *, const double *, const double *,
                                         double, const double *, double *, hipStream_t, unsigned *, int, double);
template int launch_dq_jv_diff<double, float>(const FluxLayout<float> &, const double *, const double *, double,
                                              const double *, double *, hipStream_t, unsigned *, int, double);
template int launch_dq_jv_diff<double, double>(const FluxLayout<double> &, const double *, const double *, double,
                                               const double *, double *, hipStream_t, unsigned *, int, double);

template <class TC>
size_t tet_lds_bytes(const TetLayout<TC> &L, int mode) {
  return tet_lds(L, mode);
}
template size_t tet_lds_bytes<double>(const TetLayout<double> &, int);
template size_t tet_lds_bytes<float>(const TetLayout<float> &, int);

template <class TS, class TC>
int launch_tet_residual(const TetLayout<TC> &L, const TS *y, TS *ydot, hipStream_t s, unsigned *nf) {
  return launch_tet<TS, TC, RES>(L, y, nullptr, TS(0), nullptr, nullptr, ydot, nf, s);
}
template <class TS, class TC>
int launch_tet_dq_jv(const TetLayout<TC> &L, const TS *y, const TS *v, const TS *fy, double sigma,
                     const double *sigmaDev, TS *Jv, hipStream_t s, unsigned *nf, int newton, double ngam) {
  return launch_tet<TS, TC, DQ_FY>(L, y, v, (TS)sigma, sigmaDev, fy, Jv, nf, s, newton, (TS)ngam);
}
template <class TS, class TC>
int launch_tet_dq_jv_diff(const TetLayout<TC> &L, const TS *y, const TS *v, double sigma, const double *sigmaDev,
                          TS *Jv, hipStream_t s, unsigned *nf, int newton, double ngam) {
  return launch_tet<TS, TC, DQ_DIFF>(L, y, v, (TS)sigma, sigmaDev, nullptr, Jv, nf, s, newton, (TS)ngam);
}
template <class TS, class TC>
int launch_tet_jvp(const TetLayout<TC> &L, const TS *y, const TS *v, TS *Jv, hipStream_t s, unsigned *nf, int newton,
                   double ngam) {
  return launch_tet<TS, TC, JVP>(L, y, v, TS(0), nullptr, nullptr, Jv, nf, s, newton, (TS)ngam);
}
template int launch_tet_jvp<double, double>(const TetLayout<double> &, const double *, const double *, double *,
                                            hipStream_t, unsigned *, int, double);
template int launch_tet_jvp<double, float>(const TetLayout<float> &, const double *, const double *, double *,
                                           hipStream_t, unsigned *, int, double);
template int launch_tet_residual<double, double>(const TetLayout<double> &, const double *, double *, hipStream_t, unsigned *);
template int launch_tet_residual<float, float>(const TetLayout<float> &, const float *, float *, hipStream_t, unsigned *);
template int launch_tet_dq_jv<double, double>(const TetLayout<double> &, const double *, const double *,
                                              const double *, double, const double *, double *, hipStream_t, unsigned *, int, double);
template int launch_tet_dq_jv<float, float>(const TetLayout<float> &, const float *, const float *, const float *,
                                            double, const double *, float *, hipStream_t, unsigned *, int, double);
template int launch_tet_dq_jv_diff<double, float>(const TetLayout<float> &, const double *, const double *, double,
                                                  const double *, double *, hipStream_t, unsigned *, int, double);

}  // namespace fos
