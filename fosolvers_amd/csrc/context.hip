// fos_euler context: device layout build and the CVODE-style callbacks (rhs, jtimes) plus
// findEuler on element arrays.  Preconditioner entry points live in precond.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "context.h"

namespace fos {
namespace {

// ---- WRMS norm: sigma = 1 / sqrt(mean((v_i / (rtol |y_i| + atol + rtol rms(y)))^2)) ----
// CVODE's weights 1/(rtol|y_i| + atol) plus a floor at rtol times the RMS of the state, so that
// exactly-zero components (zero transverse momentum) do not weight rounding noise by 1/atol.
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;
constexpr double SQRT_EPS32 = 3.4526698300124393e-4;  // sqrt(2^-23)

__global__ __launch_bounds__(RED_THREADS) void sumsq_partial(const double *__restrict__ y, size_t n,
                                                             double *__restrict__ part) {
  __shared__ double sm[RED_THREADS];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (size_t)gridDim.x * RED_THREADS)
    s += y[i] * y[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// floor = atol + rtol sqrt(sum / n)
__global__ __launch_bounds__(RED_THREADS) void floor_final(const double *__restrict__ part, int np, size_t n,
                                                           double rtol, double atol, double *__restrict__ out) {
  __shared__ double sm[RED_THREADS];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += RED_THREADS) s += part[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = atol + rtol * sqrt(sm[0] / (double)n);
}

// max that keeps a NaN (fmax drops it): a NaN direction must not look like a zero one
__device__ inline double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

// Per block: sum of (v_i w_i)^2 and max |v_i w_i|, w_i = 1 / (rtol |y_i| + floor)
__global__ __launch_bounds__(RED_THREADS) void wrms_partial(const double *__restrict__ v, const double *__restrict__ y,
                                                            size_t n, double rtol, const double *__restrict__ floorv,
                                                            double *__restrict__ part) {
  __shared__ double sm[RED_THREADS], sx[RED_THREADS];
  const double fl = floorv[0];
  double s = 0.0, mx = 0.0;
  for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (size_t)gridDim.x * RED_THREADS) {
    const double w = v[i] / (rtol * fabs(y[i]) + fl);
    s += w * w;
    mx = nan_max(mx, fabs(w));
  }
  sm[threadIdx.x] = s;
  sx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sm[threadIdx.x] += sm[threadIdx.x + o];
      sx[threadIdx.x] = nan_max(sx[threadIdx.x], sx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sm[0];
    part[2 * blockIdx.x + 1] = sx[0];
  }
}

// sigma[0] = 1 / ||v||_WRMS (CVODE's difference-quotient increment); sigma[2] = c / max_i |v_i w_i|,
// the increment that keeps every component's perturbation within c rtol-weighted units (the FP32
// flux difference uses c = sqrt(eps32) / rtol).  Both are 0 for v = 0: the Jv kernels then write 0.
__global__ __launch_bounds__(RED_THREADS) void wrms_final(const double *__restrict__ part, int np, size_t n, double c,
                                                          double *__restrict__ sigma) {
  __shared__ double sm[RED_THREADS], sx[RED_THREADS];
  double s = 0.0, mx = 0.0;
  for (int i = threadIdx.x; i < np; i += RED_THREADS) {
    s += part[2 * i];
    mx = nan_max(mx, part[2 * i + 1]);
  }
  sm[threadIdx.x] = s;
  sx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sm[threadIdx.x] += sm[threadIdx.x + o];
      sx[threadIdx.x] = nan_max(sx[threadIdx.x], sx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // v = 0 gives 0 (the product is 0); a non-finite v gives a NaN increment, so the product is
    // non-finite and flagged instead of silently 0
    const double nrm = sqrt(sm[0] / (double)n);
    sigma[0] = nrm == 0.0 ? 0.0 : (isfinite(nrm) ? 1.0 / nrm : NAN);
    sigma[2] = sx[0] == 0.0 ? 0.0 : (isfinite(sx[0]) ? c / sx[0] : NAN);
  }
}

// ---- findEuler on element arrays: per-pair flux, then a per-cell gather in pair order ----
template <class TC>
__device__ inline void elem_prim(const double *rho, const double *rhou, const double *rhoE, const double *p, int e,
                                 Prim<TC> &q) {
  q.rho = rho[e];
  q.rux = rhou[3 * (size_t)e];
  q.ruy = rhou[3 * (size_t)e + 1];
  q.ruz = rhou[3 * (size_t)e + 2];
  q.rE = rhoE[e];
  q.p = p[e];
  q.ir = 1.0 / q.rho;
  q.sr = sqrt(q.rho);
}

__global__ void pair_flux_kernel(int nP, int nE, const int *__restrict__ iEP, const double *__restrict__ aP,
                                 const double *__restrict__ normP, const double *rho, const double *rhou,
                                 const double *rhoE, const double *p, double g1, double *__restrict__ F) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nP) return;
  int m = iEP[2 * i], n = iEP[2 * i + 1];
  double f[5] = {0, 0, 0, 0, 0};
  if (m < nE && n < nE) {  // Euler.f90:89
    Prim<double> A, B;
    elem_prim(rho, rhou, rhoE, p, m, A);
    elem_prim(rho, rhou, rhoE, p, n, B);
    roe_flux<double>(A, B, normP[3 * i], normP[3 * i + 1], normP[3 * i + 2], aP[i], g1, f);
  }
  for (int k = 0; k < 5; k++) F[5 * (size_t)i + k] = f[k];
}

__global__ void pair_gather_kernel(int nC, const int *__restrict__ ptr, const int *__restrict__ idx,
                                   const double *__restrict__ F, double *dRho, double *dRhou, double *dRhoE) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  double a[5] = {0, 0, 0, 0, 0};
  for (int k = ptr[c]; k < ptr[c + 1]; k++) {
    int e = idx[k], i = e >> 1;
    double sg = (e & 1) ? 1.0 : -1.0;
    for (int q = 0; q < 5; q++) a[q] += sg * F[5 * (size_t)i + q];
  }
  dRho[c] = a[0];
  dRhou[3 * (size_t)c] = a[1];
  dRhou[3 * (size_t)c + 1] = a[2];
  dRhou[3 * (size_t)c + 2] = a[3];
  dRhoE[c] = a[4];
}

// element arrays of syncState (foeuler.f90:198-205); any pointer may be null
struct ElemOut {
  double *rho, *rhou, *rhoE, *u, *p, *temp, *c;
};

__device__ inline void put_elem(const ElemOut &o, size_t e, double rho, double rux, double ruy, double ruz,
                                double rE, double gamm, double r) {
  const double p = (gamm - 1.0) * (rE - 0.5 * (rux * rux + ruy * ruy + ruz * ruz) / rho);
  const double T = p / rho / r;
  if (o.rho) o.rho[e] = rho;
  if (o.rhou) { o.rhou[3 * e] = rux; o.rhou[3 * e + 1] = ruy; o.rhou[3 * e + 2] = ruz; }
  if (o.rhoE) o.rhoE[e] = rE;
  if (o.u) { o.u[3 * e] = rux / rho; o.u[3 * e + 1] = ruy / rho; o.u[3 * e + 2] = ruz / rho; }
  if (o.p) o.p[e] = p;
  if (o.temp) o.temp[e] = T;
  if (o.c) o.c[e] = sqrt(gamm * r * T);
}

// cells: the SoA state into element arrays and primitives (syncState, foeuler.f90:189-205)
__global__ void cell_state_kernel(int nC, const double *__restrict__ y, double gamm, double r, ElemOut o) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nC) return;
  const size_t n = nC;
  put_elem(o, i, y[i], y[n + i], y[2 * n + i], y[3 * n + i], y[4 * n + i], gamm, r);
}

// boundary facets: the ghost state of each boundary pair (setBC, foeuler.f90:214-329) into the
// facet element n of the pair, then its primitives
__global__ void facet_state_kernel(int nPi, int nPb, int nC, const int *__restrict__ iEP,
                                   const double *__restrict__ normP, const int *__restrict__ bndType,
                                   const double *__restrict__ bndPar, const double *__restrict__ y, double gamm,
                                   double r, ElemOut o) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nPb) return;
  const int i = nPi + k, m = iEP[2 * i], nf = iEP[2 * i + 1];
  const size_t n = nC;
  Prim<double> q;
  make_prim(y[m], y[n + m], y[2 * n + m], y[3 * n + m], y[4 * n + m], gamm - 1.0, q);
  double gr, gux, guy, guz, gE;
  ghost_state<double>(bndType[k], bndPar + 5 * (size_t)k, normP[3 * i], normP[3 * i + 1], normP[3 * i + 2], q, gamm, r,
                      gr, gux, guy, guz, gE);
  put_elem(o, nf, gr, gux, guy, guz, gE, gamm, r);
}

int check_nonneg(const char *what, long long v) {
  if (v < 0) { set_error("fos_euler_create: negative %s", what); return -1; }
  return 0;
}

}  // namespace

// reduction blocks for n values
static int red_blocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS)); }

// sigma values of v at y into sigmaDev[0..2] with partials in part[2 RED_BLOCKS]; the weight floor
// (sigmaDev[1]) is recomputed unless reuseFloor
static int wrms_sigma_into(fos_euler *ctx, const double *v, const double *y, double *sigmaDev, double *part,
                           bool reuseFloor) {
  const size_t n = 5 * (size_t)ctx->nC;
  const int nb = red_blocks(n);
  if (!reuseFloor) {
    hipLaunchKernelGGL(sumsq_partial, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, y, n, part);
    hipLaunchKernelGGL(floor_final, dim3(1), dim3(RED_THREADS), 0, ctx->stream, part, nb, n, ctx->rtol, ctx->atol,
                       sigmaDev + 1);
  }
  hipLaunchKernelGGL(wrms_partial, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, v, y, n, ctx->rtol, sigmaDev + 1,
                     part);
  hipLaunchKernelGGL(wrms_final, dim3(1), dim3(RED_THREADS), 0, ctx->stream, part, nb, n,
                     SQRT_EPS32 / std::max(ctx->rtol, 1e-300), sigmaDev);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

// ---- fos_euler_jtimes_noalloc's scratch: the WRMS reduction partials (2 RED_BLOCKS doubles) and the
// sigma values (slots 2 RED_BLOCKS .. +3), in a file-scope device array.  Each GPU holds its own copy
// until the process ends; its address on a device comes from hipGetSymbolAddress with that device
// current (the array's name in host code is only a stand-in for the symbol).
constexpr int NOALLOC_SIGMA = 2 * RED_BLOCKS;
constexpr int NOALLOC_DOUBLES = NOALLOC_SIGMA + 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

struct NoallocState {
  double *ptr = nullptr;  // the device's scratch address
  // the last call's weight floor (scratch slot NOALLOC_SIGMA + 1) was formed for this context, state
  // array and rhs epoch
  const fos_euler *ctx = nullptr;
  const double *y = nullptr;
  long long epoch = -1;
};
static std::mutex noallocMu;
static std::vector<NoallocState> noallocDev;

int resolve_noalloc_scratch(int device) {
  std::lock_guard<std::mutex> g(noallocMu);
  if ((int)noallocDev.size() <= device) noallocDev.resize(device + 1);
  if (noallocDev[device].ptr) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error("scratch of fos_euler_jtimes_noalloc: hipGetSymbolAddress failed on device %d: %s", device,
              hipGetErrorString(e));
    return -1;
  }
  noallocDev[device].ptr = static_cast<double *>(p);
  return 0;
}

int jtimes_noalloc_dev(fos_euler *ctx, const double *v, double *Jv, const double *y, const double *fy) {
  const size_t n = 5 * (size_t)ctx->nC;
  if (2 * (size_t)red_blocks(n) > (size_t)NOALLOC_SIGMA) {
    set_error("jtimes_noalloc: %d reduction partials exceed the scratch of %d", 2 * red_blocks(n), NOALLOC_SIGMA);
    return -1;
  }
  double *base, *sig;
  bool reuse;
  {
    std::lock_guard<std::mutex> g(noallocMu);
    if (ctx->device >= (int)noallocDev.size() || !noallocDev[ctx->device].ptr) {
      set_error("jtimes_noalloc: scratch not resolved on device %d", ctx->device);
      return -1;
    }
    NoallocState &st = noallocDev[ctx->device];
    base = st.ptr;
    sig = st.ptr + NOALLOC_SIGMA;
    reuse = st.ctx == ctx && st.y == y && st.epoch == ctx->rhsEpoch;
    st.ctx = ctx;
    st.y = y;
    st.epoch = ctx->rhsEpoch;
  }
  if (wrms_sigma_into(ctx, v, y, sig, base, reuse)) return -1;
  if (ctx->flags & FOS_PREC_FP32) {
    if (ctx->useTet)
      return launch_tet_dq_jv_diff<double, float>(ctx->T32, y, v, 0.0, sig + 2, Jv, ctx->stream, ctx->dNonFinite.p);
    return launch_dq_jv_diff<double, float>(ctx->L32, y, v, 0.0, sig + 2, Jv, ctx->stream, ctx->dNonFinite.p);
  }
  if (ctx->useTet)
    return launch_tet_dq_jv<double, double>(ctx->T64, y, v, fy, 0.0, sig, Jv, ctx->stream, ctx->dNonFinite.p);
  return launch_dq_jv<double, double>(ctx->L64, y, v, fy, 0.0, sig, Jv, ctx->stream, ctx->dNonFinite.p);
}

int wrms_sigma(fos_euler *ctx, const double *v, const double *y, double *sigmaDev) {
  // The floor depends on y only: reuse it while y is the same device array and no rhs call came
  // in between (a Newton iterate is fixed across the GMRES products that use it, and its fy is
  // evaluated whenever it changes).  Host-vector mode restages y each call and never reuses.
  const bool reuse = !(ctx->flags & FOS_VECTORS_ON_HOST) && ctx->floorY == y && ctx->floorEpoch == ctx->rhsEpoch;
  if (wrms_sigma_into(ctx, v, y, sigmaDev, ctx->redPart.p, reuse)) return -1;
  ctx->floorY = y;
  ctx->floorEpoch = ctx->rhsEpoch;
  return 0;
}


// Tetrahedral layout over the same chunk face lists and slots (see TetLayout).  Returns 0 when
// built, 1 when the mesh does not qualify (a face that is not a triangle, node geometry that does
// not reproduce aP/normP, more than TET_MAX_NODES nodes in a chunk, indices beyond the payload
// field, LDS), -1 on an allocation failure.
int build_tet_layout(fos_euler *ctx, const fos_fv_grid *g, const std::vector<int> &cnt, const std::vector<int2> &lr,
                     const std::vector<int> &pairOf, const std::vector<int> &bIndex) {
  const int nC = g->nC, nCh = ctx->nChunks, mxn = g->maxNNE;
  if (!g->pN || !g->iNE || !g->nNE || g->nN <= 0 || mxn < 3 || ctx->L64.nSlots > 4) return 1;
  if (nC >= TET_MAX_PAYLOAD || ctx->nPb >= TET_MAX_PAYLOAD) return 1;
  // the three nodes of pair i, ascending; false if the face is not a triangle
  auto faceNodes = [&](int i, int out[3]) -> bool {
    const int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    if (g->nNE[m] != 4) return false;
    const int *a = g->iNE + (size_t)mxn * m;
    if (n >= nC) {
      if (g->nNE[n] != 3) return false;
      for (int q = 0; q < 3; q++) out[q] = g->iNE[(size_t)mxn * n + q];
    } else {
      if (g->nNE[n] != 4) return false;
      const int *b = g->iNE + (size_t)mxn * n;
      int k = 0;
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++)
          if (a[p] == b[q]) {
            if (k == 3) return false;
            out[k++] = a[p];
          }
      if (k != 3) return false;
    }
    std::sort(out, out + 3);
    return out[0] >= 0 && out[2] < g->nN && out[0] != out[1] && out[1] != out[2];
  };
  // per chunk independent: host threads over contiguous chunk ranges, node lists combined in order
  std::vector<uint2> rec(lr.size());
  std::vector<int> nodeBeg(nCh + 1, 0), nodes;
  const int parts = host_parts((size_t)nCh, 256);
  std::vector<std::vector<int>> partNodes(parts);
  std::vector<int> partMax(parts, 0), partBad(parts, 0);
  host_parallel((size_t)nCh, parts, [&](size_t k0_, size_t k1_, int pi) {
    std::vector<int> nlist;
    std::vector<std::array<int, 3>> tri;
    std::vector<int> &pn = partNodes[pi];
    for (int k = (int)k0_; k < (int)k1_; k++) {
      const int c0 = k * CHUNK, fb = cnt[k], fe = cnt[k + 1];
      nlist.clear();
      tri.assign(fe - fb, {0, 0, 0});
      for (int f = fb; f < fe; f++) {
        int t3[3];
        if (!faceNodes(pairOf[f], t3)) { partBad[pi] = 1; return; }
        tri[f - fb] = {t3[0], t3[1], t3[2]};
        nlist.insert(nlist.end(), t3, t3 + 3);
      }
      std::sort(nlist.begin(), nlist.end());
      nlist.erase(std::unique(nlist.begin(), nlist.end()), nlist.end());
      if ((int)nlist.size() > TET_MAX_NODES) { partBad[pi] = 1; return; }
      partMax[pi] = std::max(partMax[pi], (int)nlist.size());
      nodeBeg[k + 1] = (int)nlist.size();  // count; prefix-summed below
      pn.insert(pn.end(), nlist.begin(), nlist.end());
      auto nodeLocal = [&](int nd) -> unsigned {
        return (unsigned)(std::lower_bound(nlist.begin(), nlist.end(), nd) - nlist.begin());
      };
      auto inChunk = [&](int c) { return c >= c0 && c < c0 + CHUNK; };
      for (int f = fb; f < fe; f++) {
        const int i = pairOf[f];
        const auto &t3 = tri[f - fb];
        // geometry from the nodes, as the kernel forms it, checked against aP / normP
        const double *p0 = g->pN + 3 * (size_t)t3[0], *p1 = g->pN + 3 * (size_t)t3[1], *p2 = g->pN + 3 * (size_t)t3[2];
        const double ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
        const double bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double cn = std::sqrt(cx * cx + cy * cy + cz * cz);
        const double *nP = g->normP + 3 * (size_t)i;
        if (!(cn > 0.0) || std::fabs(0.5 * cn - g->aP[i]) > 1e-9 * g->aP[i]) { partBad[pi] = 1; return; }
        const bool flip = (cx * nP[0] + cy * nP[1] + cz * nP[2]) < 0.0;
        const double sg = flip ? -1.0 : 1.0;
        if (std::fabs(sg * cx / cn - nP[0]) > 1e-9 || std::fabs(sg * cy / cn - nP[1]) > 1e-9 ||
            std::fabs(sg * cz / cn - nP[2]) > 1e-9) {
          partBad[pi] = 1;
          return;
        }
        const int L = lr[f].x, R = lr[f].y;
        unsigned kind, payload, local;
        if (R < 0) {
          kind = 3; payload = (unsigned)(-R - 1); local = (unsigned)(L - c0);
        } else if (inChunk(L) && inChunk(R)) {
          kind = 0; payload = (unsigned)(R - c0); local = (unsigned)(L - c0);
        } else if (inChunk(R)) {
          kind = 1; payload = (unsigned)L; local = (unsigned)(R - c0);
        } else {
          kind = 2; payload = (unsigned)R; local = (unsigned)(L - c0);
        }
        rec[f].x = payload | (kind << 29) | (flip ? 1u << 31 : 0u);
        rec[f].y = local | (nodeLocal(t3[0]) << 8) | (nodeLocal(t3[1]) << 16) | (nodeLocal(t3[2]) << 24);
      }
    }
  });
  int maxNodes = 0;
  for (int pi = 0; pi < parts; pi++) {
    if (partBad[pi]) return 1;
    maxNodes = std::max(maxNodes, partMax[pi]);
  }
  // ghost slots in chunk order (a wave's boundary faces then read nearly contiguous cells and
  // slots); slotPair[slot] = the boundary pair index (pair order) of the slot
  std::vector<int> slotPair;
  slotPair.reserve(ctx->nPb);
  for (size_t f = 0; f < rec.size(); f++)
    if (((rec[f].x >> 29) & 3u) == 3u) {
      slotPair.push_back((int)(rec[f].x & (TET_MAX_PAYLOAD - 1)));
      rec[f].x = (rec[f].x & ~(TET_MAX_PAYLOAD - 1)) | (unsigned)(slotPair.size() - 1);
    }
  if ((int)slotPair.size() != ctx->nPb) return 1;
  for (int k = 0; k < nCh; k++) nodeBeg[k + 1] += nodeBeg[k];
  nodes.reserve(nodeBeg[nCh]);
  for (auto &pn : partNodes) nodes.insert(nodes.end(), pn.begin(), pn.end());
  std::vector<double> n64(3 * nodes.size());
  std::vector<float> n32(3 * nodes.size());
  for (int k = 0; k < nCh; k++) {
    if (nodeBeg[k] == nodeBeg[k + 1]) continue;
    const double *o = g->pN + 3 * (size_t)nodes[nodeBeg[k]];
    for (int q = nodeBeg[k]; q < nodeBeg[k + 1]; q++)
      for (int d = 0; d < 3; d++) {
        const double x = g->pN[3 * (size_t)nodes[q] + d];
        n64[3 * (size_t)q + d] = x;
        n32[3 * (size_t)q + d] = (float)(x - o[d]);  // FP32: relative to the chunk's first node
      }
  }
  // ghost slots: cell and normal (normP) of the slot's pair for the ghost kernel, SoA
  const int nPb = ctx->nPb;
  std::vector<int> pairOfB(nPb, -1);
  for (int i = 0; i < g->nP; i++)
    if (bIndex[i] >= 0) pairOfB[bIndex[i]] = i;
  std::vector<int> bCell(nPb, 0);
  std::vector<double> bN64(3 * (size_t)nPb, 0.0);
  std::vector<float> bN32(3 * (size_t)nPb, 0.f);
  for (int q = 0; q < nPb; q++) {
    const int i = pairOfB[slotPair[q]];
    bCell[q] = g->iEP[2 * i];
    for (int d = 0; d < 3; d++) {
      bN64[(size_t)d * nPb + q] = g->normP[3 * (size_t)i + d];
      bN32[(size_t)d * nPb + q] = (float)g->normP[3 * (size_t)i + d];
    }
  }
  hipStream_t s = ctx->stream;
  if (ctx->dTetFace.upload(rec.data(), rec.size(), s) || ctx->dTetNodeBeg.upload(nodeBeg.data(), nodeBeg.size(), s) ||
      ctx->dTetNode64.upload(n64.data(), n64.size(), s) || ctx->dTetNode32.upload(n32.data(), n32.size(), s) ||
      ctx->dBndCell.upload(bCell.data(), bCell.size(), s) || ctx->dBndSlotPair.upload(slotPair.data(), slotPair.size(), s) || ctx->dBndNorm64.upload(bN64.data(), bN64.size(), s) ||
      ctx->dBndNorm32.upload(bN32.data(), bN32.size(), s) || ctx->dGhost64.alloc(16 * (size_t)nPb) ||
      ctx->dGhost32.alloc(16 * (size_t)nPb))
    return -1;
  // non-temporal face data and output once the two state vectors (80 bytes per cell) pass half of
  // the 256 MB last-level cache, i.e. once a launch's ~170 bytes per cell no longer fit it: measured
  // faster at 6M cells, slower at 1.5M
  const int streaming = (size_t)nC * 80 > ((size_t)128 << 20) ? 1 : 0;
  auto fillT = [&](auto &T, const auto &G, auto *nodeP, auto *bN, auto *gh) {
    T.streaming = streaming;
    T.nPb = nPb;
    T.bndCell = ctx->dBndCell.p;
    T.bndPair = ctx->dBndSlotPair.p;
    T.bndNorm = bN;
    T.ghost = gh;
    T.nC = nC;
    T.nChunks = nCh;
    T.maxChunkFaces = G.maxChunkFaces;
    T.maxNodes = std::max(maxNodes, 1);
    T.nSlots = G.nSlots;
    T.chunkBeg = G.chunkBeg;
    T.face = ctx->dTetFace.p;
    T.slots = G.slots;
    T.nodeBeg = ctx->dTetNodeBeg.p;
    T.node = nodeP;
    T.bndType = G.bndType;
    T.bndPar = G.bndPar;
    T.invV = G.invV;
    T.gamma = G.gamma;
    T.r = G.r;
  };
  fillT(ctx->T64, ctx->L64, ctx->dTetNode64.p, ctx->dBndNorm64.p, ctx->dGhost64.p);
  fillT(ctx->T32, ctx->L32, ctx->dTetNode32.p, ctx->dBndNorm32.p, ctx->dGhost32.p);
  // every mode the context launches must fit one workgroup's LDS
  if (!tet_lds_bytes(ctx->T64, 0) || !tet_lds_bytes(ctx->T64, 1) || !tet_lds_bytes(ctx->T32, 0) ||
      !tet_lds_bytes(ctx->T32, 1) || !tet_lds_bytes(ctx->T32, 2))
    return 1;
  return 0;
}

int build_layout(fos_euler *ctx, const fos_fv_grid *g, const fos_cond_tab *bc) {
  const int nC = g->nC, nP = g->nP, nE = g->nE;
  const int nCh = (nC + CHUNK - 1) / CHUNK;
  ctx->nChunks = nCh;
  // classify pairs, boundary index b in pair order
  std::vector<int> bIndex(nP, -1);
  int nb = 0;
  for (int i = 0; i < nP; i++) {
    int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    if (m < 0 || m >= nC || n < 0 || n >= nE || n == m) {
      set_error("fos_euler_create: pair %d has invalid elements (%d, %d)", i, m, n);
      return -1;
    }
    if (n >= nC) bIndex[i] = nb++;
  }
  ctx->nPb = nb;
  // BC per boundary pair: code and 5 parameters (condTab p(1:5, iBC)) -- foeuler.f90:145-154,214-329
  ctx->bndType.assign(nb, FOS_BC_WALL);
  ctx->bndPar.assign(5 * (size_t)nb, 0.0);
  ctx->bndIsUdf.assign(nb, 0);
  ctx->needUdf = false;
  for (int i = 0; i < nP; i++) {
    int b = bIndex[i];
    if (b < 0) continue;
    int n = g->iEP[2 * i + 1], k = -1;
    if (bc)
      for (int q = 0; q < bc->n; q++)
        if (g->gid[n] == bc->gid[q]) { k = q; break; }
    if (k < 0) continue;  // unassigned group: wall (foeuler.f90:217)
    int t = bc->t[k];
    switch (t) {
      case FOS_BC_WALL: case FOS_BC_IN_STATIC: case FOS_BC_IN_TOTAL: case FOS_BC_OUT: case FOS_BC_FAR:
        break;
      case FOS_BC_IN_STATIC_UDF: case FOS_BC_IN_TOTAL_UDF: case FOS_BC_OUT_UDF: case FOS_BC_FAR_UDF:
        ctx->bndIsUdf[b] = 1;
        ctx->needUdf = true;
        break;
      default:
        t = -1;  // 'case default' of setBC: the ghost slot keeps its last value; not supported here
        set_error("fos_euler_create: boundary condition code %d is not one of foeuler's", bc->t[k]);
        return -1;
    }
    ctx->bndType[b] = t;
    if (!ctx->bndIsUdf[b])
      for (int q = 0; q < std::min(5, bc->maxPar); q++) ctx->bndPar[5 * (size_t)b + q] = bc->p[(size_t)bc->maxPar * k + q];
  }
  // chunk face lists: interior faces, then chunk-crossing faces, then boundary faces
  std::vector<int> cnt(nCh + 1, 0);
  for (int i = 0; i < nP; i++) {
    int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    int km = m / CHUNK;
    cnt[km + 1]++;
    if (n < nC && n / CHUNK != km) cnt[n / CHUNK + 1]++;
  }
  for (int k = 0; k < nCh; k++) cnt[k + 1] += cnt[k];
  const size_t nCF = cnt[nCh];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  std::vector<int2> lr(nCF);
  std::vector<int> pairOf(nCF);
  std::vector<AreaNormal<double>> geo(nCF);
  std::vector<AreaNormal<float>> geo32(nCF);
  std::vector<short> slot8((size_t)8 * nC, (short)-1);  // up to 8 face slots per cell, then trimmed
  std::vector<unsigned char> nslot(nC, 0);
  bool over = false;
  auto addSlot = [&](int c, short code) {
    if (nslot[c] >= 8) { over = true; return; }
    slot8[(size_t)8 * c + nslot[c]++] = code;
  };
  auto put = [&](int k, int i, int L, int R) {
    int f = fill[k]++;
    lr[f] = make_int2(L, R);
    pairOf[f] = i;
    const double a = g->aP[i];
    geo[f] = {a * g->normP[3 * (size_t)i], a * g->normP[3 * (size_t)i + 1], a * g->normP[3 * (size_t)i + 2]};
    geo32[f] = {(float)geo[f].x, (float)geo[f].y, (float)geo[f].z};
    int lf = f - cnt[k];
    if (L / CHUNK == k) addSlot(L, (short)(2 * lf));
    if (R >= 0 && R / CHUNK == k) addSlot(R, (short)(2 * lf + 1));
  };
  for (int pass = 0; pass < 3; pass++)
    for (int i = 0; i < nP; i++) {
      int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
      int km = m / CHUNK;
      if (pass == 0 && n < nC && n / CHUNK == km) put(km, i, m, n);
      if (pass == 1 && n < nC && n / CHUNK != km) { put(km, i, m, n); put(n / CHUNK, i, m, n); }
      if (pass == 2 && n >= nC) put(km, i, m, -(bIndex[i] + 1));
    }
  int maxCF = 0, maxS = 0;
  for (int k = 0; k < nCh; k++) maxCF = std::max(maxCF, cnt[k + 1] - cnt[k]);
  for (int c = 0; c < nC; c++) maxS = std::max<int>(maxS, (int)nslot[c]);
  if (maxCF >= 16384) { set_error("fos_euler_create: %d faces in one chunk", maxCF); return -1; }
  const int nS = maxS <= 4 ? 4 : 8;
  if (over) { set_error("fos_euler_create: a cell has more than 8 faces"); return -1; }
  std::vector<short> slots((size_t)nS * nC, (short)-1);
  for (int c = 0; c < nC; c++)
    for (int j = 0; j < nslot[c]; j++) slots[(size_t)nS * c + j] = slot8[(size_t)8 * c + j];
  std::vector<double> iv(nC);
  std::vector<float> iv32(nC);
  for (int c = 0; c < nC; c++) {
    if (!(g->v[c] > 0.0)) { set_error("fos_euler_create: cell %d has volume %g", c, g->v[c]); return -1; }
    iv[c] = 1.0 / g->v[c];
    iv32[c] = (float)iv[c];
  }
  std::vector<float> bp32(ctx->bndPar.begin(), ctx->bndPar.end());
  hipStream_t s = ctx->stream;
  if (ctx->dChunkBeg.upload(cnt.data(), cnt.size(), s) || ctx->dLR.upload(lr.data(), nCF, s) ||
      ctx->dGeo64.upload(geo.data(), nCF, s) || ctx->dGeo32.upload(geo32.data(), nCF, s) ||
      ctx->dSlots.upload(slots.data(), slots.size(), s) || ctx->dInvV64.upload(iv.data(), nC, s) ||
      ctx->dInvV32.upload(iv32.data(), nC, s) || ctx->dBndType.upload(ctx->bndType.data(), nb, s) ||
      ctx->dBndPar64.upload(ctx->bndPar.data(), 5 * (size_t)nb, s) || ctx->dBndPar32.upload(bp32.data(), bp32.size(), s))
    return -1;
  ctx->nChunkFaces = nCF;
  auto fillL = [&](auto &L, auto *geoP, auto *ivP, auto *bpP, auto gam, auto rr) {
    L.nC = nC;
    L.nChunks = nCh;
    L.maxChunkFaces = std::max(maxCF, 1);
    L.nSlots = nS;
    L.chunkBeg = ctx->dChunkBeg.p;
    L.lr = ctx->dLR.p;
    L.geo = geoP;
    L.slots = ctx->dSlots.p;
    L.invV = ivP;
    L.bndType = ctx->dBndType.p;
    L.bndPar = bpP;
    L.gamma = gam;
    L.r = rr;
  };
  fillL(ctx->L64, ctx->dGeo64.p, ctx->dInvV64.p, ctx->dBndPar64.p, ctx->gamm, ctx->r);
  fillL(ctx->L32, ctx->dGeo32.p, ctx->dInvV32.p, ctx->dBndPar32.p, (float)ctx->gamm, (float)ctx->r);
  ctx->useTet = false;
  if (!(ctx->flags & FOS_LAYOUT_GENERIC)) {
    const int rc = build_tet_layout(ctx, g, cnt, lr, pairOf, bIndex);
    if (rc < 0) return -1;
    ctx->useTet = rc == 0;
  }
  return 0;
}

int ensure_vec(DBuf<double> &b, size_t n) { return b.n >= n ? 0 : b.alloc(n); }

// host-vector mode: stage a host vector on the device
int stage_in(fos_euler *ctx, DBuf<double> &b, const double *h) {
  size_t n = 5 * (size_t)ctx->nC;
  if (ensure_vec(b, n)) return -1;
  FOS_HIP_TRY(hipMemcpyAsync(b.p, h, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  return 0;
}
int stage_out(fos_euler *ctx, const DBuf<double> &b, double *h) {
  size_t n = 5 * (size_t)ctx->nC;
  FOS_HIP_TRY(hipMemcpyAsync(h, b.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

// The rhs is always FP64: it forms the Newton residual G, whose 1e-8 WRMS tolerance sits ~1e-14
// relative to the state.  FOS_PREC_FP32 lowers the Jacobian-vector products and the
// preconditioner apply only.
int residual_dev(fos_euler *ctx, const double *y, double *ydot) {
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("rhs: a *_UDF boundary condition is used but fos_euler_set_udf_bc was not called");
    return -1;
  }
  if (ctx->useTet) return launch_tet_residual<double, double>(ctx->T64, y, ydot, ctx->stream, ctx->dNonFinite.p);
  return launch_residual<double, double>(ctx->L64, y, ydot, ctx->stream, ctx->dNonFinite.p);
}

int jv_dev(fos_euler *ctx, const double *v, double *Jv, const double *y, const double *fy, int newton, double ngam) {
  if (wrms_sigma(ctx, v, y, ctx->dSigma.p)) return -1;
  if (ctx->flags & FOS_PREC_FP32) {
    // FP32 flux: both fluxes differenced per face, with the increment that perturbs the largest
    // weighted component by sqrt(eps32) / rtol weighted units (sigma[2]): a 1/||v||_WRMS increment
    // scaled to FP32 would push a localised v's components by O(1) relative (NaN on rho or p < 0)
    if (ctx->useTet)
      return launch_tet_dq_jv_diff<double, float>(ctx->T32, y, v, 0.0, ctx->dSigma.p + 2, Jv, ctx->stream,
                                                  ctx->dNonFinite.p, newton, ngam);
    return launch_dq_jv_diff<double, float>(ctx->L32, y, v, 0.0, ctx->dSigma.p + 2, Jv, ctx->stream,
                                            ctx->dNonFinite.p, newton, ngam);
  }
  if (ctx->useTet)
    return launch_tet_dq_jv<double, double>(ctx->T64, y, v, fy, 0.0, ctx->dSigma.p, Jv, ctx->stream,
                                            ctx->dNonFinite.p, newton, ngam);
  return launch_dq_jv<double, double>(ctx->L64, y, v, fy, 0.0, ctx->dSigma.p, Jv, ctx->stream, ctx->dNonFinite.p,
                                      newton, ngam);
}

// the analytic product J v (or v - ngam J v with newton set) on device vectors
static int jvp_dev(fos_euler *ctx, const double *v, double *out, const double *y, int newton, double ngam) {
  unsigned *nf = ctx->dNonFinite.p;
  if (ctx->flags & FOS_PREC_FP32)
    return ctx->useTet ? launch_tet_jvp<double, float>(ctx->T32, y, v, out, ctx->stream, nf, newton, ngam)
                       : launch_jvp<double, float>(ctx->L32, y, v, out, ctx->stream, nf, newton, ngam);
  return ctx->useTet ? launch_tet_jvp<double, double>(ctx->T64, y, v, out, ctx->stream, nf, newton, ngam)
                     : launch_jvp<double, double>(ctx->L64, y, v, out, ctx->stream, nf, newton, ngam);
}

int finish_call(fos_euler *ctx, const char *who) {
  if (ctx->flags & FOS_ASYNC_CALLBACKS) return 0;
  FOS_HIP_TRY(hipMemcpyAsync(ctx->hNonFinite, ctx->dNonFinite.p, sizeof(unsigned), hipMemcpyDeviceToHost,
                             ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (*ctx->hNonFinite == 0) return 0;
  *ctx->hNonFinite = 0;
  FOS_HIP_TRY(hipMemcpyAsync(ctx->dNonFinite.p, ctx->hNonFinite, sizeof(unsigned), hipMemcpyHostToDevice,
                             ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  set_error("%s: non-finite value in the result", who);
  return 1;
}

// Pair arrays on the device for the element-array entry points (fos_find_euler,
// fos_euler_sync_state), built on first use: iEP, aP, normP in pair order and the cell -> pair
// list (2 pair + side).  dPairIEP is uploaded last, so it marks a complete set.
int ensure_pair_arrays(fos_euler *ctx) {
  if (ctx->dPairIEP.p) return 0;
  const int nC = ctx->nC, nP = ctx->nP;
  std::vector<int> ptr(nC + 1, 0), idx;
  for (int i = 0; i < nP; i++) {
    int m = ctx->hIEP[2 * i], n = ctx->hIEP[2 * i + 1];
    ptr[m + 1]++;
    if (n < nC) ptr[n + 1]++;
  }
  for (int c = 0; c < nC; c++) ptr[c + 1] += ptr[c];
  idx.resize(ptr[nC]);
  std::vector<int> f(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < nP; i++) {
    int m = ctx->hIEP[2 * i], n = ctx->hIEP[2 * i + 1];
    idx[f[m]++] = 2 * i;
    if (n < nC) idx[f[n]++] = 2 * i + 1;
  }
  if (ctx->dPairA.upload(ctx->hAP.data(), ctx->hAP.size(), ctx->stream) ||
      ctx->dPairN.upload(ctx->hNormP.data(), ctx->hNormP.size(), ctx->stream) ||
      ctx->dCellPairPtr.upload(ptr.data(), ptr.size(), ctx->stream) ||
      ctx->dCellPairIdx.upload(idx.data(), idx.size(), ctx->stream) || ctx->dPairF.alloc(5 * (size_t)nP) ||
      ctx->dPairIEP.upload(ctx->hIEP.data(), ctx->hIEP.size(), ctx->stream))
    return -1;
  return 0;
}

}  // namespace fos

using namespace fos;

extern "C" {

fos_euler *fos_euler_create(const fos_fv_grid *grid, const fos_cond_tab *bc, double r, double gamma, unsigned flags,
                            void *stream) {
  if (!grid) { set_error("fos_euler_create: null grid"); return nullptr; }
  if (check_nonneg("nC", grid->nC) || check_nonneg("nP", grid->nP)) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { set_error("fos_euler_create: no HIP device"); return nullptr; }
  if (resolve_noalloc_scratch(dev)) return nullptr;
  fos_euler *ctx = new fos_euler();
  ctx->device = dev;
  ctx->stream = static_cast<hipStream_t>(stream);
  ctx->flags = flags;
  ctx->r = r;
  ctx->gamm = gamma;
  ctx->nC = grid->nC;
  ctx->nE = grid->nE;
  ctx->nP = grid->nP;
  ctx->nPi = grid->nPi;
  ctx->maxNF = grid->maxNF;
  const unsigned zero = 0;
  if (build_layout(ctx, grid, bc) || ctx->redPart.alloc(2 * RED_BLOCKS) || ctx->dSigma.alloc(4) ||
      ctx->dNonFinite.upload(&zero, 1, ctx->stream) || save_grid_host(ctx, grid)) {
    delete ctx;
    return nullptr;
  }
  if (hipHostMalloc((void **)&ctx->hNonFinite, sizeof(unsigned)) != hipSuccess) {
    ctx->hNonFinite = nullptr;
    set_error("fos_euler_create: pinned host allocation failed");
    delete ctx;
    return nullptr;
  }
  *ctx->hNonFinite = 0;
  return ctx;
}

void fos_euler_destroy(fos_euler *ctx) {
  if (!ctx) return;
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int fos_euler_set_stream(fos_euler *ctx, void *stream) {
  if (!ctx) return -1;
  ctx->stream = static_cast<hipStream_t>(stream);
  return 0;
}

int fos_euler_set_tolerances(fos_euler *ctx, double rtol, double atol) {
  if (!ctx || !(rtol >= 0) || !(atol >= 0)) return -1;
  ctx->rtol = rtol;
  ctx->atol = atol;
  ctx->rhsEpoch++;  // the cached weight floor depends on the tolerances
  return 0;
}

int fos_euler_set_udf_bc(fos_euler *ctx, const double *values) {
  if (!ctx || !values) return -1;
  for (int b = 0; b < ctx->nPb; b++)
    if (ctx->bndIsUdf[b])
      for (int q = 0; q < 5; q++) ctx->bndPar[5 * (size_t)b + q] = values[5 * (size_t)b + q];
  std::vector<float> bp32(ctx->bndPar.begin(), ctx->bndPar.end());
  FOS_HIP_TRY(hipMemcpyAsync(ctx->dBndPar64.p, ctx->bndPar.data(), sizeof(double) * ctx->bndPar.size(),
                             hipMemcpyHostToDevice, ctx->stream));
  FOS_HIP_TRY(hipMemcpyAsync(ctx->dBndPar32.p, bp32.data(), sizeof(float) * bp32.size(), hipMemcpyHostToDevice,
                             ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->haveUdf = true;
  return 0;
}

int fos_euler_rhs(double t, const double *y, double *ydot, void *user_data) {
  (void)t;  // the reference's rhs reads BC data at the module time, not the callback time
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) return -1;
  ctx->counters[0]++;
  ctx->rhsEpoch++;
  if (ctx->flags & FOS_VECTORS_ON_HOST) {
    if (stage_in(ctx, ctx->hY, y) || ensure_vec(ctx->hOut, 5 * (size_t)ctx->nC)) return -1;
    if (residual_dev(ctx, ctx->hY.p, ctx->hOut.p) || stage_out(ctx, ctx->hOut, ydot)) return -1;
    return finish_call(ctx, "rhs");
  }
  if (residual_dev(ctx, y, ydot)) return -1;
  return finish_call(ctx, "rhs");
}

int fos_euler_jtimes(const double *v, double *Jv, double t, const double *y, const double *fy, void *user_data,
                     double *tmp) {
  (void)t;
  (void)tmp;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) return -1;
  ctx->counters[1]++;
  if (ctx->flags & FOS_VECTORS_ON_HOST) {
    if (stage_in(ctx, ctx->hVec, v) || stage_in(ctx, ctx->hY, y) || stage_in(ctx, ctx->hFy, fy) ||
        ensure_vec(ctx->hOut, 5 * (size_t)ctx->nC))
      return -1;
    if (jv_dev(ctx, ctx->hVec.p, ctx->hOut.p, ctx->hY.p, ctx->hFy.p) || stage_out(ctx, ctx->hOut, Jv)) return -1;
    return finish_call(ctx, "jtimes");
  }
  if (jv_dev(ctx, v, Jv, y, fy)) return -1;
  return finish_call(ctx, "jtimes");
}

int fos_euler_jtimes_noalloc(const double *v, double *Jv, double t, const double *y, const double *fy,
                             void *user_data, double *tmp) {
  (void)t;
  (void)tmp;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) return -1;
  if (ctx->flags & FOS_VECTORS_ON_HOST) {
    set_error("jtimes_noalloc: device vectors only (host vectors need staging buffers)");
    return -1;
  }
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("jtimes_noalloc: a *_UDF boundary condition is used but fos_euler_set_udf_bc was not called");
    return -1;
  }
  ctx->counters[1]++;
  if (jtimes_noalloc_dev(ctx, v, Jv, y, fy)) return -1;
  return finish_call(ctx, "jtimes_noalloc");
}

int fos_euler_jtimes_analytic(const double *v, double *Jv, double t, const double *y, const double *fy,
                              void *user_data, double *tmp) {
  (void)t;
  (void)fy;
  (void)tmp;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) return -1;
  ctx->counters[1]++;
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("jtimes: a *_UDF boundary condition is used but fos_euler_set_udf_bc was not called");
    return -1;
  }
  const double *vd = v, *yd = y;
  double *out = Jv;
  const bool host = ctx->flags & FOS_VECTORS_ON_HOST;
  if (host) {
    if (stage_in(ctx, ctx->hVec, v) || stage_in(ctx, ctx->hY, y) || ensure_vec(ctx->hOut, 5 * (size_t)ctx->nC))
      return -1;
    vd = ctx->hVec.p;
    yd = ctx->hY.p;
    out = ctx->hOut.p;
  }
  const int rc = jvp_dev(ctx, vd, out, yd, 0, 0.0);
  if (rc || (host && stage_out(ctx, ctx->hOut, Jv))) return -1;
  return finish_call(ctx, "jtimes_analytic");
}

// The integrator's Newton matrix product in the Jv kernels' epilogue (device vectors only)
static int newton_matvec(double gamma, const double *v, double *out, const double *y, const double *fy,
                         void *user_data, bool analytic, const char *who) {
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) return -1;
  if (ctx->flags & FOS_VECTORS_ON_HOST) {
    set_error("%s: device vectors only", who);
    return -1;
  }
  if (!v || !out || !y || (!analytic && !fy) || out == v || out == y) {
    set_error("%s: bad vectors (out must not alias v or y)", who);
    return -1;
  }
  ctx->counters[1]++;
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("%s: a *_UDF boundary condition is used but fos_euler_set_udf_bc was not called", who);
    return -1;
  }
  const int rc = analytic ? jvp_dev(ctx, v, out, y, 1, gamma) : jv_dev(ctx, v, out, y, fy, 1, gamma);
  return rc ? -1 : finish_call(ctx, who);
}

int fos_euler_newton_matvec(double gamma, const double *v, double *out, double t, const double *y, const double *fy,
                            void *user_data) {
  (void)t;
  return newton_matvec(gamma, v, out, y, fy, user_data, false, "newton_matvec");
}

int fos_euler_newton_matvec_analytic(double gamma, const double *v, double *out, double t, const double *y,
                                     const double *fy, void *user_data) {
  (void)t;
  return newton_matvec(gamma, v, out, y, fy, user_data, true, "newton_matvec_analytic");
}

int fos_euler_rhs_f32(fos_euler *ctx, const float *y, float *ydot) {
  if (!ctx) return -1;
  ctx->counters[0]++;
  const int rc = ctx->useTet ? launch_tet_residual<float, float>(ctx->T32, y, ydot, ctx->stream, ctx->dNonFinite.p)
                             : launch_residual<float, float>(ctx->L32, y, ydot, ctx->stream, ctx->dNonFinite.p);
  return rc ? -1 : finish_call(ctx, "rhs_f32");
}

int fos_euler_jtimes_f32(fos_euler *ctx, const float *v, float *Jv, const float *y, const float *fy, double sigma) {
  if (!ctx) return -1;
  ctx->counters[1]++;
  const int rc =
      ctx->useTet ? launch_tet_dq_jv<float, float>(ctx->T32, y, v, fy, sigma, nullptr, Jv, ctx->stream, ctx->dNonFinite.p)
                  : launch_dq_jv<float, float>(ctx->L32, y, v, fy, sigma, nullptr, Jv, ctx->stream, ctx->dNonFinite.p);
  return rc ? -1 : finish_call(ctx, "jtimes_f32");
}

int fos_euler_nonfinite(fos_euler *ctx, int clear) {
  if (!ctx) return -1;
  FOS_HIP_TRY(hipMemcpyAsync(ctx->hNonFinite, ctx->dNonFinite.p, sizeof(unsigned), hipMemcpyDeviceToHost,
                             ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const int bad = *ctx->hNonFinite != 0;
  if (clear && bad) {
    *ctx->hNonFinite = 0;
    FOS_HIP_TRY(hipMemcpyAsync(ctx->dNonFinite.p, ctx->hNonFinite, sizeof(unsigned), hipMemcpyHostToDevice,
                               ctx->stream));
    FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return bad;
}

// ---- the callbacks with SUNDIALS' N_Vector signatures: resolve the data pointers, forward
int fos_euler_set_nvector_access(fos_euler *ctx, fos_nvector_data_fn data) {
  if (!ctx) return -1;
  ctx->nvData = data;
  return 0;
}

static fos_euler *nv_ctx(void *user_data, const char *who) {
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx) { set_error("%s: null user_data", who); return nullptr; }
  if (!ctx->nvData) { set_error("%s: no N_Vector accessor registered (fos_euler_set_nvector_access)", who); return nullptr; }
  return ctx;
}
// data pointer of an N_Vector argument; a null N_Vector stays null, a null data pointer is an error
static bool nv_data(fos_euler *ctx, void *nv, double *&out, const char *who) {
  out = nullptr;
  if (!nv) return true;
  out = ctx->nvData(nv);
  if (!out) { set_error("%s: the N_Vector accessor returned NULL", who); return false; }
  return true;
}

int fos_euler_rhs_nv(double t, void *y, void *ydot, void *user_data) {
  fos_euler *ctx = nv_ctx(user_data, "rhs_nv");
  double *py, *pyd;
  if (!ctx || !nv_data(ctx, y, py, "rhs_nv") || !nv_data(ctx, ydot, pyd, "rhs_nv")) return -1;
  return fos_euler_rhs(t, py, pyd, ctx);
}

int fos_euler_jtimes_nv(void *v, void *Jv, double t, void *y, void *fy, void *user_data, void *tmp) {
  fos_euler *ctx = nv_ctx(user_data, "jtimes_nv");
  double *pv, *pJv, *py, *pfy, *ptmp;
  if (!ctx || !nv_data(ctx, v, pv, "jtimes_nv") || !nv_data(ctx, Jv, pJv, "jtimes_nv") ||
      !nv_data(ctx, y, py, "jtimes_nv") || !nv_data(ctx, fy, pfy, "jtimes_nv") || !nv_data(ctx, tmp, ptmp, "jtimes_nv"))
    return -1;
  return fos_euler_jtimes(pv, pJv, t, py, pfy, ctx, ptmp);
}

int fos_euler_jtimes_analytic_nv(void *v, void *Jv, double t, void *y, void *fy, void *user_data, void *tmp) {
  fos_euler *ctx = nv_ctx(user_data, "jtimes_analytic_nv");
  double *pv, *pJv, *py, *pfy, *ptmp;
  if (!ctx || !nv_data(ctx, v, pv, "jtimes_analytic_nv") || !nv_data(ctx, Jv, pJv, "jtimes_analytic_nv") ||
      !nv_data(ctx, y, py, "jtimes_analytic_nv") || !nv_data(ctx, fy, pfy, "jtimes_analytic_nv") ||
      !nv_data(ctx, tmp, ptmp, "jtimes_analytic_nv"))
    return -1;
  return fos_euler_jtimes_analytic(pv, pJv, t, py, pfy, ctx, ptmp);
}

int fos_euler_psetup_nv(double t, void *y, void *fy, int jok, int *jcur, double gamma, void *user_data) {
  fos_euler *ctx = nv_ctx(user_data, "psetup_nv");
  double *py, *pfy;
  if (!ctx || !nv_data(ctx, y, py, "psetup_nv") || !nv_data(ctx, fy, pfy, "psetup_nv")) return -1;
  return fos_euler_psetup(t, py, pfy, jok, jcur, gamma, ctx);
}

int fos_euler_psolve_nv(double t, void *y, void *fy, void *r, void *z, double gamma, double delta, int lr,
                        void *user_data) {
  fos_euler *ctx = nv_ctx(user_data, "psolve_nv");
  double *py, *pfy, *pr, *pz;
  if (!ctx || !nv_data(ctx, y, py, "psolve_nv") || !nv_data(ctx, fy, pfy, "psolve_nv") ||
      !nv_data(ctx, r, pr, "psolve_nv") || !nv_data(ctx, z, pz, "psolve_nv"))
    return -1;
  return fos_euler_psolve(t, py, pfy, pr, pz, gamma, delta, lr, ctx);
}

int fos_euler_layout(const fos_euler *ctx) { return ctx && ctx->useTet ? 1 : 0; }

int fos_euler_counters(const fos_euler *ctx, long long counts[4]) {
  if (!ctx) return -1;
  for (int i = 0; i < 4; i++) counts[i] = ctx->counters[i];
  return 0;
}

int fos_find_euler(fos_euler *ctx, const double *rho, const double *rhou, const double *rhoE, const double *p,
                   double gamma, double *dRho, double *dRhou, double *dRhoE) {
  if (!ctx) return -1;
  if (ensure_pair_arrays(ctx)) return -1;
  const int nP = ctx->nP, nC = ctx->nC;
  if (nP > 0)
    hipLaunchKernelGGL(pair_flux_kernel, dim3((nP + 255) / 256), dim3(256), 0, ctx->stream, nP, ctx->nE,
                       ctx->dPairIEP.p, ctx->dPairA.p, ctx->dPairN.p, rho, rhou, rhoE, p, gamma - 1.0, ctx->dPairF.p);
  if (nC > 0)
    hipLaunchKernelGGL(pair_gather_kernel, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, ctx->dCellPairPtr.p,
                       ctx->dCellPairIdx.p, ctx->dPairF.p, dRho, dRhou, dRhoE);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

int fos_euler_sync_state(fos_euler *ctx, const double *y, double *rho, double *rhou, double *rhoE, double *u,
                         double *p, double *temp, double *c) {
  if (!ctx) return -1;
  if (ensure_pair_arrays(ctx)) return -1;
  if (ctx->needUdf && !ctx->haveUdf) {
    set_error("fos_euler_sync_state: UDF boundary values not set (fos_euler_set_udf_bc)");
    return -1;
  }
  const int nC = ctx->nC, nE = ctx->nE, nPb = ctx->nPb;
  const bool host = ctx->flags & FOS_VECTORS_ON_HOST;
  double *outs[7] = {rho, rhou, rhoE, u, p, temp, c};
  const int width[7] = {1, 3, 1, 3, 1, 1, 1};
  ElemOut o{};
  double **dst[7] = {&o.rho, &o.rhou, &o.rhoE, &o.u, &o.p, &o.temp, &o.c};
  const double *yd = y;
  if (host) {
    // device staging: the state, then the requested element arrays, each seeded with the caller's
    // contents so that facets without a boundary pair keep their values, as in the reference
    size_t need = 0;
    for (int k = 0; k < 7; k++)
      if (outs[k]) need += (size_t)width[k] * nE;
    if (stage_in(ctx, ctx->hY, y) || ensure_vec(ctx->hOut, std::max<size_t>(need, 1))) return -1;
    yd = ctx->hY.p;
    size_t off = 0;
    for (int k = 0; k < 7; k++) {
      if (!outs[k]) continue;
      *dst[k] = ctx->hOut.p + off;
      FOS_HIP_TRY(hipMemcpyAsync(*dst[k], outs[k], sizeof(double) * width[k] * nE, hipMemcpyHostToDevice,
                                 ctx->stream));
      off += (size_t)width[k] * nE;
    }
  } else {
    for (int k = 0; k < 7; k++) *dst[k] = outs[k];
  }
  if (nC > 0)
    hipLaunchKernelGGL(cell_state_kernel, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, yd, ctx->gamm, ctx->r,
                       o);
  if (nPb > 0)
    hipLaunchKernelGGL(facet_state_kernel, dim3((nPb + 255) / 256), dim3(256), 0, ctx->stream, ctx->nPi, nPb, nC,
                       ctx->dPairIEP.p, ctx->dPairN.p, ctx->dBndType.p, ctx->dBndPar64.p, yd, ctx->gamm, ctx->r, o);
  FOS_HIP_TRY(hipGetLastError());
  if (host) {
    for (int k = 0; k < 7; k++)
      if (outs[k])
        FOS_HIP_TRY(hipMemcpyAsync(outs[k], *dst[k], sizeof(double) * width[k] * nE, hipMemcpyDeviceToHost,
                                   ctx->stream));
    FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

}  // extern "C"
