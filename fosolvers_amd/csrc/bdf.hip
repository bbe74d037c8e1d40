// BDF with Newton and left-preconditioned restarted GMRES, device resident: the replacement for
// the reference's BDFNewtonKrylov over CVODE (src/libfosolvers/numerics/SUNDIALS.f90:65-74,
// 745-845).  CVODE itself is third-party and not in the reference tree; this integrator keeps its
// callback interface (CVRhsFn, CVLsJacTimesVecFn, CVLsPrecSetupFn, CVLsPrecSolveFn), its
// difference-quotient Jv and its left preconditioning, in two modes:
//  * adaptive (the default, what foeuler gets from CVodeInit + CV_ONE_STEP): CVODE's published
//    method -- the fixed-leading-coefficient BDF of orders 1..5 in Nordsieck form (Jackson and
//    Sacks-Davis 1980; Byrne and Hindmarsh 1975), CVODE's local error test, step and order
//    selection, initial-step estimate, Newton convergence-rate test and linear tolerance, with the
//    weights 1/(rtol|y| + atol);
//  * fixed steps (fos_bdf_set_dt): BDF1 then BDF2 at the prescribed step, Newton to a tolerance on
//    the WRMS norm of the BDF residual G(y) = y - sum a_k y_{n-k} - gamma f(y), with a floor of
//    rtol times the RMS of the state in the weights -- the benchmark configurations' definition.
//
// Precision: the Newton residual, its WRMS norm, the Gram-Schmidt inner products, the Nordsieck
// history and the state update are FP64; the callbacks decide the precision of the Jv products and
// the preconditioner apply.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../../include/fosolvers_amd.h"
#include "common.h"
#include "hip_util.h"

namespace fos {
namespace {

constexpr int VT = 256;       // threads per vector-kernel block
constexpr int VB = 1024;      // blocks of the reduction kernels
constexpr int DOTS = 32;      // inner products per reduction launch (plus the vector's own norm)
constexpr int QMAX = 5;       // highest BDF order (CVODE's default maximum)

struct NordCoef {
  double c[QMAX + 1];
};
enum { NORD_PREDICT, NORD_RESTORE, NORD_SCALE, NORD_CORRECT, NORD_RAISE, NORD_LOWER };

// Operations on the Nordsieck history zn[j] = h^j y^(j) / j!, j = 0..qmax, q+1 of them in use,
// stored as vectors of n at stride n.  Each element's column is held in registers:
//   PREDICT  z_{j-1} += z_j for k = 1..q, j = q..k (z A, A the Pascal matrix)
//   RESTORE  the same sweep with -=, which undoes PREDICT
//   SCALE    z_j *= c_j, j = 1..q
//   CORRECT  z_j += c_j a, j = 0..q (a = the Newton correction)
//   RAISE    order q -> q+1: t = s z_qmax (the saved correction), z_{q+1} = t, z_j += c_j t, j = 2..q
//   LOWER    order q -> q-1: z_j -= c_j z_q, j = 2..q-1
__global__ __launch_bounds__(VT) void k_nordsieck(size_t n, int op, int q, int qmax, NordCoef c, double s,
                                                  double *__restrict__ zn, const double *__restrict__ a) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    double z[QMAX + 1];
#pragma unroll
    for (int j = 0; j <= QMAX; j++) z[j] = j <= q ? zn[(size_t)j * n + i] : 0.0;
    if (op == NORD_PREDICT || op == NORD_RESTORE) {
      const double sg = op == NORD_PREDICT ? 1.0 : -1.0;
#pragma unroll
      for (int k = 1; k <= QMAX; k++)
#pragma unroll
        for (int j = QMAX; j >= k; j--)
          if (j <= q) z[j - 1] += sg * z[j];
    } else if (op == NORD_SCALE) {
#pragma unroll
      for (int j = 1; j <= QMAX; j++) z[j] *= c.c[j];
    } else if (op == NORD_CORRECT) {
      const double av = a[i];
#pragma unroll
      for (int j = 0; j <= QMAX; j++) z[j] += c.c[j] * av;
    } else if (op == NORD_RAISE) {
      const double tv = s * zn[(size_t)qmax * n + i];
#pragma unroll
      for (int j = 2; j <= QMAX; j++)
        if (j <= q) z[j] += c.c[j] * tv;
      zn[(size_t)(q + 1) * n + i] = tv;
    } else {  // NORD_LOWER
      double zq = 0.0;
#pragma unroll
      for (int j = 0; j <= QMAX; j++)
        if (j == q) zq = z[j];
#pragma unroll
      for (int j = 2; j < QMAX; j++)
        if (j < q) z[j] -= c.c[j] * zq;
    }
#pragma unroll
    for (int j = 0; j <= QMAX; j++)
      if (j <= q) zn[(size_t)j * n + i] = z[j];
  }
}

// out = sum_j c_j zn_j, j = 0..q (interpolation from the Nordsieck history)
__global__ __launch_bounds__(VT) void k_nord_comb(size_t n, int q, NordCoef c, const double *__restrict__ zn,
                                                  double *__restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    double s = 0.0;
#pragma unroll
    for (int j = QMAX; j >= 0; j--)
      if (j <= q) s += c.c[j] * zn[(size_t)j * n + i];
    out[i] = s;
  }
}

// acor += d; y = zn0 + acor
__global__ __launch_bounds__(VT) void k_acor(size_t n, const double *__restrict__ d, double *__restrict__ acor,
                                             const double *__restrict__ zn0, double *__restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    const double a = acor[i] + d[i];
    acor[i] = a;
    y[i] = zn0[i] + a;
  }
}

// squared weights w2 = 1 / (rtol |y| + atol)^2 (CVODE's ewt)
// (w2f: an FP32 copy for the FP32 Krylov basis's inner products, when given)
__global__ __launch_bounds__(VT) void k_ewt(size_t n, const double *__restrict__ y, double rtol, double atol,
                                            double *__restrict__ w2, float *__restrict__ w2f) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    const double w = 1.0 / (rtol * fabs(y[i]) + atol);
    w2[i] = w * w;
    if (w2f) w2f[i] = (float)(w * w);
  }
}

// per-block max_i |yd_i| / (hf |y_i| + atol) -> part[block]
__global__ __launch_bounds__(VT) void k_maxratio(size_t n, const double *__restrict__ y, const double *__restrict__ yd,
                                                 double hf, double atol, double *__restrict__ part) {
  __shared__ double sm[VT / 64];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT)
    m = fmax(m, fabs(yd[i]) / (hf * fabs(y[i]) + atol));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < VT / 64; w++) m = fmax(m, sm[w]);
    part[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(VT) void k_reduce_max(const double *__restrict__ part, int nb, double *__restrict__ out) {
  __shared__ double sm[VT];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += VT) m = fmax(m, part[b]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = VT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

// z = a x + b y; z may be x or y (no restrict: the integrator updates vectors in place)
__global__ __launch_bounds__(VT) void k_lin2(size_t n, double a, const double *x, double b, const double *y,
                                             double *z) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) z[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(VT) void k_lin3(size_t n, double a, const double *x, double b, const double *y,
                                             double c, const double *w, double *z) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT)
    z[i] = a * x[i] + b * y[i] + c * w[i];
}

// squared WRMS weights of y: w2 = 1 / (rtol |y| + atol + rtol rms(y))^2; sumsq = sum y^2 (device)
__global__ __launch_bounds__(VT) void k_weights(size_t n, const double *__restrict__ y, double rtol, double atol,
                                                const double *__restrict__ sumsq, double *__restrict__ w2,
                                                float *__restrict__ w2f) {
  const double fl = atol + rtol * sqrt(sumsq[0] / (double)n);
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    const double w = 1.0 / (rtol * fabs(y[i]) + fl);
    w2[i] = w * w;
    if (w2f) w2f[i] = (float)(w * w);
  }
}

// the block's sums of acc[k], k < nk, and of acc[DOTS] when self -> part[block][DOTS + 1] (the
// self sum in slot nk)
__device__ __forceinline__ void block_partials(const double (&acc)[DOTS + 1], int nk, int self,
                                               double (&sm)[DOTS + 1][VT / 64], double *__restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k <= DOTS; k++) {
    if (k < nk || (k == DOTS && self)) {
      double v = acc[k];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if (lane == 0) sm[k][wave] = v;
    }
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k < nk || (k == DOTS && self)) {
    double v = 0.0;
    for (int q = 0; q < VT / 64; q++) v += sm[k][q];
    part[(size_t)blockIdx.x * (DOTS + 1) + (k == DOTS ? nk : k)] = v;
  }
}

// The basis kernels below take NK, a compile-time bound on the number nk of basis vectors (1, 2, 4,
// 8, 16 or 32; the launches pick the smallest that holds nk), and load V_min(k, nk-1) for every
// k < NK: with a run-time nk the compiler waits on each V_k load before its FMA, so an element's
// loads were issued one round trip at a time, and the passes ran at the memory latency instead of
// the bandwidth.  Coefficients past nk are zero and sums past nk are not stored, so the results are
// those of a loop over k < nk.  Each thread takes two elements per iteration, all loads first.

// partial weighted inner products <V_k, x'>_w2 for k < nk <= NK and <x', x'>_w2 in slot nk when
// self, with x' = x - theta xs when SHIFT, else x; part[block][DOTS + 1].  Each V_k, x and w2 is
// read once per launch.  TV: the basis's storage type (the products and sums are FP64).
template <class TV, int NK, bool SHIFT>
__global__ __launch_bounds__(VT) void k_dots(size_t n, const TV *__restrict__ V, size_t ldv, int nk, int self,
                                             const double *__restrict__ x, const TV *__restrict__ xs,
                                             double theta, const TV *__restrict__ w2,
                                             double *__restrict__ part) {
  __shared__ double sm[DOTS + 1][VT / 64];
  double acc[DOTS + 1];
#pragma unroll
  for (int k = 0; k <= DOTS; k++) acc[k] = 0.0;
  const size_t kl = (size_t)(nk > 0 ? nk - 1 : 0), st = (size_t)gridDim.x * VT;
  auto elem = [&](double xv, double wv, const TV (&vv)[NK > 0 ? NK : 1]) {
    const double xi = xv * wv;
#pragma unroll
    for (int k = 0; k < NK; k++) acc[k] += (double)vv[k] * xi;
    acc[DOTS] += xv * xi;
  };
  size_t i = (size_t)blockIdx.x * VT + threadIdx.x;
  for (; i + st < n; i += 2 * st) {
    const size_t i2 = i + st;
    TV va[NK > 0 ? NK : 1], vb[NK > 0 ? NK : 1];
    double xa = x[i], xb = x[i2];
    const double wa = (double)w2[i], wb = (double)w2[i2];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      va[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i];
      vb[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i2];
    }
    if (SHIFT) {
      xa = fma(-theta, (double)xs[i], xa);
      xb = fma(-theta, (double)xs[i2], xb);
    }
    elem(xa, wa, va);
    elem(xb, wb, vb);
  }
  if (i < n) {
    TV va[NK > 0 ? NK : 1];
    double xa = x[i];
    const double wa = (double)w2[i];
#pragma unroll
    for (int k = 0; k < NK; k++) va[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i];
    if (SHIFT) xa = fma(-theta, (double)xs[i], xa);
    elem(xa, wa, va);
  }
  block_partials(acc, nk, self, sm, part);
}

// sum of x^2 -> part[block][0] (unit weights)
__global__ __launch_bounds__(VT) void k_sumsq(size_t n, const double *__restrict__ x, double *__restrict__ part) {
  __shared__ double sm[DOTS + 1][VT / 64];
  double acc[DOTS + 1];
#pragma unroll
  for (int k = 0; k <= DOTS; k++) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    const double xv = x[i];
    acc[DOTS] += xv * xv;
  }
  block_partials(acc, 0, 1, sm, part);
}

// The fixed-step Newton right-hand side r = -G, G = y - a1 y_n - a2 y_{n-1} - gamma f(y), and the
// partial sums of G^2 w2 (slot 0) and of G^2 under CVODE's unfloored weights 1/(rtol |y_n| + atol)^2
// (slot 1, the record of the converged norm), one pass instead of forming G, its norms and -G
// separately; with yp and r0 given (the step's first iteration) also copies of y and -G
__global__ __launch_bounds__(VT) void k_bdf_residual(size_t n, double a1, const double *__restrict__ yn, double a2,
                                                     const double *__restrict__ ynm1, double gamma,
                                                     const double *__restrict__ fy, const double *__restrict__ y,
                                                     const double *__restrict__ w2, double rtol, double atol,
                                                     double *__restrict__ r, double *__restrict__ part,
                                                     double *__restrict__ yp, double *__restrict__ r0) {
  __shared__ double sm[DOTS + 1][VT / 64];
  double acc[DOTS + 1];
#pragma unroll
  for (int k = 0; k <= DOTS; k++) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    const double ynv = yn[i], yv = y[i];
    const double g = yv + (-a1 * ynv + -a2 * ynm1[i] + -gamma * fy[i]);
    r[i] = -g;
    if (yp) {
      yp[i] = yv;
      r0[i] = -g;
    }
    acc[0] += g * (g * w2[i]);
    const double e = 1.0 / (rtol * fabs(ynv) + atol), e2 = e * e;
    acc[DOTS] += g * (g * e2);
  }
  block_partials(acc, 1, 1, sm, part);
}

__global__ __launch_bounds__(VT) void k_reduce(const double *__restrict__ part, int nb, int nk,
                                               double *__restrict__ out) {
  __shared__ double sm[VT];
  for (int k = 0; k < nk; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += VT) s += part[(size_t)b * (DOTS + 1) + k];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = VT / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sm[0];
    __syncthreads();
  }
}

// s = (x' - sum_{k<nk} h_k V_k) * scale, x' = x - theta xs when SHIFT (as k_dots forms it), h on the
// device; s goes to x in place, or to outV (the basis's type) and outD when outV is given.
// NK = 0: any nk, a loop over k (for nk > 32).
template <class TV, int NK, bool SHIFT>
__global__ __launch_bounds__(VT) void k_sub_comb_scale(size_t n, const TV *__restrict__ V, size_t ldv, int nk,
                                                       const double *__restrict__ h, double scale,
                                                       const TV *__restrict__ xs, double theta, double *x,
                                                       TV *__restrict__ outV, double *outD) {
  const size_t kl = (size_t)(nk > 0 ? nk - 1 : 0), st = (size_t)gridDim.x * VT;
  double hk[NK > 0 ? NK : 1];
#pragma unroll
  for (int k = 0; k < NK; k++) hk[k] = k < nk ? h[k] : 0.0;
  auto put = [&](size_t j, double s) {
    s *= scale;
    if (outV) {
      outV[j] = (TV)s;
      outD[j] = s;
    } else {
      x[j] = s;
    }
  };
  if (NK == 0) {
    for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += st) {
      double s = SHIFT ? fma(-theta, (double)xs[i], x[i]) : x[i];
      for (int k = 0; k < nk; k++) s -= h[k] * (double)V[k * ldv + i];
      put(i, s);
    }
    return;
  }
  size_t i = (size_t)blockIdx.x * VT + threadIdx.x;
  for (; i + st < n; i += 2 * st) {
    const size_t i2 = i + st;
    TV va[NK > 0 ? NK : 1], vb[NK > 0 ? NK : 1];
    double sa = x[i], sb = x[i2];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      va[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i];
      vb[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i2];
    }
    if (SHIFT) {
      sa = fma(-theta, (double)xs[i], sa);
      sb = fma(-theta, (double)xs[i2], sb);
    }
#pragma unroll
    for (int k = 0; k < NK; k++) {
      sa -= hk[k] * (double)va[k];
      sb -= hk[k] * (double)vb[k];
    }
    put(i, sa);
    put(i2, sb);
  }
  if (i < n) {
    TV va[NK > 0 ? NK : 1];
    double sa = x[i];
#pragma unroll
    for (int k = 0; k < NK; k++) va[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i];
    if (SHIFT) sa = fma(-theta, (double)xs[i], sa);
#pragma unroll
    for (int k = 0; k < NK; k++) sa -= hk[k] * (double)va[k];
    put(i, sa);
  }
}

// x += sum_{k<nk} c_k V_k, or x = sum_k c_k V_k when OVW (x is not read); NK = 0: any nk
template <class TV, int NK, bool OVW>
__global__ __launch_bounds__(VT) void k_add_comb(size_t n, const TV *__restrict__ V, size_t ldv, int nk,
                                                 const double *__restrict__ c, double *__restrict__ x) {
  const size_t kl = (size_t)(nk > 0 ? nk - 1 : 0), st = (size_t)gridDim.x * VT;
  if (NK == 0) {
    for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += st) {
      double s = OVW ? 0.0 : x[i];
      for (int k = 0; k < nk; k++) s += c[k] * (double)V[k * ldv + i];
      x[i] = s;
    }
    return;
  }
  double ck[NK > 0 ? NK : 1];
#pragma unroll
  for (int k = 0; k < NK; k++) ck[k] = k < nk ? c[k] : 0.0;
  size_t i = (size_t)blockIdx.x * VT + threadIdx.x;
  for (; i + st < n; i += 2 * st) {
    const size_t i2 = i + st;
    TV va[NK > 0 ? NK : 1], vb[NK > 0 ? NK : 1];
    double sa = OVW ? 0.0 : x[i], sb = OVW ? 0.0 : x[i2];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      va[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i];
      vb[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i2];
    }
#pragma unroll
    for (int k = 0; k < NK; k++) {
      sa += ck[k] * (double)va[k];
      sb += ck[k] * (double)vb[k];
    }
    x[i] = sa;
    x[i2] = sb;
  }
  if (i < n) {
    TV va[NK > 0 ? NK : 1];
    double sa = OVW ? 0.0 : x[i];
#pragma unroll
    for (int k = 0; k < NK; k++) va[k] = V[((size_t)k < kl ? (size_t)k : kl) * ldv + i];
#pragma unroll
    for (int k = 0; k < NK; k++) sa += ck[k] * (double)va[k];
    x[i] = sa;
  }
}

// calls f(std::integral_constant<int, NK>) with the smallest NK in {1, 2, 4, 8, 16, 32} that holds nk
// (0 for nk == 0 or nk > 32)
template <class F>
inline void with_nk(int nk, F &&f) {
  if (nk <= 0 || nk > 32) f(std::integral_constant<int, 0>());
  else if (nk <= 1) f(std::integral_constant<int, 1>());
  else if (nk <= 2) f(std::integral_constant<int, 2>());
  else if (nk <= 4) f(std::integral_constant<int, 4>());
  else if (nk <= 8) f(std::integral_constant<int, 8>());
  else if (nk <= 16) f(std::integral_constant<int, 16>());
  else f(std::integral_constant<int, 32>());
}

// vf = (float)(a x), vd = a x: the first basis vector in FP32 and the next product's FP64 input
__global__ __launch_bounds__(VT) void k_scale_split(size_t n, double a, const double *__restrict__ x,
                                                    float *__restrict__ vf, double *__restrict__ vd) {
  for (size_t i = (size_t)blockIdx.x * VT + threadIdx.x; i < n; i += (size_t)gridDim.x * VT) {
    const double s = a * x[i];
    vf[i] = (float)s;
    vd[i] = s;
  }
}

}  // namespace
}  // namespace fos

using namespace fos;

struct fos_bdf {
  int n = 0, maxl = 20, maxRestarts = 5, maxNewton = 20;
  fos_rhs_fn f = nullptr;
  fos_jtimes_fn jtimes = nullptr;
  fos_newton_matvec_fn nmv = nullptr;  // fused v - gamma J v (device vectors)
  fos_psetup_fn pset = nullptr;
  fos_psolve_fn psol = nullptr;
  fos_psetup_fn fbSet = nullptr;  // fallback preconditioner (e.g. block ILU behind a learned one)
  fos_psolve_fn fbSol = nullptr;
  int fbCap = 0;                  // GMRES iterations after which the fallback takes over
  bool onFallback = false;
  void *ud = nullptr;
  unsigned flags = 0;
  hipStream_t stream = nullptr;
  double rtol = 1e-6, atol = 1e-12, newtonTol = 1e-8, linTol = 1e-2, dt = 1e-4, t = 0.0;
  int order = 0;  // order of the next fixed step
  // adaptive mode (CVODE's method); fixedStep after fos_bdf_set_dt
  bool fixedStep = false, started = false, jcur = false;
  int qmax = QMAX, qmaxNext = QMAX, q = 1, qprime = 1, qu = 0, qwait = 2, L = 2, maxcor = 3;
  double h = 0.0, hprime = 0.0, hscale = 0.0, hu = 0.0, eta = 1.0, etamax = 1e4;
  double h0User = 0.0, hmin = 0.0, hmaxInv = 0.0;
  double tau[QMAX + 2] = {}, tq[6] = {}, l[QMAX + 1] = {};
  double rl1 = 1.0, gammaA = 0.0, gamrat = 1.0, crate = 1.0, savedTq5 = 0.0, acnrm = 0.0;
  // CVODE's constants: error-test and order-selection biases, growth and shrink limits, Newton
  // convergence constants (CRDOWN, RDIV, the 0.1 nonlinear and 0.05 linear tolerance factors)
  static constexpr double kBias1 = 6.0, kBias2 = 6.0, kBias3 = 10.0, kAddon = 1e-6, kThresh = 1.5;
  static constexpr double kEtaMxF = 0.2, kEtaMin = 0.1, kEtaCf = 0.25, kOnePsm = 1.000001;
  static constexpr double kCrDown = 0.3, kRDiv = 2.0, kNlsCoef = 0.1, kEpLiFac = 0.05;
  static constexpr double kURound = 2.220446049250313e-16;
  static constexpr int kMaxConvFails = 10, kMaxErrFails = 7;
  // preconditioner setup control as CVODE's: setup at most every msbp steps or when gamma moved by
  // more than dgmax since the last setup, a new Jacobian (jok = 0) at most every msbj steps, and a
  // fresh setup plus one retry of the step after a Newton failure on stale data
  int msbp = 20, msbj = 51;
  double dgmax = 0.3, gammaP = 0.0;
  long long nsteps = 0, nstP = -1, nstJ = -1;
  fos_bdf_stats st{};
  int lastNewton = 0, lastLin = 0;
  double ritz = 1.0;  // the last Arnoldi diagonal v^T A v: the next Gram-Schmidt shift
  double lastGnorm = 0.0, lastGnormCv = 0.0;
  // device vectors
  DBuf<double> y, yn, ynm1, ypred, r0, G, fy, w2, d, tmp, tmp2, Vb, part, dres, hcol, scal, zn, acor;
  // FOS_PREC_FP32: the Krylov basis in FP32 (Vf), with the new direction and the next product's
  // input in FP64 (wv, dv)
  DBuf<float> Vf, w2f;  // (w2f: the weights in FP32, for the FP32 basis's inner products)
  DBuf<double> wv, dv;
  std::vector<double> hostY, hostYd;  // host staging for host-vector callbacks
  double *pinned = nullptr;
  size_t npinned = 256;

  ~fos_bdf() {
    if (pinned) (void)hipHostFree(pinned);
  }
  dim3 grid() const { return dim3((unsigned)std::min<size_t>(VB, ((size_t)n + VT - 1) / VT)); }
  int nblocks() const { return (int)std::min<size_t>(VB, ((size_t)n + VT - 1) / VT); }

  // weighted inner products of nk basis vectors with x' -> device out[0..nk), and <x', x'> ->
  // out[nk] when self; x' = x - theta xs when xs is given; unit weights when unit.  The weights
  // are w2, or its FP32 copy w2f with an FP32 basis.
  template <class TV = double>
  int dots(const TV *V, int nk, const double *x, double *out, bool self = false, const TV *xs = nullptr,
           double theta = 0.0, bool unit = false) {
    const TV *wts = std::is_same<TV, float>::value ? reinterpret_cast<const TV *>(w2f.p)
                                                   : reinterpret_cast<const TV *>(w2.p);
    if (unit) {  // nk == 0, self
      hipLaunchKernelGGL(k_sumsq, grid(), dim3(VT), 0, stream, (size_t)n, x, part.p);
      hipLaunchKernelGGL(k_reduce, dim3(1), dim3(VT), 0, stream, part.p, nblocks(), 1, out);
      FOS_HIP_TRY(hipGetLastError());
      return 0;
    }
    int k0 = 0;
    do {
      const int m = std::min(DOTS, nk - k0);
      const bool last = k0 + m >= nk;
      const int sf = last && self ? 1 : 0;
      const TV *Vk = V ? V + (size_t)k0 * n : V;
      with_nk(m, [&](auto c) {
        constexpr int NK = decltype(c)::value;
        if (NK > 0 && xs)
          hipLaunchKernelGGL((k_dots<TV, NK, true>), grid(), dim3(VT), 0, stream, (size_t)n, Vk, (size_t)n, m, sf, x,
                             xs, theta, wts, part.p);
        else
          hipLaunchKernelGGL((k_dots<TV, NK, false>), grid(), dim3(VT), 0, stream, (size_t)n, Vk, (size_t)n, m, sf, x,
                             xs, theta, wts, part.p);
      });
      hipLaunchKernelGGL(k_reduce, dim3(1), dim3(VT), 0, stream, part.p, nblocks(), m + sf, out + k0);
      k0 += m;
    } while (k0 < nk);
    FOS_HIP_TRY(hipGetLastError());
    return 0;
  }
  // x = (x - theta xs - V h) * scale over nk basis vectors (h on the device), in place, or into outV
  // and outD
  template <class TV>
  void sub_comb(const TV *V, int nk, const double *h, double scale, const TV *xs, double theta, double *x,
                TV *outV, double *outD) {
    with_nk(nk, [&](auto c) {
      constexpr int NK = decltype(c)::value;
      if (xs)
        hipLaunchKernelGGL((k_sub_comb_scale<TV, NK, true>), grid(), dim3(VT), 0, stream, (size_t)n, V, (size_t)n, nk,
                           h, scale, xs, theta, x, outV, outD);
      else
        hipLaunchKernelGGL((k_sub_comb_scale<TV, NK, false>), grid(), dim3(VT), 0, stream, (size_t)n, V, (size_t)n, nk,
                           h, scale, xs, theta, x, outV, outD);
    });
  }
  // x (+)= V c over nk basis vectors (c on the device); overwrite: x is not read
  template <class TV>
  void add_comb(const TV *V, int nk, const double *c, bool overwrite, double *x) {
    with_nk(nk, [&](auto cc) {
      constexpr int NK = decltype(cc)::value;
      if (overwrite)
        hipLaunchKernelGGL((k_add_comb<TV, NK, true>), grid(), dim3(VT), 0, stream, (size_t)n, V, (size_t)n, nk, c, x);
      else
        hipLaunchKernelGGL((k_add_comb<TV, NK, false>), grid(), dim3(VT), 0, stream, (size_t)n, V, (size_t)n, nk, c, x);
    });
  }
  int wnorm2(const double *x, double *out) { return dots<double>(nullptr, 0, x, out, true); }
  // copy device scalars to host (pinned) and wait
  int fetch(const double *dsrc, int cnt, double *h) {
    if (cnt < 0 || (size_t)cnt > npinned) {
      set_error("bdf: %d scalars exceed the host staging buffer (%zu)", cnt, npinned);
      return -1;
    }
    FOS_HIP_TRY(hipMemcpyAsync(pinned, dsrc, sizeof(double) * cnt, hipMemcpyDeviceToHost, stream));
    FOS_HIP_TRY(hipStreamSynchronize(stream));
    std::memcpy(h, pinned, sizeof(double) * cnt);
    return 0;
  }

  int call_rhs(double tt, const double *yy, double *out) {
    st.nfe++;
    if (!(flags & FOS_VECTORS_ON_HOST)) return f(tt, yy, out, ud);
    FOS_HIP_TRY(hipMemcpyAsync(hostY.data(), yy, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
    FOS_HIP_TRY(hipStreamSynchronize(stream));
    int rc = f(tt, hostY.data(), hostYd.data(), ud);
    FOS_HIP_TRY(hipMemcpyAsync(out, hostYd.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
    FOS_HIP_TRY(hipStreamSynchronize(stream));
    return rc;
  }

  // Newton matrix-vector product out = v - gamma J v
  int matvec(double gamma, const double *v, double *out) {
    st.njv++;
    if (nmv && jtimes && !(flags & FOS_VECTORS_ON_HOST)) {
      const int rc = nmv(gamma, v, out, t, y.p, fy.p, ud);
      FOS_HIP_TRY(hipGetLastError());
      return rc;
    }
    int rc;
    if (jtimes && !(flags & FOS_VECTORS_ON_HOST)) {
      rc = jtimes(v, tmp2.p, t, y.p, fy.p, ud, nullptr);
    } else {
      // CVODE's difference quotient: sigma = 1 / ||v||_WRMS (step weights)
      double nv;
      if (wnorm2(v, dres.p) || fetch(dres.p, 1, &nv)) return -1;
      nv = std::sqrt(nv / n);
      if (nv == 0.0) {
        FOS_HIP_TRY(hipMemcpyAsync(out, v, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
        return 0;
      }
      const double sig = 1.0 / nv;
      hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, sig, v, 1.0, y.p, tmp.p);
      rc = call_rhs(t, tmp.p, tmp2.p);
      if (rc) return rc;
      hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0 / sig, tmp2.p, -1.0 / sig, fy.p, tmp2.p);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0, v, -gamma, tmp2.p, out);
    FOS_HIP_TRY(hipGetLastError());
    return 0;
  }

  int prec(double gamma, const double *r, double *z) {
    fos_psolve_fn ps = onFallback ? fbSol : psol;
    if (!ps) {
      if (r != z) FOS_HIP_TRY(hipMemcpyAsync(z, r, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
      return 0;
    }
    st.nps++;
    return ps(t, y.p, fy.p, r, z, gamma, 0.0, 1, ud);
  }

  // Solve (I - gamma J) d = b by left-preconditioned GMRES(maxl) with CGS2, weighted norm: to
  // ||P^-1 (b - A x)||_w2 <= absTol when absTol > 0, else to linTol times its initial value.
  // Returns 0, or a callback's nonzero code (< 0 fatal, > 0 recoverable).  With FOS_PREC_FP32 the
  // basis is stored in FP32 (the Gram-Schmidt and update passes move half the bytes); products,
  // sums, the Hessenberg matrix and the solution update stay FP64.  The basis's rounding (~1e-7)
  // bounds the orthogonality and the least-squares accuracy, far inside GMRES's relative tolerance
  // (the Newton residual that decides convergence is FP64).
  // With yAcc given, the solution is added to yAcc instead (x is then scratch): a solve that ends
  // in its first cycle adds V y to yAcc directly, without writing x and a separate pass y += x.
  int gmres(double gamma, const double *b, double *x, int &iters, bool &conv, double absTol = 0.0,
            double *yAcc = nullptr) {
    return (flags & FOS_PREC_FP32) ? gmres_t<float>(gamma, b, x, iters, conv, absTol, yAcc)
                                   : gmres_t<double>(gamma, b, x, iters, conv, absTol, yAcc);
  }
  template <class TV>
  int gmres_t(double gamma, const double *b, double *x, int &iters, bool &conv, double absTol, double *yAcc) {
    constexpr bool F32 = std::is_same<TV, float>::value;
    iters = 0;
    conv = false;
    const int cap = (fbSol && !onFallback && fbCap > 0) ? fbCap : 1 << 30;
    const int m = maxl;
    TV *V = F32 ? reinterpret_cast<TV *>(Vf.p) : reinterpret_cast<TV *>(Vb.p);
    // FP64 views: the preconditioned residual, each product's input and each new direction (with an
    // FP64 basis these are the basis vectors themselves)
    double *r0v = F32 ? wv.p : Vb.p;
    auto vin = [&](int j) -> const double * { return F32 ? dv.p : Vb.p + (size_t)j * n; };
    auto wout = [&](int j) -> double * { return F32 ? wv.p : Vb.p + (size_t)(j + 1) * n; };
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hc(m + 2), yv(m);
    // x = 0 until the first cycle's update writes it (which then does not read it); with yAcc, a
    // solve that stops here adds x (0 in the first cycle) to yAcc
    auto zero_x = [&](int restart) -> int {
      if (yAcc) {
        if (restart > 0)
          hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0, yAcc, 1.0, x, yAcc);
        FOS_HIP_TRY(hipGetLastError());
        return 0;
      }
      if (restart == 0) FOS_HIP_TRY(hipMemsetAsync(x, 0, sizeof(double) * n, stream));
      return 0;
    };
    double beta0 = -1.0;
    const bool precd = onFallback ? fbSol : psol;
    for (int restart = 0; restart <= maxRestarts; restart++) {
      // r = P^-1 (b - A x); without a preconditioner r is read where it is (b, or tmp after a
      // restart) instead of being copied
      const double *r = b;
      if (restart > 0) {
        int rc = matvec(gamma, x, tmp.p);
        if (rc) return rc;
        hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0, b, -1.0, tmp.p, tmp.p);
        r = tmp.p;
      }
      if (precd) {
        int rc = prec(gamma, r, r0v);
        if (rc) return rc;
        r = r0v;
      }
      double bb;
      if (dots<TV>(nullptr, 0, r, dres.p, true) || fetch(dres.p, 1, &bb)) return -1;  // GMRES's own weights
      double beta = std::sqrt(bb);
      if (beta0 < 0) beta0 = beta;
      const double tol = absTol > 0.0 ? absTol : linTol * beta0;
      if (beta == 0.0 || ((absTol > 0.0 || restart > 0) && beta <= tol)) { conv = true; return zero_x(restart); }
      if (iters >= cap) return zero_x(restart);
      const double ib = 1.0 / beta;
      if (F32)
        hipLaunchKernelGGL(k_scale_split, grid(), dim3(VT), 0, stream, (size_t)n, ib, r, Vf.p, dv.p);
      else
        hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, ib, r, 0.0, r, Vb.p);
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int j = 0;
      bool done = false;
      for (; j < m; j++) {
        const TV *vj = V + (size_t)j * n;
        double *w = wout(j);
        int rc;
        if (precd) {
          rc = matvec(gamma, vin(j), tmp.p);
          if (rc) return rc;
          rc = prec(gamma, tmp.p, w);
        } else {
          rc = matvec(gamma, vin(j), w);  // no preconditioner: the product is the new direction
        }
        if (rc) return rc;
        iters++;
        // Shifted Arnoldi: w' = w - theta v_j is orthogonalised, H = V^T w' + theta e_j.  The Krylov
        // spaces of A and A - theta I are the same, for any theta.  A v_j has a large component
        // along v_j (A is close to the identity with a good preconditioner; without one, the
        // upwind Jacobian's diagonal blocks dominate), and orthogonalising w itself would lose most
        // of its norm to it (||w||^2 - |h|^2 cancels), asking for a second pass almost every
        // iteration.  theta = 1 with a preconditioner (P^-1 (I - gamma J) is close to I by design);
        // without one, the previous column's diagonal v_{j-1}^T A v_{j-1} (at j = 0 the last one
        // of the previous solve), which tracks v_j^T A v_j (measured at 6M cells, Mach 2: 4 second
        // passes in 27 preconditioned iterations with theta = 1, 11 with the previous diagonal; 30
        // in 59 unpreconditioned ones with the previous diagonal, 56 with theta = 1).
        // Classical Gram-Schmidt with the DGKS test on the shifted vector: h = V^T w' and
        // <w', w'> in one pass; if ||w' - V h||^2 = <w',w'> - |h|^2 keeps at least half of
        // <w',w'>, one pass is orthogonal to working accuracy and the norm is known, so
        // w = (w' - V h) / ||.|| in one more pass; otherwise a second classical pass
        // re-orthogonalises (CGS2).
        const double theta = precd || !std::isfinite(ritz) ? 1.0 : ritz;
        std::fill(hc.begin(), hc.end(), 0.0);
        std::vector<double> hp(j + 2);
        if (dots<TV>(V, j + 1, w, hcol.p, true, vj, theta) || fetch(hcol.p, j + 2, hp.data())) return -1;
        double ww = hp[j + 1], s2 = ww;
        for (int k = 0; k <= j; k++) { hc[k] = hp[k]; s2 -= hp[k] * hp[k]; }
        const TV *shift = vj;
        if (!(s2 >= 0.5 * ww)) {
          sub_comb<TV>(V, j + 1, hcol.p, 1.0, vj, theta, w, nullptr, nullptr);
          shift = nullptr;  // w now holds w' - V h
          if (dots<TV>(V, j + 1, w, hcol.p, true) || fetch(hcol.p, j + 2, hp.data())) return -1;
          s2 = hp[j + 1];
          for (int k = 0; k <= j; k++) { hc[k] += hp[k]; s2 -= hp[k] * hp[k]; }
          st.nreorth++;
        }
        const double hh = std::sqrt(std::max(s2, 0.0));
        hc[j] += theta;
        ritz = hc[j];
        hc[j + 1] = hh;
        // the new basis vector: in place (FP64 basis), or into the FP32 basis and the next input
        sub_comb<TV>(V, j + 1, hcol.p, hh > 0.0 ? 1.0 / hh : 1.0, shift, theta, w,
                     F32 ? V + (size_t)(j + 1) * n : nullptr, F32 ? dv.p : nullptr);
        // Givens rotations on the new Hessenberg column
        for (int k = 0; k < j; k++) {
          double a = cs[k] * hc[k] + sn[k] * hc[k + 1];
          hc[k + 1] = -sn[k] * hc[k] + cs[k] * hc[k + 1];
          hc[k] = a;
        }
        double r = std::hypot(hc[j], hc[j + 1]);
        cs[j] = r > 0 ? hc[j] / r : 1.0;
        sn[j] = r > 0 ? hc[j + 1] / r : 0.0;
        hc[j] = r;
        hc[j + 1] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        for (int k = 0; k <= j; k++) H[(size_t)k * m + j] = hc[k];
        if (std::fabs(g[j + 1]) <= tol || hh == 0.0) { j++; done = true; break; }
        if (iters >= cap) { j++; break; }
      }
      // x += V y with H y = g (upper triangular j x j); the first cycle writes x = V y
      for (int k = j - 1; k >= 0; k--) {
        double s = g[k];
        for (int q = k + 1; q < j; q++) s -= H[(size_t)k * m + q] * yv[q];
        yv[k] = s / H[(size_t)k * m + k];
      }
      FOS_HIP_TRY(hipMemcpyAsync(hcol.p, yv.data(), sizeof(double) * j, hipMemcpyHostToDevice, stream));
      const bool last = done || iters >= cap || restart == maxRestarts;
      if (yAcc && last && restart == 0) {
        add_comb<TV>(V, j, hcol.p, false, yAcc);
      } else {
        add_comb<TV>(V, j, hcol.p, restart == 0, x);
        if (yAcc && last)
          hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0, yAcc, 1.0, x, yAcc);
      }
      FOS_HIP_TRY(hipGetLastError());
      FOS_HIP_TRY(hipStreamSynchronize(stream));
      if (done) { conv = true; return 0; }
      if (iters >= cap) return 0;
    }
    st.nlinfails++;
    return 0;  // not converged to linTol: Newton continues with the best iterate
  }

  // ------------------------------------------------------------------------------------------
  // Adaptive mode: CVODE's BDF (orders 1..qmax, Nordsieck history, local error control).

  int nord(int op, int qq, const double *c, double sc = 0.0, const double *a = nullptr) {
    NordCoef cc{};
    if (c)
      for (int j = 0; j <= QMAX; j++) cc.c[j] = c[j];
    hipLaunchKernelGGL(k_nordsieck, grid(), dim3(VT), 0, stream, (size_t)n, op, qq, qmax, cc, sc, zn.p, a);
    FOS_HIP_TRY(hipGetLastError());
    return 0;
  }
  double *znj(int j) { return zn.p + (size_t)j * n; }
  int wrms(const double *x, double &out) {
    double s2;
    if (wnorm2(x, dres.p) || fetch(dres.p, 1, &s2)) return -1;
    out = std::sqrt(s2 / (double)n);
    return 0;
  }
  int copy(double *dst, const double *src) {
    FOS_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
    return 0;
  }
  // zn_j *= eta^j, h = hscale eta (cvRescale)
  int rescale() {
    double c[QMAX + 1], f = 1.0;
    c[0] = 1.0;
    for (int j = 1; j <= QMAX; j++) c[j] = (f *= eta);
    if (nord(NORD_SCALE, q, c)) return -1;
    h = hscale * eta;
    hscale = h;
    return 0;
  }
  int restore(double savedT) {
    t = savedT;
    return nord(NORD_RESTORE, q, nullptr);
  }

  // Method coefficients l[0..q] and error-test constants tq[1..5] for the step h after the
  // history tau[1..q] (the fixed-leading-coefficient BDF: l(x) = (1 + x) prod_{j=2}^{q-1}
  // (1 + x h / (t_n - t_{n-j})) (1 + x / xi*), with xi* chosen so the leading coefficient is that
  // of the constant-step formula).
  void set_coeffs() {
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    l[0] = l[1] = 1.0;
    double xiInv = 1.0, xistarInv = 1.0, alpha0 = -1.0, alpha0Hat = -1.0, hsum = h;
    if (q > 1) {
      for (int j = 2; j < q; j++) {
        hsum += tau[j - 1];
        xiInv = h / hsum;
        alpha0 -= 1.0 / j;
        for (int i = j; i >= 1; i--) l[i] += l[i - 1] * xiInv;
      }
      alpha0 -= 1.0 / q;
      xistarInv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xiInv = h / hsum;
      alpha0Hat = -l[1] - xiInv;
      for (int i = q; i >= 1; i--) l[i] += l[i - 1] * xistarInv;
    }
    const double A1 = 1.0 - alpha0Hat + alpha0, A2 = 1.0 + q * A1;
    tq[2] = std::fabs(A1 / (alpha0 * A2));
    tq[5] = std::fabs(A2 * xistarInv / (l[q] * xiInv));
    if (qwait == 1) {  // constants for the order q-1 and q+1 estimates
      if (q > 1) {
        const double C = xistarInv / l[q], A3 = alpha0 + 1.0 / q, A4 = alpha0Hat + xiInv;
        tq[1] = std::fabs(C * (1.0 - A4 + A3) / A3);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xiInv = h / hsum;
      const double A5 = alpha0 - 1.0 / (q + 1), A6 = alpha0Hat - xiInv;
      tq[3] = std::fabs((1.0 - A6 + A5) / A2 / (xiInv * (q + 2) * A5));
    }
    tq[4] = kNlsCoef / tq[2];
    rl1 = 1.0 / l[1];
    gammaA = h * rl1;
    if (nsteps == 0) gammaP = gammaA;
    gamrat = nsteps > 0 ? gammaA / gammaP : 1.0;
  }

  // Order change of the history by +-1 (cvAdjustOrder for BDF); other changes truncate.
  int adjust_order(int dq) {
    if (q == 2 && dq != 1) return 0;
    double c[QMAX + 1] = {};
    if (dq == 1) {
      double alpha0 = -1.0, alpha1 = 1.0, prod = 1.0, xiold = 1.0, hsum = hscale;
      c[2] = 1.0;
      for (int j = 1; j < q; j++) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; i--) c[i] = c[i] * xiold + c[i - 1];
        xiold = xi;
      }
      return nord(NORD_RAISE, q, c, (-alpha0 - alpha1) / prod);
    }
    if (dq == -1) {
      double hsum = 0.0;
      c[2] = 1.0;
      for (int j = 1; j <= q - 2; j++) {
        hsum += tau[j];
        const double xi = hsum / hscale;
        for (int i = j + 2; i >= 2; i--) c[i] = c[i] * xi + c[i - 1];
      }
      return nord(NORD_LOWER, q, c);
    }
    return 0;
  }

  // ||(f(t + hg, y0 + hg f0) - f0) / hg||_WRMS, the second-derivative estimate of the initial
  // step selection (f0 unscaled in zn_1)
  int ydd_norm(double hg, double &out) {
    hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0, znj(0), hg, znj(1), tmp.p);
    int rc = call_rhs(t + hg, tmp.p, tmp2.p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0 / hg, tmp2.p, -1.0 / hg, znj(1), tmp2.p);
    return wrms(tmp2.p, out) ? -1 : 0;
  }

  // CVODE's initial step (cvHin): bounded by 0.1 |tOut - t0| and by 0.1 |y|/|y'| per component,
  // then h = 0.5 sqrt(2 / ||y''||) iterated up to four times from the geometric mean of the bounds.
  int initial_step(double tOut) {
    const double tdist = std::fabs(tOut - t), tround = kURound * std::max(std::fabs(t), std::fabs(tOut));
    if (!(tOut > t) || tdist < 2.0 * tround) {
      set_error("bdf step: tOut = %g must lie ahead of t = %g by more than roundoff", tOut, t);
      return -1;
    }
    const double hlb = 100.0 * tround;
    hipLaunchKernelGGL(k_maxratio, grid(), dim3(VT), 0, stream, (size_t)n, znj(0), znj(1), 0.1, atol, part.p);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(VT), 0, stream, part.p, nblocks(), dres.p);
    double hubInv;
    if (fetch(dres.p, 1, &hubInv)) return -1;
    double hub = 0.1 * tdist;
    if (hub * hubInv > 1.0) hub = 1.0 / hubInv;
    double hg = std::sqrt(hlb * hub);
    if (hub < hlb) { h = hg; return 0; }
    bool hnewOK = false;
    double hs = hg, hnew = hg, yddnrm = 0.0;
    for (int count1 = 1; count1 <= 4; count1++) {
      bool hgOK = false;
      for (int count2 = 1; count2 <= 4; count2++) {
        const int rc = ydd_norm(hg, yddnrm);
        if (rc < 0) return rc;
        if (rc == 0) { hgOK = true; break; }
        hg *= 0.2;  // recoverable rhs failure: try a smaller perturbation
      }
      if (!hgOK) {
        if (count1 <= 2) { set_error("bdf step: rhs failed repeatedly in the initial step estimate"); return -1; }
        hnew = hs;
        break;
      }
      hs = hg;
      if (hnewOK || count1 == 4) { hnew = hg; break; }
      hnew = (yddnrm * hub * hub > 2.0) ? std::sqrt(2.0 / yddnrm) : std::sqrt(hg * hub);
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) hnewOK = true;
      if (count1 > 1 && hrat > 2.0) { hnew = hg; hnewOK = true; }
      hg = hnew;
    }
    h = std::min(std::max(0.5 * hnew, hlb), hub);
    return 0;
  }

  // first call after set_iv: history, weights, f(t0, y0) and the initial step
  int start(double tOut) {
    if (!zn.p && (zn.alloc((size_t)(QMAX + 1) * n) || acor.alloc(n))) return -1;
    if (copy(znj(0), yn.p)) return -1;
    hipLaunchKernelGGL(k_ewt, grid(), dim3(VT), 0, stream, (size_t)n, znj(0), rtol, atol, w2.p, w2f.p);
    int rc = call_rhs(t, znj(0), znj(1));
    if (rc) {
      if (rc > 0) set_error("bdf step: rhs failed at the initial state (code %d)", rc);
      return rc < 0 ? rc : -1;
    }
    if (h0User > 0.0) h = h0User;
    else if ((rc = initial_step(tOut))) return rc;
    if (hmaxInv > 0.0 && h * hmaxInv > 1.0) h = 1.0 / hmaxInv;
    if (h < hmin) h = hmin;
    hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, h, znj(1), 0.0, znj(1), znj(1));
    FOS_HIP_TRY(hipGetLastError());
    hscale = hprime = h;
    q = qprime = 1;
    L = 2;
    qwait = 2;
    etamax = 1e4;  // the first step may grow by up to 10^4
    for (double &v : tau) v = 0.0;
    savedTq5 = 0.0;
    crate = 1.0;
    jcur = false;
    started = true;
    return 0;
  }

  enum { FIRST_CALL, PREV_CONV_FAIL, PREV_ERR_FAIL };
  enum { NO_FAILURES, FAIL_BAD_J, FAIL_OTHER };
  enum { NLS_OK = 0, NLS_RECOVER = 1, NLS_TRY_AGAIN = 2 };

  // Newton iteration on acor = y - zn_0: (I - gamma J) x = gamma f(y) - rl1 zn_1 - acor, to
  // CVODE's test ||x|| min(1, rate) <= 0.1 / tq_2 (rate: the correction's contraction rate), at
  // most maxcor corrections, the linear solve to 0.05 of that tolerance (absolute, weighted L2).
  int newton_iter(int &nit, int &lin) {
    double delp = 0.0;
    const double deltaLin = kEpLiFac * tq[4] * std::sqrt((double)n);
    for (int m = 0;; m++) {
      hipLaunchKernelGGL(k_lin3, grid(), dim3(VT), 0, stream, (size_t)n, gammaA, fy.p, -rl1, znj(1), -1.0, acor.p, G.p);
      FOS_HIP_TRY(hipGetLastError());
      if (m == 0) {  // the step's (predictor, Newton residual) pair for fos_bdf_last_triple
        if (copy(ypred.p, y.p) || copy(r0.p, G.p)) return -1;
      }
      int it = 0;
      bool conv = false;
      int rc = gmres(gammaA, G.p, d.p, it, conv, deltaLin);
      lin += it;
      if (rc < 0) return rc;
      if (rc == 0 && !conv && fbSol && !onFallback) {
        int jc = 0;
        st.nfallback++;
        onFallback = true;
        rc = fbSet(t, y.p, fy.p, 0, &jc, gammaA, ud);
        if (rc < 0) return rc;
        if (rc == 0) {
          rc = gmres(gammaA, G.p, d.p, it, conv, deltaLin);
          lin += it;
          if (rc < 0) return rc;
        }
      }
      nit++;
      if (rc > 0 || (!conv && m > 0)) return (!jcur && pset) ? NLS_TRY_AGAIN : NLS_RECOVER;
      double del;
      if (wrms(d.p, del)) return -1;
      hipLaunchKernelGGL(k_acor, grid(), dim3(VT), 0, stream, (size_t)n, d.p, acor.p, znj(0), y.p);
      FOS_HIP_TRY(hipGetLastError());
      if (m > 0) crate = std::max(kCrDown * crate, del / delp);
      const double dcon = del * std::min(1.0, crate) / tq[4];
      lastGnorm = dcon;
      if (dcon <= 1.0) {
        if (m == 0) acnrm = del;
        else if (wrms(acor.p, acnrm)) return -1;
        jcur = false;
        return NLS_OK;
      }
      if (m + 1 == maxcor || (m >= 1 && del > kRDiv * delp)) return (!jcur && pset) ? NLS_TRY_AGAIN : NLS_RECOVER;
      delp = del;
      rc = call_rhs(t, y.p, fy.p);
      if (rc < 0) return rc;
      if (rc > 0) return NLS_RECOVER;
    }
  }

  // the nonlinear solve of one attempt, with CVODE's preconditioner setup decisions
  int nls(int nflag, int &nit, int &lin) {
    bool callSetup = false;
    int convfail = NO_FAILURES;
    if (pset) {
      convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
      callSetup = nflag == PREV_CONV_FAIL || nflag == PREV_ERR_FAIL || nsteps == 0 || nstP < 0 ||
                  nsteps >= nstP + msbp || std::fabs(gamrat - 1.0) > dgmax;
    } else {
      crate = 1.0;
    }
    for (;;) {
      onFallback = false;
      if (copy(y.p, znj(0))) return -1;
      FOS_HIP_TRY(hipMemsetAsync(acor.p, 0, sizeof(double) * n, stream));
      int rc = call_rhs(t, y.p, fy.p);
      if (rc < 0) return rc;
      if (rc > 0) return NLS_RECOVER;
      if (callSetup) {
        const double dgamma = std::fabs(gammaA / gammaP - 1.0);
        const bool jbad = nsteps == 0 || nstJ < 0 || nsteps >= nstJ + msbj ||
                          (convfail == FAIL_BAD_J && dgamma < 0.2) || convfail == FAIL_OTHER;
        int jc = 0;
        st.npe++;
        rc = pset(t, y.p, fy.p, jbad ? 0 : 1, &jc, gammaA, ud);
        jcur = jc != 0;
        if (jcur) nstJ = nsteps;
        crate = 1.0;
        gammaP = gammaA;
        gamrat = 1.0;
        nstP = nsteps;
        if (rc < 0) return rc;
        if (rc > 0) return NLS_RECOVER;
      }
      rc = newton_iter(nit, lin);
      if (rc != NLS_TRY_AGAIN) return rc;
      callSetup = true;
      convfail = FAIL_BAD_J;
    }
  }

  // step and order for the next step after a successful one (cvPrepareNextStep, cvChooseEta,
  // cvSetEta)
  int prepare_next(double dsm) {
    if (etamax == 1.0) {  // no growth right after a failure
      qwait = std::max(qwait, 2);
      qprime = q;
      hprime = h;
      eta = 1.0;
      return 0;
    }
    const double etaq = 1.0 / (std::pow(kBias2 * dsm, 1.0 / L) + kAddon);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
    } else {
      qwait = 2;
      double etaqm1 = 0.0, etaqp1 = 0.0;
      if (q > 1) {
        double ddn;
        if (wrms(znj(q), ddn)) return -1;
        etaqm1 = 1.0 / (std::pow(kBias1 * ddn * tq[1], 1.0 / q) + kAddon);
      }
      if (q != qmax && savedTq5 != 0.0) {
        const double cquot = (tq[5] / savedTq5) * std::pow(h / tau[2], (double)L);
        hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, 1.0, acor.p, -cquot, znj(qmax), tmp.p);
        double dup;
        if (wrms(tmp.p, dup)) return -1;
        etaqp1 = 1.0 / (std::pow(kBias3 * dup * tq[3], 1.0 / (L + 1)) + kAddon);
      }
      const double etam = std::max(etaqm1, std::max(etaq, etaqp1));
      if (etam < kThresh) {
        eta = 1.0;
        qprime = q;
      } else if (etam == etaqm1) {
        eta = etaqm1;
        qprime = q - 1;
      } else if (etam == etaq) {
        eta = etaq;
        qprime = q;
      } else {
        eta = etaqp1;
        qprime = q + 1;
        if (copy(znj(qmax), acor.p)) return -1;  // the correction that builds the new column
      }
    }
    if (eta < kThresh) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = std::min(eta, etamax);
      if (hmaxInv > 0.0) eta /= std::max(1.0, std::fabs(h) * hmaxInv * eta);
      hprime = h * eta;
    }
    return 0;
  }

  // one internal step (CV_ONE_STEP): predict, Newton, local error test; on failure restore the
  // history, shrink h (and the order after repeated error-test failures) and try again
  int step_adaptive(double tOut, double *tRet, double *yOut) {
    int rc;
    if (!started) {
      if ((rc = start(tOut))) return rc;
    } else if (hprime != h) {  // step (and order) chosen after the last step
      if (qprime != q) {
        if (adjust_order(qprime - q)) return -1;
        q = qprime;
        L = q + 1;
        qwait = L;
      }
      if (rescale()) return -1;
    }
    hipLaunchKernelGGL(k_ewt, grid(), dim3(VT), 0, stream, (size_t)n, znj(0), rtol, atol, w2.p, w2f.p);
    const double savedT = t;
    int ncf = 0, nef = 0, nflag = FIRST_CALL, nit = 0, lin = 0;
    double dsm = 0.0;
    for (;;) {
      t = savedT + h;
      if (nord(NORD_PREDICT, q, nullptr)) return -1;
      set_coeffs();
      rc = nls(nflag, nit, lin);
      if (rc < 0) { restore(savedT); return rc; }
      if (rc != NLS_OK) {  // Newton failed: shrink h by 4 and try again
        st.nncfails++;
        ncf++;
        etamax = 1.0;
        if (restore(savedT)) return -1;
        if (std::fabs(h) <= hmin * kOnePsm || ncf == kMaxConvFails) {
          set_error("bdf step: Newton failed %d times at t = %g (h = %g)", ncf, t, h);
          st.nni += nit;
          st.nli += lin;
          return 1;
        }
        eta = std::max(kEtaCf, hmin / std::fabs(h));
        nflag = PREV_CONV_FAIL;
        if (rescale()) return -1;
        continue;
      }
      dsm = acnrm * tq[2];
      if (dsm <= 1.0) break;
      // local error test failed
      nef++;
      st.netf++;
      nflag = PREV_ERR_FAIL;
      if (restore(savedT)) return -1;
      if (std::fabs(h) <= hmin * kOnePsm || nef == kMaxErrFails) {
        set_error("bdf step: error test failed %d times at t = %g (h = %g)", nef, t, h);
        st.nni += nit;
        st.nli += lin;
        return 1;
      }
      etamax = 1.0;
      if (nef <= 3) {
        eta = 1.0 / (std::pow(kBias2 * dsm, 1.0 / L) + kAddon);
        eta = std::max(kEtaMin, std::max(eta, hmin / std::fabs(h)));
        if (nef >= 2) eta = std::min(eta, kEtaMxF);
        if (rescale()) return -1;
        continue;
      }
      // repeated failures: one order lower, and no order change for L steps (CVODE's
      // cvDoErrorTest after MXNEF1 = 3 failures)
      eta = std::max(kEtaMin, hmin / std::fabs(h));
      if (q > 1) {
        if (adjust_order(-1)) return -1;
        L = q;
        q--;
        qwait = L;
        if (rescale()) return -1;
        continue;
      }
      // at order 1 as well: restart the history from f at a smaller step
      h *= eta;
      hscale = h;
      qwait = 10;
      rc = call_rhs(t, znj(0), tmp.p);
      if (rc < 0) return rc;
      if (rc > 0) { set_error("bdf step: rhs failed (code %d) at t = %g", rc, t); return 1; }
      hipLaunchKernelGGL(k_lin2, grid(), dim3(VT), 0, stream, (size_t)n, h, tmp.p, 0.0, tmp.p, znj(1));
      FOS_HIP_TRY(hipGetLastError());
    }
    // accept: history update, step-size history, saved correction for the order q+1 estimate
    nsteps++;
    st.nst++;
    hu = h;
    qu = q;
    for (int i = q; i >= 2; i--) tau[i] = tau[i - 1];
    if (q == 1 && nsteps > 1) tau[2] = tau[1];
    tau[1] = h;
    if (nord(NORD_CORRECT, q, l, 0.0, acor.p)) return -1;
    qwait--;
    if (qwait == 1 && q != qmax) {
      if (copy(znj(qmax), acor.p)) return -1;
      savedTq5 = tq[5];
    }
    if (prepare_next(dsm)) return -1;
    etamax = 10.0;
    lastNewton = nit;
    lastLin = lin;
    st.nni += nit;
    st.nli += lin;
    if (tRet) *tRet = t;
    if (yOut) {
      hipMemcpyKind k = (flags & FOS_VECTORS_ON_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
      FOS_HIP_TRY(hipMemcpyAsync(yOut, znj(0), sizeof(double) * n, k, stream));
    }
    FOS_HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  }
};

extern "C" {

fos_bdf *fos_bdf_create(int nEq, fos_rhs_fn f, int maxl, fos_psetup_fn pSet, fos_psolve_fn pSol, void *user_data) {
  if (nEq <= 0 || !f) { set_error("fos_bdf_create: bad arguments"); return nullptr; }
  fos_bdf *b = new fos_bdf();
  b->n = nEq;
  b->f = f;
  b->maxl = maxl > 0 ? maxl : 5;  // SUNDIALS' default Krylov dimension for maxl = 0
  // host staging for device scalars: a Hessenberg column (maxl + 1 entries plus the norm) or the
  // integrator's scalar sets, whichever is larger
  b->npinned = std::max<size_t>(256, (size_t)b->maxl + 2);
  b->pset = pSet;
  b->psol = pSol;
  b->ud = user_data;
  const size_t n = nEq;
  if (b->y.alloc(n) || b->yn.alloc(n) || b->ynm1.alloc(n) || b->ypred.alloc(n) || b->r0.alloc(n) || b->G.alloc(n) ||
      b->fy.alloc(n) || b->w2.alloc(n) || b->d.alloc(n) || b->tmp.alloc(n) || b->tmp2.alloc(n) ||
      b->Vb.alloc((size_t)(b->maxl + 1) * n) || b->part.alloc((size_t)VB * (DOTS + 1)) || b->dres.alloc(DOTS) ||
      b->hcol.alloc(64 + b->maxl) || b->scal.alloc(4) || hipHostMalloc((void **)&b->pinned, sizeof(double) * b->npinned) != hipSuccess) {
    delete b;
    return nullptr;
  }
  return b;
}

void fos_bdf_destroy(fos_bdf *b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  delete b;
}

int fos_bdf_set_jtimes(fos_bdf *b, fos_jtimes_fn jt) {
  if (!b) return -1;
  b->jtimes = jt;
  return 0;
}

int fos_bdf_set_newton_matvec(fos_bdf *b, fos_newton_matvec_fn fn) {
  if (!b) return -1;
  b->nmv = fn;
  return 0;
}

int fos_bdf_set_options(fos_bdf *b, unsigned flags, void *stream) {
  if (!b) return -1;
  b->flags = flags;
  b->stream = static_cast<hipStream_t>(stream);
  const size_t nv = (size_t)(b->maxl + 1) * b->n;
  if ((flags & FOS_PREC_FP32) && b->Vf.n < nv &&
      (b->Vf.alloc(nv) || b->wv.alloc(b->n) || b->dv.alloc(b->n) || b->w2f.alloc(b->n))) {
    b->flags &= ~(unsigned)FOS_PREC_FP32;
    return -1;
  }
  if (flags & FOS_VECTORS_ON_HOST) {
    b->hostY.resize(b->n);
    b->hostYd.resize(b->n);
  }
  return 0;
}

int fos_bdf_set_tol(fos_bdf *b, double rtol, double atol) {
  if (!b || !(rtol >= 0) || !(atol >= 0)) return -1;
  b->rtol = rtol;
  b->atol = atol;
  return 0;
}

int fos_bdf_set_max_steps(fos_bdf *b, int maxSteps) {
  (void)maxSteps;  // fixed-step integration: no step limit to enforce (setMaxSteps(-1) in foeuler)
  return b ? 0 : -1;
}

int fos_bdf_set_newton(fos_bdf *b, double tol, int maxIter, double linTol, int maxRestarts, int jacInterval) {
  if (!b) return -1;
  if (tol > 0) b->newtonTol = tol;
  if (maxIter > 0) b->maxNewton = maxIter;
  if (linTol > 0) b->linTol = linTol;
  if (maxRestarts >= 0) b->maxRestarts = maxRestarts;
  if (jacInterval > 0) b->msbj = jacInterval;
  return 0;
}

int fos_bdf_set_lsetup_frequency(fos_bdf *b, int msbp, int msbj, double dgmax) {
  if (!b) return -1;
  if (msbp > 0) b->msbp = msbp;
  if (msbj > 0) b->msbj = msbj;
  if (dgmax > 0) b->dgmax = dgmax;
  return 0;
}

int fos_bdf_set_fallback(fos_bdf *b, int gmresCap, fos_psetup_fn pSet, fos_psolve_fn pSol) {
  if (!b || (pSol && !pSet)) return -1;
  b->fbCap = gmresCap;
  b->fbSet = pSet;
  b->fbSol = pSol;
  return 0;
}

int fos_bdf_set_dt(fos_bdf *b, double dt) {
  if (!b || !(dt > 0)) return -1;
  if (!b->fixedStep && b->started) {
    // from an adaptive run in progress: fixed steps continue from its current state with BDF1
    FOS_HIP_TRY(hipMemcpyAsync(b->yn.p, b->znj(0), sizeof(double) * b->n, hipMemcpyDeviceToDevice, b->stream));
    FOS_HIP_TRY(hipStreamSynchronize(b->stream));
    b->order = 1;
  }
  // a new step size restarts with BDF1: the constant-step BDF2 formula needs y_{n-1} at t_n - dt
  if (b->fixedStep && dt != b->dt) b->order = 1;
  b->dt = dt;
  b->fixedStep = true;
  return 0;
}

int fos_bdf_set_adaptive(fos_bdf *b, int maxOrder, double h0, double hmin, double hmax) {
  if (!b || maxOrder > QMAX || !(h0 >= 0) || !(hmin >= 0) || !(hmax >= 0) || (hmax > 0 && hmin > hmax)) {
    set_error("fos_bdf_set_adaptive: bad arguments (maxOrder <= %d, h0, hmin, hmax >= 0, hmin <= hmax)", QMAX);
    return -1;
  }
  // from fixed steps (or before the first step) the adaptive history starts afresh from the current
  // state at the next step; during an adaptive run the order limit and initial step wait for the
  // next fos_bdf_set_iv, the step bounds apply from the next step
  if (b->fixedStep) b->started = false;
  b->fixedStep = false;
  if (maxOrder > 0) b->qmaxNext = maxOrder;
  if (!b->started) b->qmax = b->qmaxNext;
  b->h0User = h0;
  b->hmin = hmin;
  b->hmaxInv = hmax > 0 ? 1.0 / hmax : 0.0;
  return 0;
}

int fos_bdf_set_iv(fos_bdf *b, double t, const double *y0) {
  if (!b || !y0) return -1;
  b->t = t;
  hipMemcpyKind k = (b->flags & FOS_VECTORS_ON_HOST) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  FOS_HIP_TRY(hipMemcpyAsync(b->yn.p, y0, sizeof(double) * b->n, k, b->stream));
  FOS_HIP_TRY(hipMemcpyAsync(b->y.p, y0, sizeof(double) * b->n, k, b->stream));
  FOS_HIP_TRY(hipStreamSynchronize(b->stream));
  b->order = 1;
  b->nsteps = 0;
  b->nstP = b->nstJ = -1;  // a new integration starts with a fresh preconditioner setup
  b->started = false;      // adaptive mode: the history and initial step are built on the next step
  b->qmax = b->qmaxNext;
  b->hu = 0.0;
  return 0;
}

int fos_bdf_step(fos_bdf *b, double tOut, double *tRet, double *yOut) {
  if (!b) return -1;
  // one internal step per call (CV_ONE_STEP in SUNDIALS.f90:836); tOut only sizes the first
  // adaptive step
  if (!b->fixedStep) return b->step_adaptive(tOut, tRet, yOut);
  const size_t n = b->n;
  const double dt = b->dt;
  const dim3 gr = b->grid(), bl(VT);
  hipStream_t s = b->stream;
  // BDF coefficients for a constant step: y - a1 y_n - a2 y_{n-1} = gamma f(y)
  double a1, a2, gamma;
  if (b->order <= 1) { a1 = 1.0; a2 = 0.0; gamma = dt; }
  else { a1 = 4.0 / 3.0; a2 = -1.0 / 3.0; gamma = 2.0 * dt / 3.0; }
  const double tNew = b->t + dt;
  const double tSave = b->t;
  int newton = 0, lin = 0, rc = 0;
  double gn = 0.0, gnCv = 0.0;
  bool converged = false;
  // preconditioner setup for this step (CVODE's lsetup test); forced after a failure on stale data
  bool force = false;
  for (int attempt = 0; attempt < 2; attempt++) {
    const bool setupNow = b->pset && (force || b->nstP < 0 || b->nsteps - b->nstP >= b->msbp ||
                                      std::fabs(gamma / b->gammaP - 1.0) > b->dgmax);
    b->onFallback = false;
    newton = 0;
    converged = false;
    // predictor: linear extrapolation from the last two states (order 2), else the last state
    if (b->order >= 2) hipLaunchKernelGGL(k_lin2, gr, bl, 0, s, n, 2.0, b->yn.p, -1.0, b->ynm1.p, b->y.p);
    else FOS_HIP_TRY(hipMemcpyAsync(b->y.p, b->yn.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    // weights of this step from y_n (CVODE's ewt, with the rms floor): sum y_n^2 under unit weights
    if (b->dots<double>(nullptr, 0, b->yn.p, b->dres.p + 4, true, nullptr, 0.0, true)) return -1;
    hipLaunchKernelGGL(k_weights, gr, bl, 0, s, n, b->yn.p, b->rtol, b->atol, b->dres.p + 4, b->w2.p, b->w2f.p);
    b->t = tNew;
    for (; newton < b->maxNewton; newton++) {
      rc = b->call_rhs(tNew, b->y.p, b->fy.p);
      if (rc) break;
      // G = y - a1 y_n - a2 y_{n-1} - gamma f(y) (FP64); the buffer G receives -G, the right-hand side
      // of (I - gamma J) d = -G, together with ||G||_w2^2
      hipLaunchKernelGGL(k_bdf_residual, gr, bl, 0, s, n, a1, b->yn.p, a2, b->order >= 2 ? b->ynm1.p : b->yn.p,
                         gamma, b->fy.p, b->y.p, b->w2.p, b->rtol, b->atol, b->G.p, b->part.p,
                         newton == 0 ? b->ypred.p : nullptr, newton == 0 ? b->r0.p : nullptr);
      hipLaunchKernelGGL(k_reduce, dim3(1), dim3(VT), 0, s, b->part.p, b->nblocks(), 2, b->dres.p);
      double g2[2];
      if (b->fetch(b->dres.p, 2, g2)) return -1;
      gn = std::sqrt(g2[0] / (double)n);
      gnCv = std::sqrt(g2[1] / (double)n);
      if (!std::isfinite(gn)) { rc = 1; break; }
      if (gn <= b->newtonTol) { converged = true; break; }
      if (newton == 0 && setupNow) {
        int jcur = 0;
        const int jok = (force || b->nstJ < 0 || b->nsteps - b->nstJ >= b->msbj) ? 0 : 1;
        b->st.npe++;
        b->onFallback = false;
        rc = b->pset(tNew, b->y.p, b->fy.p, jok, &jcur, gamma, b->ud);
        if (rc) break;
        b->gammaP = gamma;
        b->nstP = b->nsteps;
        if (jcur) b->nstJ = b->nsteps;
      }
      // (I - gamma J) d = -G; the step's (predictor, Newton residual) pair for fos_bdf_last_triple
      // was written with G at the first iteration
      int it = 0;
      bool conv = false;
      // y += d inside GMRES, unless a fallback preconditioner may replace this solve's d
      double *acc = b->fbSol ? nullptr : b->y.p;
      rc = b->gmres(gamma, b->G.p, b->d.p, it, conv, 0.0, acc);
      lin += it;
      if (rc) break;
      if (!conv && b->fbSol && !b->onFallback) {
        // the primary preconditioner did not reach the linear tolerance within the cap: set up the
        // fallback at the current iterate and solve again; it stays in use for the rest of the step
        int jcur = 0;
        b->st.nfallback++;
        b->onFallback = true;
        rc = b->fbSet(tNew, b->y.p, b->fy.p, 0, &jcur, gamma, b->ud);
        if (rc) break;
        rc = b->gmres(gamma, b->G.p, b->d.p, it, conv);
        lin += it;
        if (rc) break;
      }
      if (!acc) hipLaunchKernelGGL(k_lin2, gr, bl, 0, s, n, 1.0, b->y.p, 1.0, b->d.p, b->y.p);
      FOS_HIP_TRY(hipGetLastError());
    }
    if (converged || rc < 0 || setupNow || !b->pset) break;
    // Newton failed with a preconditioner set up at an earlier step: set up afresh and retry
    b->st.nncfails++;
    b->st.nni += newton;
    b->t = tSave;
    force = true;
  }
  // the converged ||G|| under CVODE's unfloored weights 1/(rtol |y_n| + atol), for the record
  if (converged) b->lastGnormCv = gnCv;
  b->lastNewton = newton;
  b->lastLin = lin;
  b->lastGnorm = gn;
  b->st.nni += newton;
  b->st.nli += lin;
  if (!converged) {
    b->t = tSave;
    b->st.nncfails++;
    if (rc < 0) return rc;
    // a callback's recoverable failure keeps its own message (e.g. a non-finite residual)
    if (rc == 0)
      set_error("bdf step: Newton did not reach %g (||G||_WRMS = %g after %d iterations)", b->newtonTol, gn, newton);
    return 1;
  }
  // accept: y_{n-1} <- y_n, y_n <- y by rotating the three buffers (the next predictor rewrites y)
  std::swap(b->ynm1.p, b->yn.p);
  std::swap(b->yn.p, b->y.p);
  b->order = 2;
  b->nsteps++;
  b->st.nst++;
  if (tRet) *tRet = b->t;
  if (yOut) {
    hipMemcpyKind k = (b->flags & FOS_VECTORS_ON_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    FOS_HIP_TRY(hipMemcpyAsync(yOut, b->yn.p, sizeof(double) * n, k, s));
  }
  FOS_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int fos_bdf_last_triple(fos_bdf *b, double *ypred, double *r, double *delta) {
  if (!b) return -1;
  const size_t n = b->n;
  hipStream_t s = b->stream;
  if (ypred) FOS_HIP_TRY(hipMemcpyAsync(ypred, b->ypred.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if (r) FOS_HIP_TRY(hipMemcpyAsync(r, b->r0.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  const double *cur = b->fixedStep ? b->yn.p : b->znj(0);
  if (delta) {
    if (!cur) { set_error("fos_bdf_last_triple: no step taken"); return -1; }
    hipLaunchKernelGGL(k_lin2, b->grid(), dim3(VT), 0, s, n, 1.0, cur, -1.0, b->ypred.p, delta);
  }
  FOS_HIP_TRY(hipGetLastError());
  FOS_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int fos_bdf_get_dt(fos_bdf *b, double *dt) {
  if (!b || !dt) return -1;
  *dt = b->fixedStep ? b->dt : b->hu;  // CVodeGetLastStep in adaptive mode
  return 0;
}

int fos_bdf_get_dky(fos_bdf *b, double t, int k, double *dky) {
  if (!b || !dky) return -1;
  if (b->fixedStep || !b->started || b->nsteps == 0) {
    set_error("fos_bdf_get_dky: needs an adaptive integration with at least one step");
    return -1;
  }
  if (k < 0 || k > b->q) {
    set_error("fos_bdf_get_dky: derivative order %d outside 0..%d", k, b->q);
    return -1;
  }
  // the history interpolates over the last step: t in [tn - hu, tn] (CVODE allows the same with a
  // roundoff margin)
  const double tn = b->t, tfuzz = 100.0 * 2.220446049250313e-16 * (std::fabs(tn) + std::fabs(b->hu));
  if (t < tn - b->hu - tfuzz || t > tn + tfuzz) {
    set_error("fos_bdf_get_dky: t = %g outside the last step [%g, %g]", t, tn - b->hu, tn);
    return -1;
  }
  // d^k y/dt^k (t) = sum_{j>=k} j!/(j-k)! s^(j-k) zn_j / h^k, s = (t - tn)/h
  const double h = b->h, sv = (t - tn) / h;
  NordCoef c{};
  for (int j = k; j <= b->q; j++) {
    double f = 1.0;
    for (int i = j - k + 1; i <= j; i++) f *= i;
    c.c[j] = f * std::pow(sv, j - k) / std::pow(h, k);
  }
  hipLaunchKernelGGL(k_nord_comb, b->grid(), dim3(VT), 0, b->stream, (size_t)b->n, b->q, c, b->zn.p, b->tmp.p);
  FOS_HIP_TRY(hipGetLastError());
  hipMemcpyKind kind = (b->flags & FOS_VECTORS_ON_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  FOS_HIP_TRY(hipMemcpyAsync(dky, b->tmp.p, sizeof(double) * b->n, kind, b->stream));
  FOS_HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

int fos_bdf_get_stats(fos_bdf *b, fos_bdf_stats *out) {
  if (!b || !out) return -1;
  *out = b->st;
  out->lastNewtonIters = b->lastNewton;
  out->lastLinIters = b->lastLin;
  out->lastResidualNorm = b->lastGnorm;
  out->lastResidualNormCvode = b->fixedStep ? b->lastGnormCv : b->lastGnorm;
  out->lastOrder = b->fixedStep ? (b->nsteps > 1 ? 2 : 1) : b->qu;
  out->nextOrder = b->fixedStep ? b->order : (b->hprime != b->h ? b->qprime : b->q);
  out->lastStep = b->fixedStep ? b->dt : b->hu;
  out->nextStep = b->fixedStep ? b->dt : b->hprime;
  return 0;
}

}  // extern "C"
