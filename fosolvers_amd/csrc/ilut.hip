// ILUT on a general sparse matrix behind the interface of the reference's ILU type
// (src/libfosolvers/numerics/sparse.f90:52-63: init :334-347, setCOO :249-266, fact :370-383,
// solve :386-392; linEq%setCSR / mulVec :231-247, :269-275), the preconditioner fopbc factors with
// lfil 50 and drop tolerance 1e-6 (src/fopbc/modPbc.f90:21,647-739).
//
// Design:
//  * The factorisation runs on the host: ILUT (Saad's dual-threshold incomplete LU) decides its
//    fill row by row from the values it has just produced, so there is no dependency structure to
//    schedule before it runs.  The kept entries, their storage order and the replaced zero pivots
//    follow SPARSKIT's ilut (ITSOL/ilut.f:38-381) -- the elimination in ascending column order,
//    the multiplier dropped below droptol, the U entries dropped below droptol times the row's mean
//    magnitude, the lfil largest of each kept by quick-split selection, U keeping at most lfil - 1
//    off-diagonal entries, an inverted diagonal -- so a host that compares with the reference's
//    factor sees the same one (fos_ilut_get_msr returns it in SPARSKIT's MSR arrays).
//  * The triangular solves (lusol, ITSOL/ilut.f:2272-2321) run on the GPU, level-scheduled: the
//    rows of one level depend only on rows of earlier levels, 16 lanes per row form its dot
//    product and a shuffle tree sums it; one launch per level, recorded once as a graph.
//  * The matrix-vector product (amux) runs on the GPU, 16 lanes per row as well.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/fosolvers_amd.h"
#include "hip_util.h"

namespace fos {
namespace {

constexpr int ROW_LANES = 16;  // lanes per row in the sweeps and the product
constexpr int SOLVE_THREADS = 256;

// Partial selection by quick-split (Saad, SPARSKIT ITSOL qsplit): reorders v[0..len) and c alongside
// so that the ncut entries of largest magnitude occupy the first ncut places (the ncut-th at its
// sorted place).  The partition steps are SPARSKIT's, so ties at the cut resolve the same way.
void quick_split(double *v, int *c, int len, int ncut) {
  int first = 0, last = len - 1;
  const int cut = ncut - 1;
  if (cut < first || cut > last) return;
  for (;;) {
    int mid = first;
    const double key = std::fabs(v[first]);
    for (int j = first + 1; j <= last; j++)
      if (std::fabs(v[j]) > key) {
        ++mid;
        std::swap(v[mid], v[j]);
        std::swap(c[mid], c[j]);
      }
    std::swap(v[mid], v[first]);
    std::swap(c[mid], c[first]);
    if (mid == cut) return;
    if (mid > cut) last = mid - 1;
    else first = mid + 1;
  }
}

// dot product of a row with x over 16 lanes: lane l takes entries l, l+16, ...; shuffle tree
__device__ inline double row_dot(const int *__restrict__ col, const double *__restrict__ val, int b, int e,
                                 const double *__restrict__ x, int lane) {
  double s = 0.0;
  for (int k = b + lane; k < e; k += ROW_LANES) s += val[k] * x[col[k]];
#pragma unroll
  for (int o = ROW_LANES / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, ROW_LANES);
  return s;
}

// forward sweep, one level: x_i -= sum_k L_ik x_k (unit lower triangle)
__global__ __launch_bounds__(SOLVE_THREADS) void ilut_fwd_kernel(const int *__restrict__ rows, int nrows,
                                                                 const int *__restrict__ ptr,
                                                                 const int *__restrict__ col,
                                                                 const double *__restrict__ val, double *x) {
  const int g = (blockIdx.x * SOLVE_THREADS + threadIdx.x) / ROW_LANES, lane = threadIdx.x % ROW_LANES;
  if (g >= nrows) return;  // whole 16-lane groups leave together
  const int i = rows[g];
  const double s = row_dot(col, val, ptr[i], ptr[i + 1], x, lane);
  if (lane == 0) x[i] -= s;
}

// backward sweep, one level: x_i = dinv_i (x_i - sum_k U_ik x_k)
__global__ __launch_bounds__(SOLVE_THREADS) void ilut_bwd_kernel(const int *__restrict__ rows, int nrows,
                                                                 const int *__restrict__ ptr,
                                                                 const int *__restrict__ col,
                                                                 const double *__restrict__ val,
                                                                 const double *__restrict__ dinv, double *x) {
  const int g = (blockIdx.x * SOLVE_THREADS + threadIdx.x) / ROW_LANES, lane = threadIdx.x % ROW_LANES;
  if (g >= nrows) return;
  const int i = rows[g];
  const double s = row_dot(col, val, ptr[i], ptr[i + 1], x, lane);
  if (lane == 0) x[i] = dinv[i] * (x[i] - s);
}

// y = A x (CSR)
__global__ __launch_bounds__(SOLVE_THREADS) void csr_mv_kernel(int n, const int *__restrict__ ptr,
                                                               const int *__restrict__ col,
                                                               const double *__restrict__ val,
                                                               const double *__restrict__ x, double *__restrict__ y) {
  const int i = (blockIdx.x * SOLVE_THREADS + threadIdx.x) / ROW_LANES, lane = threadIdx.x % ROW_LANES;
  if (i >= n) return;
  const double s = row_dot(col, val, ptr[i], ptr[i + 1], x, lane);
  if (lane == 0) y[i] = s;
}

unsigned groups_grid(int rows) {
  return (unsigned)std::max<long long>(1, ((long long)rows * ROW_LANES + SOLVE_THREADS - 1) / SOLVE_THREADS);
}

}  // namespace
}  // namespace fos

struct fos_ilut {
  int n = 0, maxNNz = 0, lfil = 0;
  unsigned flags = 0;
  hipStream_t stream = nullptr;
  // the matrix, CSR 0-based
  std::vector<int> ia, ja;
  std::vector<double> a;
  bool haveMatrix = false, factored = false;
  // the factor: strict L (unit diagonal) and strict U rows in storage order, the inverted diagonal
  std::vector<int> lPtr, lCol, uPtr, uCol;
  std::vector<double> lVal, uVal, dInv;
  // level schedules: rows grouped by level, forward and backward
  std::vector<int> fRows, fLev, bRows, bLev;
  // device copies (made at the first solve / product after a change)
  bool devFactor = false, devMatrix = false;
  fos::DBuf<int> dLPtr, dLCol, dUPtr, dUCol, dFRows, dBRows, dIa, dJa;
  fos::DBuf<double> dLVal, dUVal, dDInv, dA, dX, dY;
  fos::KernelChain chain;
};

namespace fos {
namespace {

int ilut_factor(fos_ilut *p, double droptol) {
  const int n = p->n, lfil = p->lfil;
  p->lPtr.assign(n + 1, 0);
  p->uPtr.assign(n + 1, 0);
  p->lCol.clear(); p->lVal.clear(); p->uCol.clear(); p->uVal.clear();
  p->dInv.assign(n, 0.0);
  // the working row: the L list (in selection order), the U list (original entries, then fill in
  // order of creation), the diagonal; pos[c] points at column c's entry (L index, or U index + n,
  // or -2 for the diagonal), -1 when absent
  std::vector<int> pos(n, -1), Lc, Uc, kc;
  std::vector<double> Lv, Uv, kv;
  Lc.reserve(n); Lv.reserve(n); Uc.reserve(n); Uv.reserve(n);
  for (int i = 0; i < n; i++) {
    const int r0 = p->ia[i], r1 = p->ia[i + 1];
    double tnorm = 0.0;
    for (int k = r0; k < r1; k++) tnorm += std::fabs(p->a[k]);
    if (tnorm == 0.0) {
      set_error("ILUT: row %d is zero", i);
      return -1;
    }
    tnorm /= (double)(float)(r1 - r0);  // the row's mean magnitude (SPARSKIT divides by a REAL count)
    Lc.clear(); Lv.clear(); Uc.clear(); Uv.clear();
    double diag = 0.0;
    pos[i] = -2;
    for (int k = r0; k < r1; k++) {
      const int c = p->ja[k];
      if (c < i) { pos[c] = (int)Lc.size(); Lc.push_back(c); Lv.push_back(p->a[k]); }
      else if (c == i) diag = p->a[k];
      else { pos[c] = n + (int)Uc.size(); Uc.push_back(c); Uv.push_back(p->a[k]); }
    }
    // eliminate the L entries in ascending column order; kept multipliers are written over the
    // front of the L list as they are produced
    int kept = 0;
    for (int jj = 0; jj < (int)Lc.size(); jj++) {
      int at = jj;
      for (int j = jj + 1; j < (int)Lc.size(); j++)
        if (Lc[j] < Lc[at]) at = j;
      if (at != jj) {
        pos[Lc[at]] = jj;
        pos[Lc[jj]] = at;
        std::swap(Lc[at], Lc[jj]);
        std::swap(Lv[at], Lv[jj]);
      }
      const int k = Lc[jj];
      pos[k] = -1;
      const double m = Lv[jj] * p->dInv[k];
      if (std::fabs(m) <= droptol) continue;
      for (int q = p->uPtr[k]; q < p->uPtr[k + 1]; q++) {
        const int c = p->uCol[q];
        const double s = m * p->uVal[q];
        const int at2 = pos[c];
        if (c == i) diag -= s;
        else if (c > i) {
          if (at2 < 0) { pos[c] = n + (int)Uc.size(); Uc.push_back(c); Uv.push_back(-s); }
          else Uv[at2 - n] -= s;
        } else {
          if (at2 < 0) { pos[c] = (int)Lc.size(); Lc.push_back(c); Lv.push_back(-s); }
          else Lv[at2] -= s;
        }
      }
      Lc[kept] = k;
      Lv[kept] = m;
      kept++;
    }
    pos[i] = -1;
    for (int c : Uc) pos[c] = -1;
    // L: the lfil largest multipliers
    const int nl = std::min(kept, lfil);
    quick_split(Lv.data(), Lc.data(), kept, nl);
    p->lCol.insert(p->lCol.end(), Lc.begin(), Lc.begin() + nl);
    p->lVal.insert(p->lVal.end(), Lv.begin(), Lv.begin() + nl);
    p->lPtr[i + 1] = (int)p->lCol.size();
    // U: drop below droptol * tnorm, then the lfil - 1 largest of the rest
    kc.clear(); kv.clear();
    for (size_t q = 0; q < Uc.size(); q++)
      if (std::fabs(Uv[q]) > droptol * tnorm) { kc.push_back(Uc[q]); kv.push_back(Uv[q]); }
    const int lenu = (int)kc.size() + 1, ncut = std::min(lenu, lfil);
    quick_split(kv.data(), kc.data(), lenu - 1, ncut);
    const int nu = std::max(ncut - 1, 0);
    p->uCol.insert(p->uCol.end(), kc.begin(), kc.begin() + nu);
    p->uVal.insert(p->uVal.end(), kv.begin(), kv.begin() + nu);
    p->uPtr[i + 1] = (int)p->uCol.size();
    if (diag == 0.0) diag = (0.0001 + droptol) * tnorm;  // a zero pivot is replaced, as ilut does
    p->dInv[i] = 1.0 / diag;
  }
  // level schedules: forward, row i after the rows its L entries name; backward, after its U rows
  std::vector<int> lev(n, 0);
  auto group = [&](std::vector<int> &rows, std::vector<int> &levPtr) {
    const int nl = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    levPtr.assign(nl + 1, 0);
    for (int i = 0; i < n; i++) levPtr[lev[i] + 1]++;
    for (int l = 0; l < nl; l++) levPtr[l + 1] += levPtr[l];
    rows.assign(n, 0);
    std::vector<int> at(levPtr.begin(), levPtr.end() - 1);
    for (int i = 0; i < n; i++) rows[at[lev[i]]++] = i;
  };
  for (int i = 0; i < n; i++) {
    int l = 0;
    for (int q = p->lPtr[i]; q < p->lPtr[i + 1]; q++) l = std::max(l, lev[p->lCol[q]] + 1);
    lev[i] = l;
  }
  group(p->fRows, p->fLev);
  for (int i = n - 1; i >= 0; i--) {
    int l = 0;
    for (int q = p->uPtr[i]; q < p->uPtr[i + 1]; q++) l = std::max(l, lev[p->uCol[q]] + 1);
    lev[i] = l;
  }
  group(p->bRows, p->bLev);
  return 0;
}

int upload_factor(fos_ilut *p) {
  if (p->devFactor) return 0;
  hipStream_t s = p->stream;
  if (p->dLPtr.upload(p->lPtr.data(), p->lPtr.size(), s) || p->dLCol.upload(p->lCol.data(), p->lCol.size(), s) ||
      p->dLVal.upload(p->lVal.data(), p->lVal.size(), s) || p->dUPtr.upload(p->uPtr.data(), p->uPtr.size(), s) ||
      p->dUCol.upload(p->uCol.data(), p->uCol.size(), s) || p->dUVal.upload(p->uVal.data(), p->uVal.size(), s) ||
      p->dDInv.upload(p->dInv.data(), p->dInv.size(), s) || p->dFRows.upload(p->fRows.data(), p->fRows.size(), s) ||
      p->dBRows.upload(p->bRows.data(), p->bRows.size(), s) || p->dX.alloc(p->n))
    return -1;
  // the sweeps as one graph of kernel nodes: forward levels, then backward levels, in place on dX
  p->chain.reset();
  for (size_t l = 0; l + 1 < p->fLev.size(); l++) {
    const int b = p->fLev[l], cnt = p->fLev[l + 1] - b;
    if (p->chain.add(ilut_fwd_kernel, dim3(groups_grid(cnt)), dim3(SOLVE_THREADS), 0, p->dFRows.p + b, cnt,
                     p->dLPtr.p, p->dLCol.p, p->dLVal.p, p->dX.p))
      break;
  }
  for (size_t l = 0; l + 1 < p->bLev.size() && !p->chain.err; l++) {
    const int b = p->bLev[l], cnt = p->bLev[l + 1] - b;
    if (p->chain.add(ilut_bwd_kernel, dim3(groups_grid(cnt)), dim3(SOLVE_THREADS), 0, p->dBRows.p + b, cnt,
                     p->dUPtr.p, p->dUCol.p, p->dUVal.p, p->dDInv.p, p->dX.p))
      break;
  }
  if (p->chain.err != hipSuccess || p->chain.instantiate()) {
    set_error("ILUT solve: graph of the sweeps could not be built (%s)", p->chain.fail());
    return -1;
  }
  p->devFactor = true;
  return 0;
}

int upload_matrix(fos_ilut *p) {
  if (p->devMatrix) return 0;
  if (p->dIa.upload(p->ia.data(), p->ia.size(), p->stream) || p->dJa.upload(p->ja.data(), p->ja.size(), p->stream) ||
      p->dA.upload(p->a.data(), p->a.size(), p->stream) || p->dX.alloc(p->n) || p->dY.alloc(p->n))
    return -1;
  p->devMatrix = true;
  return 0;
}

int set_csr(fos_ilut *p, std::vector<int> ia, std::vector<int> ja, std::vector<double> a) {
  p->ia.swap(ia);
  p->ja.swap(ja);
  p->a.swap(a);
  p->haveMatrix = true;
  p->factored = false;
  p->devFactor = false;
  p->devMatrix = false;
  return 0;
}

}  // namespace
}  // namespace fos

using namespace fos;

extern "C" {

fos_ilut *fos_ilut_create(int nEq, int maxNNz, int lFill, unsigned flags, void *stream) {
  if (nEq <= 0 || maxNNz < 0 || lFill < 0) {
    set_error("fos_ilut_create: nEq > 0, maxNNz >= 0 and lFill >= 0 required (%d, %d, %d)", nEq, maxNNz, lFill);
    return nullptr;
  }
  fos_ilut *p = new fos_ilut();
  p->n = nEq;
  p->maxNNz = maxNNz;
  p->lfil = std::min(lFill, nEq - 1);  // initILU: lFill = min(lFill, nEq - 1)
  p->flags = flags;
  p->stream = static_cast<hipStream_t>(stream);
  return p;
}

void fos_ilut_destroy(fos_ilut *p) {
  if (!p) return;
  if (p->devFactor || p->devMatrix) (void)hipStreamSynchronize(p->stream);
  delete p;
}

int fos_ilut_set_coo(fos_ilut *p, const int *iA, const int *jA, const double *A, int nNz) {
  if (!p || nNz < 0 || (nNz > 0 && (!iA || !jA || !A))) { set_error("fos_ilut_set_coo: bad arguments"); return -1; }
  if (nNz > p->maxNNz) { set_error("fos_ilut_set_coo: %d entries exceed maxNNz %d", nNz, p->maxNNz); return -1; }
  const int n = p->n;
  for (int k = 0; k < nNz; k++)
    if (iA[k] < 0 || iA[k] >= n || jA[k] < 0 || jA[k] >= n) {
      set_error("fos_ilut_set_coo: entry %d at (%d, %d) outside the %d x %d matrix", k, iA[k], jA[k], n, n);
      return -1;
    }
  // coocsr's placement: rows by a counting pass, entries of a row in input order, duplicates kept
  std::vector<int> ia(n + 1, 0), ja(nNz);
  std::vector<double> a(nNz);
  for (int k = 0; k < nNz; k++) ia[iA[k] + 1]++;
  for (int i = 0; i < n; i++) ia[i + 1] += ia[i];
  std::vector<int> at(ia.begin(), ia.end() - 1);
  for (int k = 0; k < nNz; k++) {
    const int q = at[iA[k]]++;
    ja[q] = jA[k];
    a[q] = A[k];
  }
  return set_csr(p, std::move(ia), std::move(ja), std::move(a));
}

int fos_ilut_set_csr(fos_ilut *p, const int *ia, const int *ja, const double *a) {
  if (!p || !ia || !ja || !a) { set_error("fos_ilut_set_csr: bad arguments"); return -1; }
  const int n = p->n, nnz = ia[n] - ia[0];
  if (ia[0] != 0 || nnz < 0 || nnz > p->maxNNz) {
    set_error("fos_ilut_set_csr: row pointers must start at 0 and hold at most maxNNz %d entries", p->maxNNz);
    return -1;
  }
  for (int i = 0; i < n; i++)
    if (ia[i + 1] < ia[i]) { set_error("fos_ilut_set_csr: row pointer %d decreases", i + 1); return -1; }
  for (int k = 0; k < nnz; k++)
    if (ja[k] < 0 || ja[k] >= n) { set_error("fos_ilut_set_csr: column %d of entry %d outside [0, %d)", ja[k], k, n); return -1; }
  return set_csr(p, std::vector<int>(ia, ia + n + 1), std::vector<int>(ja, ja + nnz), std::vector<double>(a, a + nnz));
}

int fos_ilut_fact(fos_ilut *p, double droptol) {
  if (!p || !p->haveMatrix) { set_error("fos_ilut_fact: no matrix set"); return -1; }
  if (!(droptol >= 0.0)) { set_error("fos_ilut_fact: drop tolerance %g < 0", droptol); return -1; }
  p->factored = false;
  p->devFactor = false;
  if (ilut_factor(p, droptol)) return -1;
  p->factored = true;
  return 0;
}

int fos_ilut_solve(fos_ilut *p, const double *rhs, double *x) {
  if (!p || !rhs || !x) { set_error("fos_ilut_solve: bad arguments"); return -1; }
  if (!p->factored) { set_error("fos_ilut_solve: called before fos_ilut_fact"); return -1; }
  if (upload_factor(p)) return -1;
  const size_t bytes = sizeof(double) * p->n;
  const bool host = p->flags & FOS_VECTORS_ON_HOST;
  FOS_HIP_TRY(hipMemcpyAsync(p->dX.p, rhs, bytes, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, p->stream));
  if (p->chain.launch(p->stream)) {
    set_error("fos_ilut_solve: graph launch failed (%s)", hipGetErrorString(p->chain.err));
    return -1;
  }
  FOS_HIP_TRY(hipMemcpyAsync(x, p->dX.p, bytes, host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, p->stream));
  if (host) FOS_HIP_TRY(hipStreamSynchronize(p->stream));
  return 0;
}

int fos_ilut_mul_vec(fos_ilut *p, const double *x, double *y) {
  if (!p || !x || !y) { set_error("fos_ilut_mul_vec: bad arguments"); return -1; }
  if (!p->haveMatrix) { set_error("fos_ilut_mul_vec: no matrix set"); return -1; }
  if (upload_matrix(p)) return -1;
  const bool host = p->flags & FOS_VECTORS_ON_HOST;
  const size_t bytes = sizeof(double) * p->n;
  const double *xd = x;
  double *yd = y;
  if (host) {
    FOS_HIP_TRY(hipMemcpyAsync(p->dX.p, x, bytes, hipMemcpyHostToDevice, p->stream));
    xd = p->dX.p;
    yd = p->dY.p;
  }
  hipLaunchKernelGGL(csr_mv_kernel, dim3(groups_grid(p->n)), dim3(SOLVE_THREADS), 0, p->stream, p->n, p->dIa.p,
                     p->dJa.p, p->dA.p, xd, yd);
  FOS_HIP_TRY(hipGetLastError());
  if (host) {
    FOS_HIP_TRY(hipMemcpyAsync(y, p->dY.p, bytes, hipMemcpyDeviceToHost, p->stream));
    FOS_HIP_TRY(hipStreamSynchronize(p->stream));
  }
  return 0;
}

int fos_ilut_get_msr(const fos_ilut *p, double *alu, int *jlu, int *ju) {
  if (!p || !p->factored) { set_error("fos_ilut_get_msr: no factor"); return -1; }
  const int n = p->n, total = n + 1 + (int)p->lCol.size() + (int)p->uCol.size();
  if (!alu && !jlu && !ju) return total;
  if (!alu || !jlu || !ju) { set_error("fos_ilut_get_msr: arrays missing"); return -1; }
  // SPARSKIT's MSR, 1-based values, Fortran index k at C index k - 1: alu(1..n) the inverted
  // diagonal, jlu(1..n+1) row starts, then per row its L entries, ju(i) the start of its U entries
  int q = n + 1;  // C index of Fortran position n + 2
  for (int i = 0; i < n; i++) {
    alu[i] = p->dInv[i];
    jlu[i] = q + 1;
    for (int k = p->lPtr[i]; k < p->lPtr[i + 1]; k++, q++) { alu[q] = p->lVal[k]; jlu[q] = p->lCol[k] + 1; }
    ju[i] = q + 1;
    for (int k = p->uPtr[i]; k < p->uPtr[i + 1]; k++, q++) { alu[q] = p->uVal[k]; jlu[q] = p->uCol[k] + 1; }
  }
  jlu[n] = q + 1;
  alu[n] = 0.0;
  return total;
}

int fos_ilut_levels(const fos_ilut *p, int *nForward, int *nBackward) {
  if (!p || !p->factored) { set_error("fos_ilut_levels: no factor"); return -1; }
  if (nForward) *nForward = (int)p->fLev.size() - 1;
  if (nBackward) *nBackward = (int)p->bLev.size() - 1;
  return 0;
}

}  // extern "C"
