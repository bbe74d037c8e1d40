// polyFvGrid%up on the GPU (src/libfosolvers/polyGrid/polyFvGrid.f90:48-57): the same fos_fv_grid
// as the host build in mesh.cpp -- element order, centroids, volumes, face pairs, neighbour table,
// pair order and pair geometry -- with the face matching done by device radix sorts.
//
//   1. elements gathered in the sorted order (cells, facets, others, each stable;
//      polyGrid.f90:88-136), centroids and cell volumes (geometry.f90:61-91);
//   2. one record per cell or facet face: its node indices sorted ascending as a two-word key,
//      the value (element, face); two stable LSD sorts (low word, then high word) group equal
//      faces with their records in (element, face) order;
//   3. each run of equal keys becomes pairs as in getNeib (polyFvGrid.f90:60-198): every cell
//      face that matches facets pairs with each of them (boundary pairs), otherwise exactly two
//      cells make one internal pair (m = lower element); more than two cells is an error;
//   4. pairs sorted by (max(m, n), m, face) -- boundary pairs fall at the tail
//      (polyFvGrid.f90:201-227) -- and their area, unit normal and centre computed from the face
//      nodes of cell m (getINF, polyGrid.f90:211-278; geometry.f90:17-58).
// Floating-point contraction is off in this file, so every value equals the host build's.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/fosolvers_amd.h"
#include "common.h"
#include "hip_util.h"

#pragma clang fp contract(off)

namespace fos {
namespace {

constexpr int TB = 256;

__device__ inline int d_nfaces(int s) {
  switch (s) {
    case FOS_TRI: case FOS_QUAD: case FOS_TRI6: return 1;
    case FOS_TET: case FOS_TET10: return 4;
    case FOS_HEX: return 6;
    default: return 0;
  }
}
__device__ inline bool d_is_cell(int s) { return s == FOS_TET || s == FOS_HEX || s == FOS_TET10; }

// local node numbers of face f of shape s (the same tables as the host build); returns the count
__device__ inline int d_face_nodes(int s, int f, int *ind) {
  switch (s) {
    case FOS_TRI: ind[0] = 0; ind[1] = 1; ind[2] = 2; return 3;
    case FOS_QUAD: ind[0] = 0; ind[1] = 1; ind[2] = 2; ind[3] = 3; return 4;
    case FOS_TRI6: for (int k = 0; k < 6; k++) ind[k] = k; return 6;
    case FOS_TET: {
      const int t[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
      for (int k = 0; k < 3; k++) ind[k] = t[f][k];
      return 3;
    }
    case FOS_HEX: {
      const int h[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6}, {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
      for (int k = 0; k < 4; k++) ind[k] = h[f][k];
      return 4;
    }
    case FOS_TET10: {
      const int t[4][6] = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 9, 7}, {0, 3, 2, 7, 8, 6}, {1, 2, 3, 5, 8, 9}};
      for (int k = 0; k < 6; k++) ind[k] = t[f][k];
      return 6;
    }
    default: return 0;
  }
}

struct D3 {
  double x, y, z;
};
__device__ inline D3 d_sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 d_cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double d_dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 d_node(const double *pN, int k) { return {pN[3 * (size_t)k], pN[3 * (size_t)k + 1], pN[3 * (size_t)k + 2]}; }
__device__ inline double d_tet_volume(D3 a, D3 b, D3 c, D3 d) {
  return fabs(d_dot(d_sub(b, a), d_cross(d_sub(c, a), d_sub(d, a)))) / 6.0;
}
__device__ inline double d_tri_area(D3 a, D3 b, D3 c) {
  D3 n = d_cross(d_sub(b, a), d_sub(c, a));
  return sqrt(d_dot(n, n)) / 2.0;
}
__device__ inline D3 d_tri_normal(D3 a, D3 b, D3 c) {
  D3 n = d_cross(d_sub(b, a), d_sub(c, b));
  double l = sqrt(d_dot(n, n));
  return {n.x / l, n.y / l, n.z / l};
}

// elements in the sorted order with their centroids
__global__ void k_gather_elems(int nE, int M, const int *__restrict__ ord, const int *__restrict__ sE0,
                               const int *__restrict__ nNE0, const int *__restrict__ iNE0,
                               const double *__restrict__ pN, int *__restrict__ sE, int *__restrict__ nNE,
                               int *__restrict__ iNE, double *__restrict__ p) {
  const int k = blockIdx.x * TB + threadIdx.x;
  if (k >= nE) return;
  const int e = ord[k], nn = nNE0[e];
  sE[k] = sE0[e];
  nNE[k] = nn;
  double s[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < M; q++) {
    const int v = iNE0[(size_t)M * e + q];
    iNE[(size_t)M * k + q] = v;
    if (q < nn)
      for (int d = 0; d < 3; d++) s[d] += pN[3 * (size_t)v + d];
  }
  for (int d = 0; d < 3; d++) p[3 * (size_t)k + d] = s[d] / nn;
}

__global__ void k_volumes(int nC, int M, const int *__restrict__ sE, const int *__restrict__ iNE,
                          const double *__restrict__ pN, double *__restrict__ v) {
  const int c = blockIdx.x * TB + threadIdx.x;
  if (c >= nC) return;
  const int s = sE[c];
  const int *nd = iNE + (size_t)M * c;
  if (s == FOS_TET || s == FOS_TET10) {
    v[c] = d_tet_volume(d_node(pN, nd[0]), d_node(pN, nd[1]), d_node(pN, nd[2]), d_node(pN, nd[3]));
  } else if (s == FOS_HEX) {  // the hex split of geometry.f90:85-91
    const int sub6[6][4] = {{0, 1, 2, 5}, {0, 2, 3, 5}, {0, 3, 4, 5}, {3, 4, 5, 6}, {2, 3, 5, 6}, {3, 4, 6, 7}};
    double acc = 0.0;
    for (int t = 0; t < 6; t++)
      acc += d_tet_volume(d_node(pN, nd[sub6[t][0]]), d_node(pN, nd[sub6[t][1]]), d_node(pN, nd[sub6[t][2]]),
                          d_node(pN, nd[sub6[t][3]]));
    v[c] = acc;
  } else {
    v[c] = 0.0;
  }
}

// one record per cell/facet face: sorted node indices as (hi, lo) = (n0 n1, n2 n3), unused slots
// all ones; faces with more than 4 nodes (second-order) key on their first 4 sorted corner
// entries plus a check in the run pass
__global__ void k_face_keys(int nE, int M, const int *__restrict__ sE, const int *__restrict__ iNE,
                            const int *__restrict__ faceBeg, uint64_t *__restrict__ khi,
                            uint64_t *__restrict__ klo, int *__restrict__ val) {
  const int e = blockIdx.x * TB + threadIdx.x;
  if (e >= nE) return;
  const int s = sE[e], nf = d_nfaces(s);
  if (!nf) return;
  for (int f = 0; f < nf; f++) {
    int ind[6], k[6];
    const int cnt = d_face_nodes(s, f, ind);
    for (int q = 0; q < cnt; q++) k[q] = iNE[(size_t)M * e + ind[q]];
    for (int a = 1; a < cnt; a++)  // insertion sort of at most 6
      for (int b = a; b > 0 && k[b - 1] > k[b]; b--) { int t = k[b]; k[b] = k[b - 1]; k[b - 1] = t; }
    uint32_t w[4];
    for (int q = 0; q < 4; q++) w[q] = q < cnt ? (uint32_t)k[q] : 0xffffffffu;
    const int r = faceBeg[e] + f;
    khi[r] = ((uint64_t)w[0] << 32) | w[1];
    klo[r] = ((uint64_t)w[2] << 32) | w[3];
    val[r] = e * 8 + f;
  }
}

// full node-list equality of two face records (second-order faces carry more than 4 nodes)
__device__ inline bool d_same_face(int va, int vb, int M, const int *sE, const int *iNE) {
  int ia[6], ib[6], ka[6], kb[6];
  const int ea = va >> 3, eb = vb >> 3;
  const int ca = d_face_nodes(sE[ea], va & 7, ia), cb = d_face_nodes(sE[eb], vb & 7, ib);
  if (ca != cb) return false;
  if (ca <= 4) return true;  // the key already holds every node
  for (int q = 0; q < ca; q++) { ka[q] = iNE[(size_t)M * ea + ia[q]]; kb[q] = iNE[(size_t)M * eb + ib[q]]; }
  for (int a = 1; a < ca; a++)
    for (int b = a; b > 0 && ka[b - 1] > ka[b]; b--) { int t = ka[b]; ka[b] = ka[b - 1]; ka[b - 1] = t; }
  for (int a = 1; a < cb; a++)
    for (int b = a; b > 0 && kb[b - 1] > kb[b]; b--) { int t = kb[b]; kb[b] = kb[b - 1]; kb[b - 1] = t; }
  for (int q = 0; q < ca; q++)
    if (ka[q] != kb[q]) return false;
  return true;
}

__device__ inline bool d_key_eq(const uint64_t *khi, const uint64_t *klo, int a, int b) {
  return khi[a] == khi[b] && klo[a] == klo[b];
}

// pairs per run of equal keys (counted at the run's first record); err[0] = 1 for a face
// shared by more than two cells, err[1] = 1 for a run mixing different second-order faces
__global__ void k_run_count(int nF, int nC, int M, const uint64_t *__restrict__ khi, const uint64_t *__restrict__ klo,
                            const int *__restrict__ val, const int *__restrict__ sE, const int *__restrict__ iNE,
                            int *__restrict__ cnt, int *__restrict__ err) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i >= nF) return;
  if (i > 0 && d_key_eq(khi, klo, i, i - 1)) { cnt[i] = 0; return; }
  int c = 0, fa = 0, j = i;
  for (; j < nF && d_key_eq(khi, klo, j, i); j++) {
    if (!d_same_face(val[i], val[j], M, sE, iNE)) err[1] = 1;
    if ((val[j] >> 3) < nC) c++;
    else fa++;
  }
  if (fa == 0 && c > 2) err[0] = 1;
  cnt[i] = fa > 0 ? c * fa : (c == 2 ? 1 : 0);
}

__global__ void k_run_write(int nF, int nC, int maxNF, const uint64_t *__restrict__ khi,
                            const uint64_t *__restrict__ klo, const int *__restrict__ val, const int *__restrict__ cnt,
                            const int *__restrict__ off, int *__restrict__ pm, int *__restrict__ pn,
                            int *__restrict__ pf, int *__restrict__ neib) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i >= nF || cnt[i] == 0) return;
  int j = i;
  while (j < nF && d_key_eq(khi, klo, j, i)) j++;
  int o = off[i];
  bool facets = false;
  for (int q = i; q < j; q++) facets |= (val[q] >> 3) >= nC;
  if (facets) {
    for (int a = i; a < j; a++) {
      const int ea = val[a] >> 3, fa = val[a] & 7;
      if (ea >= nC) continue;
      for (int b = i; b < j; b++) {
        const int eb = val[b] >> 3;
        if (eb < nC) continue;
        pm[o] = ea; pn[o] = eb; pf[o] = fa;
        o++;
        neib[(size_t)maxNF * ea + fa] = eb;
      }
    }
  } else {  // exactly two cells, in element order
    const int ea = val[i] >> 3, fa = val[i] & 7, eb = val[i + 1] >> 3, fb = val[i + 1] & 7;
    pm[o] = ea; pn[o] = eb; pf[o] = fa;
    neib[(size_t)maxNF * ea + fa] = eb;
    neib[(size_t)maxNF * eb + fb] = ea;
  }
}

__global__ void k_pair_keys(int nP, int shift, const int *__restrict__ pm, const int *__restrict__ pn,
                            const int *__restrict__ pf, uint64_t *__restrict__ key, int *__restrict__ idx) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i >= nP) return;
  const uint64_t mx = (uint64_t)max(pm[i], pn[i]);
  key[i] = (mx << shift) | ((uint64_t)pm[i] << 3) | (uint64_t)pf[i];
  idx[i] = i;
}

__global__ void k_pair_geom(int nP, int M, const int *__restrict__ perm, const int *__restrict__ pm,
                            const int *__restrict__ pn, const int *__restrict__ pf, const int *__restrict__ sE,
                            const int *__restrict__ iNE, const double *__restrict__ pN, int *__restrict__ iEP,
                            double *__restrict__ aP, double *__restrict__ normP, double *__restrict__ pP) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i >= nP) return;
  const int k = perm[i], m = pm[k];
  iEP[2 * (size_t)i] = m;
  iEP[2 * (size_t)i + 1] = pn[k];
  int ind[6];
  const int cnt = d_face_nodes(sE[m], pf[k], ind);
  D3 q[4];
  for (int t = 0; t < 4 && t < cnt; t++) q[t] = d_node(pN, iNE[(size_t)M * m + ind[t]]);
  double a = 0.0;
  D3 n{0, 0, 0}, c{0, 0, 0};
  if (cnt == 3) {
    a = d_tri_area(q[0], q[1], q[2]);
    n = d_tri_normal(q[0], q[1], q[2]);
    c = {(q[0].x + q[1].x + q[2].x) / 3.0, (q[0].y + q[1].y + q[2].y) / 3.0, (q[0].z + q[1].z + q[2].z) / 3.0};
  } else if (cnt == 4) {
    a = d_tri_area(q[0], q[1], q[2]) + d_tri_area(q[2], q[3], q[0]);
    D3 n1 = d_tri_normal(q[0], q[1], q[2]), n2 = d_tri_normal(q[2], q[3], q[0]);
    n = {n1.x + n2.x, n1.y + n2.y, n1.z + n2.z};
    double l = sqrt(d_dot(n, n));
    n = {n.x / l, n.y / l, n.z / l};
    c = {(q[0].x + q[1].x + q[2].x + q[3].x) / 4.0, (q[0].y + q[1].y + q[2].y + q[3].y) / 4.0,
         (q[0].z + q[1].z + q[2].z + q[3].z) / 4.0};
  }  // other node counts: zero area, normal and centre (polyFvGrid.f90:135-138)
  aP[i] = a;
  normP[3 * (size_t)i] = n.x; normP[3 * (size_t)i + 1] = n.y; normP[3 * (size_t)i + 2] = n.z;
  pP[3 * (size_t)i] = c.x; pP[3 * (size_t)i + 1] = c.y; pP[3 * (size_t)i + 2] = c.z;
}

__global__ void k_iota(int n, int *__restrict__ a) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i < n) a[i] = i;
}
__global__ void k_gather_u64(int n, const uint64_t *__restrict__ a, const int *__restrict__ ix,
                             uint64_t *__restrict__ b) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i < n) b[i] = a[ix[i]];
}
__global__ void k_gather_i32(int n, const int *__restrict__ a, const int *__restrict__ ix, int *__restrict__ b) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i < n) b[i] = a[ix[i]];
}

int h_nfaces(int s) {
  switch (s) {
    case FOS_TRI: case FOS_QUAD: case FOS_TRI6: return 1;
    case FOS_TET: case FOS_TET10: return 4;
    case FOS_HEX: return 6;
    default: return 0;
  }
}
bool h_is_cell(int s) { return s == FOS_TET || s == FOS_HEX || s == FOS_TET10; }
bool h_is_facet(int s) { return s == FOS_TRI || s == FOS_QUAD || s == FOS_TRI6; }
int bits_for(long long v) {
  int b = 0;
  while (b < 63 && (1ll << b) <= v) b++;
  return b;
}
unsigned grid_of(long long n) { return (unsigned)std::max<long long>(1, (n + TB - 1) / TB); }

template <class T>
int d2h(T *dst, const DBuf<T> &src, size_t n, hipStream_t s) {
  if (n) FOS_HIP_TRY(hipMemcpyAsync(dst, src.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
  return 0;
}

int grid_up_device(const fos_poly_grid *in, fos_fv_grid *out, hipStream_t s) {
  const int nE = in->nE, M = in->maxNNE, nN = in->nN;
  // element order and face-record offsets (host, O(nE))
  std::vector<int> ord;
  ord.reserve(nE);
  for (int pass = 0; pass < 3; pass++)
    for (int e = 0; e < nE; e++) {
      const int sh = in->sE[e];
      const bool take = pass == 0 ? h_is_cell(sh) : pass == 1 ? h_is_facet(sh) : !(h_is_cell(sh) || h_is_facet(sh));
      if (take) ord.push_back(e);
    }
  int nC = 0, nF = 0, maxNF = 0;
  std::vector<int> faceBeg(nE + 1, 0);
  long long nRec = 0;
  for (int k = 0; k < nE; k++) {
    const int sh = in->sE[ord[k]];
    if (h_is_cell(sh)) nC++;
    else if (h_is_facet(sh)) nF++;
    maxNF = std::max(maxNF, h_nfaces(sh));
    faceBeg[k] = (int)std::min<long long>(nRec, INT32_MAX);
    nRec += (h_is_cell(sh) || h_is_facet(sh)) ? h_nfaces(sh) : 0;
  }
  faceBeg[nE] = (int)std::min<long long>(nRec, INT32_MAX);
  if (nRec >= (1ll << 31) || (long long)nE * 8 >= (1ll << 31)) {
    set_error("fos_grid_up_gpu: mesh too large for 32-bit face records");
    return -1;
  }
  const int shift = bits_for(std::max(nC, 1)) + 3;
  if (bits_for(nE) + shift > 64) {
    set_error("fos_grid_up_gpu: mesh too large for the 64-bit pair key");
    return -1;
  }
  const int F = (int)nRec;
  out->nN = nN;
  out->nE = nE;
  out->nC = nC;
  out->nF = nF;
  out->maxNF = maxNF;
  out->maxNNE = M;
  // 1. elements, centroids, volumes
  DBuf<double> dPN, dP, dV;
  DBuf<int> dOrd, dSE0, dNNE0, dINE0, dSE, dNNE, dINE, dFaceBeg;
  if (dPN.upload(in->pN, 3 * (size_t)nN, s) || dOrd.upload(ord.data(), nE, s) || dSE0.upload(in->sE, nE, s) ||
      dNNE0.upload(in->nNE, nE, s) || dINE0.upload(in->iNE, (size_t)M * nE, s) || dSE.alloc(nE) || dNNE.alloc(nE) ||
      dINE.alloc((size_t)M * nE) || dP.alloc(3 * (size_t)nE) || dV.alloc(nC) ||
      dFaceBeg.upload(faceBeg.data(), nE + 1, s))
    return -1;
  if (nE > 0)
    hipLaunchKernelGGL(k_gather_elems, dim3(grid_of(nE)), dim3(TB), 0, s, nE, M, dOrd.p, dSE0.p, dNNE0.p, dINE0.p, dPN.p,
                       dSE.p, dNNE.p, dINE.p, dP.p);
  if (nC > 0) hipLaunchKernelGGL(k_volumes, dim3(grid_of(nC)), dim3(TB), 0, s, nC, M, dSE.p, dINE.p, dPN.p, dV.p);
  FOS_HIP_TRY(hipGetLastError());
  // 2. face records sorted by (hi, lo); records of equal keys stay in (element, face) order
  DBuf<uint64_t> dKhi, dKlo, dK2;
  DBuf<int> dVal, dIdx, dIdx2, dCnt, dOff, dErr, dNeib, dPm, dPn, dPf;
  DBuf<char> tmp;
  if (dKhi.alloc(F) || dKlo.alloc(F) || dK2.alloc(F) || dVal.alloc(F) || dIdx.alloc(F) || dIdx2.alloc(F) ||
      dCnt.alloc(F) || dOff.alloc(F) || dErr.alloc(2) || dNeib.alloc((size_t)std::max(maxNF, 1) * nC))
    return -1;
  FOS_HIP_TRY(hipMemsetAsync(dErr.p, 0, 2 * sizeof(int), s));
  if (nC > 0) FOS_HIP_TRY(hipMemsetAsync(dNeib.p, 0xff, sizeof(int) * (size_t)maxNF * nC, s));  // -1: none
  int nP = 0;
  if (F > 0) {
    hipLaunchKernelGGL(k_face_keys, dim3(grid_of(nE)), dim3(TB), 0, s, nE, M, dSE.p, dINE.p, dFaceBeg.p, dKhi.p,
                       dKlo.p, dVal.p);
    hipLaunchKernelGGL(k_iota, dim3(grid_of(F)), dim3(TB), 0, s, F, dIdx.p);
    FOS_HIP_TRY(hipGetLastError());
    size_t tb1 = 0, tb2 = 0;
    FOS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, dKlo.p, dK2.p, dIdx.p, dIdx2.p, F, 0, 64, s));
    FOS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, dCnt.p, dOff.p, F, s));
    if (tmp.alloc(std::max<size_t>(std::max(tb1, tb2), 1))) return -1;
    // lo: records ordered by lo (stable) -> dIdx2
    FOS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb1, dKlo.p, dK2.p, dIdx.p, dIdx2.p, F, 0, 64, s));
    // hi in that order, then a stable sort on hi carries the records -> dIdx
    hipLaunchKernelGGL(k_gather_u64, dim3(grid_of(F)), dim3(TB), 0, s, F, dKhi.p, dIdx2.p, dK2.p);
    FOS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb1, dK2.p, dKhi.p, dIdx2.p, dIdx.p, F, 0, 64, s));
    // lo and the records in the final order (dKhi already is)
    hipLaunchKernelGGL(k_gather_u64, dim3(grid_of(F)), dim3(TB), 0, s, F, dKlo.p, dIdx.p, dK2.p);
    hipLaunchKernelGGL(k_gather_i32, dim3(grid_of(F)), dim3(TB), 0, s, F, dVal.p, dIdx.p, dIdx2.p);
    FOS_HIP_TRY(hipGetLastError());
    const uint64_t *khi = dKhi.p, *klo = dK2.p;
    const int *val = dIdx2.p;
    // 3. runs -> pairs
    hipLaunchKernelGGL(k_run_count, dim3(grid_of(F)), dim3(TB), 0, s, F, nC, M, khi, klo, val, dSE.p, dINE.p, dCnt.p,
                       dErr.p);
    FOS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb2, dCnt.p, dOff.p, F, s));
    int last[2], errs[2];
    FOS_HIP_TRY(hipMemcpyAsync(&last[0], dOff.p + F - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FOS_HIP_TRY(hipMemcpyAsync(&last[1], dCnt.p + F - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FOS_HIP_TRY(hipMemcpyAsync(errs, dErr.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    FOS_HIP_TRY(hipStreamSynchronize(s));
    if (errs[0]) { set_error("fos_grid_up_gpu: non-manifold face shared by more than two cells"); return -1; }
    if (errs[1]) { set_error("fos_grid_up_gpu: second-order faces with equal corners but different nodes"); return -1; }
    nP = last[0] + last[1];
    if (dPm.alloc(nP) || dPn.alloc(nP) || dPf.alloc(nP)) return -1;
    hipLaunchKernelGGL(k_run_write, dim3(grid_of(F)), dim3(TB), 0, s, F, nC, maxNF, khi, klo, val, dCnt.p, dOff.p,
                       dPm.p, dPn.p, dPf.p, dNeib.p);
    FOS_HIP_TRY(hipGetLastError());
  }
  // 4. pair order by (max(m, n), m, face) and pair geometry
  DBuf<uint64_t> dPK, dPK2;
  DBuf<int> dPI, dPI2, dIEP;
  DBuf<double> dAP, dNorm, dPP;
  if (dPK.alloc(nP) || dPK2.alloc(nP) || dPI.alloc(nP) || dPI2.alloc(nP) || dIEP.alloc(2 * (size_t)nP) ||
      dAP.alloc(nP) || dNorm.alloc(3 * (size_t)nP) || dPP.alloc(3 * (size_t)nP))
    return -1;
  if (nP > 0) {
    hipLaunchKernelGGL(k_pair_keys, dim3(grid_of(nP)), dim3(TB), 0, s, nP, shift, dPm.p, dPn.p, dPf.p, dPK.p, dPI.p);
    const int endBit = std::min(64, bits_for(nE) + shift);
    size_t tb3 = 0;
    FOS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, dPK.p, dPK2.p, dPI.p, dPI2.p, nP, 0, endBit, s));
    if (tb3 > tmp.n && tmp.alloc(tb3)) return -1;
    FOS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb3, dPK.p, dPK2.p, dPI.p, dPI2.p, nP, 0, endBit, s));
    hipLaunchKernelGGL(k_pair_geom, dim3(grid_of(nP)), dim3(TB), 0, s, nP, M, dPI2.p, dPm.p, dPn.p, dPf.p, dSE.p,
                       dINE.p, dPN.p, dIEP.p, dAP.p, dNorm.p, dPP.p);
    FOS_HIP_TRY(hipGetLastError());
  }
  // host copies
  out->pN = xalloc<double>(3 * (size_t)nN);
  std::memcpy(out->pN, in->pN, sizeof(double) * 3 * (size_t)nN);
  out->gid = xalloc<int>(nE);
  for (int k = 0; k < nE; k++) out->gid[k] = in->gid[ord[k]];
  out->sE = xalloc<int>(nE);
  out->nNE = xalloc<int>(nE);
  out->iNE = xalloc<int>((size_t)M * nE);
  out->p = xalloc<double>(3 * (size_t)nE);
  out->v = xalloc<double>(nC);
  out->neib = xalloc<int>((size_t)std::max(maxNF, 1) * nC);
  out->iEP = xalloc<int>(2 * (size_t)nP);
  out->aP = xalloc<double>(nP);
  out->normP = xalloc<double>(3 * (size_t)nP);
  out->pP = xalloc<double>(3 * (size_t)nP);
  if (d2h(out->sE, dSE, nE, s) || d2h(out->nNE, dNNE, nE, s) || d2h(out->iNE, dINE, (size_t)M * nE, s) ||
      d2h(out->p, dP, 3 * (size_t)nE, s) || d2h(out->v, dV, nC, s) || d2h(out->neib, dNeib, (size_t)maxNF * nC, s) ||
      d2h(out->iEP, dIEP, 2 * (size_t)nP, s) || d2h(out->aP, dAP, nP, s) || d2h(out->normP, dNorm, 3 * (size_t)nP, s) ||
      d2h(out->pP, dPP, 3 * (size_t)nP, s))
    return -1;
  FOS_HIP_TRY(hipStreamSynchronize(s));
  out->nP = nP;
  int nPb = 0;
  for (int i = nP - 1; i >= 0 && out->iEP[2 * (size_t)i + 1] >= nC; i--) nPb++;
  out->nPb = nPb;
  out->nPi = nP - nPb;
  return 0;
}

}  // namespace
}  // namespace fos

extern "C" int fos_grid_up_gpu(const fos_poly_grid *in, fos_fv_grid *out, void *stream) {
  if (!out) { fos::set_error("fos_grid_up_gpu: null output"); return -1; }
  std::memset(out, 0, sizeof(*out));
  // every index is range-checked on the host before anything is uploaded
  if (fos::validate_poly_grid(in, "fos_grid_up_gpu")) return -1;
  const int rc = fos::grid_up_device(in, out, static_cast<hipStream_t>(stream));
  if (rc) fos_fv_grid_free(out);
  return rc;
}
