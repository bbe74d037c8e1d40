/*
 * Oracle (test infrastructure only): grid build restated from the reference.
 *   polyX%up            polyGrid/polyX.f90:59-70
 *   sortPolyGrid        polyGrid/polyGrid.f90:88-136
 *   volumes             polyGrid/polyGrid.f90:73-82, geometry/geometry.f90:61-91
 *   nF / nNF / getINF   polyGrid/polyGrid.f90:146-278
 *   getNeibPolyFvGrid   polyGrid/polyFvGrid.f90:60-198
 *   sortPolyFvGrid      polyGrid/polyFvGrid.f90:201-227 with qsortI misc/sort.f90:15-65
 * getNearPolyFvGrid (SVD gradient map) is not restated: the Euler path never reads it.
 * Parity: pinned against the compiled reference (oracle/_ref) on the 2x2x2 and
 * 40x10x10 box meshes and the duct example (tests/test_oracle_ref.py), plus the
 * mesh invariants 4nC = 2nPi + nPb, sum of volumes, closed-cell sum a*n = 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

double orc_a3p(const double *p1, const double *p2, const double *p3) {
  /* geometry.f90:17-27 */
  double a[3], b[3];
  for (int d = 0; d < 3; d++) { a[d] = p2[d] - p1[d]; b[d] = p3[d] - p1[d]; }
  double c1 = a[1] * b[2] - a[2] * b[1];
  double c2 = a[2] * b[0] - a[0] * b[2];
  double c3 = a[0] * b[1] - a[1] * b[0];
  return sqrt(c1 * c1 + c2 * c2 + c3 * c3) / 2.0;
}

void orc_n3p(const double *p1, const double *p2, const double *p3, double *n) {
  /* geometry.f90:38-49: a = p2-p1, b = p3-p2, n = a x b / |a x b| */
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; d++) { a[d] = p2[d] - p1[d]; b[d] = p3[d] - p2[d]; }
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  double nn = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int d = 0; d < 3; d++) n[d] = c[d] / nn;
}

static double a4p(const double *p1, const double *p2, const double *p3, const double *p4) {
  /* geometry.f90:30-35 */
  return orc_a3p(p1, p2, p3) + orc_a3p(p3, p4, p1);
}

static void n4p(const double *p1, const double *p2, const double *p3, const double *p4, double *n) {
  /* geometry.f90:52-58 */
  double n1[3], n2[3];
  orc_n3p(p1, p2, p3, n1);
  orc_n3p(p3, p4, p1, n2);
  for (int d = 0; d < 3; d++) n[d] = n1[d] + n2[d];
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int d = 0; d < 3; d++) n[d] /= nn;
}

double orc_v4p(const double *p1, const double *p2, const double *p3, const double *p4) {
  /* geometry.f90:61-82 (cofactor expansion, absolute value) */
  const double *P[4] = {p1, p2, p3, p4};
#define X(i) (P[(i)-1][0])
#define Y(i) (P[(i)-1][1])
#define Z(i) (P[(i)-1][2])
  double y1z2 = Y(1) * Z(2), y1z3 = Y(1) * Z(3), y1z4 = Y(1) * Z(4);
  double y2z1 = Y(2) * Z(1), y2z3 = Y(2) * Z(3), y2z4 = Y(2) * Z(4);
  double y3z1 = Y(3) * Z(1), y3z2 = Y(3) * Z(2), y3z4 = Y(3) * Z(4);
  double y4z1 = Y(4) * Z(1), y4z2 = Y(4) * Z(2), y4z3 = Y(4) * Z(3);
  double v = (X(2) * (y3z4 - y4z3) - X(3) * (y2z4 - y4z2) + X(4) * (y2z3 - y3z2))
           - (X(1) * (y3z4 - y4z3) - X(3) * (y1z4 - y4z1) + X(4) * (y1z3 - y3z1))
           + (X(1) * (y2z4 - y4z2) - X(2) * (y1z4 - y4z1) + X(4) * (y1z2 - y2z1))
           - (X(1) * (y2z3 - y3z2) - X(2) * (y1z3 - y3z1) + X(3) * (y1z2 - y2z1));
#undef X
#undef Y
#undef Z
  return fabs(v) / 6.0;
}

static int nF_of(int s) {
  /* polyGrid.f90:146-160 */
  switch (s) {
    case ORC_TRI: case ORC_QUAD: case ORC_TRI6: return 1;
    case ORC_TET: case ORC_TET10: return 4;
    case ORC_HEX: return 6;
    default: return 0;
  }
}

static int nNF_of(int s, int i) {
  /* polyGrid.f90:163-208 (i is 1-based) */
  switch (s) {
    case ORC_TRI: return i == 1 ? 3 : 0;
    case ORC_QUAD: return i == 1 ? 4 : 0;
    case ORC_TRI6: return i == 1 ? 6 : 0;
    case ORC_TET: return (i >= 1 && i <= 4) ? 3 : 0;
    case ORC_HEX: return (i >= 1 && i <= 6) ? 4 : 0;
    case ORC_TET10: return (i >= 1 && i <= 4) ? 6 : 0;
    default: return 0;
  }
}

/* getINF (polyGrid.f90:211-278): 1-based local node numbers of face i of shape s; 0 = padding */
static void getINF(int s, int i, int *ind, int len) {
  static const int tet[4][3] = {{1, 3, 2}, {1, 2, 4}, {1, 4, 3}, {2, 3, 4}};
  static const int hex[6][4] = {{2, 3, 7, 6}, {1, 5, 8, 4}, {3, 4, 8, 7}, {1, 2, 6, 5}, {5, 6, 7, 8}, {1, 4, 3, 2}};
  static const int tet10[4][6] = {{1, 3, 2, 7, 6, 5}, {1, 2, 4, 5, 10, 8}, {1, 4, 3, 8, 9, 7}, {2, 3, 4, 6, 9, 10}};
  for (int k = 0; k < len; k++) ind[k] = 0;
  switch (s) {
    case ORC_TRI: if (i == 1) for (int k = 0; k < 3; k++) ind[k] = k + 1; break;
    case ORC_QUAD: if (i == 1) for (int k = 0; k < 4; k++) ind[k] = k + 1; break;
    case ORC_TRI6: if (i == 1) for (int k = 0; k < 6; k++) ind[k] = k + 1; break;
    case ORC_TET: if (i >= 1 && i <= 4) for (int k = 0; k < 3; k++) ind[k] = tet[i - 1][k]; break;
    case ORC_HEX: if (i >= 1 && i <= 6) for (int k = 0; k < 4; k++) ind[k] = hex[i - 1][k]; break;
    case ORC_TET10: if (i >= 1 && i <= 4) for (int k = 0; k < 6; k++) ind[k] = tet10[i - 1][k]; break;
    default: break;
  }
}

static int is_cell(int s) { return s == ORC_TET || s == ORC_HEX || s == ORC_TET10; }
static int is_facet(int s) { return s == ORC_TRI || s == ORC_QUAD || s == ORC_TRI6; }

/* qsortI_actual (sort.f90:30-65), 1-based first/last on 1-based arrays a1/perm1 */
static void qsort_actual(int *a1, int *perm1, int first, int last) {
  int pivot = a1[(first + last) / 2];
  int i = first, j = last;
  for (;;) {
    while (a1[i] < pivot) i++;
    while (pivot < a1[j]) j--;
    if (i >= j) break;
    int tmp = a1[i]; a1[i] = a1[j]; a1[j] = tmp;
    tmp = perm1[i]; perm1[i] = perm1[j]; perm1[j] = tmp;
    i++;
    j--;
  }
  if (first < i - 1) qsort_actual(a1, perm1, first, i - 1);
  if (j + 1 < last) qsort_actual(a1, perm1, j + 1, last);
}

void orc_grid_free(orc_grid *g) {
  if (!g) return;
  free(g->pN); free(g->sE); free(g->nNE); free(g->iNE); free(g->gid); free(g->p); free(g->v);
  free(g->iEP); free(g->aP); free(g->normP); free(g->pP); free(g->neib);
  free(g);
}

orc_grid *orc_grid_build(int nN, const double *pN, int nE, const int *sE, const int *nNE,
                         int maxNNE, const int *iNE, const int *gid) {
  orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
  g->nN = nN; g->nE = nE; g->maxNNE = maxNNE;
  g->pN = (double *)malloc(sizeof(double) * 3 * (size_t)nN);
  memcpy(g->pN, pN, sizeof(double) * 3 * (size_t)nN);
  /* polyX%up: centroids, before sorting (polyX.f90:59-70) */
  double *p0 = (double *)malloc(sizeof(double) * 3 * (size_t)nE);
  for (int e = 0; e < nE; e++) {
    for (int d = 0; d < 3; d++) {
      double s = 0.0;
      for (int j = 0; j < nNE[e]; j++) s += pN[3 * iNE[(size_t)maxNNE * e + j] + d];
      p0[3 * e + d] = s / (double)nNE[e];
    }
  }
  /* sortPolyGrid: cells, then facets, then the rest, each stable (polyGrid.f90:88-136) */
  g->sE = (int *)malloc(sizeof(int) * nE);
  g->nNE = (int *)malloc(sizeof(int) * nE);
  g->iNE = (int *)malloc(sizeof(int) * (size_t)maxNNE * nE);
  g->gid = (int *)malloc(sizeof(int) * nE);
  g->p = (double *)malloc(sizeof(double) * 3 * (size_t)nE);
  int j = 0;
  for (int pass = 0; pass < 3; pass++) {
    for (int e = 0; e < nE; e++) {
      int take = pass == 0 ? is_cell(sE[e]) : pass == 1 ? is_facet(sE[e]) : !(is_cell(sE[e]) || is_facet(sE[e]));
      if (!take) continue;
      g->sE[j] = sE[e];
      g->nNE[j] = nNE[e];
      memcpy(g->iNE + (size_t)maxNNE * j, iNE + (size_t)maxNNE * e, sizeof(int) * maxNNE);
      memcpy(g->p + 3 * j, p0 + 3 * e, sizeof(double) * 3);
      g->gid[j] = gid[e];
      j++;
    }
    if (pass == 0) g->nC = j;
    if (pass == 1) g->nF = j - g->nC;
  }
  free(p0);
  const int nC = g->nC;
  /* volumes (polyGrid.f90:73-82) */
  g->v = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1));
  for (int c = 0; c < nC; c++) {
    const int *ie = g->iNE + (size_t)maxNNE * c;
#define PN(k) (g->pN + 3 * ie[(k)])
    if (g->sE[c] == ORC_TET || g->sE[c] == ORC_TET10) {
      g->v[c] = orc_v4p(PN(0), PN(1), PN(2), PN(3));
    } else if (g->sE[c] == ORC_HEX) {
      /* v8p geometry.f90:85-91 */
      g->v[c] = orc_v4p(PN(0), PN(1), PN(2), PN(5)) + orc_v4p(PN(0), PN(2), PN(3), PN(5)) +
                orc_v4p(PN(0), PN(3), PN(4), PN(5)) + orc_v4p(PN(3), PN(4), PN(5), PN(6)) +
                orc_v4p(PN(2), PN(3), PN(5), PN(6)) + orc_v4p(PN(3), PN(4), PN(6), PN(7));
    } else {
      g->v[c] = 0.0;
    }
#undef PN
  }
  /* getNeibPolyFvGrid (polyFvGrid.f90:60-198) */
  int *nEN = (int *)calloc(nN, sizeof(int));
  for (int e = 0; e < nE; e++)
    for (int k = 0; k < g->nNE[e]; k++) nEN[g->iNE[(size_t)maxNNE * e + k]]++;
  int maxEN = 0;
  for (int k = 0; k < nN; k++) if (nEN[k] > maxEN) maxEN = nEN[k];
  int *iEN = (int *)malloc(sizeof(int) * (size_t)(maxEN > 0 ? maxEN : 1) * nN);
  memset(nEN, 0, sizeof(int) * nN);
  for (int e = 0; e < nE; e++)
    for (int k = 0; k < g->nNE[e]; k++) {
      int nd = g->iNE[(size_t)maxNNE * e + k];
      iEN[(size_t)maxEN * nd + nEN[nd]] = e;
      nEN[nd]++;
    }
  int maxNF = 0, sumNF = 0, lenINF = 0;
  int *nFace = (int *)malloc(sizeof(int) * nE);
  for (int e = 0; e < nE; e++) {
    nFace[e] = nF_of(g->sE[e]);
    sumNF += nFace[e];
    if (nFace[e] > maxNF) maxNF = nFace[e];
    for (int f = 1; f <= nFace[e]; f++) if (nNF_of(g->sE[e], f) > lenINF) lenINF = nNF_of(g->sE[e], f);
  }
  /* neib(maxval(nFace), nC) -- maxval over all elements incl. facets */
  g->maxNF = maxNF;
  g->neib = (int *)malloc(sizeof(int) * (size_t)(maxNF > 0 ? maxNF : 1) * (nC > 0 ? nC : 1));
  for (size_t k = 0; k < (size_t)maxNF * nC; k++) g->neib[k] = -1;
  int *iEPt = (int *)malloc(sizeof(int) * 2 * (size_t)(sumNF > 0 ? sumNF : 1));
  double *aPt = (double *)malloc(sizeof(double) * (sumNF > 0 ? sumNF : 1));
  double *nPt = (double *)malloc(sizeof(double) * 3 * (sumNF > 0 ? sumNF : 1));
  double *pPt = (double *)malloc(sizeof(double) * 3 * (sumNF > 0 ? sumNF : 1));
  int *iNF = (int *)malloc(sizeof(int) * (lenINF > 0 ? lenINF : 1));
  int *iNF1 = (int *)malloc(sizeof(int) * (lenINF > 0 ? lenINF : 1));
  int nP = 0, nPi = 0, nPb = 0;
  /* node index sentinel: local 0 (padding) maps to -1, never equal to a real node */
#define MAPNODES(arr, s, f, e)                                                    \
  do {                                                                            \
    getINF((s), (f), (arr), lenINF);                                              \
    for (int q = 0; q < lenINF; q++)                                              \
      (arr)[q] = (arr)[q] > 0 ? g->iNE[(size_t)maxNNE * (e) + (arr)[q] - 1] : -1; \
  } while (0)
#define MATCHES(res)                                        \
  do {                                                      \
    res = 1;                                                \
    for (int ii = 0; ii < lenINF && res; ii++) {            \
      int any = 0;                                          \
      for (int q = 0; q < lenINF; q++)                      \
        if (iNF1[q] == iNF[ii]) { any = 1; break; }         \
      if (!any) res = 0;                                    \
    }                                                       \
  } while (0)
#define ADDGEOM(np)                                                                       \
  do {                                                                                    \
    int cnt = 0;                                                                          \
    for (int q = 0; q < lenINF; q++) if (iNF[q] >= 0) cnt++;                              \
    const double *P1 = g->pN + 3 * (iNF[0] >= 0 ? iNF[0] : 0);                           \
    if (cnt == 3) {                                                                       \
      const double *P2 = g->pN + 3 * iNF[1], *P3 = g->pN + 3 * iNF[2];                   \
      aPt[np] = orc_a3p(P1, P2, P3);                                                      \
      orc_n3p(P1, P2, P3, nPt + 3 * (np));                                                \
      for (int d = 0; d < 3; d++) pPt[3 * (np) + d] = (P1[d] + P2[d] + P3[d]) / 3.0;     \
    } else if (cnt == 4) {                                                                \
      const double *P2 = g->pN + 3 * iNF[1], *P3 = g->pN + 3 * iNF[2], *P4 = g->pN + 3 * iNF[3]; \
      aPt[np] = a4p(P1, P2, P3, P4);                                                      \
      n4p(P1, P2, P3, P4, nPt + 3 * (np));                                                \
      for (int d = 0; d < 3; d++) pPt[3 * (np) + d] = (P1[d] + P2[d] + P3[d] + P4[d]) / 4.0; \
    } else {                                                                              \
      aPt[np] = 0.0;                                                                      \
      for (int d = 0; d < 3; d++) { nPt[3 * (np) + d] = 0.0; pPt[3 * (np) + d] = 0.0; }   \
    }                                                                                     \
  } while (0)
  for (int i = 0; i < nC; i++) {
    for (int jf = 1; jf <= nFace[i]; jf++) {
      if (g->neib[(size_t)maxNF * i + jf - 1] != -1) continue;
      MAPNODES(iNF, g->sE[i], jf, i);
      int n0 = iNF[0];
      /* 1st round: facets only, no early exit (polyFvGrid.f90:111-143) */
      for (int k = 0; k < nEN[n0]; k++) {
        int m = iEN[(size_t)maxEN * n0 + k];
        if (m != i && m >= nC) {
          for (int l = 1; l <= nFace[m]; l++) {
            MAPNODES(iNF1, g->sE[m], l, m);
            int res;
            MATCHES(res);
            if (res) {
              g->neib[(size_t)maxNF * i + jf - 1] = m;
              iEPt[2 * nP] = i; iEPt[2 * nP + 1] = m;
              ADDGEOM(nP);
              nP++; nPb++;
            }
          }
        }
      }
      if (g->neib[(size_t)maxNF * i + jf - 1] != -1) continue;
      /* 2nd round: cells only; exit only the face loop (polyFvGrid.f90:148-182) */
      for (int k = 0; k < nEN[n0]; k++) {
        int m = iEN[(size_t)maxEN * n0 + k];
        if (m != i && m < nC) {
          for (int l = 1; l <= nFace[m]; l++) {
            MAPNODES(iNF1, g->sE[m], l, m);
            int res;
            MATCHES(res);
            if (res) {
              g->neib[(size_t)maxNF * i + jf - 1] = m;
              g->neib[(size_t)maxNF * m + l - 1] = i;
              iEPt[2 * nP] = i; iEPt[2 * nP + 1] = m;
              ADDGEOM(nP);
              nP++; nPi++;
              break;
            }
          }
        }
      }
    }
  }
#undef MAPNODES
#undef MATCHES
#undef ADDGEOM
  g->nP = nP; g->nPi = nPi; g->nPb = nPb;
  /* sortPolyFvGrid: key max(iEP) (1-based values), non-stable quicksort with permutation */
  int *a1 = (int *)malloc(sizeof(int) * (nP + 1));
  int *perm1 = (int *)malloc(sizeof(int) * (nP + 1));
  for (int i = 0; i < nP; i++) {
    int m = iEPt[2 * i] + 1, n = iEPt[2 * i + 1] + 1;
    a1[i + 1] = m > n ? m : n;
    perm1[i + 1] = i + 1;
  }
  if (nP > 0) qsort_actual(a1, perm1, 1, nP);
  g->iEP = (int *)malloc(sizeof(int) * 2 * (nP > 0 ? nP : 1));
  g->aP = (double *)malloc(sizeof(double) * (nP > 0 ? nP : 1));
  g->normP = (double *)malloc(sizeof(double) * 3 * (nP > 0 ? nP : 1));
  g->pP = (double *)malloc(sizeof(double) * 3 * (nP > 0 ? nP : 1));
  for (int i = 0; i < nP; i++) {
    int s = perm1[i + 1] - 1;
    g->iEP[2 * i] = iEPt[2 * s];
    g->iEP[2 * i + 1] = iEPt[2 * s + 1];
    g->aP[i] = aPt[s];
    for (int d = 0; d < 3; d++) { g->normP[3 * i + d] = nPt[3 * s + d]; g->pP[3 * i + d] = pPt[3 * s + d]; }
  }
  free(a1); free(perm1); free(iEPt); free(aPt); free(nPt); free(pPt);
  free(iNF); free(iNF1); free(nFace); free(iEN); free(nEN);
  return g;
}

void orc_map_cond(int nE, const int *gid, int nCond, const int *condGid, int *iBC) {
  /* foeuler.f90:145-154 (same as condition.f90:255-272): first matching condition */
  for (int e = 0; e < nE; e++) {
    iBC[e] = -1;
    for (int k = 0; k < nCond; k++)
      if (gid[e] == condGid[k]) { iBC[e] = k; break; }
  }
}
