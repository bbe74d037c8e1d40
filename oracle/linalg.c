/*
 * Oracle (test infrastructure only): sparse kernels restated from SPARSKIT2 as used by
 * src/libfosolvers/numerics/sparse.f90, and the block ILU(0) the foeuler preconditioner
 * is built from (pset/psol intent, src/foeuler/foeuler.f90:401-402,433).
 *   coocsr  contrib/SPARSKIT2/FORMATS/formats.f:164-230
 *   amux    contrib/SPARSKIT2/BLASSM/matvec.f:34-79
 *   ilu0    contrib/SPARSKIT2/ITSOL/ilut.f:1793-1917
 *   ilut    contrib/SPARSKIT2/ITSOL/ilut.f:38-381 (+ qsplit :2376-2434)
 *   lusol   contrib/SPARSKIT2/ITSOL/ilut.f:2272-2321
 * All SPARSKIT arrays are 1-based (index 0 unused) exactly as in the Fortran.
 * Parity: ilut+lusol and amux pinned by the reference's own KATs (src/test/ILU1.f90:20,
 * src/test/matVecMul.f90:19) and against the compiled SPARSKIT objects (oracle/_ref);
 * block ILU(0) pinned against scalar ilu0 on the block-expanded pattern (same kept-update
 * set) and, on no-fill patterns, against the exact LU solve.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void orc_coocsr(int nrow, int nnz, const double *a, const int *ir, const int *jc,
                double *ao, int *jao, int *iao) {
  for (int k = 1; k <= nrow + 1; k++) iao[k] = 0;
  for (int k = 1; k <= nnz; k++) iao[ir[k]]++;
  int k = 1;
  for (int j = 1; j <= nrow + 1; j++) { int k0 = iao[j]; iao[j] = k; k += k0; }
  for (int kk = 1; kk <= nnz; kk++) {
    int i = ir[kk], j = jc[kk];
    int iad = iao[i];
    ao[iad] = a[kk];
    jao[iad] = j;
    iao[i] = iad + 1;
  }
  for (int j = nrow; j >= 1; j--) iao[j + 1] = iao[j];
  iao[1] = 1;
}

void orc_amux(int n, const double *x, double *y, const double *a, const int *ja, const int *ia) {
  for (int i = 1; i <= n; i++) {
    double t = 0.0;
    for (int k = ia[i]; k < ia[i + 1]; k++) t = t + a[k] * x[ja[k]];
    y[i] = t;
  }
}

int orc_ilu0(int n, const double *a, const int *ja, const int *ia, double *alu, int *jlu, int *ju, int *iw) {
  int ju0 = n + 2;
  jlu[1] = ju0;
  for (int i = 1; i <= n; i++) iw[i] = 0;
  for (int ii = 1; ii <= n; ii++) {
    int js = ju0;
    for (int j = ia[ii]; j < ia[ii + 1]; j++) {
      int jcol = ja[j];
      if (jcol == ii) {
        alu[ii] = a[j];
        iw[jcol] = ii;
        ju[ii] = ju0;
      } else {
        alu[ju0] = a[j];
        jlu[ju0] = ja[j];
        iw[jcol] = ju0;
        ju0++;
      }
    }
    jlu[ii + 1] = ju0;
    int jf = ju0 - 1, jm = ju[ii] - 1;
    for (int j = js; j <= jm; j++) {
      int jrow = jlu[j];
      double tl = alu[j] * alu[jrow];
      alu[j] = tl;
      for (int jj = ju[jrow]; jj < jlu[jrow + 1]; jj++) {
        int jw = iw[jlu[jj]];
        if (jw != 0) alu[jw] = alu[jw] - tl * alu[jj];
      }
    }
    if (alu[ii] == 0.0) return ii;
    alu[ii] = 1.0 / alu[ii];
    iw[ii] = 0;
    for (int i = js; i <= jf; i++) iw[jlu[i]] = 0;
  }
  return 0;
}

static void qsplit(double *a, int *ind, int n, int ncut) {
  /* 1-based arrays a[1..n], ind[1..n] */
  int first = 1, last = n;
  if (ncut < first || ncut > last) return;
  for (;;) {
    int mid = first;
    double abskey = fabs(a[mid]);
    for (int j = first + 1; j <= last; j++) {
      if (fabs(a[j]) > abskey) {
        mid++;
        double tmp = a[mid]; a[mid] = a[j]; a[j] = tmp;
        int it = ind[mid]; ind[mid] = ind[j]; ind[j] = it;
      }
    }
    double tmp = a[mid]; a[mid] = a[first]; a[first] = tmp;
    int it = ind[mid]; ind[mid] = ind[first]; ind[first] = it;
    if (mid == ncut) return;
    if (mid > ncut) last = mid - 1; else first = mid + 1;
  }
}

int orc_ilut(int n, const double *a, const int *ja, const int *ia, int lfil, double droptol,
             double *alu, int *jlu, int *ju, int iwk, double *w, int *jw) {
  /* w[1..n+1], jw[1..2n] */
  if (lfil < 0) return -4;
  int ju0 = n + 2;
  jlu[1] = ju0;
  for (int j = 1; j <= n; j++) jw[n + j] = 0;
  for (int ii = 1; ii <= n; ii++) {
    int j1 = ia[ii], j2 = ia[ii + 1] - 1;
    double tnorm = 0.0;
    for (int k = j1; k <= j2; k++) tnorm = tnorm + fabs(a[k]);
    if (tnorm == 0.0) return -5;
    tnorm = tnorm / (double)(float)(j2 - j1 + 1); /* real() is single precision in F77 */
    int lenu = 1, lenl = 0;
    jw[ii] = ii;
    w[ii] = 0.0;
    jw[n + ii] = ii;
    for (int j = j1; j <= j2; j++) {
      int k = ja[j];
      double t = a[j];
      if (k < ii) {
        lenl++;
        jw[lenl] = k; w[lenl] = t; jw[n + k] = lenl;
      } else if (k == ii) {
        w[ii] = t;
      } else {
        lenu++;
        int jpos = ii + lenu - 1;
        jw[jpos] = k; w[jpos] = t; jw[n + k] = jpos;
      }
    }
    int jj = 0, len = 0;
    for (;;) {
      jj++;
      if (jj > lenl) break;
      int jrow = jw[jj], k = jj;
      for (int j = jj + 1; j <= lenl; j++)
        if (jw[j] < jrow) { jrow = jw[j]; k = j; }
      if (k != jj) {
        int j = jw[jj];
        jw[jj] = jw[k]; jw[k] = j;
        jw[n + jrow] = jj; jw[n + j] = k;
        double s = w[jj]; w[jj] = w[k]; w[k] = s;
      }
      jw[n + jrow] = 0;
      double fact = w[jj] * alu[jrow];
      if (fabs(fact) <= droptol) continue;
      for (int kk = ju[jrow]; kk < jlu[jrow + 1]; kk++) {
        double s = fact * alu[kk];
        int j = jlu[kk], jpos = jw[n + j];
        if (j >= ii) {
          if (jpos == 0) {
            lenu++;
            if (lenu > n) return -1;
            int i = ii + lenu - 1;
            jw[i] = j; jw[n + j] = i; w[i] = -s;
          } else {
            w[jpos] = w[jpos] - s;
          }
        } else {
          if (jpos == 0) {
            lenl++;
            if (lenl > n) return -1;
            jw[lenl] = j; jw[n + j] = lenl; w[lenl] = -s;
          } else {
            w[jpos] = w[jpos] - s;
          }
        }
      }
      len++;
      w[len] = fact;
      jw[len] = jrow;
    }
    for (int k = 1; k <= lenu; k++) jw[n + jw[ii + k - 1]] = 0;
    lenl = len;
    len = lenl < lfil ? lenl : lfil;
    qsplit(w, jw, lenl, len);
    for (int k = 1; k <= len; k++) {
      if (ju0 > iwk) return -2;
      alu[ju0] = w[k]; jlu[ju0] = jw[k]; ju0++;
    }
    ju[ii] = ju0;
    len = 0;
    for (int k = 1; k <= lenu - 1; k++) {
      if (fabs(w[ii + k]) > droptol * tnorm) {
        len++;
        w[ii + len] = w[ii + k];
        jw[ii + len] = jw[ii + k];
      }
    }
    lenu = len + 1;
    len = lenu < lfil ? lenu : lfil;
    qsplit(w + ii, jw + ii, lenu - 1, len);
    double t = fabs(w[ii]);
    if (len + ju0 > iwk) return -3;
    for (int k = ii + 1; k <= ii + len - 1; k++) {
      jlu[ju0] = jw[k]; alu[ju0] = w[k];
      t = t + fabs(w[k]);
      ju0++;
    }
    if (w[ii] == 0.0) w[ii] = (0.0001 + droptol) * tnorm;
    alu[ii] = 1.0 / w[ii];
    jlu[ii + 1] = ju0;
  }
  return 0;
}

void orc_lusol(int n, const double *y, double *x, const double *alu, const int *jlu, const int *ju) {
  for (int i = 1; i <= n; i++) {
    x[i] = y[i];
    for (int k = jlu[i]; k < ju[i]; k++) x[i] = x[i] - alu[k] * x[jlu[k]];
  }
  for (int i = n; i >= 1; i--) {
    for (int k = ju[i]; k < jlu[i + 1]; k++) x[i] = x[i] - alu[k] * x[jlu[k]];
    x[i] = alu[i] * x[i];
  }
}

/* ---------------- block ILU(0), 5x5 blocks, natural cell order ---------------- */

static void inv5(const double *A, double *Ai) {
  /* Gauss-Jordan with partial pivoting, row-major 5x5 */
  double M[5][10];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 10; j++) M[i][j] = j < 5 ? A[5 * i + j] : (j - 5 == i ? 1.0 : 0.0);
  for (int c = 0; c < 5; c++) {
    int piv = c;
    for (int r = c + 1; r < 5; r++) if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
    if (piv != c) for (int j = 0; j < 10; j++) { double t = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = t; }
    double d = 1.0 / M[c][c];
    for (int j = 0; j < 10; j++) M[c][j] *= d;
    for (int r = 0; r < 5; r++) {
      if (r == c) continue;
      double f = M[r][c];
      if (f != 0.0) for (int j = 0; j < 10; j++) M[r][j] -= f * M[c][j];
    }
  }
  for (int i = 0; i < 5; i++) for (int j = 0; j < 5; j++) Ai[5 * i + j] = M[i][j + 5];
}

static void mm5(const double *A, const double *B, double *C) { /* C = A*B row-major */
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      double s = 0.0;
      for (int k = 0; k < 5; k++) s += A[5 * i + k] * B[5 * k + j];
      C[5 * i + j] = s;
    }
}

/* rank[c] = elimination step of row c (natural order when b->order is NULL) */
static int *bilu_rank(const orc_bilu *b) {
  int *rank = (int *)malloc(sizeof(int) * (size_t)(b->nC > 0 ? b->nC : 1));
  for (int s = 0; s < b->nC; s++) rank[b->order ? b->order[s] : s] = s;
  return rank;
}

int orc_bilu_setup(orc_bilu *b, const double *JacC, const double *v, double gamma) {
  const int nC = b->nC, nb = b->maxNF, ncol = 5 * (1 + nb);
  double *A = (double *)malloc(sizeof(double) * 25 * (size_t)nC); /* diag blocks being updated */
  int *rank = bilu_rank(b);
  /* assemble M = I - gamma * JacC / v (block row c scaled by 1/v_c) */
  for (int c = 0; c < nC; c++) {
    for (int r = 0; r < 5; r++)
      for (int q = 0; q < 5; q++) {
        A[25 * (size_t)c + 5 * r + q] = (r == q ? 1.0 : 0.0) - gamma * JacC[((size_t)c * ncol + q) * 5 + r] / v[c];
        for (int j = 0; j < nb; j++)
          b->Off[25 * ((size_t)nb * c + j) + 5 * r + q] = -gamma * JacC[((size_t)c * ncol + 5 * (j + 1) + q) * 5 + r] / v[c];
      }
  }
  double T[25];
  for (int st = 0; st < nC; st++) {
    const int i = b->order ? b->order[st] : st;
    /* lower neighbours (eliminated earlier) in elimination order */
    int ord[16], no = 0;
    for (int j = 0; j < nb; j++) {
      int k = b->neib[(size_t)nb * i + j];
      if (k >= 0 && k < nC && rank[k] < st) ord[no++] = j;
    }
    for (int a = 1; a < no; a++)
      for (int q = a; q > 0 && rank[b->neib[(size_t)nb * i + ord[q]]] < rank[b->neib[(size_t)nb * i + ord[q - 1]]]; q--) {
        int t = ord[q]; ord[q] = ord[q - 1]; ord[q - 1] = t;
      }
    for (int a = 0; a < no; a++) {
      int js = ord[a], k = b->neib[(size_t)nb * i + js];
      double *Lik = b->Off + 25 * ((size_t)nb * i + js);
      mm5(Lik, b->D + 25 * (size_t)k, T); /* L_ik = A_ik * inv(U_kk) */
      memcpy(Lik, T, sizeof(T));
      /* A_ij -= L_ik U_kj for j in U-part of row k that is in pattern(i) */
      for (int jk = 0; jk < nb; jk++) {
        int j = b->neib[(size_t)nb * k + jk];
        if (j < 0 || j >= nC || rank[j] <= rank[k]) continue;
        double *dst = NULL;
        if (j == i) dst = A + 25 * (size_t)i;
        else
          for (int ji = 0; ji < nb; ji++)
            if (b->neib[(size_t)nb * i + ji] == j) { dst = b->Off + 25 * ((size_t)nb * i + ji); break; }
        if (!dst) continue;
        mm5(Lik, b->Off + 25 * ((size_t)nb * k + jk), T);
        for (int q = 0; q < 25; q++) dst[q] -= T[q];
      }
    }
    double det_guard = 0.0;
    for (int q = 0; q < 25; q++) det_guard += fabs(A[25 * (size_t)i + q]);
    if (!(det_guard > 0.0)) { free(A); free(rank); return i + 1; }
    inv5(A + 25 * (size_t)i, b->D + 25 * (size_t)i);
  }
  free(A);
  free(rank);
  return 0;
}

void orc_bilu_solve(const orc_bilu *b, const double *r, double *z) {
  const int nC = b->nC, nb = b->maxNF;
  double *y = (double *)malloc(sizeof(double) * 5 * (size_t)(nC > 0 ? nC : 1));
  int *rank = bilu_rank(b);
  for (int st = 0; st < nC; st++) {
    const int i = b->order ? b->order[st] : st;
    double t[5];
    for (int q = 0; q < 5; q++) t[q] = r[(size_t)q * nC + i];
    for (int j = 0; j < nb; j++) {
      int k = b->neib[(size_t)nb * i + j];
      if (k < 0 || k >= nC || rank[k] >= st) continue;
      const double *L = b->Off + 25 * ((size_t)nb * i + j);
      for (int q = 0; q < 5; q++) {
        double s = 0.0;
        for (int p = 0; p < 5; p++) s += L[5 * q + p] * y[5 * (size_t)k + p];
        t[q] -= s;
      }
    }
    for (int q = 0; q < 5; q++) y[5 * (size_t)i + q] = t[q];
  }
  for (int st = nC - 1; st >= 0; st--) {
    const int i = b->order ? b->order[st] : st;
    double t[5];
    for (int q = 0; q < 5; q++) t[q] = y[5 * (size_t)i + q];
    for (int j = 0; j < nb; j++) {
      int k = b->neib[(size_t)nb * i + j];
      if (k < 0 || k >= nC || rank[k] <= st) continue;
      const double *U = b->Off + 25 * ((size_t)nb * i + j);
      for (int q = 0; q < 5; q++) {
        double s = 0.0;
        for (int p = 0; p < 5; p++) s += U[5 * q + p] * z[(size_t)p * nC + k];
        t[q] -= s;
      }
    }
    const double *D = b->D + 25 * (size_t)i;
    for (int q = 0; q < 5; q++) {
      double s = 0.0;
      for (int p = 0; p < 5; p++) s += D[5 * q + p] * t[p];
      z[(size_t)q * nC + i] = s;
    }
  }
  free(y);
  free(rank);
}

void orc_bilu_matvec(const orc_grid *g, const double *JacC, const double *v, double gamma,
                     const double *x, double *y) {
  /* y = (I - gamma * JacC / v) x on SoA vectors */
  const int nC = g->nC, nb = g->maxNF, ncol = 5 * (1 + nb);
  for (int c = 0; c < nC; c++) {
    for (int r = 0; r < 5; r++) {
      double s = 0.0;
      for (int q = 0; q < 5; q++) s += JacC[((size_t)c * ncol + q) * 5 + r] * x[(size_t)q * nC + c];
      for (int j = 0; j < nb; j++) {
        int k = g->neib[(size_t)nb * c + j];
        if (k < 0 || k >= nC) continue;
        for (int q = 0; q < 5; q++) s += JacC[((size_t)c * ncol + 5 * (j + 1) + q) * 5 + r] * x[(size_t)q * nC + k];
      }
      y[(size_t)r * nC + c] = x[(size_t)r * nC + c] - gamma * s / v[c];
    }
  }
}
