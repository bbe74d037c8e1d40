/*
 * Oracle (test infrastructure only): FP64 forward, apply, loss and backward of the learned
 * per-cell preconditioner (no reference counterpart -- the reference's psol is an identity,
 * src/foeuler/foeuler.f90:434).  Same network and parameter packing as the product kernels:
 *
 *   h0 = silu(x W0 + b0)                      x: 20 features, W0 [20][128]
 *   h1 = silu([h0, mean_nb h0] W1 + b1)       W1 [256][128]
 *   h2 = silu([h1, mean_nb h1] W2 + b2)       W2 [256][128]
 *   o  = k_c (h2 W3 + b3)                     W3 [128][30], k_c = min(1, x_c[18] (1 + x_c[19]))
 *   B^ = I + o[0:25] (5x5 row-major), s = o[25:30]: weights of the 4 face neighbours (neib slot
 *   order) and of the cell itself.  x_c[18] (1 + x_c[19]) is the cell's CFL number
 *   gamma (c + |u|) / v^(1/3) from its features: below 1 the outputs shrink with it, so the operator
 *   tends to the identity as the time step shrinks (the Newton matrix I - gamma J differs from I by
 *   O(CFL)); from 1 on they are unscaled.
 *   apply (scaled space, S per variable):  t_c = B^_c (r_c / S);
 *                                         z_c = S ((1 + s_4) t_c + sum_j s_j t_nb(j))
 *   loss = mean over masked cells of |z_c / S - delta_c / S|^2
 *
 * Parameters are one flat array: W0, b0, W1, b1, W2, b2, W3, b3 (row-major [in][out]).
 * Neighbours are the cells of neib (facets and -1 skipped); a cell without neighbours has mean 0.
 * Parity: the backward is pinned by central finite differences of the loss (tests); the product's
 * BF16/MFMA kernels are checked against this FP64 oracle.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NI 20
#define NH 128
#define NO 30
#define NPAR (NI * NH + NH + 2 * NH * NH + NH + 2 * NH * NH + NH + NH * NO + NO)

int orc_mlp_nparams(void) { return NPAR; }

static double silu(double x) { return x / (1.0 + exp(-x)); }
static double dsilu(double x) {
  double s = 1.0 / (1.0 + exp(-x));
  return s * (1.0 + x * (1.0 - s));
}

typedef struct {
  const double *W0, *b0, *W1, *b1, *W2, *b2, *W3, *b3;
} params_t;

static params_t unpack(const double *p) {
  params_t q;
  q.W0 = p; p += NI * NH;
  q.b0 = p; p += NH;
  q.W1 = p; p += 2 * NH * NH;
  q.b1 = p; p += NH;
  q.W2 = p; p += 2 * NH * NH;
  q.b2 = p; p += NH;
  q.W3 = p; p += NH * NO;
  q.b3 = p;
  return q;
}

static void nbr_mean(int n, int maxnb, const int *nbr, int nc, const double *h, double *m) {
  for (int c = 0; c < n; c++) {
    int cnt = 0;
    for (int k = 0; k < NH; k++) m[(size_t)c * NH + k] = 0.0;
    for (int j = 0; j < maxnb; j++) {
      int d = nbr[(size_t)maxnb * c + j];
      if (d < 0 || d >= nc) continue;
      cnt++;
      for (int k = 0; k < NH; k++) m[(size_t)c * NH + k] += h[(size_t)d * NH + k];
    }
    if (cnt)
      for (int k = 0; k < NH; k++) m[(size_t)c * NH + k] /= cnt;
  }
}

static int nbr_count(int maxnb, const int *nbr, int nc, int c) {
  int cnt = 0;
  for (int j = 0; j < maxnb; j++) {
    int d = nbr[(size_t)maxnb * c + j];
    if (d >= 0 && d < nc) cnt++;
  }
  return cnt;
}

/* forward: O[n][30]; caches (optional) pre0/pre1/pre2 [n][128], h0/h1/h2, m0/m1 */
typedef struct {
  double *pre0, *pre1, *pre2, *h0, *h1, *h2, *m0, *m1;
} cache_t;

/* the output scale of cell c: its CFL number from features 18 (gamma c / v^(1/3)) and 19 (|u| / c),
   at most 1 */
static double kappa(const double *X, int c) {
  return fmin(1.0, X[(size_t)c * NI + 18] * (1.0 + X[(size_t)c * NI + 19]));
}

static void forward(int n, const double *X, int maxnb, const int *nbr, const params_t *P, double *O, cache_t *C) {
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NH; k++) {
      double s = P->b0[k];
      for (int i = 0; i < NI; i++) s += X[(size_t)c * NI + i] * P->W0[i * NH + k];
      C->pre0[(size_t)c * NH + k] = s;
      C->h0[(size_t)c * NH + k] = silu(s);
    }
  nbr_mean(n, maxnb, nbr, n, C->h0, C->m0);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NH; k++) {
      double s = P->b1[k];
      for (int i = 0; i < NH; i++) s += C->h0[(size_t)c * NH + i] * P->W1[i * NH + k];
      for (int i = 0; i < NH; i++) s += C->m0[(size_t)c * NH + i] * P->W1[(NH + i) * NH + k];
      C->pre1[(size_t)c * NH + k] = s;
      C->h1[(size_t)c * NH + k] = silu(s);
    }
  nbr_mean(n, maxnb, nbr, n, C->h1, C->m1);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NH; k++) {
      double s = P->b2[k];
      for (int i = 0; i < NH; i++) s += C->h1[(size_t)c * NH + i] * P->W2[i * NH + k];
      for (int i = 0; i < NH; i++) s += C->m1[(size_t)c * NH + i] * P->W2[(NH + i) * NH + k];
      C->pre2[(size_t)c * NH + k] = s;
      C->h2[(size_t)c * NH + k] = silu(s);
    }
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NO; k++) {
      double s = P->b3[k];
      for (int i = 0; i < NH; i++) s += C->h2[(size_t)c * NH + i] * P->W3[i * NO + k];
      O[(size_t)c * NO + k] = kappa(X, c) * s;
    }
}

static cache_t cache_alloc(int n) {
  cache_t C;
  size_t z = (size_t)n * NH;
  C.pre0 = calloc(z, sizeof(double)); C.pre1 = calloc(z, sizeof(double)); C.pre2 = calloc(z, sizeof(double));
  C.h0 = calloc(z, sizeof(double)); C.h1 = calloc(z, sizeof(double)); C.h2 = calloc(z, sizeof(double));
  C.m0 = calloc(z, sizeof(double)); C.m1 = calloc(z, sizeof(double));
  return C;
}
static void cache_free(cache_t *C) {
  free(C->pre0); free(C->pre1); free(C->pre2); free(C->h0); free(C->h1); free(C->h2); free(C->m0); free(C->m1);
}

void orc_mlp_forward(int n, const double *X, int maxnb, const int *nbr, const double *par, double *O) {
  params_t P = unpack(par);
  cache_t C = cache_alloc(n);
  forward(n, X, maxnb, nbr, &P, O, &C);
  cache_free(&C);
}

/* neighbour in slot j of cell c inside the batch, or -1 */
static int nb_at(int maxnb, const int *nbr, int n, int c, int j) {
  if (j >= maxnb) return -1;
  int d = nbr[(size_t)maxnb * c + j];
  return (d >= 0 && d < n) ? d : -1;
}

/* zhat_c = (1 + s_4) t_c + sum_j s_j t_nb(j) in the scaled space, t [n][5] */
static void combine(int n, const double *O, int maxnb, const int *nbr, const double *t, int c, double zh[5]) {
  const double *o = O + (size_t)c * NO;
  for (int a = 0; a < 5; a++) zh[a] = (1.0 + o[29]) * t[(size_t)c * 5 + a];
  for (int j = 0; j < 4; j++) {
    int d = nb_at(maxnb, nbr, n, c, j);
    if (d < 0) continue;
    for (int a = 0; a < 5; a++) zh[a] += o[25 + j] * t[(size_t)d * 5 + a];
  }
}

/* z = P r on SoA vectors (5 n) from network outputs O */
void orc_mlp_apply(int n, const double *O, int maxnb, const int *nbr, const double *S, const double *r, double *z) {
  double *t = malloc(sizeof(double) * 5 * (size_t)n);
  for (int c = 0; c < n; c++)
    for (int a = 0; a < 5; a++) {
      double s = r[(size_t)a * n + c] / S[a];
      for (int b = 0; b < 5; b++) s += O[(size_t)c * NO + 5 * a + b] * (r[(size_t)b * n + c] / S[b]);
      t[(size_t)c * 5 + a] = s;
    }
  for (int c = 0; c < n; c++) {
    double zh[5];
    combine(n, O, maxnb, nbr, t, c, zh);
    for (int a = 0; a < 5; a++) z[(size_t)a * n + c] = S[a] * zh[a];
  }
  free(t);
}

/* loss over masked cells and its gradient with respect to the packed parameters */
double orc_mlp_loss_grad(int n, const double *X, int maxnb, const int *nbr, const double *par, const double *S,
                         const double *r, const double *delta, const int *mask, double *grad) {
  params_t P = unpack(par);
  cache_t C = cache_alloc(n);
  double *O = calloc((size_t)n * NO, sizeof(double));
  forward(n, X, maxnb, nbr, &P, O, &C);
  int nm = 0;
  for (int c = 0; c < n; c++) nm += mask[c] != 0;
  double *t = calloc((size_t)n * 5, sizeof(double));
  double *rh = calloc((size_t)n * 5, sizeof(double));
  for (int c = 0; c < n; c++)
    for (int a = 0; a < 5; a++) rh[(size_t)c * 5 + a] = r[(size_t)a * n + c] / S[a];
  for (int c = 0; c < n; c++)
    for (int a = 0; a < 5; a++) {
      double s = rh[(size_t)c * 5 + a];
      for (int b = 0; b < 5; b++) s += O[(size_t)c * NO + 5 * a + b] * rh[(size_t)c * 5 + b];
      t[(size_t)c * 5 + a] = s;
    }
  double loss = 0.0;
  double *e = calloc((size_t)n * 5, sizeof(double));  /* dL/dzhat */
  for (int c = 0; c < n; c++) {
    if (!mask[c]) continue;
    double zh[5];
    combine(n, O, maxnb, nbr, t, c, zh);
    for (int a = 0; a < 5; a++) {
      double d = zh[a] - delta[(size_t)a * n + c] / S[a];
      loss += d * d;
      e[(size_t)c * 5 + a] = 2.0 * d / nm;
    }
  }
  loss /= nm;
  /* dL/dO (weights) and dL/dt: t_c enters zhat_c with (1 + s_4c) and zhat_d with s_jd where
     nb(j, d) = c */
  double *dO = calloc((size_t)n * NO, sizeof(double)), *dt = calloc((size_t)n * 5, sizeof(double));
  for (int c = 0; c < n; c++) {
    const double *o = O + (size_t)c * NO, *ec = e + (size_t)c * 5;
    double g4 = 0.0;
    for (int a = 0; a < 5; a++) {
      g4 += ec[a] * t[(size_t)c * 5 + a];
      dt[(size_t)c * 5 + a] += (1.0 + o[29]) * ec[a];
    }
    dO[(size_t)c * NO + 29] = g4;
    for (int j = 0; j < 4; j++) {
      int d = nb_at(maxnb, nbr, n, c, j);
      if (d < 0) continue;
      double gj = 0.0;
      for (int a = 0; a < 5; a++) {
        gj += ec[a] * t[(size_t)d * 5 + a];
        dt[(size_t)d * 5 + a] += o[25 + j] * ec[a];
      }
      dO[(size_t)c * NO + 25 + j] = gj;
    }
  }
  for (int c = 0; c < n; c++)
    for (int a = 0; a < 5; a++)
      for (int b = 0; b < 5; b++) dO[(size_t)c * NO + 5 * a + b] = dt[(size_t)c * 5 + a] * rh[(size_t)c * 5 + b];
  /* backward through the MLP */
  memset(grad, 0, sizeof(double) * NPAR);
  double *gW0 = grad, *gb0 = gW0 + NI * NH, *gW1 = gb0 + NH, *gb1 = gW1 + 2 * NH * NH, *gW2 = gb1 + NH;
  double *gb2 = gW2 + 2 * NH * NH, *gW3 = gb2 + NH, *gb3 = gW3 + NH * NO;
  size_t z = (size_t)n * NH;
  double *dh = calloc(z, sizeof(double)), *dp = calloc(z, sizeof(double)), *dm = calloc(z, sizeof(double));
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NO; k++) {
      double g = dO[(size_t)c * NO + k] * kappa(X, c);  /* through o = k_c (h2 W3 + b3) */
      gb3[k] += g;
      for (int i = 0; i < NH; i++) {
        gW3[i * NO + k] += C.h2[(size_t)c * NH + i] * g;
        dh[(size_t)c * NH + i] += P.W3[i * NO + k] * g;
      }
    }
  /* layer 2 */
  for (size_t q = 0; q < z; q++) dp[q] = dh[q] * dsilu(C.pre2[q]);
  memset(dh, 0, sizeof(double) * z);
  memset(dm, 0, sizeof(double) * z);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NH; k++) {
      double g = dp[(size_t)c * NH + k];
      gb2[k] += g;
      for (int i = 0; i < NH; i++) {
        gW2[i * NH + k] += C.h1[(size_t)c * NH + i] * g;
        gW2[(NH + i) * NH + k] += C.m1[(size_t)c * NH + i] * g;
        dh[(size_t)c * NH + i] += P.W2[i * NH + k] * g;
        dm[(size_t)c * NH + i] += P.W2[(NH + i) * NH + k] * g;
      }
    }
  for (int c = 0; c < n; c++) {
    int cnt = nbr_count(maxnb, nbr, n, c);
    if (!cnt) continue;
    for (int j = 0; j < maxnb; j++) {
      int d = nbr[(size_t)maxnb * c + j];
      if (d < 0 || d >= n) continue;
      for (int i = 0; i < NH; i++) dh[(size_t)d * NH + i] += dm[(size_t)c * NH + i] / cnt;
    }
  }
  /* layer 1 */
  for (size_t q = 0; q < z; q++) dp[q] = dh[q] * dsilu(C.pre1[q]);
  memset(dh, 0, sizeof(double) * z);
  memset(dm, 0, sizeof(double) * z);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NH; k++) {
      double g = dp[(size_t)c * NH + k];
      gb1[k] += g;
      for (int i = 0; i < NH; i++) {
        gW1[i * NH + k] += C.h0[(size_t)c * NH + i] * g;
        gW1[(NH + i) * NH + k] += C.m0[(size_t)c * NH + i] * g;
        dh[(size_t)c * NH + i] += P.W1[i * NH + k] * g;
        dm[(size_t)c * NH + i] += P.W1[(NH + i) * NH + k] * g;
      }
    }
  for (int c = 0; c < n; c++) {
    int cnt = nbr_count(maxnb, nbr, n, c);
    if (!cnt) continue;
    for (int j = 0; j < maxnb; j++) {
      int d = nbr[(size_t)maxnb * c + j];
      if (d < 0 || d >= n) continue;
      for (int i = 0; i < NH; i++) dh[(size_t)d * NH + i] += dm[(size_t)c * NH + i] / cnt;
    }
  }
  /* layer 0 */
  for (size_t q = 0; q < z; q++) dp[q] = dh[q] * dsilu(C.pre0[q]);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < NH; k++) {
      double g = dp[(size_t)c * NH + k];
      gb0[k] += g;
      for (int i = 0; i < NI; i++) gW0[i * NH + k] += X[(size_t)c * NI + i] * g;
    }
  free(dh); free(dp); free(dm); free(dO); free(dt); free(e); free(t); free(rh); free(O);
  cache_free(&C);
  return loss;
}
