// Learned per-cell preconditioner (no reference counterpart: the reference's psol is an
// identity, src/foeuler/foeuler.f90:434).  A 3x128 SiLU MLP with two rounds of one-hop message
// passing over face neighbours emits, per cell, a 5x5 block B^ = I + o[0:25] and 5 neighbour
// weights s = o[25:30] (the 4 face neighbours in neib slot order, then the cell itself):
//   t_c = B^_c (r_c / S),   z_c = S ((1 + s_4) t_c + sum_j s_j t_nb(j))
// (S = per-variable RMS of the state), a fixed linear operator for GMRES.  Trained with Adam on
// pairs (r, target) at the solver's states, loss |S^-1 (P r - target)|^2.
//
// Dense layers run on MFMA: BF16 weights and activations, FP32 accumulation.  The forward layers
// use v_mfma_f32_16x16x32_bf16 in persistent workgroups (mlp_layer_kernel); the backward data
// products (gemm_kernel) and the weight gradients (MFMA reductions over cells) use
// v_mfma_f32_16x16x16_bf16.  FP32 master weights and
// Adam moments.  Parameter packing (flat FP32): W0[20][128] b0 W1[256][128] b1 W2[256][128] b2
// W3[128][30] b3, the same as the oracle (oracle/learned.c).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "context.h"

namespace fos {

constexpr int NI = 20, NIP = 32, NH = 128, NO = 30, NOP = 32, NBR = 4;
constexpr int NPAR = NI * NH + NH + 2 * NH * NH + NH + 2 * NH * NH + NH + NH * NO + NO;
constexpr int OFF_W0 = 0, OFF_B0 = OFF_W0 + NI * NH, OFF_W1 = OFF_B0 + NH, OFF_B1 = OFF_W1 + 2 * NH * NH;
constexpr int OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + 2 * NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO;
constexpr int WG_ROWS = 2048;  // rows per weight-gradient partial

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f2bf(float f) {  // round to nearest even
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, (__bf16)f);  // v_cvt_pk_bf16_f32 (round to nearest even)
#endif
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// sigmoid by the hardware reciprocal (1 ulp) instead of the IEEE division sequence
__device__ inline float sigm_f(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }
__device__ inline float silu_f(float x) { return x * sigm_f(x); }
__device__ inline float dsilu_f(float x) {
  const float s = sigm_f(x);
  return s * (1.0f + x * (1.0f - s));
}

struct Learned {
  int nC = 0;
  bool ready = false, haveO = false;  // parameters set; network output O computed for the mesh
  long long step = 0;
  DBuf<float> par, m1, m2, grad, bias;  // bias: b0 b1 b2 b3(padded 32)
  // BF16 copies: forward operands (n-major, [out][in]) and backward-data operands ([in][out])
  DBuf<uint16_t> W0t, W1t, W2t, W3t, W1b, W2b, W3b;
  DBuf<int> nbr;          // [nC][4] cell neighbours (-1 none)
  DBuf<float> geo;        // [nC][13] log(v/vmean), area normals / v^(2/3) per neib slot
  DBuf<float> scale;      // [8]: S[5], then scratch
  // forward activations (rows = max(nC used for inference, training batch))
  size_t capF = 0, capT = 0;
  DBuf<uint16_t> X, H0, H1, H2;  // X [.][32], H0/H1 [.][256] = (h | mean_nb h), H2 [.][128]
  DBuf<float> O, t;              // O [.][32] network output, t [.][8] = B^ r^
  // training-only buffers (rows = batch)
  DBuf<float> pre0, pre1, pre2, rh, e, ds, part;
  DBuf<uint16_t> dO16, dP, dX;   // dO16 [.][32], dP [.][128], dX [.][256]
  DBuf<int> bnbr;                // batch-local neighbours [.][4]
  DBuf<double> red;              // reductions
  double *pinned = nullptr;
  ~Learned() {
    if (pinned) (void)hipHostFree(pinned);
  }
};

void destroy_learned(Learned *l) { delete l; }

namespace {

// ---------------------------------------------------------------- MFMA GEMMs

// C[r][n0 + c] = sum_k A[r][k] * Bn[n0 + c][k] (+ bias), c < NCB; A bf16 [R][lda], Bn bf16 [NC][K].
enum { EPI_SILU = 0, EPI_F32 = 1, EPI_DSILU = 2, EPI_BF16 = 3 };
template <int K, int NCB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(int R, const uint16_t *__restrict__ A, int lda,
                                                   const uint16_t *__restrict__ Bn, const float *__restrict__ bias,
                                                   float *__restrict__ pre, int ldp, uint16_t *__restrict__ outb,
                                                   int ldo, float *__restrict__ outf, int ldf) {
  constexpr int KP = K + 4;  // padded LDS row (8-byte aligned, spreads banks)
  __shared__ uint16_t sB[NCB * KP];
  const int n0 = blockIdx.y * NCB;
  for (int i = threadIdx.x; i < NCB * K / 4; i += 256) {
    const int n = (4 * i) / K, k = (4 * i) % K;
    *reinterpret_cast<uint2 *>(&sB[n * KP + k]) = *reinterpret_cast<const uint2 *>(&Bn[(size_t)(n0 + n) * K + k]);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = blockIdx.x * 64 + wave * 16;
  const int lr = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[NCB / 16];
#pragma unroll
  for (int t = 0; t < NCB / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int arow = row0 + lr;
#pragma unroll 4
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a = bf16x4{0, 0, 0, 0};
    if (arow < R) a = *reinterpret_cast<const bf16x4 *>(&A[(size_t)arow * lda + k0 + lk]);
#pragma unroll
    for (int t = 0; t < NCB / 16; t++) {
      const bf16x4 b = *reinterpret_cast<const bf16x4 *>(&sB[(t * 16 + lr) * KP + k0 + lk]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NCB / 16; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int row = row0 + lk + j, col = n0 + t * 16 + lr;
      if (row >= R) continue;
      float v = acc[t][j] + (bias ? bias[col] : 0.f);
      if (EPI == EPI_SILU) {
        if (pre) pre[(size_t)row * ldp + col] = v;
        outb[(size_t)row * ldo + col] = f2bf(silu_f(v));
      } else if (EPI == EPI_F32) {
        outf[(size_t)row * ldf + col] = v;
      } else if (EPI == EPI_BF16) {
        outb[(size_t)row * ldo + col] = f2bf(v);
      } else {  // EPI_DSILU: multiply by silu'(pre) of the forward pass, store bf16
        outb[(size_t)row * ldo + col] = f2bf(v * dsilu_f(pre[(size_t)row * ldp + col]));
      }
    }
}

// Forward layer out[r][c] = act(sum_k A[r][k] Bn[c][k] + bias[c]) for all N columns, on the gfx950
// v_mfma_f32_16x16x32_bf16 (8 bf16 per lane and operand: lane l holds row l%16 and the k-octet
// l/16 of each 32-wide k step; A and B use the same octets, so the product sums every k once).
// Persistent workgroups keep the whole weight matrix in LDS and stream 16-row tiles per wave, the
// next tile's A fragments loaded while the current tile's MFMAs run, so A is read from HBM once.
// MEAN: the k >= 128 half of A is the mean of the row's neighbours' k < 128 half (message passing
// fused into the load: the same sum order and rounding as nbr_mean_kernel, never stored).
template <int K, int N, bool MEAN, int EPI>
__global__ __launch_bounds__(256) void mlp_layer_kernel(int R, const uint16_t *__restrict__ A, int lda,
                                                        const int *__restrict__ nbr, const uint16_t *__restrict__ Bn,
                                                        const float *__restrict__ bias, float *__restrict__ pre,
                                                        int ldp, uint16_t *__restrict__ outb, int ldo,
                                                        float *__restrict__ outf, int ldf) {
  static_assert(K % 32 == 0 && N % 16 == 0 && (!MEAN || K == 256), "layer shape");
  constexpr int KS = K / 32, NT = N / 16, NOCT = K / 8;
  // LDS: row n of Bn as K/8 16-byte octets, octet o stored at o ^ (n & SW): the 16 rows one MFMA
  // operand read touches land on 16 distinct octet slots (no bank conflicts, no padding)
  constexpr int SW = NOCT < 16 ? NOCT - 1 : 15;
  extern __shared__ uint4 lds_raw[];
  uint4 *sB = lds_raw;
  for (int i = threadIdx.x; i < N * NOCT; i += blockDim.x) {
    const int n = i / NOCT, o = i % NOCT;
    sB[n * NOCT + (o ^ (n & SW))] = *reinterpret_cast<const uint4 *>(&Bn[(size_t)n * K + 8 * o]);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, kq = lane >> 4;
  // XCD-aware tiling: workgroup b runs on XCD b % 8 (round-robin dispatch; the grid is a multiple of
  // 8), and each XCD walks its own contiguous eighth of the tiles, so the rows a tile's neighbours
  // read (nearby in the chunk numbering) sit in the same XCD's L2
  const int nTiles = (R + 15) / 16, wpb = blockDim.x >> 6;
  const int xcd = blockIdx.x & 7, perX = (nTiles + 7) / 8;
  const int tBeg = min(nTiles, xcd * perX), tEnd = min(nTiles, tBeg + perX);
  const int stride = (gridDim.x >> 3) * wpb;
  constexpr int KA = MEAN ? KS / 2 : KS;  // k steps read directly from A
  // the directly read fragments (and for MEAN the neighbour list) of a tile
  auto load = [&](int tile, uint4 (&a)[KA], int4 &nb) {
    const int row = tile * 16 + lr;
    nb = make_int4(-1, -1, -1, -1);
    if (tile >= tEnd || row >= R) {
#pragma unroll
      for (int st = 0; st < KA; st++) a[st] = make_uint4(0u, 0u, 0u, 0u);
      return;
    }
    const uint16_t *ar = A + (size_t)row * lda + 8 * kq;
#pragma unroll
    for (int st = 0; st < KA; st++) a[st] = *reinterpret_cast<const uint4 *>(ar + 32 * st);
    if (MEAN) nb = *reinterpret_cast<const int4 *>(&nbr[(size_t)NBR * row]);
  };
  auto mma = [&](f32x4 (&acc)[NT], int st, uint4 av) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int n = t * 16 + lr;
      const uint4 bv = sB[n * NOCT + ((4 * st + kq) ^ (n & SW))];
      // transposed product (weights as the row operand): the lane then holds 4 consecutive output
      // columns of one row, stored as one 8-byte (bf16) or 16-byte (FP32) vector
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, bv), __builtin_bit_cast(bf16x8, av),
                                                       acc[t], 0, 0, 0);
    }
  };
  uint4 a[KA], an[KA];
  int4 nb, nbn;
  int tile = tBeg + (blockIdx.x >> 3) * wpb + wave;
  load(tile, a, nb);
  for (; tile < tEnd; tile += stride) {
    // MEAN: issue this tile's neighbour gathers first, then the next tile's loads; the MFMAs of the
    // directly read half run while both are in flight
    uint4 g[MEAN ? 4 : 1][MEAN ? KS / 2 : 1];
    if (MEAN) {
      const int d[4] = {nb.x, nb.y, nb.z, nb.w};
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int st = 0; st < KS / 2; st++)
          g[j][st] = d[j] >= 0 ? *reinterpret_cast<const uint4 *>(A + (size_t)d[j] * lda + 32 * st + 8 * kq)
                               : make_uint4(0u, 0u, 0u, 0u);
    }
    load(tile + stride, an, nbn);
    // keep the weight fragments in LDS (read per tile) instead of letting them be hoisted into
    // registers, which would cost the occupancy that hides the A loads' latency
    __asm__ volatile("" ::: "memory");
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < KA; st++) mma(acc, st, a[st]);
    if (MEAN) {
      // mean over the present neighbours, summed in slot order (as nbr_mean_kernel)
      const int d[4] = {nb.x, nb.y, nb.z, nb.w};
      const int cnt = (d[0] >= 0) + (d[1] >= 0) + (d[2] >= 0) + (d[3] >= 0);
      const float ic = cnt ? 1.0f / cnt : 0.f;
#pragma unroll
      for (int st = 0; st < KS / 2; st++) {
        float sm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (d[j] < 0) continue;
          const uint16_t *h = reinterpret_cast<const uint16_t *>(&g[j][st]);
#pragma unroll
          for (int q = 0; q < 8; q++) sm[q] += bf2f(h[q]);
        }
        uint16_t o[8];
#pragma unroll
        for (int q = 0; q < 8; q++) o[q] = f2bf(sm[q] * ic);
        mma(acc, KS / 2 + st, *reinterpret_cast<const uint4 *>(o));
      }
    }
    const int row = tile * 16 + lr;
    if (row < R) {
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const int col = t * 16 + 4 * kq;
        const float4 bb = *reinterpret_cast<const float4 *>(&bias[col]);
        const float v[4] = {acc[t][0] + bb.x, acc[t][1] + bb.y, acc[t][2] + bb.z, acc[t][3] + bb.w};
        if (EPI == EPI_SILU) {
          if (pre) *reinterpret_cast<float4 *>(&pre[(size_t)row * ldp + col]) = make_float4(v[0], v[1], v[2], v[3]);
          const uint32_t lo = f2bf(silu_f(v[0])) | ((uint32_t)f2bf(silu_f(v[1])) << 16);
          const uint32_t hi = f2bf(silu_f(v[2])) | ((uint32_t)f2bf(silu_f(v[3])) << 16);
          *reinterpret_cast<uint2 *>(&outb[(size_t)row * ldo + col]) = make_uint2(lo, hi);
        } else {
          *reinterpret_cast<float4 *>(&outf[(size_t)row * ldf + col]) = make_float4(v[0], v[1], v[2], v[3]);
        }
      }
    }
#pragma unroll
    for (int st = 0; st < KA; st++) a[st] = an[st];
    nb = nbn;
  }
}

// Weight gradient partials: part[slice][k][n] = sum_{r in slice} A[r][k] * D[r][n] for a 32x32 block
// (4 waves, one 16x16 tile each); the row index is the MFMA reduction dimension.
__global__ __launch_bounds__(256) void wgrad_kernel(int R, const uint16_t *__restrict__ A, int lda,
                                                    const uint16_t *__restrict__ D, int ldd, int K, int N,
                                                    float *__restrict__ part) {
  const int slice = blockIdx.x, kb = blockIdx.y * 32, nb = blockIdx.z * 32;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k0 = kb + (wave & 1) * 16, n0 = nb + (wave >> 1) * 16;
  const int lr = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  const int rb = slice * WG_ROWS, re = min(R, rb + WG_ROWS);
  for (int r0 = rb; r0 < re; r0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = r0 + lk + q;
      const bool in = r < re;
      a[q] = in ? (short)A[(size_t)r * lda + k0 + lr] : (short)0;
      b[q] = in ? (short)D[(size_t)r * ldd + n0 + lr] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) part[((size_t)slice * K + k0 + lk + j) * N + n0 + lr] = acc[j];
}

// grad[k][n] (flat, unpadded to Kr x Nr) = sum over slices; also bias grads from column sums of D
__global__ void wgrad_reduce(const float *__restrict__ part, int nslice, int K, int N, int Kr, int Nr,
                             float *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Kr * Nr) return;
  const int k = i / Nr, n = i % Nr;
  float s = 0.f;
  for (int sl = 0; sl < nslice; sl++) s += part[((size_t)sl * K + k) * N + n];
  g[i] = s;
}

__global__ void colsum_kernel(int R, const uint16_t *__restrict__ D, int ldd, int Nr, float *__restrict__ g) {
  __shared__ float sm[256];
  const int n = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < R; r += 256) s += bf2f(D[(size_t)r * ldd + n]);
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && n < Nr) g[n] = sm[0];
}

// ---------------------------------------------------------------- graph and elementwise kernels

// H[c][128:256] = mean over neighbours of H[nb][0:128]  (8 columns per thread)
__global__ void nbr_mean_kernel(int R, const int *__restrict__ nbr, uint16_t *__restrict__ H) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)R * 16) return;
  const int c = (int)(t >> 4), g = (int)(t & 15) * 8;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int cnt = 0;
  for (int j = 0; j < NBR; j++) {
    const int d = nbr[(size_t)NBR * c + j];
    if (d < 0) continue;
    cnt++;
    const uint4 v = *reinterpret_cast<const uint4 *>(&H[(size_t)d * 256 + g]);
    const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
    for (int q = 0; q < 8; q++) s[q] += bf2f(h[q]);
  }
  uint16_t o[8];
  const float ic = cnt ? 1.0f / cnt : 0.f;
  for (int q = 0; q < 8; q++) o[q] = f2bf(s[q] * ic);
  *reinterpret_cast<uint4 *>(&H[(size_t)c * 256 + 128 + g]) = *reinterpret_cast<const uint4 *>(o);
}

// Backward of [h, mean_nb h] -> h, then silu': dP[c][k] = (dX[c][k] + sum_nb dX[nb][128+k]/cnt_nb) silu'(pre)
__global__ void mp_backward_kernel(int R, const int *__restrict__ nbr, const uint16_t *__restrict__ dX,
                                   const float *__restrict__ pre, uint16_t *__restrict__ dP) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)R * NH) return;
  const int c = (int)(t / NH), k = (int)(t % NH);
  float s = bf2f(dX[(size_t)c * 256 + k]);
  for (int j = 0; j < NBR; j++) {
    const int d = nbr[(size_t)NBR * c + j];
    if (d < 0) continue;
    int cnt = 0;
    for (int q = 0; q < NBR; q++) cnt += nbr[(size_t)NBR * d + q] >= 0;
    s += bf2f(dX[(size_t)d * 256 + 128 + k]) / (float)cnt;
  }
  dP[(size_t)c * NH + k] = f2bf(s * dsilu_f(pre[(size_t)c * NH + k]));
}

// Features (20, zero-padded to 32): U/S, log(v/vmean), 4 area-normals / v^(2/3), gamma c / v^(1/3),
// |u|/c.  Cells [c0, c0+R) of the full-mesh state y.
__global__ void features_kernel(int R, int c0, int nC, const double *__restrict__ y, const float *__restrict__ geo,
                                const float *__restrict__ S, double gamma, double gm, uint16_t *__restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int c = c0 + i;
  double U[5];
  for (int k = 0; k < 5; k++) U[k] = y[(size_t)k * nC + c];
  uint16_t x[NIP];
  for (int k = 0; k < NIP; k++) x[k] = 0;
  for (int k = 0; k < 5; k++) x[k] = f2bf((float)(U[k] / S[k]));
  const float *gc = geo + (size_t)13 * c;
  for (int k = 0; k < 13; k++) x[5 + k] = f2bf(gc[k]);
  const double ir = 1.0 / U[0];
  const double q2 = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * ir * ir;
  const double p = (gm - 1.0) * (U[4] - 0.5 * U[0] * q2);
  const double cs = sqrt(fmax(gm * p * ir, 1e-300));
  const double vol13 = exp(gc[0] / 3.0) * (double)geo[(size_t)13 * nC];  // v^(1/3) = exp(log(v/vm)/3) vm^(1/3)
  x[18] = f2bf((float)(gamma * cs / vol13));
  x[19] = f2bf((float)(sqrt(q2) / cs));
  uint4 *dst = reinterpret_cast<uint4 *>(&X[(size_t)i * NIP]);
  const uint4 *src = reinterpret_cast<const uint4 *>(x);
  for (int k = 0; k < NIP / 8; k++) dst[k] = src[k];
}

// The output scale: O[c] *= k_c = min(1, x_c[18] (1 + x_c[19])) from the cell's (BF16) features,
// its CFL number gamma (c + |u|) / v^(1/3) capped at 1: the operator tends to the identity as the
// time step shrinks, where the Newton matrix I - gamma J differs from I by O(CFL).
__device__ inline float out_kappa(const uint16_t *__restrict__ X, int i) {
  return fminf(1.f, bf2f(X[(size_t)i * NIP + 18]) * (1.f + bf2f(X[(size_t)i * NIP + 19])));
}
__global__ void out_scale_kernel(int R, const uint16_t *__restrict__ X, float *__restrict__ O) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const float k = out_kappa(X, i);
  for (int q = 0; q < NO; q++) O[(size_t)NOP * i + q] *= k;
}

// t[c][a] = rhat_a + sum_b O[c][5a+b] rhat_b, rhat = r / S (r: SoA with stride nC, cells [c0, c0+R));
// the 128-byte output row is read as 16-byte vectors
__global__ void apply_t_kernel(int R, int c0, int nC, const float *__restrict__ O, const float *__restrict__ S,
                               const double *__restrict__ r, float *__restrict__ t, float *__restrict__ rh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  float q[5];
  for (int b = 0; b < 5; b++) q[b] = (float)(r[(size_t)b * nC + c0 + i] / S[b]);
  const float4 *o4 = reinterpret_cast<const float4 *>(O + (size_t)NOP * i);
  float o[28];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const float4 v = o4[k];
    o[4 * k] = v.x;
    o[4 * k + 1] = v.y;
    o[4 * k + 2] = v.z;
    o[4 * k + 3] = v.w;
  }
  float ta[8];
#pragma unroll
  for (int a = 0; a < 5; a++) {
    float sa = q[a];
#pragma unroll
    for (int b = 0; b < 5; b++) sa += o[5 * a + b] * q[b];
    ta[a] = sa;
  }
  ta[5] = ta[6] = ta[7] = 0.f;
  float4 *t4 = reinterpret_cast<float4 *>(t + (size_t)8 * i);
  t4[0] = make_float4(ta[0], ta[1], ta[2], ta[3]);
  t4[1] = make_float4(ta[4], 0.f, 0.f, 0.f);
  if (rh)
    for (int a = 0; a < 5; a++) rh[(size_t)8 * i + a] = q[a];
}

// z_a[c] = S_a ((1 + s4[c]) t[c][a] + sum_j s_j[c] t[nb_j(c)][a]), s = O[c][25:30]: weights of the
// 4 face neighbours (neib slot order) and of the cell itself
__global__ void apply_z_kernel(int R, int nC, const int *__restrict__ nbr, const float *__restrict__ O,
                               const float *__restrict__ S, const float *__restrict__ t, double *__restrict__ z,
                               unsigned *__restrict__ nf) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  const float4 w0 = reinterpret_cast<const float4 *>(O + (size_t)NOP * c)[6];  // o[24..27]
  const float2 w1 = reinterpret_cast<const float2 *>(O + (size_t)NOP * c)[14];  // o[28..29]
  const float sw[5] = {w0.y, w0.z, w0.w, w1.x, w1.y};
  const int4 nb = reinterpret_cast<const int4 *>(nbr)[c];
  const int d[4] = {nb.x, nb.y, nb.z, nb.w};
  const float4 *t4 = reinterpret_cast<const float4 *>(t);
  const float4 tc0 = t4[2 * (size_t)c];
  const float tc4 = t4[2 * (size_t)c + 1].x;
  float acc[5] = {(1.f + sw[4]) * tc0.x, (1.f + sw[4]) * tc0.y, (1.f + sw[4]) * tc0.z, (1.f + sw[4]) * tc0.w,
                  (1.f + sw[4]) * tc4};
#pragma unroll
  for (int j = 0; j < NBR; j++) {
    if (d[j] < 0) continue;
    const float4 a = t4[2 * (size_t)d[j]];
    const float a4 = t4[2 * (size_t)d[j] + 1].x;
    acc[0] += sw[j] * a.x;
    acc[1] += sw[j] * a.y;
    acc[2] += sw[j] * a.z;
    acc[3] += sw[j] * a.w;
    acc[4] += sw[j] * a4;
  }
  bool bad = false;
  for (int a = 0; a < 5; a++) {
    const double o = (double)S[a] * (double)acc[a];
    z[(size_t)a * nC + c] = o;
    bad |= !isfinite(o);
  }
  if (bad && nf) atomicOr(nf, 1u);
}

// loss pieces: e = 2 (zhat - dhat) / R; ds[c][j] = e_c . t_nb(j) (j < 4), ds[c][4] = e_c . t_c (the
// weight gradients); part[block] = sum (zhat - dhat)^2
__global__ void loss_e_kernel(int R, int c0, int nC, const int *__restrict__ nbr, const float *__restrict__ O,
                              const float *__restrict__ S, const float *__restrict__ t, const double *__restrict__ delta,
                              float *__restrict__ e, float *__restrict__ ds, double *__restrict__ part) {
  __shared__ double sm[256];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (c < R) {
    const float *o = O + (size_t)NOP * c;
    float zh[5];
    for (int a = 0; a < 5; a++) zh[a] = (1.f + o[29]) * t[(size_t)8 * c + a];
    int nb[NBR];
    for (int j = 0; j < NBR; j++) {
      nb[j] = nbr[(size_t)NBR * c + j];
      if (nb[j] < 0) continue;
      for (int a = 0; a < 5; a++) zh[a] += o[25 + j] * t[(size_t)8 * nb[j] + a];
    }
    float ea[5];
    for (int a = 0; a < 5; a++) {
      const float dd = zh[a] - (float)(delta[(size_t)a * nC + c0 + c] / S[a]);
      l += (double)dd * dd;
      ea[a] = 2.0f * dd / (float)R;
      e[(size_t)8 * c + a] = ea[a];
    }
    for (int j = 0; j < NBR; j++) {
      float g = 0.f;
      if (nb[j] >= 0)
        for (int a = 0; a < 5; a++) g += ea[a] * t[(size_t)8 * nb[j] + a];
      ds[(size_t)8 * c + j] = g;
    }
    float g4 = 0.f;
    for (int a = 0; a < 5; a++) g4 += ea[a] * t[(size_t)8 * c + a];
    ds[(size_t)8 * c + 4] = g4;
  }
  sm[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// dt_c = (1 + s4_c) e_c + sum over neighbours d of c of s_j(d) e_d, j the slot of c in d's list;
// dO[5a+b] = dt_a rhat_b, dO[25+j] = ds[c][j], both times k_c (the output scale) for the last
// layer's gradient; bf16 dO [R][32]
__global__ void loss_dO_kernel(int R, const int *__restrict__ nbr, const float *__restrict__ O,
                               const float *__restrict__ e, const float *__restrict__ rh, const float *__restrict__ ds,
                               const uint16_t *__restrict__ X, uint16_t *__restrict__ dO) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  float dt[5];
  const float sc = 1.f + O[(size_t)NOP * c + 29];
  for (int a = 0; a < 5; a++) dt[a] = sc * e[(size_t)8 * c + a];
  for (int j = 0; j < NBR; j++) {
    const int d = nbr[(size_t)NBR * c + j];
    if (d < 0) continue;
    for (int q = 0; q < NBR; q++) {
      if (nbr[(size_t)NBR * d + q] != c) continue;
      const float w = O[(size_t)NOP * d + 25 + q];
      for (int a = 0; a < 5; a++) dt[a] += w * e[(size_t)8 * d + a];
    }
  }
  uint16_t *o = dO + (size_t)NOP * c;
  const float k = out_kappa(X, c);
  for (int a = 0; a < 5; a++)
    for (int b = 0; b < 5; b++) o[5 * a + b] = f2bf(k * dt[a] * rh[(size_t)8 * c + b]);
  for (int j = 0; j < 5; j++) o[25 + j] = f2bf(k * ds[(size_t)8 * c + j]);
  o[30] = o[31] = 0;
}

// batch-local neighbour table for cells [c0, c0+R)
__global__ void batch_nbr_kernel(int R, int c0, const int *__restrict__ nbr, int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  for (int j = 0; j < NBR; j++) {
    const int d = nbr[(size_t)NBR * (c0 + i) + j];
    out[(size_t)NBR * i + j] = (d >= c0 && d < c0 + R) ? d - c0 : -1;
  }
}

__global__ void adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 operand copies from the FP32 master parameters
__global__ void pack_kernel(const float *__restrict__ p, uint16_t *W0t, uint16_t *W1t, uint16_t *W2t, uint16_t *W3t,
                            uint16_t *W1b, uint16_t *W2b, uint16_t *W3b, float *bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NH * NIP) {  // W0t [128][32]
    const int n = i / NIP, k = i % NIP;
    W0t[i] = k < NI ? f2bf(p[OFF_W0 + k * NH + n]) : 0;
  }
  if (i < NH * 256) {  // W1t, W2t [128][256]; W1b, W2b [256][128]
    const int n = i / 256, k = i % 256;
    W1t[i] = f2bf(p[OFF_W1 + k * NH + n]);
    W2t[i] = f2bf(p[OFF_W2 + k * NH + n]);
    const int kk = i / NH, nn = i % NH;
    W1b[i] = f2bf(p[OFF_W1 + kk * NH + nn]);
    W2b[i] = f2bf(p[OFF_W2 + kk * NH + nn]);
  }
  if (i < NOP * NH) {  // W3t [32][128]; W3b [128][32]
    const int n = i / NH, k = i % NH;
    W3t[i] = n < NO ? f2bf(p[OFF_W3 + k * NO + n]) : 0;
    const int kk = i / NOP, nn = i % NOP;
    W3b[i] = nn < NO ? f2bf(p[OFF_W3 + kk * NO + nn]) : 0;
  }
  if (i < 3 * NH + NOP) {
    if (i < NH) bias[i] = p[OFF_B0 + i];
    else if (i < 2 * NH) bias[i] = p[OFF_B1 + i - NH];
    else if (i < 3 * NH) bias[i] = p[OFF_B2 + i - 2 * NH];
    else bias[i] = (i - 3 * NH) < NO ? p[OFF_B3 + i - 3 * NH] : 0.f;
  }
}

// per-variable state scale S = rms(rho), rms(|rhou|)/sqrt(3) (x3), rms(rhoE), floored at 1e-6 rms(rhoE)
__global__ void scale_partial(int nC, const double *__restrict__ y, double *__restrict__ part) {
  __shared__ double sm[3][256];
  double a = 0, b = 0, c = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nC; i += gridDim.x * 256) {
    a += y[i] * y[i];
    for (int k = 1; k < 4; k++) b += y[(size_t)k * nC + i] * y[(size_t)k * nC + i];
    c += y[(size_t)4 * nC + i] * y[(size_t)4 * nC + i];
  }
  sm[0][threadIdx.x] = a;
  sm[1][threadIdx.x] = b;
  sm[2][threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 3; k++) sm[k][threadIdx.x] += sm[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ void scale_final(int nb, int nC, const double *__restrict__ part, float *__restrict__ S) {
  if (threadIdx.x) return;
  double a = 0, b = 0, c = 0;
  for (int i = 0; i < nb; i++) { a += part[3 * i]; b += part[3 * i + 1]; c += part[3 * i + 2]; }
  const double ra = sqrt(a / nC), rb = sqrt(b / (3.0 * nC)), rc = sqrt(c / nC);
  const double fl = 1e-6 * rc + 1e-300;
  S[0] = (float)fmax(ra, fl);
  S[1] = S[2] = S[3] = (float)fmax(rb, fmax(1e-3 * ra, fl));
  S[4] = (float)fmax(rc, fl);
}

template <int K, int NCB, int EPI>
void gemm(hipStream_t s, int R, int NC, const uint16_t *A, int lda, const uint16_t *Bn, const float *bias, float *pre,
          int ldp, uint16_t *outb, int ldo, float *outf, int ldf) {
  dim3 g((R + 63) / 64, NC / NCB);
  hipLaunchKernelGGL((gemm_kernel<K, NCB, EPI>), g, dim3(256), 0, s, R, A, lda, Bn, bias, pre, ldp, outb, ldo, outf, ldf);
}

template <int K, int N, bool MEAN, int EPI>
void layer(hipStream_t s, int R, const uint16_t *A, int lda, const int *nbr, const uint16_t *Bn, const float *bias,
           float *pre, int ldp, uint16_t *outb, int ldo, float *outf, int ldf) {
  const size_t lds = sizeof(uint16_t) * N * K;
  const int tiles = (R + 15) / 16;
  const int per = std::max(1, (int)(160 * 1024 / lds));  // workgroups per CU that the LDS allows
  const int grid = 8 * std::max(1, std::min((tiles + 31) / 32, 32 * std::min(per, 4)));  // a multiple of 8
  hipLaunchKernelGGL((mlp_layer_kernel<K, N, MEAN, EPI>), dim3(grid), dim3(256), lds, s, R, A, lda, nbr, Bn, bias,
                     pre, ldp, outb, ldo, outf, ldf);
}

int ensure_fwd(Learned *L, size_t n) {
  if (L->capF >= n) return 0;
  L->capF = n;
  return (L->X.alloc(n * NIP) || L->H0.alloc(n * 256) || L->H1.alloc(n * 256) || L->H2.alloc(n * NH) ||
          L->O.alloc(n * NOP) || L->t.alloc(n * 8)) ? -1 : 0;
}

int ensure_train(Learned *L, size_t n) {
  if (ensure_fwd(L, n)) return -1;
  if (L->capT >= n) return 0;
  L->capT = n;
  const size_t ns = (n + WG_ROWS - 1) / WG_ROWS;
  return (L->pre0.alloc(n * NH) || L->pre1.alloc(n * NH) || L->pre2.alloc(n * NH) || L->rh.alloc(n * 8) ||
          L->e.alloc(n * 8) || L->ds.alloc(n * 8) || L->part.alloc(ns * 256 * NH) || L->dO16.alloc(n * NOP) ||
          L->dP.alloc(n * NH) || L->dX.alloc(n * 256) || L->bnbr.alloc(n * NBR) ||
          L->red.alloc(std::max<size_t>(1024, (n + 255) / 256 + 8))) ? -1 : 0;
}

int compute_scale(fos_euler *ctx, Learned *L, const double *y) {
  const int nb = std::max(1, std::min(256, (ctx->nC + 255) / 256));
  hipLaunchKernelGGL(scale_partial, dim3(nb), dim3(256), 0, ctx->stream, ctx->nC, y, L->red.p);
  hipLaunchKernelGGL(scale_final, dim3(1), dim3(64), 0, ctx->stream, nb, ctx->nC, L->red.p, L->scale.p);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

// network forward on R cells starting at c0 (nbr: neighbour table with indices local to the range)
int forward(fos_euler *ctx, Learned *L, const double *y, int c0, int R, const int *nbr, double gamma, bool keepPre) {
  hipStream_t s = ctx->stream;
  const float *bias = L->bias.p;
  const dim3 g16((unsigned)(((size_t)R * 16 + 255) / 256));
  hipLaunchKernelGGL(features_kernel, dim3((R + 255) / 256), dim3(256), 0, s, R, c0, ctx->nC, y, L->geo.p,
                     L->scale.p, gamma, ctx->gamm, L->X.p);
  // inference fuses the neighbour means into the next layer's loads; training stores them (the
  // weight gradients read [h, mean h])
  layer<NIP, NH, false, EPI_SILU>(s, R, L->X.p, NIP, nullptr, L->W0t.p, bias, keepPre ? L->pre0.p : nullptr, NH, L->H0.p,
                                  256, nullptr, 0);
  if (keepPre) {
    hipLaunchKernelGGL(nbr_mean_kernel, g16, dim3(256), 0, s, R, nbr, L->H0.p);
    layer<256, NH, false, EPI_SILU>(s, R, L->H0.p, 256, nbr, L->W1t.p, bias + NH, L->pre1.p, NH, L->H1.p, 256, nullptr,
                                    0);
    hipLaunchKernelGGL(nbr_mean_kernel, g16, dim3(256), 0, s, R, nbr, L->H1.p);
    layer<256, NH, false, EPI_SILU>(s, R, L->H1.p, 256, nbr, L->W2t.p, bias + 2 * NH, L->pre2.p, NH, L->H2.p, NH,
                                    nullptr, 0);
  } else {
    layer<256, NH, true, EPI_SILU>(s, R, L->H0.p, 256, nbr, L->W1t.p, bias + NH, nullptr, NH, L->H1.p, 256, nullptr, 0);
    layer<256, NH, true, EPI_SILU>(s, R, L->H1.p, 256, nbr, L->W2t.p, bias + 2 * NH, nullptr, NH, L->H2.p, NH, nullptr,
                                   0);
  }
  layer<NH, NOP, false, EPI_F32>(s, R, L->H2.p, NH, nullptr, L->W3t.p, bias + 3 * NH, nullptr, 0, nullptr, 0, L->O.p,
                                 NOP);
  hipLaunchKernelGGL(out_scale_kernel, dim3((R + 255) / 256), dim3(256), 0, s, R, L->X.p, L->O.p);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

int repack(fos_euler *ctx, Learned *L) {
  hipLaunchKernelGGL(pack_kernel, dim3((NH * 256 + 255) / 256), dim3(256), 0, ctx->stream, L->par.p, L->W0t.p,
                     L->W1t.p, L->W2t.p, L->W3t.p, L->W1b.p, L->W2b.p, L->W3b.p, L->bias.p);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

int create_learned(fos_euler *ctx) {
  if (ctx->learned) return 0;
  if (ctx->maxNF > NBR) {
    set_error("learned preconditioner: cells with %d faces (tetrahedra only)", ctx->maxNF);
    return -1;
  }
  Learned *L = new Learned();  // attached to the context only once complete
  const int nC = ctx->nC, mf = ctx->maxNF;
  L->nC = nC;
  std::vector<int> nbr((size_t)NBR * std::max(nC, 1), -1);
  for (int c = 0; c < nC; c++)
    for (int j = 0; j < mf; j++) {
      const int e = ctx->hNeib[(size_t)mf * c + j];
      nbr[(size_t)NBR * c + j] = (e >= 0 && e < nC) ? e : -1;
    }
  // static geometry features: log(v/vmean) and the 4 outward area vectors / v^(2/3) in neib order
  double vm = 0;
  for (int c = 0; c < nC; c++) vm += ctx->hV[c];
  vm /= std::max(nC, 1);
  std::vector<float> geo((size_t)13 * nC + 1, 0.f);
  for (int c = 0; c < nC; c++) geo[(size_t)13 * c] = (float)std::log(ctx->hV[c] / vm);
  for (int i = 0; i < ctx->nP; i++) {
    const int m = ctx->hIEP[2 * i], n = ctx->hIEP[2 * i + 1];
    const double a = ctx->hAP[i];
    for (int side = 0; side < 2; side++) {
      const int c = side ? n : m, o = side ? m : n;
      if (c >= nC) continue;
      for (int j = 0; j < mf; j++)
        if (ctx->hNeib[(size_t)mf * c + j] == o) {
          const double sc = (side ? -a : a) / std::pow(ctx->hV[c], 2.0 / 3.0);
          for (int d = 0; d < 3; d++) geo[(size_t)13 * c + 1 + 3 * j + d] = (float)(sc * ctx->hNormP[3 * (size_t)i + d]);
          break;
        }
    }
  }
  geo[(size_t)13 * nC] = (float)std::cbrt(vm);
  hipStream_t s = ctx->stream;
  if (L->nbr.upload(nbr.data(), nbr.size(), s) || L->geo.upload(geo.data(), geo.size(), s) || L->par.alloc(NPAR) ||
      L->m1.alloc(NPAR) || L->m2.alloc(NPAR) || L->grad.alloc(NPAR) || L->bias.alloc(3 * NH + NOP) ||
      L->W0t.alloc(NH * NIP) || L->W1t.alloc(NH * 256) || L->W2t.alloc(NH * 256) || L->W3t.alloc(NOP * NH) ||
      L->W1b.alloc(256 * NH) || L->W2b.alloc(256 * NH) || L->W3b.alloc(NH * NOP) || L->scale.alloc(8) ||
      L->red.alloc(1024) || hipHostMalloc((void **)&L->pinned, 4096) != hipSuccess) {
    if (!L->pinned) set_error("learned preconditioner: pinned host buffer allocation failed");
    delete L;
    return -1;
  }
  ctx->learned = L;
  return 0;
}

int set_params(fos_euler *ctx, const float *host) {
  Learned *L = ctx->learned;
  FOS_HIP_TRY(hipMemcpyAsync(L->par.p, host, sizeof(float) * NPAR, hipMemcpyHostToDevice, ctx->stream));
  FOS_HIP_TRY(hipMemsetAsync(L->m1.p, 0, sizeof(float) * NPAR, ctx->stream));
  FOS_HIP_TRY(hipMemsetAsync(L->m2.p, 0, sizeof(float) * NPAR, ctx->stream));
  L->step = 0;
  if (repack(ctx, L)) return -1;
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  L->ready = true;
  return 0;
}

// forward + loss + backward for cells [c0, c0+R) of a snapshot; the gradient is left in L->grad
int loss_grad(fos_euler *ctx, const double *y, const double *r, const double *delta, double gamma, int c0, int R,
              double *loss) {
  Learned *L = ctx->learned;
  hipStream_t s = ctx->stream;
  L->haveO = false;
  if (ensure_train(L, R) || compute_scale(ctx, L, y)) return -1;
  hipLaunchKernelGGL(batch_nbr_kernel, dim3((R + 255) / 256), dim3(256), 0, s, R, c0, L->nbr.p, L->bnbr.p);
  const int *nb = L->bnbr.p;
  if (forward(ctx, L, y, c0, R, nb, gamma, true)) return -1;
  const int gb = (R + 255) / 256;
  hipLaunchKernelGGL(apply_t_kernel, dim3(gb), dim3(256), 0, s, R, c0, ctx->nC, L->O.p, L->scale.p, r, L->t.p,
                     L->rh.p);
  hipLaunchKernelGGL(loss_e_kernel, dim3(gb), dim3(256), 0, s, R, c0, ctx->nC, nb, L->O.p, L->scale.p, L->t.p, delta,
                     L->e.p, L->ds.p, L->red.p);
  hipLaunchKernelGGL(loss_dO_kernel, dim3(gb), dim3(256), 0, s, R, nb, L->O.p, L->e.p, L->rh.p, L->ds.p, L->X.p,
                     L->dO16.p);
  std::vector<double> parts(gb);
  FOS_HIP_TRY(hipMemcpyAsync(parts.data(), L->red.p, sizeof(double) * gb, hipMemcpyDeviceToHost, s));
  const int ns = (R + WG_ROWS - 1) / WG_ROWS;
  float *g = L->grad.p;
  const dim3 gr((unsigned)(((size_t)R * NH + 255) / 256));
  // layer 3: dW3 = H2^T dO, db3 = sum dO; dP = (dO W3^T) silu'(pre2)
  hipLaunchKernelGGL(wgrad_kernel, dim3(ns, NH / 32, NOP / 32), dim3(256), 0, s, R, L->H2.p, NH, L->dO16.p, NOP, NH,
                     NOP, L->part.p);
  hipLaunchKernelGGL(wgrad_reduce, dim3((NH * NO + 255) / 256), dim3(256), 0, s, L->part.p, ns, NH, NOP, NH, NO,
                     g + OFF_W3);
  hipLaunchKernelGGL(colsum_kernel, dim3(NOP), dim3(256), 0, s, R, L->dO16.p, NOP, NO, g + OFF_B3);
  gemm<NOP, 64, EPI_DSILU>(s, R, NH, L->dO16.p, NOP, L->W3b.p, nullptr, L->pre2.p, NH, L->dP.p, NH, nullptr, 0);
  // layers 2 and 1: dW = [h, m]^T dP, db = sum dP; dX = dP W^T; dP <- mp-backward(dX) silu'(pre)
  const uint16_t *Hin[2] = {L->H1.p, L->H0.p};
  const uint16_t *Wb[2] = {L->W2b.p, L->W1b.p};
  float *preIn[2] = {L->pre1.p, L->pre0.p};
  const int offW[2] = {OFF_W2, OFF_W1}, offB[2] = {OFF_B2, OFF_B1};
  for (int l = 0; l < 2; l++) {
    hipLaunchKernelGGL(wgrad_kernel, dim3(ns, 256 / 32, NH / 32), dim3(256), 0, s, R, Hin[l], 256, L->dP.p, NH, 256,
                       NH, L->part.p);
    hipLaunchKernelGGL(wgrad_reduce, dim3((256 * NH + 255) / 256), dim3(256), 0, s, L->part.p, ns, 256, NH, 256, NH,
                       g + offW[l]);
    hipLaunchKernelGGL(colsum_kernel, dim3(NH), dim3(256), 0, s, R, L->dP.p, NH, NH, g + offB[l]);
    gemm<NH, 64, EPI_BF16>(s, R, 256, L->dP.p, NH, Wb[l], nullptr, nullptr, 0, L->dX.p, 256, nullptr, 0);
    hipLaunchKernelGGL(mp_backward_kernel, gr, dim3(256), 0, s, R, nb, L->dX.p, preIn[l], L->dP.p);
  }
  // layer 0: dW0 = X^T dP, db0
  hipLaunchKernelGGL(wgrad_kernel, dim3(ns, NIP / 32, NH / 32), dim3(256), 0, s, R, L->X.p, NIP, L->dP.p, NH, NIP, NH,
                     L->part.p);
  hipLaunchKernelGGL(wgrad_reduce, dim3((NI * NH + 255) / 256), dim3(256), 0, s, L->part.p, ns, NIP, NH, NI, NH,
                     g + OFF_W0);
  hipLaunchKernelGGL(colsum_kernel, dim3(NH), dim3(256), 0, s, R, L->dP.p, NH, NH, g + OFF_B0);
  FOS_HIP_TRY(hipGetLastError());
  FOS_HIP_TRY(hipStreamSynchronize(s));
  double l2 = 0.0;
  for (double v : parts) l2 += v;
  if (loss) *loss = l2 / R;
  return 0;
}

}  // namespace

int learned_psetup(fos_euler *ctx, const double *y, double gamma) {
  Learned *L = ctx->learned;
  if (!L || !L->ready) {
    set_error("learned preconditioner: no parameters (fos_learned_init or fos_learned_set_params)");
    return -1;
  }
  L->haveO = false;  // training overwrites O with batch rows; the mesh output is recomputed here
  if (ensure_fwd(L, ctx->nC) || compute_scale(ctx, L, y)) return -1;
  if (forward(ctx, L, y, 0, ctx->nC, L->nbr.p, gamma, false)) return -1;
  L->haveO = true;
  return 0;
}

int learned_psolve(fos_euler *ctx, const double *r, double *z) {
  Learned *L = ctx->learned;
  if (!L->haveO) {
    set_error("psolve_learned: called before psetup_learned");
    return -1;
  }
  const int nC = ctx->nC, gb = (nC + 255) / 256;
  hipLaunchKernelGGL(apply_t_kernel, dim3(gb), dim3(256), 0, ctx->stream, nC, 0, nC, L->O.p, L->scale.p, r, L->t.p,
                     nullptr);
  hipLaunchKernelGGL(apply_z_kernel, dim3(gb), dim3(256), 0, ctx->stream, nC, nC, L->nbr.p, L->O.p, L->scale.p, L->t.p,
                     z, ctx->dNonFinite.p);
  FOS_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fos

using namespace fos;

extern "C" {

int fos_learned_nparams(void) { return NPAR; }

int fos_learned_init(fos_euler *ctx, uint64_t seed) {
  if (!ctx || create_learned(ctx)) return -1;
  // Glorot-uniform hidden layers, zero output layer: the untrained network is the identity
  std::mt19937_64 rng(seed);
  std::vector<float> p(NPAR, 0.f);
  auto fill = [&](int off, int fin, int fout) {
    const double a = std::sqrt(6.0 / (fin + fout));
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < fin * fout; i++) p[off + i] = (float)u(rng);
  };
  fill(OFF_W0, NI, NH);
  fill(OFF_W1, 2 * NH, NH);
  fill(OFF_W2, 2 * NH, NH);
  return set_params(ctx, p.data());
}

int fos_learned_set_params(fos_euler *ctx, const float *params) {
  if (!ctx || !params || create_learned(ctx)) return -1;
  return set_params(ctx, params);
}

int fos_learned_get_params(fos_euler *ctx, float *params) {
  if (!ctx || !params || !ctx->learned) return -1;
  FOS_HIP_TRY(hipMemcpyAsync(params, ctx->learned->par.p, sizeof(float) * NPAR, hipMemcpyDeviceToHost, ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

int fos_learned_loss_grad(fos_euler *ctx, const double *y, const double *r, const double *delta, double gamma,
                          int batchStart, int batchSize, double *loss, float *grad) {
  if (!ctx || !ctx->learned || !ctx->learned->ready) { set_error("learned: not initialised"); return -1; }
  if (batchStart < 0 || batchSize <= 0 || batchStart + batchSize > ctx->nC) { set_error("learned: bad batch"); return -1; }
  if (loss_grad(ctx, y, r, delta, gamma, batchStart, batchSize, loss)) return -1;
  if (grad) {
    FOS_HIP_TRY(hipMemcpyAsync(grad, ctx->learned->grad.p, sizeof(float) * NPAR, hipMemcpyDeviceToHost, ctx->stream));
    FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int fos_learned_train_step(fos_euler *ctx, const double *y, const double *r, const double *delta, double gamma,
                           int batchStart, int batchSize, double lr, double *loss) {
  if (fos_learned_loss_grad(ctx, y, r, delta, gamma, batchStart, batchSize, loss, nullptr)) return -1;
  Learned *L = ctx->learned;
  L->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)L->step), c2 = 1.f - std::pow(b2, (float)L->step);
  hipLaunchKernelGGL(adam_kernel, dim3((NPAR + 255) / 256), dim3(256), 0, ctx->stream, NPAR, L->par.p, L->grad.p,
                     L->m1.p, L->m2.p, (float)lr, b1, b2, 1e-8f, c1, c2);
  return repack(ctx, L);
}

int fos_learned_inputs(fos_euler *ctx, const double *y, double gamma, float *X, float *S) {
  if (!ctx || create_learned(ctx)) return -1;
  Learned *L = ctx->learned;
  const int nC = ctx->nC;
  L->haveO = false;
  if (ensure_fwd(L, nC) || compute_scale(ctx, L, y)) return -1;
  hipLaunchKernelGGL(features_kernel, dim3((nC + 255) / 256), dim3(256), 0, ctx->stream, nC, 0, nC, y, L->geo.p,
                     L->scale.p, gamma, ctx->gamm, L->X.p);
  FOS_HIP_TRY(hipGetLastError());
  std::vector<uint16_t> xb((size_t)nC * NIP);
  float sc[8];
  FOS_HIP_TRY(hipMemcpyAsync(xb.data(), L->X.p, sizeof(uint16_t) * xb.size(), hipMemcpyDeviceToHost, ctx->stream));
  FOS_HIP_TRY(hipMemcpyAsync(sc, L->scale.p, sizeof(float) * 8, hipMemcpyDeviceToHost, ctx->stream));
  FOS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (X)
    for (int c = 0; c < nC; c++)
      for (int k = 0; k < NI; k++) {
        uint32_t u = (uint32_t)xb[(size_t)c * NIP + k] << 16;
        std::memcpy(&X[(size_t)c * NI + k], &u, 4);
      }
  if (S)
    for (int k = 0; k < 5; k++) S[k] = sc[k];
  return 0;
}

int fos_euler_psetup_learned(double t, const double *y, const double *fy, int jok, int *jcur, double gamma,
                             void *user_data) {
  (void)t; (void)fy; (void)jok;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx || !jcur) return -1;
  ctx->counters[2]++;
  *jcur = 1;
  return learned_psetup(ctx, y, gamma);
}

int fos_euler_psolve_learned(double t, const double *y, const double *fy, const double *r, double *z, double gamma,
                             double delta, int lr, void *user_data) {
  (void)t; (void)y; (void)fy; (void)gamma; (void)delta; (void)lr;
  fos_euler *ctx = static_cast<fos_euler *>(user_data);
  if (!ctx || !ctx->learned || !ctx->learned->ready) { set_error("psolve_learned: not set up"); return -1; }
  ctx->counters[3]++;
  if (learned_psolve(ctx, r, z)) return -1;
  return finish_call(ctx, "psolve_learned");
}

}  // extern "C"
