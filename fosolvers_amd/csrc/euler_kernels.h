// Device layout and kernel launchers for the first-order Roe residual and the fused
// difference-quotient Jacobian-vector product.
#pragma once
#include <hip/hip_runtime.h>

#include "hip_util.h"

namespace fos {

// Cells are processed in chunks of CHUNK consecutive cells, one workgroup per chunk.
constexpr int CHUNK = 256;
constexpr int THREADS = 256;

// Everything a residual launch reads, for one arithmetic precision TC.
template <class TC>
struct FluxLayout {
  int nC = 0, nChunks = 0, maxChunkFaces = 0, nSlots = 4;
  const int *chunkBeg = nullptr;                 // [nChunks+1] chunk-face ranges
  const int2 *lr = nullptr;                      // [nCF] (L, R); R < 0: boundary pair -(R+1)
  const typename Vec4<TC>::type *geo = nullptr;  // [nCF] (nx, ny, nz, area)
  const short *slots = nullptr;                  // [nSlots nC] 2*localFace + side, -1 = none
  const TC *invV = nullptr;                      // [nC] 1/v
  const int *bndType = nullptr;                  // [nPb] BC code
  const TC *bndPar = nullptr;                    // [5 nPb] BC parameters
  TC gamma = 1.4, r = 1.0;
};

// ydot = R(y) (SoA, 5 nC).  TS: storage type of the vectors, TC: arithmetic type.
template <class TS, class TC>
int launch_residual(const FluxLayout<TC> &L, const TS *y, TS *ydot, hipStream_t s);

// Jv = (R(y + sigma v) - fy) / sigma, fused (one pass over the mesh).  sigma is read from
// sigmaDev (a device scalar) when that is not null.
template <class TS, class TC>
int launch_dq_jv(const FluxLayout<TC> &L, const TS *y, const TS *v, const TS *fy, double sigma,
                 const double *sigmaDev, TS *Jv, hipStream_t s);

// Jv = (R(y + sigma v) - R(y)) / sigma with both residuals formed in the same pass (face-wise
// flux differences): consistent in reduced precision, where subtracting an FP64 fy would not be.
template <class TS, class TC>
int launch_dq_jv_diff(const FluxLayout<TC> &L, const TS *y, const TS *v, double sigma, const double *sigmaDev,
                      TS *Jv, hipStream_t s);

}  // namespace fos
