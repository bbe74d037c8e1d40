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
  const AreaNormal<TC> *geo = nullptr;           // [nCF] area-weighted normal a n
  const short *slots = nullptr;                  // [nSlots nC] 2*localFace + side, -1 = none
  const TC *invV = nullptr;                      // [nC] 1/v
  const int *bndType = nullptr;                  // [nPb] BC code
  const TC *bndPar = nullptr;                    // [5 nPb] BC parameters
  TC gamma = 1.4, r = 1.0;
};

// Tetrahedral-mesh layout: the same chunk face lists and cell slots as FluxLayout, but each face is
// one 8-byte record and its area and unit normal are recomputed from the chunk's node coordinates
// (staged in LDS) instead of being read per face.
//   rec.x: bits 0-28 payload, bits 29-30 kind, bit 31 flip the normal
//     kind 0  both cells in the chunk: payload = R (chunk-local)
//     kind 1  L outside the chunk:     payload = L (global)
//     kind 2  R outside the chunk:     payload = R (global)
//     kind 3  boundary face:           payload = boundary pair index b
//   rec.y: bits 0-7 the chunk-local cell (L, or R for kind 1), bits 8-15, 16-23, 24-31 the face's
//          three nodes (chunk-local, ascending global id)
constexpr int TET_MAX_NODES = 256;  // 8-bit node fields
constexpr int TET_MAX_PAYLOAD = 1 << 29;
template <class TC>
struct TetLayout {
  int nC = 0, nChunks = 0, maxChunkFaces = 0, maxNodes = 0, nSlots = 4;
  const int *chunkBeg = nullptr;  // [nChunks+1] chunk-face ranges (same faces as FluxLayout)
  const uint2 *face = nullptr;    // [nCF] records
  const short *slots = nullptr;   // [nSlots nC] 2*localFace + side, -1 = none (FluxLayout's)
  const int *nodeBeg = nullptr;   // [nChunks+1]
  const TC *node = nullptr;       // [3 per node] coordinates (FP32: relative to the chunk's first node)
  const int *bndType = nullptr;   // [nPb] BC code (pair order)
  const TC *bndPar = nullptr;     // [5 nPb] BC parameters (pair order)
  const TC *invV = nullptr;       // [nC] 1/v
  TC gamma = 1.4, r = 1.0;
  // the state vectors exceed the last-level cache: once-read face data and the output go through
  // non-temporal loads and stores
  int streaming = 0;
  // boundary pairs: the ghost states are formed by a small kernel ahead of the face pass (they need
  // the boundary-condition code, which would otherwise set the face pass's register count)
  // ghost slots run in chunk order (a chunk's boundary faces read consecutive slots)
  int nPb = 0;
  const int *bndPair = nullptr;   // [nPb] slot -> boundary pair index (pair order: bndType, bndPar)
  const int *bndCell = nullptr;   // [nPb] the slot's cell
  const TC *bndNorm = nullptr;    // [3][nPb] the slot's unit normal (normP, out of the cell)
  TC *ghost = nullptr;            // [16][nPb] ghost primitives of the current launch
};

// Every launcher takes an optional device word nf: a non-finite value among the launch's outputs
// sets it to 1 (it is never cleared by a launch).  The Jv launchers also take (newton, ngam): with
// newton set they store the Newton matrix product v - ngam J v instead of J v (the integrator's
// GMRES product, one pass instead of two).

// LDS bytes a tet-layout launch needs (0 if the mode does not fit one workgroup).
template <class TC>
size_t tet_lds_bytes(const TetLayout<TC> &L, int mode);

template <class TS, class TC>
int launch_tet_residual(const TetLayout<TC> &L, const TS *y, TS *ydot, hipStream_t s, unsigned *nf = nullptr);
template <class TS, class TC>
int launch_tet_dq_jv(const TetLayout<TC> &L, const TS *y, const TS *v, const TS *fy, double sigma,
                     const double *sigmaDev, TS *Jv, hipStream_t s, unsigned *nf = nullptr,
    int newton = 0, double ngam = 0.0);
template <class TS, class TC>
int launch_tet_dq_jv_diff(const TetLayout<TC> &L, const TS *y, const TS *v, double sigma, const double *sigmaDev,
                          TS *Jv, hipStream_t s, unsigned *nf = nullptr,
    int newton = 0, double ngam = 0.0);

// Jv = R'(y) v exactly: the residual's directional derivative (dual-number arithmetic through the
// primitives, the ghost states and the Roe flux), one pass over the mesh, no sigma and no fy.
template <class TS, class TC>
int launch_tet_jvp(const TetLayout<TC> &L, const TS *y, const TS *v, TS *Jv, hipStream_t s, unsigned *nf = nullptr,
    int newton = 0, double ngam = 0.0);
template <class TS, class TC>
int launch_jvp(const FluxLayout<TC> &L, const TS *y, const TS *v, TS *Jv, hipStream_t s, unsigned *nf = nullptr,
    int newton = 0, double ngam = 0.0);

// ydot = R(y) (SoA, 5 nC).  TS: storage type of the vectors, TC: arithmetic type.
template <class TS, class TC>
int launch_residual(const FluxLayout<TC> &L, const TS *y, TS *ydot, hipStream_t s, unsigned *nf = nullptr);

// Jv = (R(y + sigma v) - fy) / sigma, fused (one pass over the mesh).  sigma is read from
// sigmaDev (a device scalar) when that is not null.
template <class TS, class TC>
int launch_dq_jv(const FluxLayout<TC> &L, const TS *y, const TS *v, const TS *fy, double sigma,
                 const double *sigmaDev, TS *Jv, hipStream_t s, unsigned *nf = nullptr,
    int newton = 0, double ngam = 0.0);

// Jv = (R(y + sigma v) - R(y)) / sigma with both residuals formed in the same pass (face-wise
// flux differences): consistent in reduced precision, where subtracting an FP64 fy would not be.
template <class TS, class TC>
int launch_dq_jv_diff(const FluxLayout<TC> &L, const TS *y, const TS *v, double sigma, const double *sigmaDev,
                      TS *Jv, hipStream_t s, unsigned *nf = nullptr,
    int newton = 0, double ngam = 0.0);

}  // namespace fos
