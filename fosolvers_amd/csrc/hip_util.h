// HIP helpers shared by the device sources of libfosolvers_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <tuple>
#include <utility>

#include "common.h"

#define FOS_HIP_TRY(expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      fos::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return -1;                                                                            \
    }                                                                                       \
  } while (0)

namespace fos {

template <class T>
struct Vec4;
template <>
struct Vec4<double> { using type = double4; };
template <>
struct Vec4<float> { using type = float4; };

// Area-weighted face normal a n (3 values; 24 bytes in FP64 instead of 32 for (n, a))
template <class T>
struct AreaNormal {
  T x, y, z;
};

// device buffer with RAII release
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * count);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
      return -1;
    }
    n = count;
    return 0;
  }
  int upload(const T *h, size_t count, hipStream_t s) {
    if (alloc(count)) return -1;
    if (count == 0) return 0;
    hipError_t e = hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("upload failed: %s", hipGetErrorString(e));
      return -1;
    }
    return 0;
  }
};

// blockIdx -> work item mapping that gives each of the 8 XCDs a contiguous range of work items
// (workgroups are dealt round-robin to XCDs, and each XCD has its own L2).
__device__ inline int xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  int per = nblocks / nx, rem = nblocks % nx;
  int xcd = bid % nx, k = bid / nx;
  // XCD x owns per (+1 if x < rem) consecutive items
  int base = xcd * per + (xcd < rem ? xcd : rem);
  int item = base + k;
  // blocks beyond the balanced range (only when the count per XCD is uneven) fall back
  int cnt = per + (xcd < rem ? 1 : 0);
  return k < cnt ? item : bid;
}


// A fixed chain of dependent kernel launches recorded once as a graph of kernel nodes (built
// explicitly, no stream capture, so it works on any stream including the default one) and replayed
// with one hipGraphLaunch.  The level-scheduled sweeps and the factorisation are hundreds of short
// dependent kernels whose launch gaps would otherwise exceed their run time.
struct KernelChain {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  hipGraphNode_t last = nullptr;
  KernelChain() = default;
  KernelChain(const KernelChain &) = delete;
  KernelChain &operator=(const KernelChain &) = delete;
  ~KernelChain() { reset(); }
  void reset() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
    e = nullptr;
    g = nullptr;
    last = nullptr;
  }
  bool ready() const { return e != nullptr; }
  // one kernel node depending on the previous one; the argument values are copied into the node
  template <class... P, class... A>
  int add(void (*k)(P...), dim3 grid, dim3 block, size_t lds, A... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<P...> vals(static_cast<P>(a)...);
    return add_node(reinterpret_cast<void *>(k), grid, block, lds, vals, std::index_sequence_for<P...>{});
  }
  template <class T, size_t... I>
  int add_node(void *f, dim3 grid, dim3 block, size_t lds, T &vals, std::index_sequence<I...>) {
    void *args[] = {reinterpret_cast<void *>(&std::get<I>(vals))...};
    hipKernelNodeParams p{};
    p.func = f;
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = (unsigned)lds;
    p.kernelParams = args;
    p.extra = nullptr;
    if (!g && (err = hipGraphCreate(&g, 0)) != hipSuccess) return -1;
    hipGraphNode_t n = nullptr;
    if ((err = hipGraphAddKernelNode(&n, g, last ? &last : nullptr, last ? 1 : 0, &p)) != hipSuccess) return -1;
    last = n;
    return 0;
  }
  int instantiate() {
    if (!g) return -1;
    err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err != hipSuccess) e = nullptr;
    return err == hipSuccess ? 0 : -1;
  }
  int launch(hipStream_t s) { return (err = hipGraphLaunch(e, s)) == hipSuccess ? 0 : -1; }
  // the error of the failed call; the chain is dropped so that the next call records it afresh
  const char *fail() {
    const char *why = hipGetErrorString(err);
    reset();
    return why;
  }
  hipError_t err = hipSuccess;
};

}  // namespace fos
