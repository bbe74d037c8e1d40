// Internal helpers shared by the host-side sources of libfosolvers_amd.
#pragma once
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fosolvers_amd.h"

namespace fos {

// Thread-local message for fos_last_error().
void set_error(const char *fmt, ...);

template <class T>
T *xalloc(size_t n) {
  T *p = static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
  if (!p) {
    std::fprintf(stderr, "[E] fosolvers_amd: host allocation of %zu bytes failed\n", sizeof(T) * n);
    std::abort();
  }
  return p;
}

// Checks an element list before any of its indices is used (host or device): sizes non-negative,
// arrays present, every element's node count equal to its shape's (or within [0, maxNNE] for
// shapes the grid build skips), every named node inside [0, nN).  Returns 0, or -1 with
// fos_last_error naming `who`, the element and the fault.
int validate_poly_grid(const fos_poly_grid *g, const char *who);

// Host threads for the structure builds: at most 16, and one per `grain` items.
inline int host_parts(size_t n, size_t grain) {
  const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  return (int)std::max<size_t>(1, std::min(hw, (n + grain - 1) / grain));
}

// f(lo, hi, part) over `parts` contiguous ranges of [0, n) in increasing order (part 0 first), one
// host thread each; the ranges are fixed by n and parts alone, so results that are combined in part
// order do not depend on the thread schedule.
template <class F>
void host_parallel(size_t n, int parts, F f) {
  auto lo = [&](int p) { return n * (size_t)p / (size_t)parts; };
  if (parts <= 1) {
    f((size_t)0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(parts - 1);
  for (int p = 1; p < parts; p++) th.emplace_back([&, p] { f(lo(p), lo(p + 1), p); });
  f(lo(0), lo(1), 0);
  for (auto &t : th) t.join();
}

}  // namespace fos
