// Internal helpers shared by the host-side sources of libfosolvers_amd.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>

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

}  // namespace fos
