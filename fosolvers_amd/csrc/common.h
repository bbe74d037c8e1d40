// Internal helpers shared by the host-side sources of libfosolvers_amd.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>

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

}  // namespace fos
