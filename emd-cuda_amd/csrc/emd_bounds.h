// emd_bounds.h -- checked arrays for builds with EMD_BOUNDS_CHECK defined (host emulation
// of the sift kernels; the gfx950 build never defines it and uses plain pointers).
//
// A Span carries the bounds of the array it was carved from (one knot-tile or work-area
// array of the workgroup's LDS or global region); indexing outside them, or outside after
// pointer arithmetic, reports the array and the index and aborts.  Converting a Span to a
// raw pointer (for a callee that takes one) keeps the address and drops the check.
#pragma once
#include <cstddef>
#include <cstdio>
#include <cstdlib>

namespace emdk {

[[noreturn]] inline void bounds_fail(const char* name, long idx, long size) {
  std::fprintf(stderr, "EMD_BOUNDS_CHECK: %s[%ld] outside [0, %ld)\n", name, idx, size);
  std::fflush(stderr);
  std::abort();
}

template <class T>
struct Span {
  T* p = nullptr;
  T* lo = nullptr;
  T* hi = nullptr;
  const char* name = "";

  Span() = default;
  Span(std::nullptr_t) {}
  Span(T* p_, T* lo_, T* hi_, const char* n) : p(p_), lo(lo_), hi(hi_), name(n) {}

  T& operator[](long i) const {
    T* q = p + i;
    if (q < lo || q >= hi) bounds_fail(name, (long)(q - lo), (long)(hi - lo));
    return *q;
  }
  Span operator+(long i) const { return Span(p + i, lo, hi, name); }
  Span operator-(long i) const { return Span(p - i, lo, hi, name); }
  operator T*() const { return p; }
  explicit operator bool() const { return p != nullptr; }
};

}  // namespace emdk

#define EMD_ARR(T) ::emdk::Span<T>
#define EMD_SPAN(T, ptr, count, name) ::emdk::Span<T>((ptr), (ptr), (ptr) + (count), (name))
