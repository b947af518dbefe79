// Constants and helpers shared by the host setup (g++) and the device code
// (hipcc).
#pragma once

#include <memory>
#include <utility>
#include <vector>

namespace ugmg {
constexpr int kSlice = 64;   // rows per SELL slice = one wavefront
constexpr int kBlock = 256;  // threads per block of the row kernels

// Allocator that leaves new elements uninitialised: the large host arrays
// of the set-up are written in full by parallel loops (or copies), which
// then also take the first touch of their pages.
template <class T>
struct DefaultInit : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  DefaultInit() = default;
  template <class U>
  DefaultInit(const DefaultInit<U> &) noexcept {}
  template <class U, class... Args>
  void construct(U *p, Args &&...args) {
    if constexpr (sizeof...(Args) == 0) ::new ((void *)p) U;
    else ::new ((void *)p) U(std::forward<Args>(args)...);
  }
};
template <class T>
using hvec = std::vector<T, DefaultInit<T>>;
}  // namespace ugmg
