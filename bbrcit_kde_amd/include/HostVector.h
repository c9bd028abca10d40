#ifndef BBRCIT_KDE_AMD_HOSTVECTOR_H_
#define BBRCIT_KDE_AMD_HOSTVECTOR_H_

// Host buffers that resize() does not zero-fill: large scratch arrays are
// then first touched by the threads that fill them, page by page in
// parallel, instead of by one thread writing zeros that are overwritten.
#include <memory>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

namespace bbrcit {

// An allocator whose value-construction is default-initialisation (for
// trivial types: no initialisation); every other construction is unchanged.
template <typename T>
struct NoInitAllocator : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAllocator<U>;
  };
  NoInitAllocator() = default;
  template <typename U>
  NoInitAllocator(const NoInitAllocator<U> &) noexcept {}
  template <typename U>
  void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void *>(p)) U;
  }
  template <typename U, typename... Args>
  void construct(U *p, Args &&...args) {
    ::new (static_cast<void *>(p)) U(std::forward<Args>(args)...);
  }
};

template <typename T>
using HostVector = std::vector<T, NoInitAllocator<T>>;

}  // namespace bbrcit

#endif  // BBRCIT_KDE_AMD_HOSTVECTOR_H_
