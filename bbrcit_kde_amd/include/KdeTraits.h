// KdeTraits.h -- typed constants used across the library.
// Same names as the reference's include/KdeTraits.h:13-47 (ConstantTraits);
// the device point traits of KdeTraits.h:49-77 are not needed here: device
// data is laid out by the C ABI (bbrcit_hip.h).
#ifndef BBRCIT_AMD_KDETRAITS_H_
#define BBRCIT_AMD_KDETRAITS_H_

namespace bbrcit {

template <typename T>
struct ConstantTraits {
  using Type = T;
  static constexpr T zero() { return T(0); }
  static constexpr T one() { return T(1); }
};

}  // namespace bbrcit

#endif
