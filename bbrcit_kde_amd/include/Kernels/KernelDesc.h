// KernelDesc.h -- the C ABI's descriptor of each kernel class.
#ifndef BBRCIT_AMD_KERNELDESC_H_
#define BBRCIT_AMD_KERNELDESC_H_

#include <bbrcit_hip.h>

namespace bbrcit {

inline bk_kernel_desc make_kernel_desc(int kind, int dim, double hx, double hy) {
  bk_kernel_desc d;
  d.kind = kind;
  d.dim = dim;
  d.hx = hx;
  d.hy = hy;
  return d;
}

}  // namespace bbrcit

#endif
