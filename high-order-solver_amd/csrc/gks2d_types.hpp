// gks2d_types.hpp — plain data shared by the kernel builds and the C ABI layer.
#pragma once

namespace gksd {

// snapshot of the reference's gas / collision-time globals
// (basic_function.cpp:3-13)
struct Gas {
  int K;
  double gamma, mu, nu, pr, c1, c2;
  int tau_type;
  int prandtl_fix;
  double inv_gm1;  // 1 / (gamma - 1), used by the fast build
};

}  // namespace gksd
