// gks3d_device_types.hpp — plain data of the 3D path shared by the kernel
// builds and the C ABI layer.
#pragma once

namespace gksd3 {

// gas and collision-time parameters (the reference's globals K, Gamma, Mu,
// Nu, c1_euler, c2_euler, tau_type)
struct Gas3 {
  int K;
  double gamma, mu, nu, c1, c2;
  int tau_type;  // 0 Euler, 1 Navier-Stokes
};

}  // namespace gksd3
