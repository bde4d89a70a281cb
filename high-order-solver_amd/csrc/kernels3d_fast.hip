// Fast build of the 3D kernels: FMA contraction within this file
// (-ffp-contract=fast-honor-pragmas) and reciprocal multiplication where the
// WENO leaves and the moments divide by constants, the cell size or a shared
// denominator.
#define GKS_VARIANT fast
#define GKS_FAST 1
#include "gks3d_kernels.hpp"
