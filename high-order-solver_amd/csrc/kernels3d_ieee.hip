// IEEE build of the 3D kernels: no FMA contraction (-ffp-contract=off) and
// every division kept, i.e. the operation order of the reference CPU code.
#define GKS_VARIANT ieee
#define GKS_FAST 0
#include "gks3d_kernels.hpp"
