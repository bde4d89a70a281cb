// IEEE build of the 1D kernels: no FMA contraction (-ffp-contract=off) and
// every division kept, i.e. the operation order of the reference CPU code.
#define GKS_VARIANT ieee
#define GKS_FAST 0
#include "gks1d_kernels.hpp"
