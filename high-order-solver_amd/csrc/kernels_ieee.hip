// IEEE build of the kernels: no FMA contraction (-ffp-contract=off) and every
// division kept, i.e. the operation order of the reference CPU code.
#define GKS_VARIANT ieee
#define GKS_FAST 0
#include "gks2d_kernels.hpp"
