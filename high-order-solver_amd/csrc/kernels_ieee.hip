// IEEE build of the kernels: no FMA contraction (-ffp-contract=off), the
// operation order of the reference CPU code.
#define GKS_VARIANT ieee
#include "gks2d_kernels.hpp"
