// Fast build of the 1D kernels: FMA contraction within this file
// (-ffp-contract=fast-honor-pragmas) and reciprocal multiplication where the
// shared WENO leaves divide by constants or the cell size.
#define GKS_VARIANT fast
#define GKS_FAST 1
#include "gks1d_kernels.hpp"
