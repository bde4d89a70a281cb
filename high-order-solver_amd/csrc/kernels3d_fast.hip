// Fast build of the 3D kernels: the same source with FMA contraction
// (-ffp-contract=fast-honor-pragmas).
#define GKS_VARIANT fast
#include "gks3d_kernels.hpp"
