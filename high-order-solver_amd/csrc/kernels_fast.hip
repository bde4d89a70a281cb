// Fast build of the kernels: FMA contraction of this source's expressions
// (-ffp-contract=fast-honor-pragmas; plain "fast" would also fuse inside the
// device library's erfc/exp, whose error terms need separate roundings) and
// reciprocal multiplication for divisions by constants, cell sizes and shared
// denominators.  Same source as the IEEE build.
#define GKS_VARIANT fast
#define GKS_FAST 1
#include "gks2d_kernels.hpp"
