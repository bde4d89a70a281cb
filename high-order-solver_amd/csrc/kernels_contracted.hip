// Contracted build of the kernels (-ffp-contract=fast): same source, FMAs
// formed by the compiler.
#define GKS_VARIANT contracted
#include "gks2d_kernels.hpp"
