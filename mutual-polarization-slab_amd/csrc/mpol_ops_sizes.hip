// Launch sequence of the polydisperse operator (one scalar grid), FP32 and FP64 (its own translation
// unit so the kernel instantiations compile in parallel).
#include "mpol_internal.h"

namespace mpol {
template int run_operator_sizes<float>(mpol_ctx*, const double*, double*, double*, int);
template int run_operator_sizes<double>(mpol_ctx*, const double*, double*, double*, int);
}  // namespace mpol
