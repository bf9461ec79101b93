// Launch sequence of the dipole operator (3 grid components), FP32 and FP64 (its own translation
// unit so the kernel instantiations compile in parallel).
#include "mpol_internal.h"

namespace mpol {
template int run_operator<float, kOpDipole>(mpol_ctx*, const double*, double*, double*, double*, int, double,
                                            hipEvent_t*, int);
template int run_operator<double, kOpDipole>(mpol_ctx*, const double*, double*, double*, double*, int, double,
                                             hipEvent_t*, int);
}  // namespace mpol
