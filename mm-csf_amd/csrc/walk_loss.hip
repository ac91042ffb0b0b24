// Chunked root walk for the CP-completion loss and residuals.
#include "walk_kernel.hpp"

namespace mmcsf {
template void walk_launch<kLossResid>(const Csf &, const Factors &, const WalkArgs &, hipStream_t);
template void walk_preload<kLossResid>();
}  // namespace mmcsf
