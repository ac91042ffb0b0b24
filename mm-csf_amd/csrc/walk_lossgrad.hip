// Chunked root walk for the CP-completion loss and gradient.
#include "walk_kernel.hpp"

namespace mmcsf {
template void walk_launch<kLossGrad>(const Csf &, const Factors &, const WalkArgs &, hipStream_t);
template void walk_preload<kLossGrad>();
}  // namespace mmcsf
