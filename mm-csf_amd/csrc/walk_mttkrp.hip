// Chunked root walk for the root-mode MTTKRP.
#include "walk_kernel.hpp"

namespace mmcsf {
template void walk_launch<kMttkrp>(const Csf &, const Factors &, const WalkArgs &, hipStream_t);
template void walk_preload<kMttkrp>();
}  // namespace mmcsf
