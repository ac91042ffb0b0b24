// Chunked root walk for the leaf-mode TTM.
#include "walk_kernel.hpp"

namespace mmcsf {
template void walk_launch<kTtm>(const Csf &, const Factors &, const WalkArgs &, hipStream_t);
template void walk_preload<kTtm>();
}  // namespace mmcsf
