// Chunked root walk for the weighted CP gradient (vector-Jacobian product).
#include "walk_kernel.hpp"

namespace mmcsf {
template void walk_launch<kWeightGrad>(const Csf &, const Factors &, const WalkArgs &, hipStream_t);
template void walk_preload<kWeightGrad>();
}  // namespace mmcsf
