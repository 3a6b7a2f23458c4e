// Multi-rhs coarse stencil kernels instantiated for nbasis 32.
#include "mrhs_kernels.hpp"

namespace mgc {
template void mrhs_launch<32>(int, int, int, int, int, const uint4*, const float*, const void*,
                               const float*, float*, const int*, int, int, int, int, hipStream_t);
template void mrhs_launch_batched<32>(int, int, int, int, const MrhsBatch&, int, int, int,
                                       hipStream_t);
}  // namespace mgc
