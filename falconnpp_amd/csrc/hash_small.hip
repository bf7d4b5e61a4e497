// Explicit instantiation of the hashing kernels for fhtDim in {2, 4, 8, 16, 32}.
#include "hash_core.h"

namespace fpp {
namespace hashk {
FPP_INSTANTIATE_HASH(2)
FPP_INSTANTIATE_HASH(4)
FPP_INSTANTIATE_HASH(8)
FPP_INSTANTIATE_HASH(16)
FPP_INSTANTIATE_HASH(32)
}  // namespace hashk
}  // namespace fpp
