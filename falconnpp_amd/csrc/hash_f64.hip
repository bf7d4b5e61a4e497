// Explicit instantiation of the hashing kernels for fhtDim in {64}.
#include "hash_core.h"

namespace fpp {
namespace hashk {
FPP_INSTANTIATE_HASH(64)
}  // namespace hashk
}  // namespace fpp
