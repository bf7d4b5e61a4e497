// Explicit instantiation of the hashing kernels for fhtDim in {512}.
#include "hash_core.h"

namespace fpp {
namespace hashk {
FPP_INSTANTIATE_HASH(512)
}  // namespace hashk
}  // namespace fpp
