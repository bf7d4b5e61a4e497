// Explicit instantiation of the hashing kernels for fhtDim in {256}.
#include "hash_core.h"

namespace fpp {
namespace hashk {
FPP_INSTANTIATE_HASH(256)
}  // namespace hashk
}  // namespace fpp
