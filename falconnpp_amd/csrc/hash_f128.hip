// Explicit instantiation of the hashing kernels for fhtDim in {128}.
#include "hash_core.h"

namespace fpp {
namespace hashk {
FPP_INSTANTIATE_HASH(128)
}  // namespace hashk
}  // namespace fpp
