// Explicit instantiation of the hashing kernels for fhtDim in {1024}.
#include "hash_core.h"

namespace fpp {
namespace hashk {
FPP_INSTANTIATE_HASH(1024)
}  // namespace hashk
}  // namespace fpp
