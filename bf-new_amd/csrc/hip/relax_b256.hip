// relax_b256.hip -- relaxation kernel variants of 256-thread workgroups.
#include "relax_kernels.hpp"

namespace bfn {
BFN_RELAX_SHAPE(256, 8)
}  // namespace bfn
