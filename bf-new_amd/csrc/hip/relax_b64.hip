// relax_b64.hip -- relaxation kernel variants of 64-thread workgroups.
#include "relax_kernels.hpp"

namespace bfn {
BFN_RELAX_SHAPE(64, 32)
}  // namespace bfn
