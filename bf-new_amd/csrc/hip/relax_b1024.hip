// relax_b1024.hip -- relaxation kernel variants of 1024-thread workgroups.
#include "relax_kernels.hpp"

namespace bfn {
BFN_RELAX_SHAPE(1024, 2)
}  // namespace bfn
