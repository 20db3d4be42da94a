// relax_b512.hip -- relaxation kernel variants of 512-thread workgroups.
#include "relax_kernels.hpp"

namespace bfn {
BFN_RELAX_SHAPE(512, 4)
}  // namespace bfn
