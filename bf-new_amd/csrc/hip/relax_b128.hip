// relax_b128.hip -- relaxation kernel variants of 128-thread workgroups.
#include "relax_kernels.hpp"

namespace bfn {
BFN_RELAX_SHAPE(128, 16)
}  // namespace bfn
