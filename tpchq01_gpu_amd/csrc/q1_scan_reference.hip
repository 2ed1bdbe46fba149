// The scan kernels of the reference layout (all placements and block sizes, the any-value
// and wide kernels, the batched kernel), instantiated in a translation unit of their own.
#include "q1_scan.h"

namespace q1 {
Q1_DEFINE_LAYOUT_KERNELS(Q1_LAYOUT_REFERENCE, reference)
}  // namespace q1
