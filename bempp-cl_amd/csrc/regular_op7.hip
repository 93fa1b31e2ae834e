// Regular-pair kernels of operator kLaplaceHS (see regular_impl.h).
#include "regular_impl.h"

namespace bem {
int regular_launch_7(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s) {
    return launch_op<kLaplaceHS>(a, precision, kahan, tile, out_dtype, s);
}
}  // namespace bem
