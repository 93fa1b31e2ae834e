// Regular-pair kernels of operator kHelmholtzHS (see regular_impl.h).
#include "regular_impl.h"

namespace bem {
int regular_launch_8(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s) {
    return launch_op<kHelmholtzHS>(a, precision, kahan, tile, out_dtype, s);
}
}  // namespace bem
