// Regular-pair kernels assembling kHelmholtzSL and the double layer of the same family in one pass
// (see regular_impl.h, DUAL).
#include "regular_impl.h"

namespace bem {
int regular_launch_dual_3(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s) {
    return launch_op<kHelmholtzSL, true>(a, precision, kahan, tile, out_dtype, s);
}
}  // namespace bem
