// Regular-pair kernels of operator kHelmholtzADL (see regular_impl.h).
#include "regular_impl.h"

namespace bem {
int regular_launch_5(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s) {
    return launch_op<kHelmholtzADL>(a, precision, kahan, tile, out_dtype, s);
}
}  // namespace bem
