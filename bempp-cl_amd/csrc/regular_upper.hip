// Upper-triangle direct kernels for the symmetric DP0 single layers (kLaplaceSL, kHelmholtzSL);
// see regular_impl.h, UPPER.
#include "regular_impl.h"

namespace bem {
int regular_launch_upper_0(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s) {
    return launch_op<kLaplaceSL, false, false, true>(a, precision, kahan, false, out_dtype, s);
}
int regular_launch_upper_3(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s) {
    return launch_op<kHelmholtzSL, false, false, true>(a, precision, kahan, false, out_dtype, s);
}
}  // namespace bem
