// Symmetric regular-pair kernels (every unordered pair once) assembling kLaplaceSL and the
// Laplace double layer in one pass (see regular_impl.h, SYM).
#include "regular_impl.h"

namespace bem {
int regular_launch_sym_0(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s) {
    return launch_op<kLaplaceSL, true, true>(a, precision, kahan, true, out_dtype, s);
}
}  // namespace bem
