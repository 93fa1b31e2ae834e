// Symmetric regular-pair kernels (every unordered pair once) assembling kLaplaceHS alone
// (see regular_impl.h, SYM without DUAL); the caller adds the transpose.
#include "regular_impl.h"

namespace bem {
int regular_launch_symsl_7(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s) {
    return launch_op<kLaplaceHS, false, true>(a, precision, kahan, true, out_dtype, s);
}
}  // namespace bem
