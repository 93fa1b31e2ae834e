// Symmetric regular-pair kernels (every unordered pair once) assembling kHelmholtzHS alone
// (see regular_impl.h, SYM without DUAL); the caller adds the transpose.
#include "regular_impl.h"

namespace bem {
int regular_launch_symsl_8(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s) {
    return launch_op<kHelmholtzHS, false, true>(a, precision, kahan, true, out_dtype, s);
}
}  // namespace bem
