// Symmetric regular-pair kernels (every unordered pair once) assembling kMaxwellEFIE alone on one
// RWG space (the EFIE's regular part is symmetric: the same quadrature on both elements and a
// kernel symmetric in the two RWG functions; see regular_impl.h, SYM without DUAL); the caller
// adds the transpose.
#include "regular_impl.h"

namespace bem {
int regular_launch_symsl_6(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s) {
    return launch_op<kMaxwellEFIE, false, true>(a, precision, kahan, true, out_dtype, s);
}
}  // namespace bem
