// Far pairs of the Maxwell magnetic field operator in mixed precision, through the tiled regular
// kernel (owner reduction of the RWG dof blocks; see regular_impl.h). Near pairs go to the FP64
// near pass (bem_near_assemble: mfie_near_launch), singular pairs to mfie_singular_launch.
#include "regular_impl.h"

namespace bem {
int regular_launch_mfie(const RegularArgs &a, int out_dtype, hipStream_t s) {
    if (a.tb.start == nullptr || a.sb.start == nullptr) {
        set_error("the mixed magnetic field pass needs dof blocks");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (a.rule.n != 6) {
        set_error("the magnetic field operator uses the 6-point regular rule (order 4), got %d points", a.rule.n);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (out_dtype == BEM_C64) return launch3<kMaxwellMFIE, 3, 3, 6, true, float>(a, 0, true, s);
    if (out_dtype == BEM_C128) return launch3<kMaxwellMFIE, 3, 3, 6, true, double>(a, 0, true, s);
    set_error("the magnetic field operator needs complex output");
    return BEM_ERR_INVALID_ARGUMENT;
}
}  // namespace bem
