// Mixed-precision, many-field evaluation.
#include "common.h"
#include "devmath.h"
#include "plan.h"

namespace sal {

void eval_mixed(sal_geom g, const sal_eval_opts& o, const double* area, const double* ssh, int R, double* gx,
                double* gy, int accumulate, hipStream_t s) {
    SAL_REQUIRE(false, SAL_ERR_STATE, "mixed precision not available in this build");
}

}  // namespace sal
