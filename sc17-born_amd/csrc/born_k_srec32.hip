// Background-recompute step launchers of the exact adjoint, FP32 levels.
#include "born_ops.h"

namespace born {

BORN_SREC_STEP(template, 0, float);
BORN_SREC_STEP(template, 1, float);

}  // namespace born
