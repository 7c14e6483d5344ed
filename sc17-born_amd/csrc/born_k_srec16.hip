// Background-recompute step launchers of the exact adjoint, FP16 levels.
#include "born_ops.h"

namespace born {

BORN_SREC_STEP(template, 0, __half);
BORN_SREC_STEP(template, 1, __half);

}  // namespace born
