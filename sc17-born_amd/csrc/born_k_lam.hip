// Adjoint-state step launchers of the exact adjoint, FP32 and FP16 levels.
#include "born_ops.h"

namespace born {

BORN_LAM_STEP(template, float);
BORN_LAM_STEP(template, __half);

}  // namespace born
