// Shot-batch exact-adjoint step launchers (background recompute, adjoint state), FP32 levels.
#include "born_ops.h"

namespace born {

BORN_BATCH_SREC_STEP(template, 0, float);
BORN_BATCH_SREC_STEP(template, 1, float);
BORN_BATCH_LAM_STEP(template, float);

}  // namespace born
