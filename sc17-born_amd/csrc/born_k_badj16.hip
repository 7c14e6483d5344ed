// Shot-batch exact-adjoint step launchers (background recompute, adjoint state), FP16 levels.
#include "born_ops.h"

namespace born {

BORN_BATCH_SREC_STEP(template, 0, __half);
BORN_BATCH_SREC_STEP(template, 1, __half);
BORN_BATCH_LAM_STEP(template, __half);

}  // namespace born
