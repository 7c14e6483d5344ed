// Shot-batch exact-adjoint step launchers (background recompute, adjoint state), 24-bit levels.
#include "born_ops.h"

namespace born {

BORN_BATCH_SREC_STEP(template, 0, Split);
BORN_BATCH_SREC_STEP(template, 1, Split);
BORN_BATCH_LAM_STEP(template, Split);

}  // namespace born
