// Shot-batch forward step launchers, FP32 levels.
#include "born_ops.h"

namespace born {

BORN_BATCH_FWD_STEP(template, 0, float);
BORN_BATCH_FWD_STEP(template, 1, float);

}  // namespace born
