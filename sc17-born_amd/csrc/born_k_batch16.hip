// Shot-batch forward step launchers, FP16 levels.
#include "born_ops.h"

namespace born {

BORN_BATCH_FWD_STEP(template, 0, __half);
BORN_BATCH_FWD_STEP(template, 1, __half);

}  // namespace born
