// Born forward step launchers, FP32 levels (both scatter terms).
#include "born_ops.h"

namespace born {

BORN_FWD_STEP(template, 0, float);
BORN_FWD_STEP(template, 1, float);

}  // namespace born
