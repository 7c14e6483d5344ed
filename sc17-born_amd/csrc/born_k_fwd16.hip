// Born forward step launchers, FP16 levels (both scatter terms).
#include "born_ops.h"

namespace born {

BORN_FWD_STEP(template, 0, __half);
BORN_FWD_STEP(template, 1, __half);

}  // namespace born
