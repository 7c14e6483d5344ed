// Born forward step launchers, 24-bit levels (FP16 + refinement byte).
#include "born_ops.h"

namespace born {

BORN_FWD_STEP(template, 0, Split);
BORN_FWD_STEP(template, 1, Split);

}  // namespace born
