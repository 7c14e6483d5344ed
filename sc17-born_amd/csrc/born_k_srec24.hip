// Background-recompute step launchers of the exact adjoint, 24-bit levels.
#include "born_ops.h"

namespace born {

BORN_SREC_STEP(template, 0, Split);
BORN_SREC_STEP(template, 1, Split);
BORN_LAM_STEP(template, Split);

}  // namespace born
