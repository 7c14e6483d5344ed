// Shot-batch forward step launchers, 24-bit levels.
#include "born_ops.h"

namespace born {

BORN_BATCH_FWD_STEP(template, 0, Split);
BORN_BATCH_FWD_STEP(template, 1, Split);

}  // namespace born
