/* The kernels read the thread-local flag SQ_AB ("compute the |.| companion sums"), which the reference
 * defines in driver/eval.c:8.  libsqk_ref.so is loaded without the driver, so it carries the definition. */
#include "common.h"
__thread int SQ_AB = 1;
extern "C" void sqk_set_ab(int on){ SQ_AB = on; }
