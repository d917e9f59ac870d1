// Window kernel: the 512-thread launch shapes (kernel: window_lm.hpp).
#include "window_lm.hpp"

namespace sfmba {
SFMBA_WINDOW_SHAPE(float, 512, 2)
SFMBA_WINDOW_SHAPE(float, 512, 1)
SFMBA_WINDOW_SHAPE(double, 512, 2)
SFMBA_WINDOW_SHAPE(double, 512, 1)
}  // namespace sfmba
