// Window kernel: the 256-thread launch shapes (kernel: window_lm.hpp).
#include "window_lm.hpp"

namespace sfmba {
SFMBA_WINDOW_SHAPE(float, 256, 2)
SFMBA_WINDOW_SHAPE(float, 256, 1)
SFMBA_WINDOW_SHAPE(double, 256, 2)
SFMBA_WINDOW_SHAPE(double, 256, 1)
}  // namespace sfmba
