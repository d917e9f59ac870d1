// Window kernel: the 256-thread launch shapes (kernel: window_lm.hpp).
#include "window_lm.hpp"

namespace sfmba {
SFMBA_WINDOW_SHAPE(float, 256, 8)
SFMBA_WINDOW_SHAPE(float, 256, 4)
SFMBA_WINDOW_SHAPE(double, 256, 8)
SFMBA_WINDOW_SHAPE(double, 256, 4)
}  // namespace sfmba
