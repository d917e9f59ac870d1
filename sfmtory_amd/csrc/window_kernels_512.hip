// Window kernel: the 512-thread launch shapes (kernel: window_lm.hpp).
#include "window_lm.hpp"

namespace sfmba {
SFMBA_WINDOW_SHAPE(float, 512, 8)
SFMBA_WINDOW_SHAPE(float, 512, 4)
SFMBA_WINDOW_SHAPE(double, 512, 8)
SFMBA_WINDOW_SHAPE(double, 512, 4)
}  // namespace sfmba
