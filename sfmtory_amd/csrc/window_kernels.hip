// Window kernel: the 1024-thread launch shapes, the shape dispatch and the
// final-state gather of a batch (the kernel itself is window_lm.hpp).
#include "window_lm.hpp"

namespace sfmba {

SFMBA_WINDOW_SHAPE(float, 1024, 2)
SFMBA_WINDOW_SHAPE(float, 1024, 1)
SFMBA_WINDOW_SHAPE(double, 1024, 2)
SFMBA_WINDOW_SHAPE(double, 1024, 1)

template <typename T>
void launch_window_lm(const View<T>* views, uint32_t n_win, uint32_t max_grp_dim, const WindowParams& prm,
                      WindowResult* res, double* outer_costs, int shape, hipStream_t s) {
  if (n_win == 0) return;
  switch (shape) {
    case 0: launch_window_shape<T, 1024, 2>(views, n_win, max_grp_dim, prm, res, outer_costs, s); break;
    case 1: launch_window_shape<T, 1024, 1>(views, n_win, max_grp_dim, prm, res, outer_costs, s); break;
    case 2: launch_window_shape<T, 512, 2>(views, n_win, max_grp_dim, prm, res, outer_costs, s); break;
    case 3: launch_window_shape<T, 512, 1>(views, n_win, max_grp_dim, prm, res, outer_costs, s); break;
    case 4: launch_window_shape<T, 256, 2>(views, n_win, max_grp_dim, prm, res, outer_costs, s); break;
    case 5: launch_window_shape<T, 256, 1>(views, n_win, max_grp_dim, prm, res, outer_costs, s); break;
    default: throw Error(SFMBA_ERR_INVALID_ARGUMENT, "window launch shape out of range");
  }
  HIP_CHECK(hipGetLastError());
}

template void launch_window_lm<float>(const View<float>*, uint32_t, uint32_t, const WindowParams&,
                                      WindowResult*, double*, int, hipStream_t);
template void launch_window_lm<double>(const View<double>*, uint32_t, uint32_t, const WindowParams&,
                                       WindowResult*, double*, int, hipStream_t);

// One workgroup per window: its final state buffer to the compact output.
__global__ __launch_bounds__(256) void k_window_gather(const WindowResult* __restrict__ res,
                                                       const uint64_t* __restrict__ state,
                                                       const uint64_t* __restrict__ nd,
                                                       const uint64_t* __restrict__ out_off,
                                                       const double* __restrict__ base, double* __restrict__ out) {
  const uint32_t w = blockIdx.x;
  const uint64_t n = nd[w];
  const double* src = base + state[w] + (uint64_t)res[w].cur * n;
  double* dst = out + out_off[w];
  for (uint64_t i = threadIdx.x; i < n; i += 256) dst[i] = src[i];
}

void launch_window_gather(uint32_t n_win, const WindowResult* res, const uint64_t* state, const uint64_t* nd,
                          const uint64_t* out_off, const double* base, double* out, hipStream_t s) {
  if (n_win == 0) return;
  hipLaunchKernelGGL(k_window_gather, dim3(n_win), dim3(256), 0, s, res, state, nd, out_off, base, out);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sfmba
