// Batched solve of many small problems by one persistent launch
// (window_kernels.hip): parameters and per-problem results shared by the
// host code and the kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine_device.hpp"

namespace sfmba {

struct WindowParams {
  uint64_t max_iterations;
  double initial_lambda;
  double loss_scale_px;
  int32_t robust_loss;
  int32_t pad;
};

struct WindowResult {
  double initial_cost, final_cost;
  uint64_t iterations_run, num_linearizations, num_factorizations, num_cholesky_failures,
      num_lu_solves, num_rejected_steps, num_outer;
  int32_t cur;  // which state buffer holds the result
  int32_t pad;
};

// Largest reduced system the window kernel solves (its back-substitution
// keeps the solution in 16 registers per lane).
constexpr uint32_t WINDOW_MAX_DIM = 1024;

// One workgroup per problem; outer_costs holds max_iterations doubles per
// problem.
template <typename T>
void launch_window_lm(const View<T>* views, uint32_t n_win, uint32_t max_grp_dim,
                      const WindowParams& prm, WindowResult* res, double* outer_costs,
                      hipStream_t s);

}  // namespace sfmba
