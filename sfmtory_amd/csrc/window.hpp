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

// Launch shapes of the window kernel: threads per window (one workgroup)
// and the Schur reduction's batch depth (load instructions per record side
// and batch, schur_chunk_rec's INS).  The autotuner (engine_host.hip) times
// them on a batch; without a tuned choice the default per precision is
// used (the best of a measured sweep over 256 windows of 10-40 cameras on
// gfx950).
struct WindowShape {
  int threads, batch;
};
constexpr int N_WINDOW_SHAPES = 6;
constexpr WindowShape WINDOW_SHAPES[N_WINDOW_SHAPES] = {{1024, 2}, {1024, 1}, {512, 2},
                                                        {512, 1},  {256, 2},  {256, 1}};
template <typename T>
constexpr int window_default_shape() {
  return sizeof(T) == 4 ? 0 : 2;
}

// One workgroup per problem; outer_costs holds max_iterations doubles per
// problem; shape indexes WINDOW_SHAPES.
template <typename T>
void launch_window_lm(const View<T>* views, uint32_t n_win, uint32_t max_grp_dim,
                      const WindowParams& prm, WindowResult* res, double* outer_costs, int shape,
                      hipStream_t s);

template <typename T, int WGS, int NP>
void launch_window_shape(const View<T>* views, uint32_t n_win, uint32_t max_grp_dim, const WindowParams& prm,
                         WindowResult* res, double* outer_costs, hipStream_t s);

// Copies each window's final state (buffer res[w].cur of its two state
// buffers, nd[w] doubles each, starting at state[w]) to out + out_off[w].
void launch_window_gather(uint32_t n_win, const WindowResult* res, const uint64_t* state,
                          const uint64_t* nd, const uint64_t* out_off, const double* base, double* out,
                          hipStream_t s);

}  // namespace sfmba
