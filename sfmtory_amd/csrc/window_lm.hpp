// Persistent batched solver for many small problems (local-BA windows):
// the kernel template, instantiated per launch shape (threads per window,
// batch depth of the Schur reduction) by window_kernels*.hip.
//
// One workgroup runs the whole Levenberg-Marquardt solve of
// one problem: the reference's bundle_adjust (crates/sfm-ba/src/lib.rs:
// 747-1232) with the same schedule, acceptance test, damping and
// Cholesky -> LU -> damping-bump chain as the single-problem path.  All
// windows of a batch run in one launch, one workgroup each, so small
// windows no longer pay per-kernel launch latency and host round trips.
// The per-element work (linearization, point blocks, Schur reduction,
// updates, cost) is the shared code of engine_device.hpp; the reduced
// system, at most a few hundred unknowns per window, is factored in place
// by the workgroup:
//   - blocked right-looking Cholesky, 32-column panels staged in LDS;
//   - the augmented rhs row carries the forward solve;
//   - back-substitution by one wavefront with the solution in registers;
//   - LU with partial pivoting (nalgebra's rule) when Cholesky fails (FP64).
// Storage per window: S column-major with ld = D + 1, row D the rhs.
#pragma once
#include <hip/hip_runtime.h>

#include "engine_device.hpp"
#include "window.hpp"

namespace sfmba {

// The helpers stride by blockDim.x (the launch shape's thread count).
constexpr int MAXWAVE = 16;
#define WG ((uint32_t)blockDim.x)
#define NWAVE ((int)(blockDim.x >> 6))
constexpr int PB = 32;          // Cholesky panel width
constexpr int PANEL_ROWS = 416; // LDS panel capacity (rows); larger windows use global memory

struct WgShared {
  double red0[MAXWAVE], red1[MAXWAVE];
  double bcast[4];
  int ibcast[4];
};

// Deterministic workgroup sum of (a, b); every thread gets the result.
__device__ inline void wg_sum2(WgShared& sh, double& a, double& b) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  __syncthreads();
  if (lane == 0) {
    sh.red0[wave] = a;
    sh.red1[wave] = b;
  }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0;
  for (int w = 0; w < NWAVE; ++w) {
    s0 += sh.red0[w];
    s1 += sh.red1[w];
  }
  a = s0;
  b = s1;
  __syncthreads();
}

template <typename T>
__device__ inline void wg_cost(const View<T>& v, int which, int loss, double scale, WgShared& sh,
                               double& cost, double& n_front) {
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t o = threadIdx.x; o < v.n_obs; o += WG) {
    double term, front;
    obs_cost(v.obs_img, v.obs_pt, v.obs_xy, v.cam_of_img, v.cam_model, v.poses[which], v.cams[which],
             v.pts[which], o, loss, scale, 0, term, front);
    s0 += term;
    s1 += front;
  }
  wg_sum2(sh, s0, s1);
  cost = s0;
  n_front = s1;
}

// In-place Cholesky of the augmented lower triangle (rows 0..D, columns
// 0..D-1, ld = D + 1): afterwards S holds L and row D holds L^{-1} rhs.
// Returns false (uniformly) when a pivot is not > 0 (or NaN).
__device__ inline bool wg_cholesky(double* S, uint32_t D, double* panel, WgShared& sh) {
  const uint32_t ld = D + 1;
  const int t = threadIdx.x;
  bool ok = true;
  for (uint32_t k0 = 0; k0 < D; k0 += PB) {
    const uint32_t pb = min((uint32_t)PB, D - k0);
    const uint32_t m = ld - k0;  // rows k0 .. D
    const bool in_lds = m <= PANEL_ROWS;
    double* P = in_lds ? panel : S + (size_t)k0 * ld + k0;
    const uint32_t lp = in_lds ? m : ld;
    if (in_lds) {
      for (uint32_t e = t; e < pb * m; e += WG) {
        const uint32_t c = e / m, r = e - c * m;
        P[c * lp + r] = S[(size_t)(k0 + c) * ld + k0 + r];
      }
      __syncthreads();
    }
    for (uint32_t j = 0; j < pb; ++j) {
      const double d = P[j * lp + j];
      if (!(d > 0.0)) ok = false;  // nalgebra's Cholesky::new test (diag <= 0 or NaN)
      const double l = sqrt(d);
      __syncthreads();
      for (uint32_t r = j + 1 + t; r < m; r += WG) P[j * lp + r] /= l;
      if (t == 0) P[j * lp + j] = l;
      __syncthreads();
      // columns j+1 .. pb-1 of the panel, rows >= the column
      const uint32_t nc = pb - j - 1;
      for (uint32_t e = t; e < nc * m; e += WG) {
        const uint32_t c = j + 1 + e / m, r = e % m;
        if (r >= c) P[c * lp + r] -= P[j * lp + r] * P[j * lp + c];
      }
      __syncthreads();
    }
    if (in_lds) {
      for (uint32_t e = t; e < pb * m; e += WG) {
        const uint32_t c = e / m, r = e - c * m;
        if (r >= c) S[(size_t)(k0 + c) * ld + k0 + r] = P[c * lp + r];
      }
    }
    // Trailing update of columns k0+pb .. D-1, rows >= column (rhs row D
    // included), 4x4 register tiles.
    const uint32_t c0 = k0 + pb;
    if (c0 < D) {
      const uint32_t nt = D - c0;                 // trailing columns
      const uint32_t tr = (nt + 1 + 3) / 4;       // row tiles (rows c0 .. D)
      const uint32_t tc = (nt + 3) / 4;           // column tiles
      for (uint32_t e = t; e < tr * tc; e += WG) {
        const uint32_t ti = e / tc, tj = e - ti * tc;
        if (ti < tj) continue;
        const uint32_t rb = c0 + 4 * ti, cb = c0 + 4 * tj;
        double acc[4][4];
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (uint32_t j = 0; j < pb; ++j) {
          double lr[4], lc[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const uint32_t r = rb + a, c = cb + a;
            lr[a] = r <= D ? P[j * lp + (r - k0)] : 0.0;
            lc[a] = c < D ? P[j * lp + (c - k0)] : 0.0;
          }
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += lr[a] * lc[b];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const uint32_t r = rb + a, c = cb + b;
            if (r <= D && c < D && r >= c) S[(size_t)c * ld + r] -= acc[a][b];
          }
      }
    }
    __syncthreads();
  }
  // a failed pivot anywhere fails the whole factorization
  double f = ok ? 0.0 : 1.0, z = 0.0;
  wg_sum2(sh, f, z);
  return f == 0.0;
}

// Solve L^T x = y (y = row D of the factored augmented matrix) by one
// wavefront: lane l keeps x_i for i = l (mod 64) in registers.
constexpr int XREG = 16;  // up to 1024 unknowns per window
__device__ inline void wg_backsolve(const double* S, uint32_t D, double* x) {
  const uint32_t ld = D + 1;
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    double xr[XREG];
#pragma unroll
    for (int q = 0; q < XREG; ++q) xr[q] = 0.0;
    for (int j = (int)D - 1; j >= 0; --j) {
      const double* col = S + (size_t)j * ld;
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < XREG; ++q) {
        const int i = lane + 64 * q;
        if (i > j && i < (int)D) s += col[i] * xr[q];
      }
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      const double xj = (col[D] - s) / col[j];
#pragma unroll
      for (int q = 0; q < XREG; ++q)
        if (lane + 64 * q == j) xr[q] = xj;
    }
#pragma unroll
    for (int q = 0; q < XREG; ++q) {
      const int i = lane + 64 * q;
      if (i < (int)D) x[i] = xr[q];
    }
  }
  __syncthreads();
}

// LU with partial pivoting of the full symmetric matrix given by S's lower
// triangle (mirrored first), then the solve (nalgebra LU::new + solve:
// pivot = first maximal |a|, a zero column is skipped, solve fails on a zero
// U diagonal).  x receives the solution.  Returns false (uniformly) on
// failure.
__device__ inline bool wg_lu_solve(double* S, uint32_t D, double* x, WgShared& sh) {
  const uint32_t ld = D + 1;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (uint32_t e = t; e < D * D; e += WG) {
    const uint32_t c = e / D, r = e - c * D;
    if (r < c) S[(size_t)c * ld + r] = S[(size_t)r * ld + c];
  }
  for (uint32_t i = t; i < D; i += WG) x[i] = S[(size_t)i * ld + D];
  __syncthreads();
  for (uint32_t i = 0; i < D; ++i) {
    // pivot: first maximal |a_ri|, r >= i
    double best = -1.0;
    uint32_t bi = D;
    for (uint32_t r = i + t; r < D; r += WG) {
      const double a = fabs(S[(size_t)i * ld + r]);
      if (a > best) {
        best = a;
        bi = r;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double b2 = __shfl_xor(best, o);
      const uint32_t i2 = __shfl_xor(bi, o);
      if (b2 > best || (b2 == best && i2 < bi)) {
        best = b2;
        bi = i2;
      }
    }
    if (lane == 0) {
      sh.red0[wave] = best;
      sh.red1[wave] = (double)bi;
    }
    __syncthreads();
    if (t == 0) {
      double b = -1.0;
      uint32_t p = D;
      for (int w = 0; w < NWAVE; ++w) {
        const uint32_t pw = (uint32_t)sh.red1[w];
        if (sh.red0[w] > b || (sh.red0[w] == b && pw < p)) {
          b = sh.red0[w];
          p = pw;
        }
      }
      sh.ibcast[0] = (int)p;
    }
    __syncthreads();
    const uint32_t p = (uint32_t)sh.ibcast[0];
    const double diag = p < D ? S[(size_t)i * ld + p] : 0.0;
    __syncthreads();
    if (diag == 0.0) continue;  // zero column: nothing to eliminate
    if (p != i) {
      for (uint32_t c = t; c < D; c += WG) {
        const double a = S[(size_t)c * ld + i];
        S[(size_t)c * ld + i] = S[(size_t)c * ld + p];
        S[(size_t)c * ld + p] = a;
      }
      if (t == 0) {
        const double tmp = x[i];
        x[i] = x[p];
        x[p] = tmp;
      }
      __syncthreads();
    }
    const double inv = 1.0 / diag;
    for (uint32_t r = i + 1 + t; r < D; r += WG) S[(size_t)i * ld + r] *= inv;
    __syncthreads();
    const uint32_t m = D - i - 1;
    for (uint32_t e = t; e < m * m; e += WG) {
      const uint32_t c = i + 1 + e / m, r = i + 1 + e % m;
      const double f = S[(size_t)c * ld + i];
      if (f != 0.0) S[(size_t)c * ld + r] += -f * S[(size_t)i * ld + r];
    }
    __syncthreads();
  }
  // forward (unit L), then backward (U); serial on thread 0 of each step's
  // column, parallel over rows
  for (uint32_t i = 0; i < D; ++i) {
    const double c = x[i];
    __syncthreads();
    for (uint32_t r = i + 1 + t; r < D; r += WG) x[r] += -c * S[(size_t)i * ld + r];
    __syncthreads();
  }
  bool ok = true;
  for (int i = (int)D - 1; i >= 0; --i) {
    const double diag = S[(size_t)i * ld + i];
    if (diag == 0.0) {
      ok = false;
      break;
    }
    const double c = x[i] / diag;
    __syncthreads();
    if (t == 0) x[i] = c;
    for (int r = t; r < i; r += WG) x[r] += -c * S[(size_t)i * ld + r];
    __syncthreads();
  }
  return ok;
}

// The reduced system at damping lam from the incidence records: point
// inverses, Y of every record (four lanes per incidence), then the
// record-based reduction (engine_device.hpp schur_chunk_rec), one chunk per
// destination, each wavefront with its own part of the LDS stage.
template <typename T, int DB, int INS>
__device__ inline void wg_assemble(const View<T>& v, double lam, T* stage) {
  for (uint32_t p = threadIdx.x; p < v.n_pt; p += WG) point_prepare(v, p, lam);
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < 4 * v.n_inc; t += WG) compute_Y<T, DB>(v, t >> 2, t & 3u);
  __syncthreads();
  const uint32_t wave = threadIdx.x >> 6;
  T* lds = stage + wave * RecShape<T, DB, INS>::LDS;
  for (uint32_t ci = wave; ci < v.n_chunks; ci += NWAVE)
    schur_chunk_rec<T, DB, true, INS>(v, v.chunks[ci], lam, nullptr, lds);
  __syncthreads();
}

template <typename T, int DB>
__device__ inline void wg_step_and_cost(const View<T>& v, int cur, int loss, double scale,
                                        WgShared& sh, double& trial, double& n_front) {
  for (uint32_t p = threadIdx.x; p < v.n_pt; p += WG) backsub_point_rec<T, DB>(v, p, cur);
  for (uint32_t i = threadIdx.x; i < v.n_img; i += WG) update_pose(v, i, cur);
  for (uint32_t c = threadIdx.x; c < v.n_cam; c += WG) update_cam(v, c, cur);
  __syncthreads();
  wg_cost(v, 1 - cur, loss, scale, sh, trial, n_front);
}

// WGS threads per window, records of DB rows, INS load instructions per
// record side and batch of the Schur reduction.  The Cholesky's LDS panel
// doubles as the reduction's stage (the two never run at once).
template <typename T, int DB, int WGS, int INS>
__global__ __launch_bounds__(WGS) void k_window_lm(const View<T>* __restrict__ views,
                                                  WindowParams prm, WindowResult* __restrict__ res,
                                                  double* __restrict__ outer_costs) {
  static_assert((WGS / 64) * RecShape<T, DB, INS>::LDS * sizeof(T) <= PANEL_ROWS * PB * sizeof(double),
                "the reduction's stage fits the panel");
  __shared__ WgShared sh;
  __shared__ __attribute__((aligned(16))) double panel[PANEL_ROWS * PB];
  T* stage = reinterpret_cast<T*>(panel);
  const View<T>& v = views[blockIdx.x];
  WindowResult& R = res[blockIdx.x];
  double* oc = outer_costs + (size_t)blockIdx.x * prm.max_iterations;
  const bool mixed = sizeof(T) == 4;
  const int loss = prm.robust_loss;
  const double scale = prm.loss_scale_px;

  int cur = 0;
  double cost, n_front;
  wg_cost(v, 0, loss, scale, sh, cost, n_front);
  const double initial_cost = cost;
  uint64_t iterations = 0, n_lin = 0, n_fact = 0, n_chol_fail = 0, n_lu = 0, n_rej = 0, n_outer = 0;

  if (v.D != 0) {
    double lambda = prm.initial_lambda, nu = 2.0;
    for (uint64_t outer = 0; outer < prm.max_iterations; ++outer) {
      ++iterations;
      const double cost_before = cost;
      if (n_front == 0.0) break;  // no observation in front of its camera (lib.rs:855-857)
      for (uint32_t o = threadIdx.x; o < v.n_obs; o += WG) linearize_obs<T, DB>(v, o, cur, loss, scale);
      __syncthreads();
      for (uint32_t p = threadIdx.x; p < v.n_pt; p += WG) point_accumulate(v, p);
      __syncthreads();
      ++n_lin;
      bool accepted = false;
      for (int inner = 0; inner < 8; ++inner) {
        ++n_fact;
        double trial = 0.0, trial_front = 0.0;
        bool have_step = false, stepped = false;
        // FP32-assembled systems are damped at least at their resolution
        const double lam0 = mixed ? fmax(lambda, MIXED_MIN_LAMBDA) : lambda;
        wg_assemble<T, DB, INS>(v, lam0, stage);
        if (wg_cholesky(v.S, v.D, panel, sh)) {
          have_step = true;
        } else {
          ++n_chol_fail;
          if (mixed) {
            // no step: the damping bump takes over (an LU solve of an
            // FP32-assembled system that rounding made indefinite would
            // carry no information)
          } else {
            // not positive definite: rebuild S and solve by LU (lib.rs:1083-1093)
            wg_assemble<T, DB, INS>(v, lam0, stage);
            ++n_lu;
            if (wg_lu_solve(v.S, v.D, v.delta, sh)) {
              wg_step_and_cost<T, DB>(v, cur, loss, scale, sh, trial, trial_front);
              have_step = stepped = true;
            }
          }
        }
        if (have_step && !stepped) {
          wg_backsolve(v.S, v.D, v.delta);
          wg_step_and_cost<T, DB>(v, cur, loss, scale, sh, trial, trial_front);
        }
        if (!have_step) {
          lambda *= nu;
          nu *= 2.0;
          continue;
        }
        if (isfinite(trial) && trial < cost) {
          cur = 1 - cur;
          cost = trial;
          n_front = trial_front;
          lambda = fmax(lambda / 3.0, 1e-12);
          nu = 2.0;
          accepted = true;
          break;
        }
        ++n_rej;
        lambda *= nu;
        nu *= 2.0;
      }
      if (threadIdx.x == 0 && n_outer < prm.max_iterations) oc[n_outer] = cost;
      ++n_outer;
      if (!accepted) break;
      const double rel = (cost_before - cost) / fmax(cost_before, 1e-12);
      if (rel < 1e-6) break;
    }
  }
  if (threadIdx.x == 0) {
    R.initial_cost = initial_cost;
    R.final_cost = cost;
    R.iterations_run = iterations;
    R.num_linearizations = n_lin;
    R.num_factorizations = n_fact;
    R.num_cholesky_failures = n_chol_fail;
    R.num_lu_solves = n_lu;
    R.num_rejected_steps = n_rej;
    R.num_outer = n_outer;
    R.cur = cur;
  }
}

// Launch of one shape, dispatched on the record size (the batch's largest
// block, rec_db_for).
template <typename T, int WGS, int INS>
void launch_window_shape(const View<T>* views, uint32_t n_win, uint32_t max_grp_dim, const WindowParams& prm,
                         WindowResult* res, double* outer_costs, hipStream_t s) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(n_win), dim3(WGS), 0, s, views, prm, res, outer_costs);
  };
  switch (rec_db_for(max_grp_dim)) {
    case 8: go(k_window_lm<T, 8, WGS, INS>); break;
    case 12: go(k_window_lm<T, 12, WGS, INS>); break;
    default: go(k_window_lm<T, 16, WGS, INS>); break;
  }
}

#undef WG
#undef NWAVE

#define SFMBA_WINDOW_SHAPE(T, WGS, INS)                                                                      \
  template void launch_window_shape<T, WGS, INS>(const View<T>*, uint32_t, uint32_t, const WindowParams&,  \
                                                 WindowResult*, double*, hipStream_t);

}  // namespace sfmba
