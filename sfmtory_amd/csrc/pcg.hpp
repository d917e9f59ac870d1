// Block-Jacobi PCG on the reduced camera system (pcg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfmba {

struct PcgBuffers {
  double* Minv = nullptr;      // inverses of the diagonal parameter blocks (sum of d^2)
  uint32_t* moff = nullptr;    // per block: offset into Minv
  uint32_t* row_grp = nullptr; // per unknown: its parameter block
  double *r = nullptr, *r2 = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double* symv_part = nullptr; // pcg_symv_part_doubles(D)
  double* dot_part = nullptr;  // pcg_dot_part_doubles(D)
  double* scal = nullptr;      // 16 doubles: 2 dot products, then the iteration's device state
  double* host = nullptr;      // 8 pinned host doubles
  float* S32 = nullptr;        // mixed storage: FP32 copy of the lower triangle, ld x Dp
  double* x_best = nullptr;    // mixed storage: the iterate of the smallest true residual, D
};

struct PcgResult {
  bool ok = false;        // false: breakdown (a block or the system not positive definite)
  int iterations = 0;
  double rel_residual = 0.0;
};

size_t pcg_symv_part_doubles(uint32_t D);
size_t pcg_dot_part_doubles(uint32_t D);

// Solves S x = b (b = row Dp of S) to |r| <= tolerance |b| or max_iterations.
// The iterations run on a device-side state; the host reads it once per
// batch of iterations.  x has D entries.  fp32_matrix: the iterations'
// matvecs read an FP32 copy of S (B.S32, half the bytes), and before every
// batch the true residual b - S x is recomputed in FP64 and replaces the
// recurrence's, keeping the search direction while the two stay close (a
// reliable update); convergence is judged on the true residual, and the
// iteration goes on with the FP64 matrix once the rounded one's error
// dominates or the true residual grows (from the best iterate so far).
PcgResult pcg_solve(const PcgBuffers& B, const double* S, size_t ld, uint32_t D, uint32_t Dp,
                    uint32_t n_grp, const uint32_t* grp_off, const uint32_t* grp_dim, int max_iterations,
                    double tolerance, double* x, int* flags, double* host_scal, hipStream_t s,
                    bool fp32_matrix = false);

}  // namespace sfmba
