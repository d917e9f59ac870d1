// Dense FP64 solver for the reduced camera system (gfx950).
//
// Replaces the reference's nalgebra Cholesky::new(S).solve(rhs) with the LU
// fallback (crates/sfm-ba/src/lib.rs:1076-1094).
//
// Storage: S is column-major with leading dimension ld = Dp + NB, Dp = D
// rounded up to NB; only the lower triangle is referenced.  Row Dp holds the
// right-hand side, so the factorization of the augmented matrix
// [S ; rhs^T] yields L and, in row Dp, y = L^{-1} rhs (the forward solve
// rides along with the trailing updates).
//
// Blocked right-looking Cholesky, per diagonal block k (NB = 128):
//   k_potrf_block   one workgroup: L_kk = chol(A_kk) in LDS and its
//                   inverse Linv_kk; a pivot that is not > 0 (or NaN) sets
//                   flags[0] (nalgebra's failure test: diag <= 0 or NaN);
//   k_trsm_panel    L_ik = A_ik Linv_kk^T for every row block i > k (and
//                   the rhs block), 128x128x128 FP64 MFMA tiles;
//   k_syrk_update   A_ij -= L_ik L_jk^T on the trailing lower triangle,
//                   128x128 tiles, v_mfma_f64_16x16x4_f64.
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace sfmba {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int KC = 16;  // K chunk staged through LDS

// acc(128x128 tile, 4 waves of 64x64) += sign * A(128 x K) * B(128 x K)^T,
// A and B column-major (rows contiguous) with leading dims lda, ldb.
// The v_mfma_f64_16x16x4_f64 operand layout: lane l supplies A[l%16][l/16]
// and B^T[l/16][l%16]; result register v of lane l is D[l/16 + 4v][l%16]
// (checked on gfx950 with a one-wavefront probe of the builtin).
__device__ inline void tile_mma(double4_t acc[4][4], const double* __restrict__ A, size_t lda,
                                const double* __restrict__ B, size_t ldb, int K, double sign,
                                double* As, double* Bs) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  for (int k0 = 0; k0 < K; k0 += KC) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < (KC * 128) / 256; ++it) {
      const int idx = t + 256 * it;
      const int k = idx >> 7, row = idx & 127;
      As[k * 128 + row] = A[(size_t)(k0 + k) * lda + row];
      Bs[k * 128 + row] = B[(size_t)(k0 + k) * ldb + row];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      double a[4], b[4];
      const int kr = kk + (lane >> 4);
#pragma unroll
      for (int m = 0; m < 4; ++m) a[m] = sign * As[kr * 128 + wr * 64 + m * 16 + (lane & 15)];
#pragma unroll
      for (int n = 0; n < 4; ++n) b[n] = Bs[kr * 128 + wc * 64 + n * 16 + (lane & 15)];
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[n], acc[m][n], 0, 0, 0);
    }
  }
}

// Load / store a 128x128 tile into the MFMA accumulator layout.
__device__ inline void tile_load(double4_t acc[4][4], const double* C, size_t ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int col = wc * 64 + n * 16 + (lane & 15);
      const int row = wr * 64 + m * 16 + (lane >> 4);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[m][n][v] = C[(size_t)col * ldc + row + 4 * v];
    }
}

// Stores only entries that carry data: global column < D, and global row
// < D (the system) or == Dp (the right-hand side).  Padding and the spare
// rows of the rhs block keep their initial values, so a failed (NaN)
// factorization cannot leak into storage the next assembly does not rewrite
// (NaN * 0 is NaN).
__device__ inline void tile_store(const double4_t acc[4][4], double* C, size_t ldc, bool lower_only,
                                  uint32_t row0, uint32_t col0, uint32_t D, uint32_t Dp) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int col = wc * 64 + n * 16 + (lane & 15);
      const int row = wr * 64 + m * 16 + (lane >> 4);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t g = row0 + row + 4 * v;
        if ((!lower_only || row + 4 * v >= col) && (g < D || g == Dp) && col0 + col < D)
          C[(size_t)col * ldc + row + 4 * v] = acc[m][n][v];
      }
    }
}

// ---------------------------------------------------------------- potrf
// Packed lower-triangular index inside an NB block.
__device__ inline int pk(int r, int c) { return c * NB - (c * (c - 1)) / 2 + (r - c); }
constexpr int PACKED = NB * (NB + 1) / 2;

__global__ __launch_bounds__(256) void k_potrf_block(double* __restrict__ S, size_t ld, int k,
                                                     uint32_t D, double* __restrict__ Linv,
                                                     int* __restrict__ flags) {
  __shared__ double Lp[PACKED];
  __shared__ double Xp[PACKED];
  __shared__ double piv;
  const int t = threadIdx.x;
  double* A = S + (size_t)k * NB * ld + (size_t)k * NB;
  for (int idx = t; idx < NB * NB; idx += 256) {
    const int c = idx >> 7, r = idx & 127;
    if (r >= c) Lp[pk(r, c)] = A[(size_t)c * ld + r];
  }
  __syncthreads();
  // Right-looking unblocked Cholesky in LDS.
  for (int j = 0; j < NB; ++j) {
    if (t == 0) {
      const double d = Lp[pk(j, j)];
      if (!(d > 0.0) && (uint32_t)(k * NB + j) < D) atomicOr(&flags[0], 1);
      piv = sqrt(d);
      Lp[pk(j, j)] = piv;
    }
    __syncthreads();
    const double p = piv;
    for (int r = j + 1 + t; r < NB; r += 256) Lp[pk(r, j)] = Lp[pk(r, j)] / p;
    __syncthreads();
    const int m = NB - 1 - j;  // trailing size
    const int cnt = m * (m + 1) / 2;
    for (int e = t; e < cnt; e += 256) {
      // e -> (rr >= cc) in the trailing triangle, column-major packed order
      int cc = (int)((2 * m + 1 - sqrt((double)(2 * m + 1) * (2 * m + 1) - 8.0 * e)) * 0.5);
      if (cc < 0) cc = 0;
      while (cc > 0 && cc * m - (cc * (cc - 1)) / 2 > e) --cc;
      while ((cc + 1) * m - ((cc + 1) * cc) / 2 <= e) ++cc;
      const int rr = cc + (e - (cc * m - (cc * (cc - 1)) / 2));
      const int r = j + 1 + rr, c = j + 1 + cc;
      Lp[pk(r, c)] -= Lp[pk(r, j)] * Lp[pk(c, j)];
    }
    __syncthreads();
  }
  for (int idx = t; idx < NB * NB; idx += 256) {
    const int c = idx >> 7, r = idx & 127;
    if (r >= c && (uint32_t)(k * NB + r) < D) A[(size_t)c * ld + r] = Lp[pk(r, c)];
  }
  // Inverse: X = L^{-1}, row by row; column c of row i handled by thread
  // pair (2c, 2c+1) splitting the inner sum.
  for (int i = 0; i < NB; ++i) {
    const int c = t >> 1, half = t & 1;
    double s = 0.0;
    if (c < i) {
      for (int kk = c + half; kk < i; kk += 2) s += Lp[pk(i, kk)] * Xp[pk(kk, c)];
    }
    s += __shfl_xor(s, 1);
    if (half == 0 && c <= i) {
      const double lii = Lp[pk(i, i)];
      Xp[pk(i, c)] = (c == i) ? 1.0 / lii : -s / lii;
    }
    __syncthreads();
  }
  double* X = Linv + (size_t)k * NB * NB;
  for (int idx = t; idx < NB * NB; idx += 256) {
    const int c = idx >> 7, r = idx & 127;
    const bool pad = (uint32_t)(k * NB + r) >= D || (uint32_t)(k * NB + c) >= D;
    X[idx] = pad ? (r == c ? 1.0 : 0.0) : (r >= c ? Xp[pk(r, c)] : 0.0);
  }
}

// L_ik = A_ik Linv^T for row blocks i = k+1 .. nrb-1 (nrb includes the rhs block).
__global__ __launch_bounds__(256) void k_trsm_panel(double* __restrict__ S, size_t ld, int k,
                                                    const double* __restrict__ Linv, uint32_t D,
                                                    uint32_t Dp) {
  __shared__ double As[KC * 128];
  __shared__ double Bs[KC * 128];
  const int i = k + 1 + blockIdx.x;
  double* A = S + (size_t)k * NB * ld + (size_t)i * NB;
  const double* X = Linv + (size_t)k * NB * NB;
  double4_t acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = double4_t{0.0, 0.0, 0.0, 0.0};
  tile_mma(acc, A, ld, X, NB, NB, 1.0, As, Bs);
  __syncthreads();
  tile_store(acc, A, ld, false, (uint32_t)i * NB, (uint32_t)k * NB, D, Dp);
}

// A_ij -= L_ik L_jk^T for k < j <= i, j < nblk, i <= nblk (i == nblk: rhs block).
__global__ __launch_bounds__(256) void k_syrk_update(double* __restrict__ S, size_t ld, int k,
                                                     int nblk, uint32_t D, uint32_t Dp) {
  __shared__ double As[KC * 128];
  __shared__ double Bs[KC * 128];
  const int j = k + 1 + blockIdx.x;
  const int i = k + 1 + blockIdx.y;
  if (j > i || j >= nblk) return;
  double* C = S + (size_t)j * NB * ld + (size_t)i * NB;
  const double* Ai = S + (size_t)k * NB * ld + (size_t)i * NB;
  const double* Aj = S + (size_t)k * NB * ld + (size_t)j * NB;
  double4_t acc[4][4];
  tile_load(acc, C, ld);
  tile_mma(acc, Ai, ld, Aj, ld, NB, -1.0, As, Bs);
  tile_store(acc, C, ld, i == j, (uint32_t)i * NB, (uint32_t)j * NB, D, Dp);
}

// ------------------------------------------------------ triangular solves
// Backward solve L^T x = y, column-block k (from the last): x_k = Linv_k^T y_k
// (every workgroup recomputes it), workgroup 0 stores x_k, and all
// workgroups apply y_c -= L[block k, c] . x_k for c < k*NB.
__global__ __launch_bounds__(256) void k_backsolve_step(const double* __restrict__ S, size_t ld,
                                                        const double* __restrict__ Linv, int k,
                                                        double* __restrict__ y,
                                                        double* __restrict__ x) {
  __shared__ double xs[NB];
  __shared__ double ys[NB];
  const int t = threadIdx.x;
  if (t < NB) ys[t] = y[(size_t)k * NB + t];
  __syncthreads();
  const double* X = Linv + (size_t)k * NB * NB;
  if (t < NB) {
    // x[t] = sum_r X[r][t] y[r], r >= t
    double s = 0.0;
    for (int r = t; r < NB; ++r) s += X[(size_t)t * NB + r] * ys[r];
    xs[t] = s;
  }
  __syncthreads();
  if (blockIdx.x == 0 && t < NB) x[(size_t)k * NB + t] = xs[t];
  // update: columns c in [0, k*NB), 4 waves, each lane group handles a column
  const int ncols = k * NB;
  const int lane = t & 63, wave = t >> 6;
  for (int c = (blockIdx.x * 4 + wave); c < ncols; c += gridDim.x * 4) {
    const double* col = S + (size_t)c * ld + (size_t)k * NB;
    double s = col[lane] * xs[lane] + col[lane + 64] * xs[lane + 64];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) y[c] -= s;
  }
}

// Forward solve L y = b in place: y_k = Linv_k b_k, then b_r -= L[r, block k] y_k
// for rows r below.
__global__ __launch_bounds__(256) void k_forwardsolve_step(const double* __restrict__ S, size_t ld,
                                                           const double* __restrict__ Linv, int k,
                                                           uint32_t Dp, double* __restrict__ b,
                                                           double* __restrict__ y) {
  __shared__ double bs[NB];
  __shared__ double ys[NB];
  const int t = threadIdx.x;
  if (t < NB) bs[t] = b[(size_t)k * NB + t];
  __syncthreads();
  const double* X = Linv + (size_t)k * NB * NB;
  if (t < NB) {
    double s = 0.0;
    for (int c = 0; c <= t; ++c) s += X[(size_t)c * NB + t] * bs[c];
    ys[t] = s;
  }
  __syncthreads();
  const uint32_t r0 = (uint32_t)(k + 1) * NB;
  // Every block re-derives y_k; block 0 stores it.  Only rows r >= r0 of b
  // are written, disjoint from block k, which every block reads.
  if (blockIdx.x == 0 && t < NB) y[(size_t)k * NB + t] = ys[t];
  for (uint32_t r = r0 + blockIdx.x * 256 + t; r < Dp; r += gridDim.x * 256) {
    double s = 0.0;
    for (int c = 0; c < NB; ++c) s += S[(size_t)(k * NB + c) * ld + r] * ys[c];
    b[r] -= s;
  }
}

// r = b - A x, A symmetric stored full (column-major, ld).
__global__ __launch_bounds__(256) void k_gemv_residual(const double* __restrict__ A, size_t ld,
                                                       uint32_t n, const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ r) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (uint32_t c = 0; c < n; ++c) s += A[(size_t)c * ld + i] * x[c];
  r[i] = b[i] - s;
}

// dst (full symmetric) <- lower triangle of src; dst and src share ld.
__global__ __launch_bounds__(256) void k_mirror_copy(const double* __restrict__ src,
                                                     double* __restrict__ dst, size_t ld,
                                                     uint32_t n) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  const uint32_t c = blockIdx.y;
  if (r >= n) return;
  dst[(size_t)c * ld + r] = r >= c ? src[(size_t)c * ld + r] : src[(size_t)r * ld + c];
}

__global__ void k_extract_row(const double* __restrict__ S, size_t ld, uint32_t n, uint32_t row,
                              double* __restrict__ out) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c < n) out[c] = S[(size_t)c * ld + row];
}

__global__ void k_vec_add(double* __restrict__ x, const double* __restrict__ d, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] += d[i];
}

__global__ void k_copy(const double* __restrict__ s, double* __restrict__ d, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// ------------------------------------------------------------ LU fallback
// nalgebra LU::new: partial pivoting on the first maximal |a|, a zero
// column is skipped; solve fails on a zero diagonal of U.
__global__ __launch_bounds__(1024) void k_lu_pivot(double* __restrict__ A, size_t ld, uint32_t n,
                                                   int i, double* __restrict__ b,
                                                   int* __restrict__ skip) {
  __shared__ double bv[1024];
  __shared__ int bi[1024];
  const int t = threadIdx.x;
  double best = -1.0;
  int bidx = n;
  for (uint32_t r = i + t; r < n; r += 1024) {
    const double v = fabs(A[(size_t)i * ld + r]);
    if (v > best) {
      best = v;
      bidx = r;
    }
  }
  bv[t] = best;
  bi[t] = bidx;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (t < h) {
      const double v2 = bv[t + h];
      const int i2 = bi[t + h];
      if (v2 > bv[t] || (v2 == bv[t] && i2 < bi[t])) {
        bv[t] = v2;
        bi[t] = i2;
      }
    }
    __syncthreads();
  }
  const int piv = bi[0];
  const double diag = piv < (int)n ? A[(size_t)i * ld + piv] : 0.0;
  if (diag == 0.0) {
    if (t == 0) skip[i] = 1;
    return;
  }
  if (t == 0) skip[i] = 0;
  if (piv != i) {
    for (uint32_t c = t; c < n; c += 1024) {
      const double a = A[(size_t)c * ld + i];
      A[(size_t)c * ld + i] = A[(size_t)c * ld + piv];
      A[(size_t)c * ld + piv] = a;
    }
    if (t == 0) {
      const double tmp = b[i];
      b[i] = b[piv];
      b[piv] = tmp;
    }
  }
  __syncthreads();
  const double inv = 1.0 / diag;
  for (uint32_t r = i + 1 + t; r < n; r += 1024) A[(size_t)i * ld + r] *= inv;
}

__global__ __launch_bounds__(256) void k_lu_update(double* __restrict__ A, size_t ld, uint32_t n,
                                                   int i, const int* __restrict__ skip) {
  if (skip[i]) return;
  const uint32_t r = i + 1 + blockIdx.x * 256 + threadIdx.x;
  const uint32_t c = i + 1 + blockIdx.y;
  if (r >= n || c >= n) return;
  const double f = A[(size_t)c * ld + i];
  if (f == 0.0) return;
  A[(size_t)c * ld + r] += -f * A[(size_t)i * ld + r];
}

__global__ __launch_bounds__(1024) void k_lu_solve(const double* __restrict__ A, size_t ld,
                                                   uint32_t n, double* __restrict__ b,
                                                   int* __restrict__ flags) {
  const int t = threadIdx.x;
  for (uint32_t i = 0; i < n; ++i) {
    const double c = b[i];
    __syncthreads();
    for (uint32_t r = i + 1 + t; r < n; r += 1024) b[r] += -c * A[(size_t)i * ld + r];
    __syncthreads();
  }
  for (int i = (int)n - 1; i >= 0; --i) {
    const double diag = A[(size_t)i * ld + i];
    if (diag == 0.0) {
      if (t == 0) flags[1] = 1;
      return;
    }
    const double c = b[i] / diag;
    __syncthreads();
    if (t == 0) b[i] = c;
    for (int r = t; r < i; r += 1024) b[r] += -c * A[(size_t)i * ld + r];
    __syncthreads();
  }
}

// ------------------------------------------------------------- launchers
void dense_cholesky(double* S, double* Linv, uint32_t Dp, uint32_t ld, uint32_t D, int* flags,
                    hipStream_t s) {
  const int nblk = (int)(Dp / NB);
  for (int k = 0; k < nblk; ++k) {
    hipLaunchKernelGGL(k_potrf_block, dim3(1), dim3(256), 0, s, S, (size_t)ld, k, D, Linv, flags);
    const int nbelow = nblk - k;  // row blocks k+1 .. nblk (rhs block included)
    hipLaunchKernelGGL(k_trsm_panel, dim3(nbelow), dim3(256), 0, s, S, (size_t)ld, k, Linv, D, Dp);
    const int ntrail = nblk - 1 - k;
    if (ntrail > 0)
      hipLaunchKernelGGL(k_syrk_update, dim3(ntrail, ntrail + 1), dim3(256), 0, s, S, (size_t)ld,
                         k, nblk, D, Dp);
  }
  HIP_CHECK(hipGetLastError());
}

void dense_extract_row(const double* S, uint32_t Dp, uint32_t ld, uint32_t row, double* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_extract_row, dim3((Dp + 255) / 256), dim3(256), 0, s, S, (size_t)ld, Dp, row,
                     out);
  HIP_CHECK(hipGetLastError());
}

void dense_backsolve(const double* S, const double* Linv, uint32_t Dp, uint32_t ld, double* y,
                     double* x, hipStream_t s) {
  // y is consumed as scratch (its blocks are reduced in place); x receives
  // the solution.  Separate buffers: a workgroup writing x_k must not race
  // with another one still reading y_k.
  const int nblk = (int)(Dp / NB);
  for (int k = nblk - 1; k >= 0; --k) {
    const int ncols = k * NB;
    int grid = (ncols + 4 * 16 - 1) / (4 * 16);
    if (grid < 1) grid = 1;
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(k_backsolve_step, dim3(grid), dim3(256), 0, s, S, (size_t)ld, Linv, k, y, x);
  }
  HIP_CHECK(hipGetLastError());
}

void dense_forwardsolve(const double* S, const double* Linv, uint32_t Dp, uint32_t ld, double* b,
                        double* y, hipStream_t s) {
  // b is consumed as scratch (rows below block k are reduced), y receives
  // the solution.
  const int nblk = (int)(Dp / NB);
  for (int k = 0; k < nblk; ++k) {
    const uint32_t rows = Dp - (uint32_t)(k + 1) * NB;
    int grid = (int)((rows + 255) / 256);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_forwardsolve_step, dim3(grid), dim3(256), 0, s, S, (size_t)ld, Linv, k, Dp,
                       b, y);
  }
  HIP_CHECK(hipGetLastError());
}

void dense_symv_residual(const double* A, uint32_t Dp, uint32_t ld, const double* x,
                         const double* b, double* r, hipStream_t s) {
  hipLaunchKernelGGL(k_gemv_residual, dim3((Dp + 255) / 256), dim3(256), 0, s, A, (size_t)ld, Dp,
                     x, b, r);
  HIP_CHECK(hipGetLastError());
}

void dense_mirror_copy(const double* src, double* dst, uint32_t Dp, uint32_t ld, hipStream_t s) {
  hipLaunchKernelGGL(k_mirror_copy, dim3((Dp + 255) / 256, Dp), dim3(256), 0, s, src, dst,
                     (size_t)ld, Dp);
  HIP_CHECK(hipGetLastError());
}

void dense_vec_add(double* x, const double* d, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_vec_add, dim3((n + 255) / 256), dim3(256), 0, s, x, d, n);
  HIP_CHECK(hipGetLastError());
}

void dense_copy(const double* src, double* dst, size_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
  HIP_CHECK(hipGetLastError());
}

void dense_lu_solve(double* S, uint32_t Dp, uint32_t ld, uint32_t D, double* x, int* piv,
                    int* flags, hipStream_t s) {
  (void)D;
  dense_extract_row(S, Dp, ld, Dp, x, s);
  // Full symmetric matrix from the lower triangle, in place.
  dense_mirror_copy(S, S, Dp, ld, s);
  for (uint32_t i = 0; i < Dp; ++i) {
    hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(1024), 0, s, S, (size_t)ld, Dp, (int)i, x, piv);
    const uint32_t m = Dp - i - 1;
    if (m > 0)
      hipLaunchKernelGGL(k_lu_update, dim3((m + 255) / 256, m), dim3(256), 0, s, S, (size_t)ld, Dp,
                         (int)i, piv);
  }
  hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(1024), 0, s, S, (size_t)ld, Dp, x, flags);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sfmba
