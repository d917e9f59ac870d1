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

constexpr int KC = 16;   // K chunk staged through LDS
constexpr int PW = 16;   // panel width inside a diagonal block
constexpr int SMEM_DOUBLES = NB * NB + NB * PW + (NB - PW) * PW;  // block, its X11 blocks, trtri scratch

// acc(128x128 tile, 4 waves of 64x64) += sign * A(128 x K) * B(128 x K)^T,
// A and B column-major (rows contiguous) with leading dims lda, ldb.
// The v_mfma_f64_16x16x4_f64 operand layout: lane l supplies A[l%16][l/16]
// and B^T[l/16][l%16]; result register v of lane l is D[l/16 + 4v][l%16]
// (checked on gfx950 with a one-wavefront probe of the builtin).
// Software-pipelined: the next K chunk's global loads are in flight while
// the current chunk's MFMAs run.
__device__ inline void tile_mma(double4_t acc[4][4], const double* __restrict__ A, size_t lda,
                                const double* __restrict__ B, size_t ldb, int K, double sign,
                                double* As, double* Bs) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  constexpr int PER = (KC * 128) / 256;
  double ra[PER], rb[PER];
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const int idx = t + 256 * it;
    const int k = idx >> 7, row = idx & 127;
    ra[it] = A[(size_t)k * lda + row];
    rb[it] = B[(size_t)k * ldb + row];
  }
  for (int k0 = 0; k0 < K; k0 += KC) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int idx = t + 256 * it;
      As[idx] = ra[it];
      Bs[idx] = rb[it];
    }
    __syncthreads();
    if (k0 + KC < K) {
#pragma unroll
      for (int it = 0; it < PER; ++it) {
        const int idx = t + 256 * it;
        const int k = idx >> 7, row = idx & 127;
        ra[it] = A[(size_t)(k0 + KC + k) * lda + row];
        rb[it] = B[(size_t)(k0 + KC + k) * ldb + row];
      }
    }
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

// Load a 128x128 tile into the MFMA accumulator layout.
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

// The accumulator tile into a column-major NB x NB LDS block, with the
// identity on padding rows/columns (global index >= D).
__device__ inline void tile_to_lds(const double4_t acc[4][4], double* L, uint32_t base, uint32_t D) {
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
        const int r = row + 4 * v;
        const bool pad = base + r >= D || base + col >= D;
        L[col * NB + r] = pad ? (r == col ? 1.0 : 0.0) : acc[m][n][v];
      }
    }
}

// ------------------------------------------------------- diagonal block
// Broadcast lane l's double to the whole wavefront (v_readlane).
__device__ __attribute__((always_inline)) inline double readlane_f64(double v, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// 1/sqrt(d): the hardware estimate refined by one Newton step (a few ulp).
__device__ __attribute__((always_inline)) inline double rsqrt_f64(double d) {
  const double y = __builtin_amdgcn_rsq(d);
  return y * (1.5 - 0.5 * d * y * y);
}

// Cholesky of a PW x PW block and its inverse, on wave 0 of the workgroup:
// lane r < PW holds row r in registers; pivots and multipliers travel by
// v_readlane, and each pivot costs one reciprocal square root (no divide),
// which keeps the 16-step dependent chain short.  L11 is written back to A
// (column-major, stride NB) and X11 = L11^{-1} (column-major, stride PW)
// to Xd.  A pivot that is not > 0 (or NaN) on a real row sets flags[0].
__device__ __attribute__((always_inline)) inline void wave_potrf16(double* A, int j0, double* Xd,
                                                                    uint32_t base, uint32_t D,
                                                                    int* flags) {
  const int lane = threadIdx.x & 63;
  const int r = lane & (PW - 1);
  double row[PW];
  double ipiv[PW];  // 1 / L[j][j]
#pragma unroll
  for (int c = 0; c < PW; ++c) row[c] = (c <= r) ? A[(j0 + c) * NB + j0 + r] : 0.0;
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    const double d = readlane_f64(row[j], j);
    if (lane == 0 && !(d > 0.0) && base + j0 + j < D) atomicOr(&flags[0], 1);
    const double ip = rsqrt_f64(d);
    ipiv[j] = ip;
    row[j] = (r == j) ? d * ip : (r > j ? row[j] * ip : row[j]);
#pragma unroll
    for (int c = j + 1; c < PW; ++c) {
      const double lcj = readlane_f64(row[j], c);
      if (r >= c) row[c] -= row[j] * lcj;
    }
  }
  if (lane < PW) {
#pragma unroll
    for (int c = 0; c < PW; ++c)
      if (c <= r) A[(j0 + c) * NB + j0 + r] = row[c];
  }
  // X11, lane c computes column c: x_i = (delta_ic - sum_{k<i} L[i][k] x_k) / L[i][i],
  // L entries broadcast from the owning lane.
  const int c = lane & (PW - 1);
  double x[PW];
#pragma unroll
  for (int i = 0; i < PW; ++i) {
    double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) s -= readlane_f64(row[k], i) * x[k];
    x[i] = s * ipiv[i];
  }
  if (lane < PW) {
#pragma unroll
    for (int i = 0; i < PW; ++i) Xd[c * PW + i] = x[i];
  }
}

// Cholesky of the lower NB x NB block in LDS (column-major, A[c*NB + r]),
// in PW-wide panels: the register factorization above, the panel below it
// as L21 = A21 X11^T (independent loads, no dependent chain), and a
// 4x4-register-tiled rank-PW update of the trailing lower triangle.  The
// inverses X11 of the diagonal sub-blocks are kept in Xd (NB/PW blocks of
// PW x PW) for the inverse pass.  A pivot that is not > 0 (or NaN) on a
// real row (base + j < D) sets flags[0] (nalgebra's Cholesky::new failure
// test).
__device__ __attribute__((always_inline)) inline void lds_potrf(double* A, double* Xd, uint32_t base,
                                                                 uint32_t D, int* flags) {
  const int t = threadIdx.x, wave = t >> 6;
  for (int j0 = 0; j0 < NB; j0 += PW) {
    double* X11 = Xd + (j0 / PW) * PW * PW;
    if (wave == 0) wave_potrf16(A, j0, X11, base, D, flags);
    __syncthreads();
    const int h = NB - j0 - PW;
    if (h == 0) break;
    // L21 = A21 X11^T: four threads per row, each producing outputs
    // c = q, q+4, q+8, q+12 (x_c = sum_{k <= c} a_k X11[c][k]).
    // The trip count is uniform across the workgroup (barrier inside).
    for (int e0 = 0; e0 < 4 * h; e0 += blockDim.x) {
      const int e = e0 + t;
      const bool valid = e < 4 * h;
      const int rr = j0 + PW + (valid ? (e >> 2) : 0), q = e & 3;
      double o[4];
      if (valid) {
        double a[PW];
#pragma unroll
        for (int k = 0; k < PW; ++k) a[k] = A[(j0 + k) * NB + rr];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int cc = q + 4 * m;
          double sum = 0.0;
#pragma unroll
          for (int k = 0; k < PW; ++k)
            if (k <= cc) sum += a[k] * X11[k * PW + cc];
          o[m] = sum;
        }
      }
      __syncthreads();  // every thread of the row has read it before it is overwritten
      if (valid) {
#pragma unroll
        for (int m = 0; m < 4; ++m) A[(j0 + q + 4 * m) * NB + rr] = o[m];
      }
    }
    __syncthreads();
    const int T = h / 4;
    const int ntiles = T * (T + 1) / 2;
    for (int e = t; e < ntiles; e += blockDim.x) {
      int ti = (int)((sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
      while (ti * (ti + 1) / 2 > e) --ti;
      while ((ti + 1) * (ti + 2) / 2 <= e) ++ti;
      const int tj = e - ti * (ti + 1) / 2;
      const int r0 = j0 + PW + 4 * ti, c0 = j0 + PW + 4 * tj;
      double acc[4][4];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = A[(c0 + q) * NB + r0 + p];
#pragma unroll 4
      for (int k = 0; k < PW; ++k) {
        double lr[4], lc[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          lr[p] = A[(j0 + k) * NB + r0 + p];
          lc[p] = A[(j0 + k) * NB + c0 + p];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[p][q] -= lr[p] * lc[q];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (ti != tj || p >= q) A[(c0 + q) * NB + r0 + p] = acc[p][q];
    }
    __syncthreads();
  }
}

// In-place inverse of the lower triangle L held in LDS, block columns right
// to left: A21 <- -(X22 A21) X11 with X22 the already inverted trailing
// triangle and X11 kept by lds_potrf, then the diagonal block <- X11.
// tmp holds (NB - PW) x PW doubles.
__device__ __attribute__((always_inline)) inline void lds_trtri(double* A, const double* Xd, double* tmp) {
  const int t = threadIdx.x;
  for (int j0 = NB - PW; j0 >= 0; j0 -= PW) {
    const double* X11 = Xd + (j0 / PW) * PW * PW;
    const int h = NB - j0 - PW;
    if (h > 0) {
      for (int e = t; e < h * PW; e += blockDim.x) {
        const int c = e / h, rr = e - c * h;
        tmp[c * h + rr] = A[(j0 + c) * NB + j0 + PW + rr];
      }
      __syncthreads();
      // A21 <- X22 * A21, 4x4 output tiles (X22 lower, already inverted)
      const int nt = (h / 4) * (PW / 4);
      for (int e = t; e < nt; e += blockDim.x) {
        const int ti = e / (PW / 4), tj = e - ti * (PW / 4);
        const int r0 = 4 * ti, c0 = 4 * tj;
        double acc[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
#pragma unroll 4
        for (int k = 0; k < r0 + 4; ++k) {
          double xr[4], tc[4];
#pragma unroll
          for (int p = 0; p < 4; ++p) xr[p] = k <= r0 + p ? A[(j0 + PW + k) * NB + j0 + PW + r0 + p] : 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) tc[q] = tmp[(c0 + q) * h + k];
#pragma unroll
          for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[p][q] += xr[p] * tc[q];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) A[(j0 + c0 + q) * NB + j0 + PW + r0 + p] = acc[p][q];
      }
      __syncthreads();
      // A21 <- -A21 X11: y_c = -sum_{k >= c} a_k X11[k][c], in place from the first column
      if (t < h) {
        const int rr = j0 + PW + t;
        double a[PW];
#pragma unroll
        for (int k = 0; k < PW; ++k) a[k] = A[(j0 + k) * NB + rr];
#pragma unroll
        for (int c = 0; c < PW; ++c) {
          double s = 0.0;
#pragma unroll
          for (int k = c; k < PW; ++k) s += a[k] * X11[c * PW + k];
          a[c] = -s;
        }
#pragma unroll
        for (int k = 0; k < PW; ++k) A[(j0 + k) * NB + rr] = a[k];
      }
    }
    for (int e = t; e < PW * PW; e += blockDim.x) {
      const int c = e / PW, rr = e - c * PW;
      if (rr >= c) A[(j0 + c) * NB + j0 + rr] = X11[c * PW + rr];
    }
    __syncthreads();
  }
}

// Factor the diagonal block held in LDS, store L (real rows) into S and
// L^{-1} (identity on padding) into Linv.
__device__ __attribute__((always_inline)) inline void factor_diag_block(double* smem, double* __restrict__ S,
                                                                         size_t ld, int k, uint32_t D,
                                                                         double* __restrict__ Linv,
                                                                         int* __restrict__ flags) {
  const int t = threadIdx.x;
  const uint32_t base = (uint32_t)k * NB;
  double* A = smem;
  double* Xd = smem + NB * NB;
  double* tmp = Xd + NB * PW;
  __syncthreads();
  lds_potrf(A, Xd, base, D, flags);
  double* G = S + (size_t)k * NB * ld + (size_t)k * NB;
  for (int idx = t; idx < NB * NB; idx += blockDim.x) {
    const int c = idx >> 7, r = idx & 127;
    if (r >= c && base + r < D) G[(size_t)c * ld + r] = A[c * NB + r];
  }
  __syncthreads();
  lds_trtri(A, Xd, tmp);
  // Linv holds, per block, L^{-1} (column-major) followed by its transpose.
  double* X = Linv + (size_t)k * 2 * NB * NB;
  for (int idx = t; idx < NB * NB; idx += blockDim.x) {
    const int c = idx >> 7, r = idx & 127;
    const bool pad = base + r >= D || base + c >= D;
    X[idx] = pad ? (r == c ? 1.0 : 0.0) : (r >= c ? A[c * NB + r] : 0.0);
    const bool padT = base + c >= D || base + r >= D;
    X[NB * NB + idx] = padT ? (r == c ? 1.0 : 0.0) : (c >= r ? A[r * NB + c] : 0.0);
  }
}

__global__ __launch_bounds__(256) void k_potrf_block(double* __restrict__ S, size_t ld, int k,
                                                     uint32_t D, double* __restrict__ Linv,
                                                     int* __restrict__ flags) {
  __shared__ double smem[SMEM_DOUBLES];
  const uint32_t base = (uint32_t)k * NB;
  const double* G = S + (size_t)k * NB * ld + (size_t)k * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int c = idx >> 7, r = idx & 127;
    const bool pad = base + r >= D || base + c >= D;
    smem[c * NB + r] = pad ? (r == c ? 1.0 : 0.0) : (r >= c ? G[(size_t)c * ld + r] : 0.0);
  }
  factor_diag_block(smem, S, ld, k, D, Linv, flags);
}

// L_ik = A_ik Linv_k^T for row blocks i = k+1 .. nblk (nblk: the rhs block).
__global__ __launch_bounds__(256) void k_trsm_panel(double* __restrict__ S, size_t ld, int k,
                                                    const double* __restrict__ Linv, uint32_t D,
                                                    uint32_t Dp) {
  __shared__ double As[KC * 128];
  __shared__ double Bs[KC * 128];
  const int i = k + 1 + blockIdx.x;
  double* A = S + (size_t)k * NB * ld + (size_t)i * NB;
  const double* X = Linv + (size_t)k * 2 * NB * NB;
  double4_t acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = double4_t{0.0, 0.0, 0.0, 0.0};
  tile_mma(acc, A, ld, X, NB, NB, 1.0, As, Bs);
  __syncthreads();
  tile_store(acc, A, ld, false, (uint32_t)i * NB, (uint32_t)k * NB, D, Dp);
}

// A_ij -= L_ik L_jk^T for k < j <= i, j < nblk, i <= nblk (i == nblk: rhs
// block).  The workgroup of the next diagonal tile (k+1, k+1) factors it
// right away (lookahead), so the next step's panel solve can start.
__global__ __launch_bounds__(256) void k_syrk_update(double* __restrict__ S, size_t ld, int k,
                                                     int nblk, uint32_t D, uint32_t Dp,
                                                     double* __restrict__ Linv,
                                                     int* __restrict__ flags) {
  __shared__ double smem[SMEM_DOUBLES];
  // Workgroup (0, 0) is the next diagonal tile (k+1, k+1): dispatched first,
  // so its factorization overlaps the other tiles' updates.
  const int j = k + 1 + blockIdx.x;
  const int i = k + 1 + blockIdx.y;
  if (j > i || j >= nblk) return;
  double* C = S + (size_t)j * NB * ld + (size_t)i * NB;
  const double* Ai = S + (size_t)k * NB * ld + (size_t)i * NB;
  const double* Aj = S + (size_t)k * NB * ld + (size_t)j * NB;
  double4_t acc[4][4];
  tile_load(acc, C, ld);
  tile_mma(acc, Ai, ld, Aj, ld, NB, -1.0, smem, smem + KC * 128);
  if (i == k + 1 && j == k + 1) {
    __syncthreads();
    tile_to_lds(acc, smem, (uint32_t)j * NB, D);
    factor_diag_block(smem, S, ld, k + 1, D, Linv, flags);
    return;
  }
  tile_store(acc, C, ld, i == j, (uint32_t)i * NB, (uint32_t)j * NB, D, Dp);
}

// ------------------------------------------------------ triangular solves
// Backward solve L^T x = y, block k (last to first), in two launches:
// x_k = Linv_k^T y_k (one workgroup), then y_c -= L[block k, c] . x_k for
// every column c < k*NB.
__global__ __launch_bounds__(128) void k_backsolve_block(const double* __restrict__ Linv, int k,
                                                         const double* __restrict__ y,
                                                         double* __restrict__ x) {
  __shared__ double ys[NB];
  const int t = threadIdx.x;
  ys[t] = y[(size_t)k * NB + t];
  __syncthreads();
  // x[t] = sum_r Linv[r][t] y[r] = sum_r LinvT[t][r] y[r]; LinvT column-major
  // makes LinvT[t][r] = XT[r*NB + t] contiguous across threads.
  const double* XT = Linv + (size_t)k * 2 * NB * NB + NB * NB;
  double s = 0.0;
#pragma unroll 8
  for (int r = 0; r < NB; ++r) s += XT[(size_t)r * NB + t] * ys[r];
  x[(size_t)k * NB + t] = s;
}

__global__ __launch_bounds__(256) void k_backsolve_update(const double* __restrict__ S, size_t ld, int k,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ y) {
  __shared__ double xs[NB];
  const int t = threadIdx.x;
  if (t < NB) xs[t] = x[(size_t)k * NB + t];
  __syncthreads();
  const int ncols = k * NB;
  const int lane = t & 63, wave = t >> 6;
  for (int c = blockIdx.x * 4 + wave; c < ncols; c += gridDim.x * 4) {
    const double* col = S + (size_t)c * ld + (size_t)k * NB;
    double s = col[lane] * xs[lane] + col[lane + 64] * xs[lane + 64];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) y[c] -= s;
  }
}

// Forward solve L y = b: y_k = Linv_k b_k (one workgroup), then
// b_r -= L[r, block k] . y_k for the rows below.
__global__ __launch_bounds__(128) void k_forward_block(const double* __restrict__ Linv, int k,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ y) {
  __shared__ double bs[NB];
  const int t = threadIdx.x;
  bs[t] = b[(size_t)k * NB + t];
  __syncthreads();
  const double* X = Linv + (size_t)k * 2 * NB * NB;
  double s = 0.0;
#pragma unroll 8
  for (int c = 0; c < NB; ++c) s += X[(size_t)c * NB + t] * bs[c];
  y[(size_t)k * NB + t] = s;
}

__global__ __launch_bounds__(256) void k_forward_update(const double* __restrict__ S, size_t ld, int k,
                                                        uint32_t Dp, const double* __restrict__ y,
                                                        double* __restrict__ b) {
  __shared__ double ys[NB];
  const int t = threadIdx.x;
  if (t < NB) ys[t] = y[(size_t)k * NB + t];
  __syncthreads();
  const uint32_t r = (uint32_t)(k + 1) * NB + blockIdx.x * 256 + t;
  if (r >= Dp) return;
  double s = 0.0;
  for (int c = 0; c < NB; ++c) s += S[(size_t)(k * NB + c) * ld + r] * ys[c];
  b[r] -= s;
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
  hipLaunchKernelGGL(k_potrf_block, dim3(1), dim3(256), 0, s, S, (size_t)ld, 0, D, Linv, flags);
  for (int k = 0; k < nblk; ++k) {
    const int nbelow = nblk - k;  // row blocks k+1 .. nblk (rhs block included)
    hipLaunchKernelGGL(k_trsm_panel, dim3(nbelow), dim3(256), 0, s, S, (size_t)ld, k, Linv, D, Dp);
    const int ntrail = nblk - 1 - k;
    if (ntrail > 0)
      hipLaunchKernelGGL(k_syrk_update, dim3(ntrail, ntrail + 1), dim3(256), 0, s, S, (size_t)ld, k,
                         nblk, D, Dp, Linv, flags);
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
  // the solution.
  const int nblk = (int)(Dp / NB);
  for (int k = nblk - 1; k >= 0; --k) {
    hipLaunchKernelGGL(k_backsolve_block, dim3(1), dim3(NB), 0, s, Linv, k, y, x);
    const int ncols = k * NB;
    if (ncols > 0) {
      int grid = (ncols + 15) / 16;
      if (grid > 2048) grid = 2048;
      hipLaunchKernelGGL(k_backsolve_update, dim3(grid), dim3(256), 0, s, S, (size_t)ld, k, x, y);
    }
  }
  HIP_CHECK(hipGetLastError());
}

void dense_forwardsolve(const double* S, const double* Linv, uint32_t Dp, uint32_t ld, double* b,
                        double* y, hipStream_t s) {
  // b is consumed as scratch (rows below block k are reduced), y receives
  // the solution.
  const int nblk = (int)(Dp / NB);
  for (int k = 0; k < nblk; ++k) {
    hipLaunchKernelGGL(k_forward_block, dim3(1), dim3(NB), 0, s, Linv, k, b, y);
    const uint32_t rows = Dp - (uint32_t)(k + 1) * NB;
    if (rows > 0)
      hipLaunchKernelGGL(k_forward_update, dim3((rows + 255) / 256), dim3(256), 0, s, S, (size_t)ld, k, Dp, y,
                         b);
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
