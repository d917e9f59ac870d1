// Block-Jacobi preconditioned conjugate gradients on the reduced camera
// system (the iterative alternative to the dense factorization for large
// systems; no counterpart in the reference, which always factors).
//
// S is the FP64 reduced system as the assembly leaves it: column-major,
// lower triangle only, leading dimension ld, right-hand side in row Dp.
//   k_bj_invert      inverse of every parameter block's diagonal block
//                    (Gauss-Jordan, one wavefront per block, no pivoting: the
//                    blocks of a damped normal matrix are positive definite;
//                    a pivot that is not > 0 flags a breakdown)
//   k_symv_fused     one pass over the lower triangle L: a workgroup reads
//                    a 256-row x 1024-column tile once and forms both its
//                    row products (L x, per-chunk partials) and the column
//                    products of its strictly lower part ((L - diag)^T x,
//                    one wavefront reduction per column, the four
//                    wavefronts combined in a fixed order, per-row-block
//                    partials)
//   k_symv_finish    y = the row partials in chunk order + the column
//                    partials in row-block order
//   k_dot2 / k_dot2_final
//                    two dot products, deterministic two-level reduction
// Every reduction has a fixed order, so a solve is reproducible.
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.hpp"
#include "pcg.hpp"

namespace sfmba {

constexpr int SYMV_ROWS = 256;
constexpr int SYMV_COLS = 1024;
constexpr int DOT_BLOCK = 256;
constexpr int DOT_ITEMS = 8;

__global__ __launch_bounds__(256) void k_bj_invert(uint32_t n_grp, const uint32_t* __restrict__ grp_off,
                                                   const uint32_t* __restrict__ grp_dim,
                                                   const uint32_t* __restrict__ moff,
                                                   const double* __restrict__ S, size_t ld,
                                                   double* __restrict__ Minv, int* __restrict__ flags) {
  __shared__ double A[4][16 * 16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g = blockIdx.x * 4 + wave;
  if (g >= n_grp) return;  // whole wavefront; no workgroup barrier below
  double* a = A[wave];
  const uint32_t d = grp_dim[g], off = grp_off[g];
  for (uint32_t e = lane; e < d * d; e += 64) {
    const uint32_t r = e % d, c = e / d;
    a[c * 16 + r] = r >= c ? S[(size_t)(off + c) * ld + off + r] : S[(size_t)(off + r) * ld + off + c];
  }
  bool ok = true;
  for (uint32_t k = 0; k < d; ++k) {
    // every lane reads the old matrix before any lane writes (one wavefront)
    const double p = a[k * 16 + k];
    if (!(p > 0.0)) ok = false;
    double nv[4];
    uint32_t ne = 0;
    for (uint32_t e = lane; e < d * d; e += 64, ++ne) {
      const uint32_t i = e % d, j = e / d;
      const double rk = (j == k ? 1.0 : a[j * 16 + k]) / p;  // row k, scaled
      nv[ne] = (i == k) ? rk : ((j == k ? 0.0 : a[j * 16 + i]) - a[k * 16 + i] * rk);
    }
    __asm__ volatile("" ::: "memory");
    ne = 0;
    for (uint32_t e = lane; e < d * d; e += 64, ++ne) a[(e / d) * 16 + e % d] = nv[ne];
    __asm__ volatile("" ::: "memory");
  }
  if (!ok && lane == 0) atomicOr(&flags[0], 1);
  double* m = Minv + moff[g];
  for (uint32_t e = lane; e < d * d; e += 64) m[e] = a[(e / d) * 16 + e % d];
}

// z = M^{-1} r: one thread per unknown.
__global__ __launch_bounds__(256) void k_bj_apply(uint32_t D, const uint32_t* __restrict__ row_grp,
                                                  const uint32_t* __restrict__ grp_off,
                                                  const uint32_t* __restrict__ grp_dim,
                                                  const uint32_t* __restrict__ moff,
                                                  const double* __restrict__ Minv,
                                                  const double* __restrict__ r, double* __restrict__ z) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const uint32_t g = row_grp[i], off = grp_off[g], d = grp_dim[g];
  const uint32_t ri = i - off;
  const double* m = Minv + moff[g];
  double s = 0.0;
  for (uint32_t c = 0; c < d; ++c) s += m[c * d + ri] * r[off + c];
  z[i] = s;
}

// prow[cc * D + i]: row i's products over column chunk cc (columns <= i);
// pcol[rb * D + j]: column j's products over the rows > j of row block rb.
// Tiles right of the diagonal are skipped; every partial the finish reads
// belongs to a tile that is not.
__global__ __launch_bounds__(SYMV_ROWS) void k_symv_fused(uint32_t D, const double* __restrict__ S, size_t ld,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ prow, double* __restrict__ pcol) {
  __shared__ double xs[SYMV_COLS];
  __shared__ double cs[SYMV_ROWS / 64][SYMV_COLS];
  const uint32_t rb = blockIdx.x, cc = blockIdx.y;
  const uint32_t r0 = rb * SYMV_ROWS, c0 = cc * SYMV_COLS;
  if (c0 > r0 + SYMV_ROWS - 1) return;  // chunk right of the diagonal: nothing
  const uint32_t c1 = min(D, c0 + SYMV_COLS);
  const uint32_t nc = c1 - c0;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (uint32_t j = t; j < nc; j += SYMV_ROWS) xs[j] = x[c0 + j];
  for (uint32_t j = t; j < SYMV_COLS; j += SYMV_ROWS)
#pragma unroll
    for (int q = 0; q < SYMV_ROWS / 64; ++q) cs[q][j] = 0.0;
  __syncthreads();
  const uint32_t i = r0 + t;
  const bool row_ok = i < D;
  const double xi = row_ok ? x[i] : 0.0;
  // columns past this wavefront's last row hold no lower-triangle entries
  const uint32_t wlast = r0 + 64 * w + 63;
  const uint32_t jend = min(c1, wlast + 1);
  double acc = 0.0;
  // Eight columns at a time: eight independent loads, then the column sums
  // of the eight by recursive halving (lanes l and l ^ 32 / 16 / 8 split
  // and add the eight values, then xor 4 / 2 / 1 finish the sum): lane l
  // ends with column 4 b5 + 2 b4 + b3 (bits of l) summed over the wavefront.
  for (uint32_t j0 = c0; j0 < jend; j0 += 8) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t j = j0 + k;
      const double a = (row_ok && j < jend && j <= i) ? S[(size_t)j * ld + i] : 0.0;
      acc += a * (j < jend ? xs[j - c0] : 0.0);
      v[k] = (j < i) ? a * xi : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool up = lane & 32;
      const double send = up ? v[k] : v[k + 4], keep = up ? v[k + 4] : v[k];
      v[k] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const bool up = lane & 16;
      const double send = up ? v[k] : v[k + 2], keep = up ? v[k + 2] : v[k];
      v[k] = keep + __shfl_xor(send, 16);
    }
    {
      const bool up = lane & 8;
      const double send = up ? v[0] : v[1], keep = up ? v[1] : v[0];
      v[0] = keep + __shfl_xor(send, 8);
    }
    v[0] += __shfl_xor(v[0], 4);
    v[0] += __shfl_xor(v[0], 2);
    v[0] += __shfl_xor(v[0], 1);
    const uint32_t jc = j0 + ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
    if ((lane & 7) == 0 && jc < jend) cs[w][jc - c0] = v[0];
  }
  if (row_ok) prow[(size_t)cc * D + i] = acc;
  __syncthreads();
  for (uint32_t j = t; j < nc; j += SYMV_ROWS) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < SYMV_ROWS / 64; ++q) v += cs[q][j];
    pcol[(size_t)rb * D + c0 + j] = v;
  }
}

__global__ __launch_bounds__(256) void k_symv_finish(uint32_t D, const double* __restrict__ prow,
                                                     const double* __restrict__ pcol, double* __restrict__ y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const uint32_t ncc = i / SYMV_COLS + 1;  // chunks left of / at the diagonal
  double s = 0.0;
  for (uint32_t cc = 0; cc < ncc; ++cc) s += prow[(size_t)cc * D + i];
  const uint32_t nrb = (D + SYMV_ROWS - 1) / SYMV_ROWS;
  for (uint32_t rb = i / SYMV_ROWS; rb < nrb; ++rb) s += pcol[(size_t)rb * D + i];
  y[i] = s;
}

// part[2 b], part[2 b + 1] = block b's sums of a.b and c.d.
__global__ __launch_bounds__(DOT_BLOCK) void k_dot2(uint32_t n, const double* __restrict__ a,
                                                    const double* __restrict__ b, const double* __restrict__ c,
                                                    const double* __restrict__ d, double* __restrict__ part) {
  __shared__ double r0[DOT_BLOCK], r1[DOT_BLOCK];
  double s0 = 0.0, s1 = 0.0;
  const uint32_t base = blockIdx.x * DOT_BLOCK * DOT_ITEMS + threadIdx.x;
#pragma unroll
  for (int it = 0; it < DOT_ITEMS; ++it) {
    const uint32_t i = base + it * DOT_BLOCK;
    if (i < n) {
      s0 += a[i] * b[i];
      if (c) s1 += c[i] * d[i];
    }
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int h = DOT_BLOCK / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r0[threadIdx.x] += r0[threadIdx.x + h];
      r1[threadIdx.x] += r1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

__global__ __launch_bounds__(1024) void k_dot2_final(uint32_t n_part, const double* __restrict__ part,
                                                     double* __restrict__ out) {
  __shared__ double r0[1024], r1[1024];
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t i = threadIdx.x; i < n_part; i += 1024) {
    s0 += part[2 * i];
    s1 += part[2 * i + 1];
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r0[threadIdx.x] += r0[threadIdx.x + h];
      r1[threadIdx.x] += r1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r0[0];
    out[1] = r1[0];
  }
}

// x += alpha p, r -= alpha q
__global__ __launch_bounds__(256) void k_pcg_xr(uint32_t n, double alpha, const double* __restrict__ p,
                                                const double* __restrict__ q, double* __restrict__ x,
                                                double* __restrict__ r) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_pcg_p(uint32_t n, double beta, const double* __restrict__ z,
                                               double* __restrict__ p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + beta * p[i];
}

// r = b (row Dp of S), x = 0
__global__ __launch_bounds__(256) void k_pcg_init(uint32_t n, const double* __restrict__ S, size_t ld,
                                                  uint32_t Dp, double* __restrict__ r, double* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = S[(size_t)i * ld + Dp];
  x[i] = 0.0;
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t pcg_symv_part_doubles(uint32_t D) { return ((size_t)cdiv(D, SYMV_COLS) + cdiv(D, SYMV_ROWS)) * D; }
size_t pcg_dot_part_doubles(uint32_t D) { return 2 * (size_t)cdiv(D, DOT_BLOCK * DOT_ITEMS) + 2; }

static void symv(const PcgBuffers& B, const double* S, size_t ld, uint32_t D, const double* x, double* y,
                 hipStream_t s) {
  double* prow = B.symv_part;
  double* pcol = B.symv_part + (size_t)cdiv(D, SYMV_COLS) * D;
  hipLaunchKernelGGL(k_symv_fused, dim3(cdiv(D, SYMV_ROWS), cdiv(D, SYMV_COLS)), dim3(SYMV_ROWS), 0, s, D, S,
                     ld, x, prow, pcol);
  hipLaunchKernelGGL(k_symv_finish, dim3(cdiv(D, 256)), dim3(256), 0, s, D, prow, pcol, y);
}

static void dot2(const PcgBuffers& B, uint32_t n, const double* a, const double* b, const double* c,
                 const double* d, hipStream_t s) {
  const uint32_t nb = cdiv(n, DOT_BLOCK * DOT_ITEMS);
  hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(DOT_BLOCK), 0, s, n, a, b, c, d, B.dot_part);
  hipLaunchKernelGGL(k_dot2_final, dim3(1), dim3(1024), 0, s, nb, B.dot_part, B.scal);
}

PcgResult pcg_solve(const PcgBuffers& B, const double* S, size_t ld, uint32_t D, uint32_t Dp,
                    uint32_t n_grp, const uint32_t* grp_off, const uint32_t* grp_dim, int max_iterations,
                    double tolerance, double* x, int* flags, double* host_scal, hipStream_t s) {
  PcgResult res;
  if (D == 0) {
    res.ok = true;
    return res;
  }
  const uint32_t g256 = cdiv(D, 256);
  HIP_CHECK(hipMemsetAsync(flags, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_bj_invert, dim3(cdiv(n_grp, 4)), dim3(256), 0, s, n_grp, grp_off, grp_dim, B.moff, S,
                     ld, B.Minv, flags);
  hipLaunchKernelGGL(k_pcg_init, dim3(g256), dim3(256), 0, s, D, S, ld, Dp, B.r, x);
  hipLaunchKernelGGL(k_bj_apply, dim3(g256), dim3(256), 0, s, D, B.row_grp, grp_off, grp_dim, B.moff, B.Minv,
                     B.r, B.z);
  dot2(B, D, B.r, B.r, B.r, B.z, s);  // |b|^2, r.z
  HIP_CHECK(hipMemcpyAsync(host_scal, B.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(host_scal + 2, flags, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (reinterpret_cast<const int*>(host_scal + 2)[0] != 0) return res;  // a block is not positive definite
  const double bn = sqrt(host_scal[0]);
  double rz = host_scal[1];
  if (bn == 0.0) {
    res.ok = true;
    return res;
  }
  HIP_CHECK(hipMemcpyAsync(B.p, B.z, (size_t)D * sizeof(double), hipMemcpyDeviceToDevice, s));
  for (int it = 0; it < max_iterations; ++it) {
    symv(B, S, ld, D, B.p, B.q, s);
    dot2(B, D, B.p, B.q, nullptr, nullptr, s);
    HIP_CHECK(hipMemcpyAsync(host_scal, B.scal, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const double pq = host_scal[0];
    if (!(pq > 0.0) || !std::isfinite(rz)) return res;  // breakdown: not positive definite
    const double alpha = rz / pq;
    hipLaunchKernelGGL(k_pcg_xr, dim3(g256), dim3(256), 0, s, D, alpha, B.p, B.q, x, B.r);
    hipLaunchKernelGGL(k_bj_apply, dim3(g256), dim3(256), 0, s, D, B.row_grp, grp_off, grp_dim, B.moff,
                       B.Minv, B.r, B.z);
    dot2(B, D, B.r, B.r, B.r, B.z, s);
    HIP_CHECK(hipMemcpyAsync(host_scal, B.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    res.iterations = it + 1;
    const double rn = sqrt(host_scal[0]);
    res.rel_residual = rn / bn;
    if (!std::isfinite(rn)) return res;
    if (rn <= tolerance * bn) {
      res.ok = true;
      return res;
    }
    const double rz_new = host_scal[1];
    const double beta = rz_new / rz;
    rz = rz_new;
    hipLaunchKernelGGL(k_pcg_p, dim3(g256), dim3(256), 0, s, D, beta, B.z, B.p);
  }
  // iteration cap: the step so far is still a descent step of the damped
  // model; it is used as is
  res.ok = true;
  return res;
}

}  // namespace sfmba
