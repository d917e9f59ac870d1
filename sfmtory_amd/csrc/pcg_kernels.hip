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
//   k_symv_finish_dot
//                    y = the row partials in chunk order + the column
//                    partials in row-block order, and block partials of p.y
//   k_dot2 / k_dot2_final
//                    two dot products, deterministic two-level reduction
// The iteration runs without the host: alpha, beta, the convergence test
// and the breakdown test live in a device-side state (PcgState) that every
// iteration kernel reads; once it is no longer running the remaining
// enqueued kernels return at once.  The host enqueues iterations in
// batches and reads the state once per batch.
// Every reduction has a fixed order, so a solve is reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "pcg.hpp"

namespace sfmba {

constexpr int SYMV_ROWS = 256;
// Column chunk of a matvec tile: 1,024 columns above SYMV_SMALL_D unknowns;
// below, 256, so that a system of a few thousand unknowns still spreads
// over enough workgroups (D = 7,994: 528 tiles of 0.5 MB instead of 144 of
// 2 MB, which left most CUs idle).
constexpr int SYMV_COLS_LARGE = 1024;
constexpr int SYMV_COLS_SMALL = 256;
constexpr uint32_t SYMV_SMALL_D = 16384;
static inline int symv_cols(uint32_t D) { return D > SYMV_SMALL_D ? SYMV_COLS_LARGE : SYMV_COLS_SMALL; }
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
template <int SYMV_COLS>
__global__ __launch_bounds__(SYMV_ROWS) void k_symv_fused(uint32_t D, const double* __restrict__ S, size_t ld,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ prow, double* __restrict__ pcol,
                                                          const int* __restrict__ stopped) {
  __shared__ double xs[SYMV_COLS];
  __shared__ double cs[SYMV_ROWS / 64][SYMV_COLS];
  if (stopped && *stopped != 0) return;  // the device-driven iteration has finished
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
  // The next eight columns' loads are issued before this group's sums.
  auto load8 = [&](uint32_t j0, double (&a)[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t j = j0 + k;
      a[k] = (row_ok && j < jend && j <= i) ? S[(size_t)j * ld + i] : 0.0;
    }
  };
  double an[8];
  if (c0 < jend) load8(c0, an);
  for (uint32_t j0 = c0; j0 < jend; j0 += 8) {
    double v[8], a8[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a8[k] = an[k];
    if (j0 + 8 < jend) load8(j0 + 8, an);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t j = j0 + k;
      const double a = a8[k];
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

// The same product with an FP32 matrix (FP64 products and sums): two rows
// per lane, so each load moves 8 bytes a lane as in the FP64 kernel, and
// half the lanes per row block (two wavefronts).  Same partial layout as
// k_symv_fused.  ld is even (a float2 at an even row is aligned); row
// D (ld > D) may be read and is discarded.
template <int SYMV_COLS>
__global__ __launch_bounds__(SYMV_ROWS / 2) void k_symv_fused32(uint32_t D, const float* __restrict__ S, size_t ld,
                                                                const double* __restrict__ x,
                                                                double* __restrict__ prow,
                                                                double* __restrict__ pcol,
                                                                const int* __restrict__ stopped) {
  constexpr int NT = SYMV_ROWS / 2, NW = NT / 64;
  __shared__ double xs[SYMV_COLS];
  __shared__ double cs[NW][SYMV_COLS];
  if (stopped && *stopped != 0) return;  // the device-driven iteration has finished
  const uint32_t rb = blockIdx.x, cc = blockIdx.y;
  const uint32_t r0 = rb * SYMV_ROWS, c0 = cc * SYMV_COLS;
  if (c0 > r0 + SYMV_ROWS - 1) return;  // chunk right of the diagonal: nothing
  const uint32_t c1 = min(D, c0 + SYMV_COLS);
  const uint32_t nc = c1 - c0;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (uint32_t j = t; j < nc; j += NT) xs[j] = x[c0 + j];
  for (uint32_t j = t; j < SYMV_COLS; j += NT)
#pragma unroll
    for (int q = 0; q < NW; ++q) cs[q][j] = 0.0;
  __syncthreads();
  const uint32_t i0 = r0 + 2 * t, i1 = i0 + 1;
  const bool ok0 = i0 < D, ok1 = i1 < D;
  const double x0 = ok0 ? x[i0] : 0.0, x1 = ok1 ? x[i1] : 0.0;
  const uint32_t wlast = r0 + 128 * w + 127;
  const uint32_t jend = min(c1, wlast + 1);
  double acc0 = 0.0, acc1 = 0.0;
  auto load8 = [&](uint32_t j0, float2 (&a)[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t j = j0 + k;
      a[k] = (ok0 && j < jend && j <= i1) ? *reinterpret_cast<const float2*>(S + (size_t)j * ld + i0)
                                          : make_float2(0.f, 0.f);
    }
  };
  float2 an[8];
  if (c0 < jend) load8(c0, an);
  for (uint32_t j0 = c0; j0 < jend; j0 += 8) {
    double v[8];
    float2 a8[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a8[k] = an[k];
    if (j0 + 8 < jend) load8(j0 + 8, an);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t j = j0 + k;
      const double a0 = j <= i0 ? (double)a8[k].x : 0.0;
      const double a1 = ok1 ? (double)a8[k].y : 0.0;
      const double xj = j < jend ? xs[j - c0] : 0.0;
      acc0 += a0 * xj;
      acc1 += a1 * xj;
      v[k] = (j < i0 ? a0 * x0 : 0.0) + (j < i1 ? a1 * x1 : 0.0);
    }
    // column sums over the wavefront, as in k_symv_fused
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
  if (ok0) prow[(size_t)cc * D + i0] = acc0;
  if (ok1) prow[(size_t)cc * D + i1] = acc1;
  __syncthreads();
  for (uint32_t j = t; j < nc; j += NT) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) v += cs[q][j];
    pcol[(size_t)rb * D + c0 + j] = v;
  }
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

// r = b (row Dp of S), x = 0
__global__ __launch_bounds__(256) void k_pcg_init(uint32_t n, const double* __restrict__ S, size_t ld,
                                                  uint32_t Dp, double* __restrict__ r, double* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = S[(size_t)i * ld + Dp];
  x[i] = 0.0;
}

// ------------------------------------------------- device-driven iteration
// status: 0 running, 1 converged, 2 breakdown (p^T S p not positive, or
// r.z not finite), 3 residual not finite, 4 zero right-hand side, 5 a
// diagonal block not positive definite.  rz[2]: r.z of the current iteration and of the next (by iteration
// parity), so no kernel reads a value another block of it writes.
struct PcgState {
  double bn, rn;
  double rz[2];
  double rn_rec;  // mixed storage, last replacement: the recurrence's |r| before it
  double gap;     // and |r_true - r_recurrence|
  int status, iterations;
};

__device__ __forceinline__ bool pcg_stopped(const PcgState* st) { return st->status != 0; }

// Block sum in a fixed order (256 threads).
__device__ __forceinline__ double block_sum256(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double t = red[0];
  __syncthreads();
  return t;
}

// Fixed-order sum of n block partials (stride apart), in every block that
// needs it (each gets the same value): strided per thread, then a tree.
__device__ __forceinline__ double sum_partials256(uint32_t n, const double* __restrict__ part, uint32_t stride,
                                                  double* red) {
  double a = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += 256) a += part[(size_t)i * stride];
  return block_sum256(a, red);
}

// After the initial residual: |b| and r.z from the two dot products, the
// block-inverse breakdown flag, and p = z.
__global__ __launch_bounds__(256) void k_pcg_start(uint32_t D, const double* __restrict__ dots,
                                                   const int* __restrict__ flags, const double* __restrict__ z,
                                                   double* __restrict__ p, PcgState* __restrict__ st) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D) p[i] = z[i];
  if (i == 0) {
    st->bn = sqrt(dots[0]);
    st->rn = st->bn;
    st->rz[0] = dots[1];
    st->rz[1] = 0.0;
    st->iterations = 0;
    st->status = flags[0] != 0 ? 5 : (st->bn == 0.0 ? 4 : 0);
  }
}

// q = S p from the row and column partials, and block partials of p.q.
__global__ __launch_bounds__(256) void k_symv_finish_dot(uint32_t D, uint32_t cols, const double* __restrict__ prow,
                                                         const double* __restrict__ pcol,
                                                         const double* __restrict__ p, double* __restrict__ q,
                                                         double* __restrict__ part, const PcgState* __restrict__ st) {
  __shared__ double red[256];
  if (pcg_stopped(st)) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double pq = 0.0;
  if (i < D) {
    const uint32_t ncc = i / cols + 1;
    double s = 0.0;
    for (uint32_t cc = 0; cc < ncc; ++cc) s += prow[(size_t)cc * D + i];
    const uint32_t nrb = (D + SYMV_ROWS - 1) / SYMV_ROWS;
    for (uint32_t rb = i / SYMV_ROWS; rb < nrb; ++rb) s += pcol[(size_t)rb * D + i];
    q[i] = s;
    pq = p[i] * s;
  }
  const double t = block_sum256(pq, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Iteration it, first half: alpha = r.z / p.q (every block sums the p.q
// partials itself, in the same order), or a breakdown; x += alpha p,
// r_out = r_in - alpha q, z = M^{-1} r_out (the block's rows of r_out
// recomputed from r_in and q, so no other thread's write is read), and
// block partials of r.r and r.z into part2.
__global__ __launch_bounds__(256) void k_pcg_update(uint32_t D, int it, uint32_t n_part,
                                                    const double* __restrict__ part_pq,
                                                    const uint32_t* __restrict__ row_grp,
                                                    const uint32_t* __restrict__ grp_off,
                                                    const uint32_t* __restrict__ grp_dim,
                                                    const uint32_t* __restrict__ moff, const double* __restrict__ Minv,
                                                    const double* __restrict__ p, const double* __restrict__ q,
                                                    const double* __restrict__ r_in, double* __restrict__ r_out,
                                                    double* __restrict__ x, double* __restrict__ z,
                                                    double* __restrict__ part2, PcgState* __restrict__ st) {
  __shared__ double red[256];
  if (pcg_stopped(st)) return;
  const double pq = sum_partials256(n_part, part_pq, 1, red);
  const double rz = st->rz[it & 1];
  if (!(pq > 0.0) || !isfinite(rz)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->status = 2;
    return;
  }
  const double alpha = rz / pq;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rzi = 0.0;
  if (i < D) {
    x[i] += alpha * p[i];
    const double ri = r_in[i] - alpha * q[i];
    r_out[i] = ri;
    const uint32_t g = row_grp[i], off = grp_off[g], d = grp_dim[g];
    const double* m = Minv + moff[g];
    double zi = 0.0;
    for (uint32_t c = 0; c < d; ++c) zi += m[c * d + (i - off)] * (r_in[off + c] - alpha * q[off + c]);
    z[i] = zi;
    rr = ri * ri;
    rzi = ri * zi;
  }
  const double t0 = block_sum256(rr, red);
  const double t1 = block_sum256(rzi, red);
  if (threadIdx.x == 0) {
    part2[2 * blockIdx.x] = t0;
    part2[2 * blockIdx.x + 1] = t1;
  }
}

// Iteration it, second half: the convergence test on |r|, then
// beta = r_new.z_new / r.z and p = z + beta p (every block sums the
// partials itself); block 0 records the state.
__global__ __launch_bounds__(256) void k_pcg_next(uint32_t D, int it, uint32_t n_part,
                                                  const double* __restrict__ part2, double tolerance,
                                                  const double* __restrict__ z, double* __restrict__ p,
                                                  PcgState* __restrict__ st) {
  __shared__ double red[256];
  if (pcg_stopped(st)) return;
  const double rr = sum_partials256(n_part, part2, 2, red);
  const double rz_new = sum_partials256(n_part, part2 + 1, 2, red);
  const double rn = sqrt(rr);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    st->iterations = it + 1;
    st->rn = rn;
  }
  if (!isfinite(rn) || rn <= tolerance * st->bn) {
    if (lead) st->status = isfinite(rn) ? 1 : 3;
    return;
  }
  const double beta = rz_new / st->rz[it & 1];
  if (lead) st->rz[(it + 1) & 1] = rz_new;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D) p[i] = z[i] + beta * p[i];
}

// The FP32 copy of the lower triangle (rows >= column, below D) for the
// mixed-storage matvec: columns strided over y, rows over x.
__global__ __launch_bounds__(256) void k_lower_to_fp32(uint32_t D, const double* __restrict__ S, size_t ld,
                                                       float* __restrict__ S32) {
  for (uint32_t c = blockIdx.y; c < D; c += gridDim.y)
    for (uint32_t r = c + blockIdx.x * 256 + threadIdx.x; r < D; r += gridDim.x * 256)
      S32[(size_t)c * ld + r] = (float)S[(size_t)c * ld + r];
}

// Residual replacement of the mixed-storage iteration, before iteration
// st->iterations: r = b - q with q = S x in FP64 (into the r buffer that
// iteration reads, over the recurrence's), z = M^{-1} r, and block partials
// of r.r, r.z and |r - r_recurrence|^2 (stride 3).  Runs while the
// iteration is running or has just converged on its recurrence residual
// (status 0 or 1).
__global__ __launch_bounds__(256) void k_pcg_replace(uint32_t D, const double* __restrict__ S, size_t ld,
                                                     uint32_t Dp, const double* __restrict__ q,
                                                     const uint32_t* __restrict__ row_grp,
                                                     const uint32_t* __restrict__ grp_off,
                                                     const uint32_t* __restrict__ grp_dim,
                                                     const uint32_t* __restrict__ moff,
                                                     const double* __restrict__ Minv, double* __restrict__ r_even,
                                                     double* __restrict__ r_odd, double* __restrict__ z,
                                                     double* __restrict__ part3, const PcgState* __restrict__ st) {
  __shared__ double red[256];
  if (st->status > 1) return;
  double* r = (st->iterations & 1) ? r_odd : r_even;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rz = 0.0, gg = 0.0;
  if (i < D) {
    const double ri = S[(size_t)i * ld + Dp] - q[i];
    gg = (ri - r[i]) * (ri - r[i]);
    r[i] = ri;
    const uint32_t g = row_grp[i], off = grp_off[g], d = grp_dim[g];
    const double* m = Minv + moff[g];
    double zi = 0.0;
    for (uint32_t c = 0; c < d; ++c) zi += m[c * d + (i - off)] * (S[(size_t)(off + c) * ld + Dp] - q[off + c]);
    z[i] = zi;
    rr = ri * ri;
    rz = ri * zi;
  }
  const double t0 = block_sum256(rr, red);
  const double t1 = block_sum256(rz, red);
  const double t2 = block_sum256(gg, red);
  if (threadIdx.x == 0) {
    part3[3 * blockIdx.x] = t0;
    part3[3 * blockIdx.x + 1] = t1;
    part3[3 * blockIdx.x + 2] = t2;
  }
}

// After a replacement: the true residual decides convergence (a
// recurrence-residual convergence is confirmed or undone); the recurrence's
// |r| and its distance from the true residual are kept, and r.z of the next
// iteration is the replaced residual's.
__global__ __launch_bounds__(256) void k_pcg_replace_fin(uint32_t n_part, const double* __restrict__ part3,
                                                         double tolerance, PcgState* __restrict__ st) {
  __shared__ double red[256];
  if (st->status > 1) return;
  const double rr = sum_partials256(n_part, part3, 3, red);
  const double rz = sum_partials256(n_part, part3 + 1, 3, red);
  const double gg = sum_partials256(n_part, part3 + 2, 3, red);
  if (threadIdx.x == 0) {
    const double rn = sqrt(rr);
    st->rn_rec = st->rn;
    st->gap = sqrt(gg);
    st->rn = rn;
    st->rz[st->iterations & 1] = rz;
    if (!isfinite(rn)) st->status = 3;
    else st->status = rn <= tolerance * st->bn ? 1 : 0;
  }
}

// The FP64 matvec S x for a replacement; runs unless the iteration broke
// down (status > 1).
__global__ __launch_bounds__(256) void k_symv_finish_plain(uint32_t D, uint32_t cols, const double* __restrict__ prow,
                                                           const double* __restrict__ pcol, double* __restrict__ y,
                                                           const PcgState* __restrict__ st) {
  if (st->status > 1) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const uint32_t ncc = i / cols + 1;
  double s = 0.0;
  for (uint32_t cc = 0; cc < ncc; ++cc) s += prow[(size_t)cc * D + i];
  const uint32_t nrb = (D + SYMV_ROWS - 1) / SYMV_ROWS;
  for (uint32_t rb = i / SYMV_ROWS; rb < nrb; ++rb) s += pcol[(size_t)rb * D + i];
  y[i] = s;
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t pcg_symv_part_doubles(uint32_t D) { return ((size_t)cdiv(D, symv_cols(D)) + cdiv(D, SYMV_ROWS)) * D; }
size_t pcg_dot_part_doubles(uint32_t D) {
  // the dot products' block partials, and four per 256 unknowns for the
  // iteration's own block partials (p.q, then r.r and r.z, or a
  // replacement's r.r, r.z and gap)
  return std::max<size_t>(2 * (size_t)cdiv(D, DOT_BLOCK * DOT_ITEMS) + 2, 4 * (size_t)cdiv(D, 256));
}

static void dot2(const PcgBuffers& B, uint32_t n, const double* a, const double* b, const double* c,
                 const double* d, hipStream_t s) {
  const uint32_t nb = cdiv(n, DOT_BLOCK * DOT_ITEMS);
  hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(DOT_BLOCK), 0, s, n, a, b, c, d, B.dot_part);
  hipLaunchKernelGGL(k_dot2_final, dim3(1), dim3(1024), 0, s, nb, B.dot_part, B.scal);
}

// Iterations enqueued between two reads of the device state.
constexpr int PCG_BATCH = 16;

static void symv_launch(uint32_t D, int cols, const double* S, size_t ld, const double* x, double* prow,
                        double* pcol, const int* stopped, hipStream_t s) {
  const dim3 grid(cdiv(D, SYMV_ROWS), cdiv(D, cols));
  if (cols == SYMV_COLS_LARGE)
    hipLaunchKernelGGL(k_symv_fused<SYMV_COLS_LARGE>, grid, dim3(SYMV_ROWS), 0, s, D, S, ld, x, prow, pcol, stopped);
  else
    hipLaunchKernelGGL(k_symv_fused<SYMV_COLS_SMALL>, grid, dim3(SYMV_ROWS), 0, s, D, S, ld, x, prow, pcol, stopped);
}

static void symv_launch(uint32_t D, int cols, const float* S, size_t ld, const double* x, double* prow,
                        double* pcol, const int* stopped, hipStream_t s) {
  const dim3 grid(cdiv(D, SYMV_ROWS), cdiv(D, cols));
  if (cols == SYMV_COLS_LARGE)
    hipLaunchKernelGGL(k_symv_fused32<SYMV_COLS_LARGE>, grid, dim3(SYMV_ROWS / 2), 0, s, D, S, ld, x, prow, pcol,
                       stopped);
  else
    hipLaunchKernelGGL(k_symv_fused32<SYMV_COLS_SMALL>, grid, dim3(SYMV_ROWS / 2), 0, s, D, S, ld, x, prow, pcol,
                       stopped);
}

PcgResult pcg_solve(const PcgBuffers& B, const double* S, size_t ld, uint32_t D, uint32_t Dp,
                    uint32_t n_grp, const uint32_t* grp_off, const uint32_t* grp_dim, int max_iterations,
                    double tolerance, double* x, int* flags, double* host_scal, hipStream_t s, bool fp32_matrix) {
  PcgResult res;
  if (D == 0) {
    res.ok = true;
    return res;
  }
  static_assert(sizeof(PcgState) <= 8 * sizeof(double), "PCG state in B.scal + 8");
  PcgState* st = reinterpret_cast<PcgState*>(B.scal + 8);
  const int* stopped = &st->status;
  const uint32_t g256 = cdiv(D, 256);
  double* prow = B.symv_part;
  const int cols = symv_cols(D);
  double* pcol = B.symv_part + (size_t)cdiv(D, cols) * D;
  HIP_CHECK(hipMemsetAsync(flags, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_bj_invert, dim3(cdiv(n_grp, 4)), dim3(256), 0, s, n_grp, grp_off, grp_dim, B.moff, S,
                     ld, B.Minv, flags);
  hipLaunchKernelGGL(k_pcg_init, dim3(g256), dim3(256), 0, s, D, S, ld, Dp, B.r, x);
  hipLaunchKernelGGL(k_bj_apply, dim3(g256), dim3(256), 0, s, D, B.row_grp, grp_off, grp_dim, B.moff, B.Minv,
                     B.r, B.z);
  dot2(B, D, B.r, B.r, B.r, B.z, s);  // |b|^2, r.z
  hipLaunchKernelGGL(k_pcg_start, dim3(g256), dim3(256), 0, s, D, B.scal, flags, B.z, B.p, st);
  if (fp32_matrix)
    hipLaunchKernelGGL(k_lower_to_fp32, dim3(std::min<uint32_t>(cdiv(D, 256), 64), std::min<uint32_t>(D, 16384)),
                       dim3(256), 0, s, D, S, ld, B.S32);
  HIP_CHECK(hipGetLastError());
  PcgState h{};
  auto read_state = [&]() {
    HIP_CHECK(hipMemcpyAsync(host_scal, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(&h, host_scal, sizeof(h));
  };
  // mixed storage: the true residual b - S x, in FP64, replaces the
  // recurrence's before the next iteration and decides convergence
  auto replace = [&]() {
    symv_launch(D, cols, S, ld, x, prow, pcol, nullptr, s);
    hipLaunchKernelGGL(k_symv_finish_plain, dim3(g256), dim3(256), 0, s, D, (uint32_t)cols, prow, pcol, B.q, st);
    hipLaunchKernelGGL(k_pcg_replace, dim3(g256), dim3(256), 0, s, D, S, ld, Dp, B.q, B.row_grp, grp_off, grp_dim,
                       B.moff, B.Minv, B.r, B.r2, B.z, B.dot_part + g256, st);
    hipLaunchKernelGGL(k_pcg_replace_fin, dim3(1), dim3(256), 0, s, g256, B.dot_part + g256, tolerance, st);
    HIP_CHECK(hipGetLastError());
  };
  // the iterations read the FP32 copy, with a replacement after every
  // batch, until its rounding dominates the residual (the replaced residual
  // is more than half its distance from the recurrence's) or the true
  // residual grows past 4x the smallest so far (the rounded matrix is not
  // positive definite on what is left); then the plain FP64 iteration from
  // the best iterate, converging on its recurrence residual like the FP64
  // solver (a true residual can sit above a tight tolerance at FP64's own
  // rounding floor)
  bool use32 = fp32_matrix;
  double best_rn = INFINITY;
  const size_t xbytes = (size_t)D * sizeof(double);
  int it = 0;
  while (it < max_iterations) {
    const int n = std::min(PCG_BATCH, max_iterations - it);
    for (int k = 0; k < n; ++k, ++it) {
      // r alternates between two buffers (the update reads one, writes the other)
      const double* r_in = (it & 1) ? B.r2 : B.r;
      double* r_out = (it & 1) ? B.r : B.r2;
      if (use32) symv_launch(D, cols, B.S32, ld, B.p, prow, pcol, stopped, s);
      else symv_launch(D, cols, S, ld, B.p, prow, pcol, stopped, s);
      hipLaunchKernelGGL(k_symv_finish_dot, dim3(g256), dim3(256), 0, s, D, (uint32_t)cols, prow, pcol, B.p, B.q,
                         B.dot_part, st);
      hipLaunchKernelGGL(k_pcg_update, dim3(g256), dim3(256), 0, s, D, it, g256, B.dot_part, B.row_grp, grp_off,
                         grp_dim, B.moff, B.Minv, B.p, B.q, r_in, r_out, x, B.z, B.dot_part + g256, st);
      hipLaunchKernelGGL(k_pcg_next, dim3(g256), dim3(256), 0, s, D, it, g256, B.dot_part + g256, tolerance, B.z,
                         B.p, st);
    }
    HIP_CHECK(hipGetLastError());
    const bool replaced = use32;
    if (replaced) replace();
    read_state();
    if (!replaced) {
      if (h.status != 0) break;
      continue;
    }
    // the iterations after a recurrence convergence or a breakdown did nothing
    it = h.iterations;
    bool restart = false;
    if (use32 && (h.status == 2 || (h.status == 0 && h.rn > 4.0 * best_rn))) {
      // p.Sp <= 0 with the rounded matrix, or a growing residual: back to
      // the best iterate, on with the FP64 matrix
      use32 = false;
      restart = true;
      if (h.status == 2) HIP_CHECK(hipMemsetAsync(&st->status, 0, sizeof(int), s));
      if (best_rn < INFINITY) HIP_CHECK(hipMemcpyAsync(x, B.x_best, xbytes, hipMemcpyDeviceToDevice, s));
      replace();
      read_state();
    } else if (h.status == 0) {
      if (use32 && h.gap > 0.5 * h.rn) use32 = false;
      // the direction is kept while the replaced residual stays close to
      // the recurrence's; also restarted when the recurrence converged and
      // the true residual did not (its last direction was not updated)
      restart = h.gap > 0.25 * h.rn || h.rn_rec <= tolerance * h.bn;
    }
    if (h.status != 0) break;
    if (use32 && h.rn < best_rn) {
      best_rn = h.rn;
      HIP_CHECK(hipMemcpyAsync(B.x_best, x, xbytes, hipMemcpyDeviceToDevice, s));
    }
    if (restart) HIP_CHECK(hipMemcpyAsync(B.p, B.z, xbytes, hipMemcpyDeviceToDevice, s));
  }
  if (use32 && h.status == 0 && h.rn > best_rn) {
    // the cap, still on the rounded matrix: the best iterate is the result
    HIP_CHECK(hipMemcpyAsync(x, B.x_best, xbytes, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    h.rn = best_rn;
  }
  res.iterations = h.iterations;
  res.rel_residual = h.bn > 0.0 ? h.rn / h.bn : 0.0;
  // converged, zero right-hand side, or the iteration cap (the step so far
  // is still a descent step of the damped model, used as is); a breakdown
  // or a non-finite residual gives no step
  res.ok = h.status == 0 || h.status == 1 || h.status == 4;
  return res;
}

}  // namespace sfmba
