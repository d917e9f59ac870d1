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
//   diagonal block  one workgroup factors L_kk in LDS: 16-column panels on
//                   one wavefront in registers, MFMA trailing updates, then
//                   the inverses X11 of its 16x16 diagonal blocks; a pivot
//                   that is not > 0 (or NaN) sets flags[0] (nalgebra's
//                   failure test: diag <= 0 or NaN);
//   k_trsm_panel    L_ik = A_ik L_kk^{-T} for every row block i > k (and
//                   the rhs row) by blocked substitution with the X11, one
//                   wavefront per 16-row strip, all in MFMA registers;
//   k_syrk_persist  A_ij -= L_ik L_jk^T on the trailing lower triangle,
//                   128x128 tiles on v_mfma_f64_16x16x4_f64, one
//                   eight-wavefront workgroup per CU walking its tiles,
//                   panels staged through LDS by global_load_lds (double
//                   buffered), C loaded while the products run; its first
//                   workgroup updates and factors the next diagonal block;
//                   a last round of tiles that would leave most workgroups
//                   idle is split into column halves.
// While the trailing matrix is large, block columns go in pairs: the update
// of column block k+1 alone, its factorization and panel solve, then one
// update of the rest with both panels (K = 2 NB), which halves the trailing
// matrix's HBM traffic per FLOP.  Near the end, where the diagonal chain
// bounds each step, single steps keep that chain short.
// L_kk^{-1} (for the triangular solves) is formed for all blocks at once
// after the factorization, as panel solves of the identity.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "engine.hpp"

namespace sfmba {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int KC = 16;   // K chunk staged through LDS
constexpr int KS = 144;  // LDS row stride of a staged K chunk (2 KS mod 64 = 32 banks)
constexpr int PW = 16;   // panel width inside a diagonal block
// Block columns go in pairs while at least this many trailing block columns
// remain: the trailing lower triangle is then larger than the 256 MB MALL
// (64 blocks of 128 doubles: 8192^2 / 2 x 8 bytes), so its HBM traffic, not
// the chain of diagonal factorizations, bounds each step.  Measured on
// MI355X: D = 44,994 factors in 571 ms in pairs; D = 7,994 (always below the
// threshold) in 6.7 ms in single steps against 6.9 - 7.4 ms with pairs from
// 40 - 16 trailing block columns on.
constexpr int PAIR_MIN_TRAIL = 64;
// The diagonal block in LDS is stored as its 36 lower 16x16 tiles, tile
// (ti, tj), ti >= tj, at (ti (ti + 1) / 2 + tj) * 256, each column-major
// with stride 16: every MFMA operand or accumulator access (16 rows of 4
// columns of one tile) is then 64 consecutive doubles.  128 doubles after
// the tiles hold 1/L[j][j], then four 16-double broadcast buffers.  75 KB:
// two workgroups fit in a CU's LDS.
constexpr int DIAG_TILES = 36;
constexpr int DIAG_IP = DIAG_TILES * 256;
constexpr int DIAG_COL = DIAG_IP + NB;  // per-wavefront 16-double broadcast buffers
constexpr int SMEM_DIAG = DIAG_COL + 4 * PW;

__device__ __attribute__((always_inline)) inline int tile_index(int ti, int tj) {
  return ti * (ti + 1) / 2 + tj;
}
// LDS offset of element (r, c), r >= c (tile-wise), of the diagonal block.
__device__ __attribute__((always_inline)) inline int doff(int r, int c) {
  return tile_index(r >> 4, c >> 4) * 256 + (c & 15) * 16 + (r & 15);
}

// The v_mfma_f64_16x16x4_f64 operand layout: lane l supplies A[l%16][l/16]
// and B^T[l/16][l%16]; result register v of lane l is D[l/16 + 4v][l%16]
// (checked on gfx950 with a one-wavefront probe of the builtin).
typedef __attribute__((address_space(3))) void lds_void;

// One chunk (KC columns of 128 rows) of the panel A straight from global
// memory into LDS (global_load_lds_dwordx4: lane l moves 16 bytes to dst +
// 16 l), column k of the chunk to offset k KS; wavefront w of eight moves
// columns w and w + 8.
__device__ __attribute__((always_inline)) inline void diag_stage(const double* __restrict__ A, size_t lda, int k0,
                                                                  double* As) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
  for (int k = wave; k < KC; k += 8)
    __builtin_amdgcn_global_load_lds((const void*)(A + (size_t)(k0 + k) * lda + 2 * lane), (lds_void*)(As + k * KS),
                                     16, 0, 0);
}

// The diagonal tile's update, lower triangle only, on eight wavefronts (two
// per SIMD: gfx950 reaches its FP64 MFMA rate only so): its 36 16x16 tiles
// (row tile ti >= column tile tj) in four groups of nine, group p = W / 2
// holding row tiles 7 - p (column tiles 0..7-p) and p (column tiles 0..p),
// the first five of a group on wavefront 2p, the last four on 2p + 1.
// Element v of lane l of tile (ti, tj) is C[ti*16 + l%16][tj*16 + l/16 + 4v]
// (the tile is computed transposed: the column tile's panel rows are the A
// operand).  W is a template argument so that every accumulator index is a
// constant; callers switch on the wavefront index.
template <int W>
struct DiagSplit {
  static constexpr int P = W >> 1, NH = 8 - P;
  static constexpr int FIRST = (W & 1) ? 5 : 0, COUNT = (W & 1) ? 4 : 5;
  static __device__ __attribute__((always_inline)) inline int row_tile(int u) { return FIRST + u < NH ? 7 - P : P; }
  static __device__ __attribute__((always_inline)) inline int col_tile(int u) {
    return FIRST + u < NH ? FIRST + u : FIRST + u - NH;
  }
  // one K step of four: row points at the step's panel row for this lane
  static __device__ __attribute__((always_inline)) inline void step(double4_t (&acc)[5], const double* row) {
#pragma unroll
    for (int u = 0; u < COUNT; ++u)
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-row[col_tile(u) * 16], row[row_tile(u) * 16], acc[u], 0, 0, 0);
  }
  static __device__ __attribute__((always_inline)) inline void load(double4_t (&acc)[5], const double* C, size_t ldc) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < COUNT; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        acc[u][v] = C[(size_t)(col_tile(u) * 16 + (lane >> 4) + 4 * v) * ldc + row_tile(u) * 16 + (lane & 15)];
  }
  // every entry of the lower 16x16 tiles into the tiled LDS block, the
  // identity on padding
  static __device__ __attribute__((always_inline)) inline void to_lds(const double4_t (&acc)[5], double* L,
                                                                      uint32_t base, uint32_t D) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < COUNT; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row_tile(u) * 16 + (lane & 15), c = col_tile(u) * 16 + (lane >> 4) + 4 * v;
        const bool pad = base + r >= D || base + c >= D;
        L[tile_index(row_tile(u), col_tile(u)) * 256 + 64 * v + lane] = pad ? (r == c ? 1.0 : 0.0) : acc[u][v];
      }
  }
};

#define SFMBA_DIAG_SWITCH(w, CALL)       \
  switch (w) {                           \
    case 0: DiagSplit<0>::CALL; break;   \
    case 1: DiagSplit<1>::CALL; break;   \
    case 2: DiagSplit<2>::CALL; break;   \
    case 3: DiagSplit<3>::CALL; break;   \
    case 4: DiagSplit<4>::CALL; break;   \
    case 5: DiagSplit<5>::CALL; break;   \
    case 6: DiagSplit<6>::CALL; break;   \
    default: DiagSplit<7>::CALL; break;  \
  }

// The diagonal workgroup's update of its tile (C -= A A^T over K columns of
// the panel A, eight wavefronts) and the result into the tiled LDS block at
// As, which also holds the two staging buffers until then.
__device__ inline void diag_tile_update(const double* __restrict__ C, size_t ldc, const double* __restrict__ A,
                                        size_t lda, int K, double* As, uint32_t base, uint32_t D) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double4_t acc[5];
  diag_stage(A, lda, 0, As);
  SFMBA_DIAG_SWITCH(w, load(acc, C, ldc))
  __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += KC) {
    if (k0 + KC < K) diag_stage(A, lda, k0 + KC, As + (buf ^ 1) * KC * KS);
    const double* Ac = As + buf * KC * KS;
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const double* row = Ac + (kk + (lane >> 4)) * KS + (lane & 15);
      SFMBA_DIAG_SWITCH(w, step(acc, row))
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    buf ^= 1;
  }
  SFMBA_DIAG_SWITCH(w, to_lds(acc, As, base, D))
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

// Right-looking factorization of the PW-column panel j0 of the LDS block,
// all rows j0 .. NB-1 at once, one row per lane on all four wavefronts:
// lanes 0..15 of every wavefront hold the panel's diagonal tile (each
// wavefront factors its own copy, so the multipliers never cross
// wavefronts), lanes 16..63 of wavefront w hold rows j0 + 16 + 48 w + ...
// The pivot travels by v_readlane, the column of multipliers through the
// wavefront's 16-double LDS buffer (written by the owning lanes, read as a
// broadcast: LDS operations of one wavefront complete in order), and each
// pivot costs one reciprocal square root; 1/L[j][j] is kept after the
// tiles.  A pivot that is not > 0 (or NaN) on a real row (base + j < D)
// sets flags[0], nalgebra's Cholesky::new failure test.
__device__ __attribute__((always_inline)) inline void panel_factor(double* A, int j0, uint32_t base,
                                                                    uint32_t D, int* flags) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane < PW ? j0 + lane : j0 + PW + 48 * wave + (lane - PW);
  const bool has = r < NB;
  const bool writer = has && (lane >= PW || wave == 0);
  // element (r, j0 + c) of the panel sits at the row's base + 16 c
  const int o = has ? doff(r, j0) : 0;
  double* colbuf = A + DIAG_COL + PW * wave;
  double a[PW];
#pragma unroll
  for (int c = 0; c < PW; ++c) a[c] = has ? A[o + 16 * c] : 0.0;
  double my_d = 0.0, my_ip = 0.0;  // lane j < PW: pivot j and 1/L[j][j]
  double d = readlane_f64(a[0], 0);
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    const double ip = rsqrt_f64(d);
    if (lane == j) {
      my_d = d;
      my_ip = ip;
    }
    a[j] *= ip;  // lane j holds the pivot itself: d * ip = L[j][j]
    if (j + 1 < PW) {
      // The next pivot, a[j+1] - L[j+1][j]^2 on lane j+1, needs no other
      // lane; the multiplier L[j+1][j] reaches the other lanes by
      // v_readlane and the rest of the column through the LDS buffer.
      d = readlane_f64(a[j + 1] - a[j] * a[j], j + 1);
      const double l1 = readlane_f64(a[j], j + 1);
      a[j + 1] -= a[j] * l1;
      if (j + 2 < PW) {
        // compiler barriers keep each step's buffer accesses in program order
        __asm__ volatile("" ::: "memory");
        if (lane < PW) colbuf[lane] = a[j];
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int c = j + 2; c < PW; ++c) a[c] -= a[j] * colbuf[c];
        __asm__ volatile("" ::: "memory");
      }
    }
  }
  if (wave == 0 && lane < PW) {
    if (!(my_d > 0.0) && base + j0 + lane < D) atomicOr(&flags[0], 1);
    A[DIAG_IP + j0 + lane] = my_ip;
  }
  if (writer) {
#pragma unroll
    for (int c = 0; c < PW; ++c) A[o + 16 * c] = a[c];
  }
}

// Rank-PW update of the trailing lower triangle after panel p, one 16x16
// tile per wave at a time on v_mfma_f64_16x16x4f64, computed transposed
// (D'[n][i] = C[r0+i][c0+n]) so that operands and accumulators are read and
// written as 64 consecutive doubles of a tile.
template <int NW>
__device__ __attribute__((always_inline)) inline void lds_trailing_update(double* A, int p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int T = NB / PW - 1 - p;
  const int ntiles = T * (T + 1) / 2;
  // two tiles per wave at a time, their MFMA chains interleaved
  for (int e0 = wave; e0 < ntiles; e0 += 2 * NW) {
    const int e1 = e0 + NW;
    const bool two = e1 < ntiles;
    double* C[2];
    const double* Li[2];
    const double* Lj[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = (u == 0 || !two) ? e0 : e1;
      int ti = 0;
      while ((ti + 1) * (ti + 2) / 2 <= e) ++ti;
      const int tj = e - ti * (ti + 1) / 2;
      const int I = p + 1 + ti, J = p + 1 + tj;
      C[u] = A + tile_index(I, J) * 256 + lane;
      Li[u] = A + tile_index(I, p) * 256 + lane;
      Lj[u] = A + tile_index(J, p) * 256 + lane;
    }
    double4_t acc[2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = C[u][64 * v];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-Lj[u][64 * s], Li[u][64 * s], acc[u], 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) C[0][64 * v] = acc[0][v];
    if (two) {
#pragma unroll
      for (int v = 0; v < 4; ++v) C[1][64 * v] = acc[1][v];
    }
  }
}

// Factor the diagonal block k held in LDS (tiled, lower; padding rows and
// columns hold the identity), store L (lower, real rows) into S, and write
// the panel solve's operands for this block to Top (TOP_BLOCK doubles):
//   block pq = p(p-1)/2 + q, q < p:  -L tile (p, q) (lane order = tile order:
//                                    operand s, lane l -> -L[16p + l%16][16q + 4s + l/16])
//   block 28 + p:                    X11_p = L_pp^{-1}, column-major 16x16
//                                    (operand s, lane l -> X11_p[l%16][4s + l/16])
template <int NW>
__device__ __attribute__((always_inline)) inline void diag_factor(double* A, double* __restrict__ S,
                                                                   size_t ld, int k, uint32_t D,
                                                                   double* __restrict__ Top,
                                                                   int* __restrict__ flags) {
  const int t = threadIdx.x;
  const uint32_t base = (uint32_t)k * NB;
  double* T = Top + (size_t)k * TOP_BLOCK;
  __syncthreads();
  double* G = S + (size_t)k * NB * ld + (size_t)k * NB;
#pragma unroll 1
  for (int p = 0; p < NB / PW; ++p) {
    if (t < 256) panel_factor(A, PW * p, base, D, flags);
    __syncthreads();
    if (p + 1 < NB / PW) {
      lds_trailing_update<NW>(A, p);
      __syncthreads();
    }
  }
  // Threads 0..127: X11_p by forward substitution, one thread per (p,
  // column c), straight to Top.  The other threads meanwhile store L (lower
  // triangle, real rows) into S and the -L tiles into Top, 16 bytes at a
  // time for the 28 off-diagonal tiles.
  if (t >= 128) {
    constexpr int NT = NW * 64 - 128;
    const int u = t - 128;
#pragma unroll 4
    for (int idx = u; idx < 28 * 128; idx += NT) {
      // Top tile it = ti (ti - 1) / 2 + tj, ti > tj
      const int it = idx >> 7;
      int ti = 1;
      while ((ti + 1) * ti / 2 <= it) ++ti;
      const int tj = it - ti * (ti - 1) / 2;
      const int e = 2 * (idx & 127), rr = e & 15, cc = e >> 4;
      const double2 v = *reinterpret_cast<const double2*>(A + tile_index(ti, tj) * 256 + e);
      *reinterpret_cast<double2*>(T + it * 256 + e) = make_double2(-v.x, -v.y);
      const uint32_t r = base + PW * ti + rr;
      double* g = G + (size_t)(PW * tj + cc) * ld + PW * ti + rr;
      if (r + 1 < D) {
        *reinterpret_cast<double2*>(g) = v;
      } else if (r < D) {
        g[0] = v.x;
      }
    }
#pragma unroll 2
    for (int idx = u; idx < 8 * 256; idx += NT) {
      const int ti = idx >> 8, e = idx & 255, rr = e & 15, cc = e >> 4;
      if (rr < cc || base + PW * ti + rr >= D) continue;
      G[(size_t)(PW * ti + cc) * ld + PW * ti + rr] = A[tile_index(ti, ti) * 256 + e];
    }
  } else {
    const int p = t >> 4, c = t & 15;
    const double* Lpp = A + tile_index(p, p) * 256;
    const double* ip = A + DIAG_IP + PW * p;
    double x[PW];
#pragma unroll
    for (int i = 0; i < PW; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int m = 0; m < i; ++m) s -= Lpp[m * 16 + i] * x[m];
      x[i] = s * ip[i];
      __asm__ volatile("" ::: "memory");  // bounds the loads kept in flight
    }
    double* X = T + (28 + p) * 256 + c * 16;
#pragma unroll
    for (int i = 0; i < PW; ++i) X[i] = x[i];
  }
}

__global__ __launch_bounds__(256) void k_potrf_block(double* __restrict__ S, size_t ld, int k,
                                                     uint32_t D, double* __restrict__ Top,
                                                     int* __restrict__ flags) {
  __shared__ double smem[SMEM_DIAG];
  const uint32_t base = (uint32_t)k * NB;
  const double* G = S + (size_t)k * NB * ld + (size_t)k * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int c = idx >> 7, r = idx & 127;
    if ((r >> 4) < (c >> 4)) continue;
    const bool pad = base + r >= D || base + c >= D;
    smem[doff(r, c)] = pad ? (r == c ? 1.0 : 0.0) : (r >= c ? G[(size_t)c * ld + r] : 0.0);
  }
  diag_factor<4>(smem, S, ld, k, D, Top, flags);
}

// One 16-row strip of the panel solve X = A L_kk^{-T}, one wavefront,
// entirely in registers: for each 16-column sub-panel p,
//   R_p = A_p - sum_{q<p} X_q L_pq^T,   X_p = R_p X11_p^T,
// computed transposed on v_mfma_f64_16x16x4f64, whose accumulator layout
// for X_p^T is exactly the B-operand layout X_p^T needs in the later
// sub-panels (lane l, register v: X_p[l%16][l/16 + 4v]).  The L_pq and
// X11_p operands come from Top in lane order (one coalesced 512-byte load
// per operand): op(j) is this lane's value of operand j = 4 block + s of
// block k's Top.  Lane l's row of the strip is row l % 16: load(c) gives
// its A entry in column c, store(c, x) takes its X entry.
template <typename Load, typename Op, typename Store>
__device__ __attribute__((always_inline)) inline void panel_strip(Load load, Op op, Store store) {
  const int lk = (threadIdx.x & 63) >> 4;
  double xb[8][4];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    double4_t r;
#pragma unroll
    for (int v = 0; v < 4; ++v) r[v] = load(PW * p + lk + 4 * v);
#pragma unroll
    for (int q = 0; q < p; ++q)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        r = __builtin_amdgcn_mfma_f64_16x16x4f64(op((p * (p - 1) / 2 + q) * 4 + s), xb[q][s], r, 0, 0, 0);
    double4_t x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
      x = __builtin_amdgcn_mfma_f64_16x16x4f64(op((28 + p) * 4 + s), r[s], x, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      xb[p][v] = x[v];
      store(PW * p + lk + 4 * v, x[v]);
    }
  }
}

// L_ik = A_ik L_kk^{-T} in place for row blocks i = k+1 .. nblk-1 and the
// rhs row (block nblk, first 16 rows), one wavefront per strip.
__global__ __launch_bounds__(64) void k_trsm_panel(double* __restrict__ S, size_t ld, int k,
                                                   const double* __restrict__ Top, uint32_t D,
                                                   uint32_t Dp) {
  const int lane = threadIdx.x;
  const int i = k + 1 + (int)(blockIdx.x >> 3), rt = blockIdx.x & 7;
  const uint32_t row0 = (uint32_t)i * NB + 16 * rt;
  const uint32_t grow = row0 + (lane & 15);
  const bool row_ok = grow < D || grow == Dp;
  double* Acol = S + (size_t)k * NB * ld + grow;
  const uint32_t col0 = (uint32_t)k * NB;
  // every operand is loaded up front: the chain then waits on each in turn
  // as it arrives instead of on eight rounds of loads issued as it goes
  const double* T = Top + (size_t)k * TOP_BLOCK + lane;
  double t[TOP_BLOCK / 64];
#pragma unroll
  for (int j = 0; j < (int)(TOP_BLOCK / 64); ++j) t[j] = T[j * 64];
  panel_strip([&](int c) { return Acol[(size_t)c * ld]; }, [&](int j) { return t[j]; },
              [&](int c, double x) {
                if (row_ok && col0 + (uint32_t)c < D) Acol[(size_t)c * ld] = x;
              });
}

// Tile t of a trailing update, in column-major order of its lower
// triangle: column c = 0 .. (block column j0 + c), rows r = c .. nrow - 1
// (block row j0 + r; the last row is the rhs block).  Uniform per
// workgroup, at most one iteration per block column.
__device__ __attribute__((always_inline)) inline void trail_tile(int t, int nrow, int& r, int& c) {
  c = 0;
  while (t >= nrow - c) {
    t -= nrow - c;
    ++c;
  }
  r = c + t;
}

// The trailing update with one workgroup of eight wavefronts per CU
// (persistent form).  gfx950 issues FP64 MFMAs at its full rate only from
// two wavefronts per SIMD (one wavefront alone reaches about half), so the
// 128x128 tile is split over eight wavefronts of 64x32 (wr = w / 4 picks
// 64 columns, wc = w % 4 picks 32 rows).  The tile is computed transposed,
// D = A B^T with A the column panel and B the row panel, so that every
// global access to C covers 16 consecutive rows (128 bytes): lane l,
// accumulator (m, n), register v holds C[r][c] with r = 32 wc + 16 n +
// l % 16, c = 64 wr + 16 m + l / 16 + 4 v.  Chunks of PK panel columns are
// staged double buffered; the products are summed from zero while the
// tile's C arrives in registers during the first four steps, and the sum
// is subtracted from C at the end.
struct Persist {
  static constexpr int PK = 32;            // panel columns per step
  static constexpr int BUF = 2 * PK * KS;  // one staged chunk (A then B)
  static constexpr int SMEM = 2 * BUF > SMEM_DIAG ? 2 * BUF : SMEM_DIAG;

  static __device__ __attribute__((always_inline)) inline int row(int n, int lane) {
    return ((threadIdx.x >> 6) & 3) * 32 + n * 16 + (lane & 15);
  }
  // NM blocks of 16 columns per wavefront from column cb: NM = 4 covers the
  // whole tile (cb = 0), NM = 2 one half of it (cb = 0 or 64), NM = 1 one
  // quarter (cb = 0, 32, 64 or 96).
  template <int NM>
  static __device__ __attribute__((always_inline)) inline int col(int m, int v, int lane, int cb) {
    return cb + (threadIdx.x >> 8) * (16 * NM) + m * 16 + (lane >> 4) + 4 * v;
  }

  // Chunk k0 of panels A (column) and B (row) into LDS: row k of the chunk
  // at offset k KS; the wavefront index made uniform so that the row
  // addresses are scalar.
  static __device__ __attribute__((always_inline)) inline void stage(const double* __restrict__ A,
                                                                     const double* __restrict__ B, size_t ld, int k0,
                                                                     double* As) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
    for (int k = wave; k < PK; k += 8) {
      __builtin_amdgcn_global_load_lds((const void*)(A + (size_t)(k0 + k) * ld + 2 * lane),
                                       (lds_void*)(As + k * KS), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(B + (size_t)(k0 + k) * ld + 2 * lane),
                                       (lds_void*)(As + PK * KS + k * KS), 16, 0, 0);
    }
  }

  // acc += A B^T over the staged chunk; the operands of the next four
  // columns are read from LDS before the products of the current four.
  template <int NM>
  static __device__ __attribute__((always_inline)) inline void mma(double4_t (&acc)[NM][2], const double* Ac, int cb) {
    const int lane = threadIdx.x & 63;
    const int wr = threadIdx.x >> 8, wc = (threadIdx.x >> 6) & 3;
    const double* pa = Ac + (lane >> 4) * KS + cb + wr * (16 * NM) + (lane & 15);
    const double* pb = Ac + PK * KS + (lane >> 4) * KS + wc * 32 + (lane & 15);
    double a[2][NM], b[2][2];
#pragma unroll
    for (int m = 0; m < NM; ++m) a[0][m] = pa[m * 16];
#pragma unroll
    for (int n = 0; n < 2; ++n) b[0][n] = pb[n * 16];
#pragma unroll
    for (int s = 0; s < PK / 4; ++s) {
      const int cur = s & 1;
      if (s + 1 < PK / 4) {
#pragma unroll
        for (int m = 0; m < NM; ++m) a[cur ^ 1][m] = pa[(s + 1) * 4 * KS + m * 16];
#pragma unroll
        for (int n = 0; n < 2; ++n) b[cur ^ 1][n] = pb[(s + 1) * 4 * KS + n * 16];
      }
#pragma unroll
      for (int m = 0; m < NM; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[cur][m], b[cur][n], acc[m][n], 0, 0, 0);
    }
  }

  // C group G (8 of the 8 NM values per lane: m = G) into the prefetch
  // registers.
  template <int NM, int G>
  static __device__ __attribute__((always_inline)) inline void c_group(double (&cr)[NM][2][4], const double* C,
                                                                      size_t ld, int cb) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int n = 0; n < 2; ++n) cr[G][n][v] = C[(size_t)col<NM>(G, v, lane, cb) * ld + row(n, lane)];
  }

  // One step: stage the next chunk (A = An, B = Bn; none if An is null),
  // load C group G (G >= 0), run the products of the current buffer, wait
  // for the staged chunk (a C group issued after it may stay in flight:
  // loads of a wavefront complete in issue order), meet the other
  // wavefronts.  Each wavefront's LDS reads of the current buffer
  // are complete at the barrier, so the next step may stage into it.
  // (s_barrier alone: the fence of __syncthreads() would wait for the C
  // group too.)
  template <int NM, int G>
  static __device__ __attribute__((always_inline)) inline void step(double4_t (&acc)[NM][2], double (&cr)[NM][2][4],
                                                                    const double* C, size_t ld, const double* An,
                                                                    const double* Bn, int k_next, double* smem,
                                                                    int& buf, int cb) {
    if (An) stage(An, Bn, ld, k_next, smem + (buf ^ 1) * BUF);
    if constexpr (G >= 0) {
      __asm__ volatile("" ::: "memory");
      c_group<NM, G>(cr, C, ld, cb);
      __asm__ volatile("" ::: "memory");
    }
    mma<NM>(acc, smem + buf * BUF, cb);
    if constexpr (G >= 0) {
      static_assert(G < NM, "NM C groups of eight loads");
      __asm__ volatile("s_waitcnt vmcnt(8) lgkmcnt(0)" ::: "memory");
    } else {
      __asm__ volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    }
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    buf ^= 1;
  }

  // Stores only entries that carry data: column < D, row < D or == Dp (the
  // rhs), the lower triangle on a diagonal tile.  Padding and the spare rows
  // of the rhs block keep their initial values, so a failed (NaN)
  // factorization cannot leak into storage the next assembly does not
  // rewrite (NaN * 0 is NaN).
  template <int NM>
  static __device__ __attribute__((always_inline)) inline void store(const double4_t (&acc)[NM][2], double* C,
                                                                     size_t ldc, bool lower_only, uint32_t row0,
                                                                     uint32_t col0, uint32_t D, uint32_t Dp, int cb) {
    const int lane = threadIdx.x & 63;
    const bool full = !lower_only && row0 + NB <= D && col0 + NB <= D;
#pragma unroll
    for (int m = 0; m < NM; ++m)
#pragma unroll
      for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = row(n, lane), c = col<NM>(m, v, lane, cb);
          const uint32_t g = row0 + r;
          if (full || ((!lower_only || r >= c) && (g < D || g == Dp) && col0 + c < D))
            C[(size_t)c * ldc + r] = acc[m][n][v];
        }
  }
};

// One tile of a trailing update (NM = 4) or one column part of it (NM = 2:
// a half, columns cb .. cb + 63; NM = 1: a quarter, cb .. cb + 31): C - A_i A_j^T over the K / PK chunks, the first
// of which is staged in buffer buf; the next work item's first chunk (An =
// Ajn, Bn = Ain; none if Ajn is null) is staged during the last step.
template <int NM>
__device__ __attribute__((always_inline)) inline void persist_tile(double* C, size_t ld, const double* Aj,
                                                                    const double* Ai, const double* Ajn,
                                                                    const double* Ain, int nch, double* smem,
                                                                    int& buf, bool lower_only, uint32_t row0,
                                                                    uint32_t col0, uint32_t D, uint32_t Dp, int cb) {
  constexpr int PK = Persist::PK;
  double4_t acc[NM][2];
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int n = 0; n < 2; ++n) acc[m][n] = double4_t{0.0, 0.0, 0.0, 0.0};
  double cr[NM][2][4];
  // C arrives in NM groups during the first NM steps (nch >= 4)
  Persist::step<NM, 0>(acc, cr, C, ld, Aj, Ai, 1 * PK, smem, buf, cb);
  Persist::step<NM, (1 < NM ? 1 : -1)>(acc, cr, C, ld, Aj, Ai, 2 * PK, smem, buf, cb);
  Persist::step<NM, (2 < NM ? 2 : -1)>(acc, cr, C, ld, Aj, Ai, 3 * PK, smem, buf, cb);
  Persist::step<NM, (3 < NM ? 3 : -1)>(acc, cr, C, ld, 4 < nch ? Aj : Ajn, 4 < nch ? Ai : Ain,
                                       4 < nch ? 4 * PK : 0, smem, buf, cb);
#pragma unroll 1
  for (int ch = 4; ch < nch; ++ch) {
    const bool inner = ch + 1 < nch;
    Persist::step<NM, -1>(acc, cr, C, ld, inner ? Aj : Ajn, inner ? Ai : Ain, inner ? (ch + 1) * PK : 0, smem, buf,
                          cb);
  }
  // C - sum, all before the first store: the stores then issue back to
  // back (a difference left next to its predicated store would wait for
  // every earlier store too).
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int n = 0; n < 2; ++n) {
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[m][n][v] = cr[m][n][v] - acc[m][n][v];
      __asm__ volatile("" : "+v"(acc[m][n]));
    }
  Persist::store<NM>(acc, C, ld, lower_only, row0, col0, D, Dp, cb);
}

// A_ij -= sum over the panels kp .. kp + K/NB - 1 of L_ip L_jp^T for
// j0 <= j <= i, j < nblk, i <= nblk (i == nblk: the rhs block); ncol == 1
// updates block column j0 only.  One workgroup per CU walks its tiles in
// turn; the next tile's first chunk is staged during the last step of the
// current one, and the stores of one tile drain during the next one's
// first step.  Workgroup 0 updates the diagonal tile (j0, j0) with all
// eight wavefronts and factors it with the first four (lookahead: the next
// panel solve can start as soon as this launch ends); the other four end
// there (s_barrier waits for the wavefronts still running).  K is a
// multiple of 128, so a tile takes at least four steps of PK = 32 columns.
// Measured at D = 7,994 on MI355X: 6.4 ms per factorization against 7.6 ms
// with two four-wavefront workgroups per CU, one tile each.
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_syrk_persist(
    double* __restrict__ S, size_t ld, int kp, int K, int j0, int nblk, int ncol, uint32_t D, uint32_t Dp,
    double* __restrict__ Top, int* __restrict__ flags) {
  __shared__ double smem[Persist::SMEM];
  const double* P = S + (size_t)kp * NB * ld;
  if (blockIdx.x == 0) {
    const double* Aj = P + (size_t)j0 * NB;
    diag_tile_update(S + (size_t)j0 * NB * ld + (size_t)j0 * NB, ld, Aj, ld, K, smem, (uint32_t)j0 * NB, D);
    diag_factor<8>(smem, S, ld, j0, D, Top, flags);
    return;
  }
  // Workgroup b >= 1 takes tiles b, b + G1, ... of tiles 1 .. ntiles - 1
  // (G1 = gridDim.x - 1).  When the last round would leave more than half
  // (three quarters) of the workgroups idle, its tiles go as column halves
  // (quarters) to two (four) workgroups each (same products, same order:
  // the same results).
  const int nrow = nblk - j0 + 1;
  const int ntiles = ncol * nrow - ncol * (ncol - 1) / 2;
  const int G1 = (int)gridDim.x - 1, b = (int)blockIdx.x;
  const int rounds = (ntiles - 1) / G1, rem = ntiles - 1 - rounds * G1;
  const int parts = rem == 0 ? 1 : 4 * rem <= G1 ? 4 : 2 * rem <= G1 ? 2 : 1;
  const int nfull = rounds + (parts == 1 && b <= rem ? 1 : 0);
  const int nitem = nfull + (parts > 1 && b <= parts * rem ? 1 : 0);
  if (nitem == 0) return;
  // work item q: block row and column of its tile, and the column part
  // (-1: the whole tile)
  auto item = [&](int q, int& ti, int& tj, int& half) {
    int t;
    if (q < nfull) {
      t = b + q * G1;
      half = -1;
    } else {
      t = rounds * G1 + 1 + (b - 1) / parts;
      half = (b - 1) % parts;
    }
    int r, c;
    trail_tile(t, nrow, r, c);
    ti = j0 + r;
    tj = j0 + c;
  };
  int i, j, half;
  item(0, i, j, half);
  const double* Ai = P + (size_t)i * NB;
  const double* Aj = P + (size_t)j * NB;
  Persist::stage(Aj, Ai, ld, 0, smem);
  __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  int buf = 0;
  const int nch = K / Persist::PK;
  for (int q = 0; q < nitem; ++q) {
    const bool more = q + 1 < nitem;
    int in = 0, jn = 0, halfn = -1;
    if (more) item(q + 1, in, jn, halfn);
    const double* Ain = P + (size_t)in * NB;
    const double* Ajn = more ? P + (size_t)jn * NB : nullptr;
    double* C = S + (size_t)j * NB * ld + (size_t)i * NB;
    if (half < 0)
      persist_tile<4>(C, ld, Aj, Ai, Ajn, Ain, nch, smem, buf, i == j, (uint32_t)i * NB, (uint32_t)j * NB, D, Dp, 0);
    else if (parts == 2)
      persist_tile<2>(C, ld, Aj, Ai, Ajn, Ain, nch, smem, buf, i == j, (uint32_t)i * NB, (uint32_t)j * NB, D, Dp,
                      64 * half);
    else
      persist_tile<1>(C, ld, Aj, Ai, Ajn, Ain, nch, smem, buf, i == j, (uint32_t)i * NB, (uint32_t)j * NB, D, Dp,
                      32 * half);
    i = in;
    j = jn;
    half = halfn;
    Ai = Ain;
    Aj = Ajn;
  }
}

// L^{-1} of every diagonal block, all blocks at once after the
// factorization (only the triangular solves use it), as the panel solve of
// the identity: strip s of I L_kk^{-T} is rows 16 s .. 16 s + 15 of
// L_kk^{-T}, that is columns of L_kk^{-1}.  One workgroup per block, one
// strip per wavefront, the block's operands staged in LDS first (read from
// memory by each step of the chains, 63 blocks took 59 us).  Linv holds,
// per block, L^{-1} (column-major, identity on padding) followed by its
// transpose.
__global__ __launch_bounds__(512) void k_block_inverse(uint32_t D, const double* __restrict__ Top,
                                                       double* __restrict__ Linv) {
  __shared__ __attribute__((aligned(16))) double Tl[TOP_BLOCK];
  const int lane = threadIdx.x & 63, strip = threadIdx.x >> 6;
  const int k = blockIdx.x;
  {
    const double2* src = reinterpret_cast<const double2*>(Top + (size_t)k * TOP_BLOCK);
    double2* dst = reinterpret_cast<double2*>(Tl);
    static_assert(TOP_BLOCK % (2 * 512) == 0, "whole rounds of 512 threads");
#pragma unroll
    for (int it = 0; it < (int)TOP_BLOCK / (2 * 512); ++it) dst[it * 512 + threadIdx.x] = src[it * 512 + threadIdx.x];
  }
  __syncthreads();
  const int rho = 16 * strip + (lane & 15);  // row of L^{-T}, column of L^{-1}
  const bool pad_rho = (uint32_t)k * NB + rho >= D;
  double* X = Linv + (size_t)k * LINV_BLOCK;
  const double* T = Tl + lane;
  panel_strip([&](int c) { return c == rho ? 1.0 : 0.0; }, [&](int j) { return T[j * 64]; },
              [&](int c, double x) {
                if (pad_rho || (uint32_t)k * NB + c >= D) x = c == rho ? 1.0 : 0.0;
                X[(size_t)rho * NB + c] = x;                    // L^{-1}[c][rho]
                X[(size_t)NB * NB + (size_t)c * NB + rho] = x;  // the transpose
              });
}

// ------------------------------------------------------ triangular solves
// Back substitution x = L^-T y in one launch.  Workgroup w owns block
// j = nblk - 1 - w: it applies every solved block x_k (k > j) to its part of
// y as soon as x_k is published, y_j -= L_kj^T x_k, then solves
// x_j = L_jj^-T y_j and publishes it (x_j stored, then chain[j] = epoch with
// release semantics at device scope).  A workgroup waits only on blocks
// owned by workgroups dispatched before it, so the waits cannot deadlock;
// each wait is bounded (about a quarter of a second), and on timeout
// status is set and the workgroup stops waiting for the rest of the launch
// (the host reports the failure), so a broken hand-over costs one time-out.
//
// The critical path per block is one flag hand-over, one multiply of an
// L block already held in registers (it is loaded while the workgroup
// waits), a 32-column reduction and one solve against L_jj^-T held in LDS.
constexpr uint32_t CHAIN_SPIN_LIMIT = 1u << 18;
constexpr int CHAIN_COLS = NB / 4;  // columns of L_kj per wavefront

// Lane l of wavefront w holds rows l and l + 64 of columns w*32 .. w*32+31
// of block (row k, column j); each load is 512 contiguous bytes.
__device__ __forceinline__ void chain_load(const double* __restrict__ S, size_t ld, int j, int k, int wave,
                                           int lane, double (&a)[CHAIN_COLS], double (&b)[CHAIN_COLS]) {
  const double* base = S + (size_t)(j * NB + wave * CHAIN_COLS) * ld + (size_t)k * NB + lane;
#pragma unroll
  for (int c = 0; c < CHAIN_COLS; ++c) {
    a[c] = base[(size_t)c * ld];
    b[c] = base[(size_t)c * ld + 64];
  }
}

// One step of a recursive-halving reduction of H*2 values per lane across
// lane pairs (l, l ^ O): lanes with bit O clear keep the lower H values,
// lanes with it set the upper H, each summed over the pair.
template <int O, int H>
__device__ __forceinline__ void halve(double (&p)[CHAIN_COLS], int lane) {
  const bool up = (lane & O) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double send = up ? p[i] : p[i + H];
    const double keep = up ? p[i + H] : p[i];
    p[i] = keep + __shfl_xor(send, O);
  }
}

__global__ __launch_bounds__(256) void k_backsolve_chain(const double* __restrict__ S, size_t ld,
                                                         const double* __restrict__ Linv,
                                                         const double* __restrict__ y, double* __restrict__ x,
                                                         uint32_t* __restrict__ chain, uint32_t epoch, int nblk,
                                                         int* __restrict__ status) {
  __shared__ double xt[NB * NB];  // L_jj^-T, column-major
  __shared__ double ys[NB];
  __shared__ double xs[NB];
  __shared__ double part[2 * NB];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int j = nblk - 1 - (int)blockIdx.x;
  double a[CHAIN_COLS], b[CHAIN_COLS];
  if (j + 1 < nblk) chain_load(S, ld, j, nblk - 1, wave, lane, a, b);
  {
    const double2* src = reinterpret_cast<const double2*>(Linv + (size_t)j * LINV_BLOCK + NB * NB);
    double2* dst = reinterpret_cast<double2*>(xt);
#pragma unroll 8
    for (int e = t; e < NB * NB / 2; e += 256) dst[e] = src[e];
  }
  if (t < NB) ys[t] = y[(size_t)j * NB + t];
  bool timed_out = false;  // thread 0 only
  for (int k = nblk - 1; k > j; --k) {
    if (t == 0 && !timed_out) {
      uint32_t spins = 0;
      while (__hip_atomic_load(&chain[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > CHAIN_SPIN_LIMIT) {
          __hip_atomic_fetch_or(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          timed_out = true;
          break;
        }
      }
    }
    __syncthreads();
    if (t < NB) xs[t] = __hip_atomic_load(&x[(size_t)k * NB + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const double x0 = xs[lane], x1 = xs[lane + 64];
    double p[CHAIN_COLS];
#pragma unroll
    for (int c = 0; c < CHAIN_COLS; ++c) p[c] = a[c] * x0 + b[c] * x1;
    if (k - 1 > j) chain_load(S, ld, j, k - 1, wave, lane, a, b);
    // column (lane >> 1) & 31 of this wavefront's 32 ends in lanes 2m, 2m+1
    halve<32, 16>(p, lane);
    halve<16, 8>(p, lane);
    halve<8, 4>(p, lane);
    halve<4, 2>(p, lane);
    halve<2, 1>(p, lane);
    const double v = p[0] + __shfl_xor(p[0], 1);
    if ((lane & 1) == 0) ys[wave * CHAIN_COLS + (lane >> 1)] -= v;
  }
  __syncthreads();
  // x[c] = sum_r L_jj^-1[r][c] y[r] = sum_r xt[r*NB + c] y[r]; threads
  // t and t + 128 take the two halves of r.
  {
    const int c = t & (NB - 1), r0 = (t >> 7) * 64;
    double acc = 0.0;
#pragma unroll 16
    for (int r = r0; r < r0 + 64; ++r) acc += xt[r * NB + c] * ys[r];
    part[t] = acc;
  }
  __syncthreads();
  if (t < NB) {
    __hip_atomic_store(&x[(size_t)j * NB + t], part[t] + part[t + NB], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    // every storing wavefront completes its stores at device scope before
    // the barrier that precedes the flag
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  }
  __syncthreads();
  if (t == 0) __hip_atomic_store(&chain[j], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
  const double* X = Linv + (size_t)k * LINV_BLOCK;
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
// Compute units of the current device (one persistent workgroup each).
static int compute_units() {
  int dev = 0, n = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(2, n);
}

void dense_cholesky(double* S, double* Linv, uint32_t Dp, uint32_t ld, uint32_t D, int* flags,
                    hipStream_t s, KernelTimer* kt) {
  const int nblk = (int)(Dp / NB);
  double* Top = Linv + (size_t)nblk * LINV_BLOCK;
  const int n_cu = compute_units();
  auto trsm = [&](int k) {
    // 8 strips per row block k+1 .. nblk-1, one for the rhs row
    const int strips = 8 * (nblk - 1 - k) + 1;
    hipLaunchKernelGGL(k_trsm_panel, dim3(strips), dim3(64), 0, s, S, (size_t)ld, k, Top, D, Dp);
  };
  // Trailing update with panels kp .. kp + np - 1 of block columns j0 ..
  // (all of them, or j0 only); factors diagonal block j0.
  auto syrk = [&](int kp, int np, int j0, bool column_only) {
    const int ncol = column_only ? 1 : nblk - j0;
    hipEvent_t e0 = kt ? kt->before(s) : nullptr;
    {
      const int nrow = nblk - j0 + 1;
      const int ntiles = ncol * nrow - ncol * (ncol - 1) / 2;
      hipLaunchKernelGGL(k_syrk_persist, dim3(std::min(ntiles, n_cu)), dim3(512), 0, s, S, (size_t)ld, kp, np * NB,
                         j0, nblk, ncol, D, Dp, Top, flags);
    }
    if (kt) {
      // algorithmic FLOP: the panel columns that carry data (nk) times the
      // lower-triangle entries updated (rows >= j0 NB, below D) and the rhs
      // row, plus the factorization of the diagonal block
      const double nk = std::max(0.0, std::min<double>(np * NB, (double)D - (double)kp * NB));
      const double m = std::max(0.0, (double)D - (double)j0 * NB);
      const double nb = std::min<double>(NB, m);
      const double entries = column_only ? nb * m - nb * (nb - 1.0) / 2.0 + nb : m * (m + 1.0) / 2.0 + m;
      kt->after(e0, s, 2.0 * nk * entries + nb * nb * nb / 3.0);
    }
  };
  hipLaunchKernelGGL(k_potrf_block, dim3(1), dim3(256), 0, s, S, (size_t)ld, 0, D, Top, flags);
  int k = 0;
  while (k < nblk) {
    trsm(k);
    const int ntrail = nblk - 1 - k;
    if (ntrail >= PAIR_MIN_TRAIL) {
      syrk(k, 1, k + 1, true);   // block column k+1 with panel k; factors diagonal k+1
      trsm(k + 1);
      syrk(k, 2, k + 2, false);  // the rest with panels k and k+1; factors diagonal k+2
      k += 2;
    } else {
      if (ntrail > 0) syrk(k, 1, k + 1, false);
      k += 1;
    }
  }
  hipLaunchKernelGGL(k_block_inverse, dim3(nblk), dim3(512), 0, s, D, Top, Linv);
  HIP_CHECK(hipGetLastError());
}

void dense_extract_row(const double* S, uint32_t Dp, uint32_t ld, uint32_t row, double* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_extract_row, dim3((Dp + 255) / 256), dim3(256), 0, s, S, (size_t)ld, Dp, row,
                     out);
  HIP_CHECK(hipGetLastError());
}

void dense_backsolve(const double* S, const double* Linv, uint32_t Dp, uint32_t ld, const double* y,
                     double* x, uint32_t* chain, uint32_t epoch, int* status, hipStream_t s) {
  const int nblk = (int)(Dp / NB);
  if (nblk == 0) return;
  hipLaunchKernelGGL(k_backsolve_chain, dim3(nblk), dim3(256), 0, s, S, (size_t)ld, Linv, y, x, chain, epoch,
                     nblk, status);
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
