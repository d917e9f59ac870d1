// Grid-wide kernels of one LM try: one thread per observation, point or
// incidence, one wavefront per Schur chunk.  The per-element work lives in
// engine_device.hpp, shared with the persistent window solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "engine_device.hpp"

namespace sfmba {

// One lane per observation, records and point terms stored through LDS:
// each lane leaves its observation's record pieces (X2 and R2, then its J2
// rows: P pieces of two values) in its wavefront's stage, then the
// wavefront stores the pieces of 64 consecutive observations, lane l taking
// piece l of each group of 64, so that one store instruction covers
// consecutive pieces of neighbouring records instead of one piece in each
// of 64 records; the point terms (nine values per observation) follow
// through the same stage and are summed there into the point blocks of the
// points seen whole by the wavefront.  An observation of two incidences
// writes its records itself.  Against four
// lanes per observation that each compute the whole Jacobian and store a
// quarter of its pieces: FP64 0.50 -> 0.47 ms, FP32 0.35 -> 0.27 ms per
// linearization of the 1,000-camera scene (with the point accumulation).
// The FP64 form keeps two wavefronts per SIMD (187 registers); asking for
// three spilled the point terms (0.57 ms).
template <typename T, int DB>
__global__ __launch_bounds__(256) void k_linearize_staged(
    View<T> v, int cur, int loss, double scale) {
  using V2 = typename Vec2<T>::type;
  constexpr uint32_t P = 4 + DB;
  constexpr size_t RS2 = rec_values(DB) / 2;
  __shared__ __attribute__((aligned(16))) V2 stage[4 * 64 * P];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  V2* sw = stage + w * 64 * P;
  static_assert(LINEARIZE_SPAN == 64, "one wavefront per span");
  uint32_t inc = NONE;  // the staged record (all DB rows: those past the block's are zero)
  uint32_t pb = 0, pe = 0, p = 0;  // the observation's point and its observations
  bool valid = false, act = false;
  T pc[9];
  if (o < v.n_obs) {
    valid = linearize_obs<T, DB>(v, o, cur, loss, scale, sw + lane * P, pc);
    const uint32_t i0 = v.obs_inc[2 * o];
    act = valid && i0 != NONE;
    if (i0 != NONE && v.obs_inc[2 * o + 1] == NONE) inc = i0;
    p = v.obs_pt[o];
    pb = v.pt_ptr[p];
    pe = v.pt_ptr[p + 1];
  }
  // the wavefront's stage is complete (one wavefront writes and reads it)
  __asm__ volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
  // 16-byte units: two pieces of an FP32 record, one of an FP64 record
  using U = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  constexpr uint32_t PPU = sizeof(U) / sizeof(V2);
  static_assert(P % PPU == 0 && RS2 % PPU == 0, "records in whole 16-byte units");
  constexpr uint32_t PU = P / PPU;
  constexpr size_t RSU = RS2 / PPU;
  U* recu = reinterpret_cast<U*>(v.rec);
  const U* su = reinterpret_cast<const U*>(sw);
#pragma unroll
  for (uint32_t j = 0; j < PU; ++j) {
    const uint32_t idx = j * 64 + lane;
    const uint32_t slot = idx / PU, piece = idx - slot * PU;
    const uint32_t sinc = (uint32_t)__shfl((int)inc, (int)slot);
    if (sinc != NONE) recu[(size_t)sinc * RSU + piece] = su[idx];
  }
  // Point terms: through the stage, then summed here for the points whose
  // observations all lie in this span (the lane of a point's last
  // observation adds the terms of its valid observations in order) and
  // stored in obs_pc for the others (k_point_accumulate sums those).
  static_assert(64 * 9 <= 2 * 64 * P, "the point terms fit the stage");
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  T* st = reinterpret_cast<T*>(sw);
#pragma unroll
  for (int k = 0; k < 9; ++k) st[lane * 9 + k] = pc[k];
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  const uint32_t o0 = o - lane;
  const bool in_span = o < v.n_obs && pb / LINEARIZE_SPAN == (pe - 1) / LINEARIZE_SPAN;
  const uint64_t split = __ballot(o < v.n_obs && !in_span);
  const uint64_t vmask = __ballot(valid), amask = __ballot(act);
  if (split != 0) {
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) {
      const uint32_t idx = k * 64 + lane;  // value idx of the span's terms
      if ((split >> (idx / 9)) & 1u) v.obs_pc[9 * (size_t)o0 + idx] = st[idx];
    }
  }
  if (in_span && o + 1 == pe) {
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = 0.0;
    for (uint32_t j = pb - o0; j <= lane; ++j) {
      if (!((vmask >> j) & 1u)) continue;
#pragma unroll
      for (int k = 0; k < 9; ++k) a[k] += (double)st[j * 9 + k];
    }
    const uint32_t h = pb - o0;
    const uint64_t range = (lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1)) & ~((1ull << h) - 1);
#pragma unroll
    for (int k = 0; k < 6; ++k) v.pt_C[6 * (size_t)p + k] = a[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) v.pt_g[3 * (size_t)p + k] = a[6 + k];
    v.pt_active[p] = (amask & range) != 0 ? 1 : 0;
  }
}

// The points of a list (those the linearization leaves: no observation, or
// observations in two spans), LPP lanes per point (neighbouring lanes read
// neighbouring observations).
constexpr int ACCUMULATE_LANES = 4;
template <typename T>
__global__ __launch_bounds__(256) void k_point_accumulate(View<T> v, const uint32_t* __restrict__ list,
                                                          uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = t / ACCUMULATE_LANES;
  if (i < n) point_accumulate<T, ACCUMULATE_LANES>(v, list[i], t % ACCUMULATE_LANES);
}

template <typename T>
__global__ __launch_bounds__(256) void k_point_prepare(View<T> v, double lambda) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < v.n_pt) point_prepare(v, p, lambda);
}

// Y and z of every incidence record, four lanes per incidence (the groups
// of four lanes of a wavefront are whole incidences or past the end).
template <typename T, int DB>
__global__ __launch_bounds__(256) void k_compute_Y(View<T> v) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t a = t >> 2;
  if (a < v.n_inc) compute_Y<T, DB>(v, a, t & 3u);
}

// The Schur reduction from incidence records, one wavefront per chunk,
// chunks in destination-tile order (the wavefronts in flight work on
// neighbouring destinations and share their records in L2 and the MALL).
// Four wavefronts per workgroup, each with its own LDS stage.
template <typename T, int DB>
__global__ __launch_bounds__(256) void k_schur_rec(View<T> v, double lambda, double* __restrict__ parts) {
  constexpr uint32_t PER_WAVE = RecShape<T, DB>::LDS;
  __shared__ __attribute__((aligned(16))) T stage[4 * PER_WAVE];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t ci = blockIdx.x * 4 + w;
  if (ci >= v.n_chunks) return;
  schur_chunk_rec<T, DB>(v, v.chunks[ci], lambda, parts, stage + w * PER_WAVE);
}

// Destinations split over several chunks: sum the partials in chunk order
// (sums, then on diagonal blocks the damping diagonal and the rhs).
__global__ __launch_bounds__(256) void k_schur_reduce(
    uint32_t n_multi, const uint32_t* __restrict__ multi_dest,
    const uint32_t* __restrict__ dest_part0, const uint32_t* __restrict__ dest_nparts,
    const uint32_t* __restrict__ grp_off, const uint32_t* __restrict__ grp_dim,
    const double* __restrict__ parts, double lambda, double* __restrict__ S, uint32_t ld,
    uint32_t Dp) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t m = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (m >= n_multi) return;
  const uint32_t ga = multi_dest[m] >> 16, gb = multi_dest[m] & 0xFFFFu;
  const uint32_t da = grp_dim[ga], db = grp_dim[gb], ne = da * db;
  const bool diag = ga == gb;
  const uint32_t p0 = dest_part0[m], np = dest_nparts[m];
  const uint32_t oa = grp_off[ga], ob = grp_off[gb];
  for (uint32_t e = lane; e < ne; e += 64) {
    const uint32_t r = e % da, c = e / da;
    if (diag && r < c) continue;
    double sv = 0.0;
    for (uint32_t k = 0; k < np; ++k) sv += parts[(size_t)(p0 + k) * PART_STRIDE + e];
    if (diag && r == c) {
      double bd = 0.0;
      for (uint32_t k = 0; k < np; ++k) bd += parts[(size_t)(p0 + k) * PART_STRIDE + 196 + r];
      sv += lambda * bd;
    }
    S[(size_t)(ob + c) * ld + oa + r] = sv;
  }
  if (diag && lane < da) {
    double rv = 0.0;
    for (uint32_t k = 0; k < np; ++k) rv += parts[(size_t)(p0 + k) * PART_STRIDE + 392 + lane];
    S[(size_t)(oa + lane) * ld + Dp] = rv;
  }
}

// Point back-substitution in two passes.  First E_a^T delta_g = X2^T
// (J2_a^T delta_g) per incidence a: LPI lanes per incidence, one J2 row per
// lane, summed across the lanes.  The record's row and X2 are loaded before
// the block's offset and size are known (rows past the block's are zero),
// so the record loads do not wait on the chain incidence -> block -> delta.
template <typename T, int DB, int LPI>
__global__ __launch_bounds__(256) void k_backsub_inc(View<T> v, double* __restrict__ inc_part) {
  using V2 = typename Vec2<T>::type;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t a = t / LPI, r = t % LPI;
  double t0 = 0.0, t1 = 0.0;
  T x2[6];
  if (a < v.n_inc) {
    const T* rc = v.rec + (size_t)a * rec_values(DB);
    V2 row = {T(0), T(0)};
    if (r < (uint32_t)DB) row = *reinterpret_cast<const V2*>(rc + 8 + 2 * r);
    if (r == 0)
#pragma unroll
      for (int j = 0; j < 6; ++j) x2[j] = rc[j];
    const uint32_t g = v.inc_grp[a];
    const double d = r < v.grp_dim[g] ? v.delta[v.grp_off[g] + r] : 0.0;
    t0 = (double)row.x * d;
    t1 = (double)row.y * d;
  }
  // every lane of the wavefront takes part (LPI divides 64)
#pragma unroll
  for (int o = LPI / 2; o > 0; o >>= 1) {
    t0 += __shfl_xor(t0, o);
    t1 += __shfl_xor(t1, o);
  }
  if (a < v.n_inc && r == 0) {
#pragma unroll
    for (int j = 0; j < 3; ++j) inc_part[3 * (size_t)a + j] = (double)x2[j] * t0 + (double)x2[3 + j] * t1;
  }
}

// Then per point: dp = -Cinv (g + sum over its incidences), added to the
// point (a point left out of the Schur complement keeps its position).
template <typename T>
__global__ __launch_bounds__(256) void k_backsub(View<T> v, int cur, const double* __restrict__ inc_part) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.n_pt) return;
  // the point's own values are loaded before its flag is tested, so that
  // the incidence sums wait on one load (the incidence range), not two
  const double* X = v.pts[cur] + 3 * (size_t)p;
  double* Y = v.pts[1 - cur] + 3 * (size_t)p;
  const double x0 = X[0], x1 = X[1], x2 = X[2];
  const uint8_t ok = v.pt_ok[p];
  const uint32_t a0 = v.pt_inc_ptr[p], a1 = v.pt_inc_ptr[p + 1];
  double acc0 = v.pt_g[3 * (size_t)p], acc1 = v.pt_g[3 * (size_t)p + 1], acc2 = v.pt_g[3 * (size_t)p + 2];
  const double* ci = v.pt_cinv + 9 * (size_t)p;
  double c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) c[k] = ci[k];
  if (!ok) {
    Y[0] = x0;
    Y[1] = x1;
    Y[2] = x2;
    return;
  }
  for (uint32_t a = a0; a < a1; ++a) {
    acc0 += inc_part[3 * (size_t)a];
    acc1 += inc_part[3 * (size_t)a + 1];
    acc2 += inc_part[3 * (size_t)a + 2];
  }
  Y[0] = x0 + -(c[0] * acc0 + c[1] * acc1 + c[2] * acc2);
  Y[1] = x1 + -(c[3] * acc0 + c[4] * acc1 + c[5] * acc2);
  Y[2] = x2 + -(c[6] * acc0 + c[7] * acc1 + c[8] * acc2);
}

template <typename T>
__global__ __launch_bounds__(256) void k_update_poses(View<T> v, int cur) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < v.n_img) update_pose(v, i, cur);
}

template <typename T>
__global__ __launch_bounds__(256) void k_update_cams(View<T> v, int cur) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < v.n_cam) update_cam(v, c, cur);
}

// ------------------------------------------------------------------- cost
// mode 0: sum w(|r|) |r|^2 (total_cost); mode 1: sum |r| (the numerator of
// mean_reprojection_error).  Both also count the observations in front of
// their camera.  Deterministic two-level reduction.

__global__ __launch_bounds__(COST_BLOCK) void k_cost(
    uint32_t n_obs, const uint32_t* __restrict__ obs_img, const uint32_t* __restrict__ obs_pt,
    const double2* __restrict__ obs_xy, const uint32_t* __restrict__ cam_of_img,
    const int32_t* __restrict__ cam_model, const double* __restrict__ poses,
    const double* __restrict__ cams, const double* __restrict__ pts, int loss, double scale,
    int mode, double* __restrict__ part) {
  __shared__ double red0[COST_BLOCK];
  __shared__ double red1[COST_BLOCK];
  double s0 = 0.0, s1 = 0.0;
  const uint32_t base = blockIdx.x * COST_BLOCK * COST_ITEMS + threadIdx.x;
  // every item's observation record first, so that the items' dependent
  // loads (image -> camera, pose, point) do not each wait for it in turn
  uint32_t img[COST_ITEMS], pt[COST_ITEMS];
  double2 xy[COST_ITEMS];
#pragma unroll
  for (int it = 0; it < COST_ITEMS; ++it) {
    const uint32_t o = base + it * COST_BLOCK;
    if (o < n_obs) {
      img[it] = obs_img[o];
      pt[it] = obs_pt[o];
      xy[it] = obs_xy[o];
    }
  }
#pragma unroll
  for (int it = 0; it < COST_ITEMS; ++it) {
    const uint32_t o = base + it * COST_BLOCK;
    if (o < n_obs) {
      double term, front;
      obs_cost_at(img[it], pt[it], xy[it], cam_of_img, cam_model, poses, cams, pts, loss, scale, mode, term, front);
      s0 += term;
      s1 += front;
    }
  }
  red0[threadIdx.x] = s0;
  red1[threadIdx.x] = s1;
  __syncthreads();
  for (int h = COST_BLOCK / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red0[threadIdx.x] += red0[threadIdx.x + h];
      red1[threadIdx.x] += red1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red0[0];
    part[2 * blockIdx.x + 1] = red1[0];
  }
}

__global__ __launch_bounds__(1024) void k_cost_final(uint32_t n_part, const double* __restrict__ part,
                                                     double* __restrict__ out) {
  __shared__ double red0[1024];
  __shared__ double red1[1024];
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t i = threadIdx.x; i < n_part; i += 1024) {
    s0 += part[2 * i];
    s1 += part[2 * i + 1];
  }
  red0[threadIdx.x] = s0;
  red1[threadIdx.x] = s1;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red0[threadIdx.x] += red0[threadIdx.x + h];
      red1[threadIdx.x] += red1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red0[0];
    out[1] = red1[0];
  }
}

// -------------------------------------------------- Jacobian check kernel
template <typename T>
__global__ __launch_bounds__(256) void k_jacobians_debug(
    uint32_t n_obs, const uint32_t* __restrict__ obs_img, const uint32_t* __restrict__ obs_pt,
    const double2* __restrict__ obs_xy, const uint32_t* __restrict__ cam_of_img,
    const int32_t* __restrict__ cam_model, const double* __restrict__ poses,
    const double* __restrict__ cams, const double* __restrict__ pts, double* __restrict__ r,
    double* __restrict__ jp_out, double* __restrict__ jx_out, double* __restrict__ jc_out,
    uint8_t* __restrict__ valid) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_obs) return;
  const uint32_t img = obs_img[o], pt = obs_pt[o], cam = cam_of_img[img];
  const double* pose = poses + 7 * (size_t)img;
  const double* P = cams + 8 * (size_t)cam;
  double c0, c1, c2;
  transform_point(pose, pts[3 * (size_t)pt], pts[3 * (size_t)pt + 1], pts[3 * (size_t)pt + 2], c0,
                  c1, c2);
  if (c2 <= 1e-9) {
    valid[o] = 0;
    for (int k = 0; k < 2; ++k) r[2 * (size_t)o + k] = 0.0;
    for (int k = 0; k < 12; ++k) jp_out[12 * (size_t)o + k] = 0.0;
    for (int k = 0; k < 6; ++k) jx_out[6 * (size_t)o + k] = 0.0;
    for (int k = 0; k < 16; ++k) jc_out[16 * (size_t)o + k] = 0.0;
    return;
  }
  valid[o] = 1;
  const double2 xy = obs_xy[o];
  double r0, r1;
  linearization_residual(cam_model[cam], P, c0, c1, c2, xy.x, xy.y, r0, r1);
  r[2 * (size_t)o] = r0;
  r[2 * (size_t)o + 1] = r1;
  T du[3], dv[3], jc0[8], jc1[8];
  model_jacobian<T>(cam_model[cam], P, (T)c0, (T)c1, (T)c2, du, dv, jc0, jc1);
  T R[9];
  rotation_matrix<T>(pose, R);
  T jp[12], jx[6];
  pose_point_jacobians<T>(R, (T)(c0 - pose[4]), (T)(c1 - pose[5]), (T)(c2 - pose[6]), du, dv, jp, jx);
  for (int k = 0; k < 12; ++k) jp_out[12 * (size_t)o + k] = (double)jp[k];
  for (int k = 0; k < 6; ++k) jx_out[6 * (size_t)o + k] = (double)jx[k];
  for (int k = 0; k < 8; ++k) {
    jc_out[16 * (size_t)o + k] = (double)jc0[k];
    jc_out[16 * (size_t)o + 8 + k] = (double)jc1[k];
  }
}

// Residual norm and in-front flag of observations given in the caller's
// order (the outlier filter of filter_and_reoptimize,
// crates/sfm-reconstruction/src/lib.rs:1734-1746): |project(R X + t) - x|,
// in_front = z > 1e-9.
__global__ __launch_bounds__(256) void k_residual_norms(
    uint32_t n_obs, const sfmba_observation* __restrict__ obs, const uint64_t* __restrict__ cam_of_img,
    const int32_t* __restrict__ cam_model, const double* __restrict__ poses,
    const double* __restrict__ cams, const double* __restrict__ pts, double* __restrict__ norm,
    uint8_t* __restrict__ in_front) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_obs) return;
  const sfmba_observation ob = obs[o];
  const uint64_t cam = cam_of_img[ob.image_idx];
  double c0, c1, c2;
  const double* X = pts + 3 * ob.point_idx;
  transform_point(poses + 7 * ob.image_idx, X[0], X[1], X[2], c0, c1, c2);
  if (c2 <= 1e-9) {
    in_front[o] = 0;
    norm[o] = 0.0;
    return;
  }
  double r0, r1;
  project_residual(cam_model[cam], cams + 8 * cam, c0, c1, c2, ob.x, ob.y, r0, r1);
  in_front[o] = 1;
  norm[o] = sqrt(r0 * r0 + r1 * r1);
}

// ------------------------------------------------------------- launchers
static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }
// Lanes given to one incidence by the row-per-lane kernels: the power of two
// at or above the largest group dimension.
static inline uint32_t lanes_per_incidence(const Layout& L) {
  if (L.max_grp_dim > 16) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "group dimension above 16");
  return L.max_grp_dim <= 2 ? 2 : L.max_grp_dim <= 4 ? 4 : L.max_grp_dim <= 8 ? 8 : 16;
}

template <typename T>
View<T> make_view(const Layout& L, Work<T>& W) {
  View<T> v{};
  v.n_img = L.n_img;
  v.n_cam = L.n_cam;
  v.n_pt = L.n_pt;
  v.n_obs = L.n_obs;
  v.n_inc = L.n_inc;
  v.n_chunks = L.n_chunks;
  v.D = L.D;
  v.Dp = L.Dp;
  v.ld = L.ld;
  v.obs_img = L.obs_img;
  v.obs_pt = L.obs_pt;
  v.obs_xy = L.obs_xy;
  v.obs_inc = L.obs_inc;
  v.pt_ptr = L.pt_ptr;
  v.inc_grp = L.inc_grp;
  v.inc_pt = L.inc_pt;
  v.pt_inc_ptr = L.pt_inc_ptr;
  v.cam_of_img = L.cam_of_img;
  v.cam_model = L.cam_model;
  v.img_slot = L.img_slot;
  v.cam_slot = L.cam_slot;
  v.cam_free = L.cam_free;
  v.cam_nfree = L.cam_nfree;
  v.grp_off = L.grp_off;
  v.grp_dim = L.grp_dim;
  v.pairs = reinterpret_cast<const uint2*>(L.pairs);
  v.chunks = L.chunks;
  v.rec = W.rec;
  v.obs_pc = W.obs_pc;
  v.obs_valid = W.obs_valid;
  v.pt_C = W.pt_C;
  v.pt_g = W.pt_g;
  v.pt_active = W.pt_active;
  v.pt_cinv = W.pt_cinv;
  v.pt_ok = W.pt_ok;
  v.S = W.S;
  v.delta = W.delta;
  for (int k = 0; k < 2; ++k) {
    v.poses[k] = W.poses[k];
    v.cams[k] = W.cams[k];
    v.pts[k] = W.pts[k];
  }
  return v;
}

template <typename T>
void launch_linearize(const Layout& L, Work<T>& W, int cur, int loss, double scale, hipStream_t s) {
  if (L.n_obs == 0) return;
  const View<T> v = make_view(L, W);
  const dim3 grid(blocks_for(L.n_obs, 256));
  switch (L.rec_db) {
    case 8: hipLaunchKernelGGL((k_linearize_staged<T, 8>), grid, dim3(256), 0, s, v, cur, loss, scale); break;
    case 12: hipLaunchKernelGGL((k_linearize_staged<T, 12>), grid, dim3(256), 0, s, v, cur, loss, scale); break;
    default: hipLaunchKernelGGL((k_linearize_staged<T, 16>), grid, dim3(256), 0, s, v, cur, loss, scale); break;
  }
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_point_accumulate(const Layout& L, Work<T>& W, hipStream_t s) {
  if (L.n_pt_fix == 0) return;
  hipLaunchKernelGGL(k_point_accumulate<T>, dim3(blocks_for((uint64_t)L.n_pt_fix * ACCUMULATE_LANES, 256)), dim3(256),
                     0, s, make_view(L, W), (const uint32_t*)L.pt_fix, L.n_pt_fix);
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_point_prepare(const Layout& L, Work<T>& W, double lambda, hipStream_t s) {
  if (L.n_pt == 0) return;
  const View<T> v = make_view(L, W);
  hipLaunchKernelGGL(k_point_prepare<T>, dim3(blocks_for(L.n_pt, 256)), dim3(256), 0, s, v, lambda);
  HIP_CHECK(hipGetLastError());
  if (L.n_inc == 0) return;
  const dim3 grid(blocks_for(4 * (uint64_t)L.n_inc, 256));
  switch (L.rec_db) {
    case 8: hipLaunchKernelGGL((k_compute_Y<T, 8>), grid, dim3(256), 0, s, v); break;
    case 12: hipLaunchKernelGGL((k_compute_Y<T, 12>), grid, dim3(256), 0, s, v); break;
    default: hipLaunchKernelGGL((k_compute_Y<T, 16>), grid, dim3(256), 0, s, v); break;
  }
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_schur(const Layout& L, Work<T>& W, double lambda, hipStream_t s) {
  const View<T> v = make_view(L, W);
  if (L.n_chunks) {
    if (L.max_grp_dim > 16) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "group dimension above 16");
    const dim3 grid(blocks_for(L.n_chunks, 4));
    switch (L.rec_db) {
      case 8: hipLaunchKernelGGL((k_schur_rec<T, 8>), grid, dim3(256), 0, s, v, lambda, W.parts); break;
      case 12: hipLaunchKernelGGL((k_schur_rec<T, 12>), grid, dim3(256), 0, s, v, lambda, W.parts); break;
      default: hipLaunchKernelGGL((k_schur_rec<T, 16>), grid, dim3(256), 0, s, v, lambda, W.parts); break;
    }
    HIP_CHECK(hipGetLastError());
  }
  if (L.n_multi) {
    hipLaunchKernelGGL(k_schur_reduce, dim3(blocks_for(L.n_multi, 4)), dim3(256), 0, s, L.n_multi,
                       L.multi_dest, L.dest_part0, L.dest_nparts, L.grp_off, L.grp_dim, W.parts,
                       lambda, W.S, L.ld, L.Dp);
    HIP_CHECK(hipGetLastError());
  }
}

template <typename T>
void launch_backsub_update(const Layout& L, Work<T>& W, int cur, hipStream_t s) {
  const View<T> v = make_view(L, W);
  if (L.n_pt) {
    if (L.n_inc) {
      const uint32_t lpi = lanes_per_incidence(L);
      const dim3 grid(blocks_for((uint64_t)L.n_inc * lpi, 256));
      const dim3 blk(256);
      double* ip = W.inc_part;
      if (L.rec_db == 8) {
        switch (lpi) {
          case 2: hipLaunchKernelGGL((k_backsub_inc<T, 8, 2>), grid, blk, 0, s, v, ip); break;
          case 4: hipLaunchKernelGGL((k_backsub_inc<T, 8, 4>), grid, blk, 0, s, v, ip); break;
          default: hipLaunchKernelGGL((k_backsub_inc<T, 8, 8>), grid, blk, 0, s, v, ip); break;
        }
      } else if (L.rec_db == 12) {
        hipLaunchKernelGGL((k_backsub_inc<T, 12, 16>), grid, blk, 0, s, v, ip);
      } else {
        hipLaunchKernelGGL((k_backsub_inc<T, 16, 16>), grid, blk, 0, s, v, ip);
      }
      HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_backsub<T>, dim3(blocks_for(L.n_pt, 256)), dim3(256), 0, s, v, cur, W.inc_part);
    HIP_CHECK(hipGetLastError());
  }
  if (L.n_img) {
    hipLaunchKernelGGL(k_update_poses<T>, dim3(blocks_for(L.n_img, 256)), dim3(256), 0, s, v, cur);
    HIP_CHECK(hipGetLastError());
  }
  if (L.n_cam) {
    hipLaunchKernelGGL(k_update_cams<T>, dim3(blocks_for(L.n_cam, 256)), dim3(256), 0, s, v, cur);
    HIP_CHECK(hipGetLastError());
  }
}

void launch_cost(const Layout& L, const double* poses, const double* cams, const double* pts,
                 int loss, double scale, double* part, double* out, int mode, hipStream_t s) {
  const uint32_t nb = blocks_for(L.n_obs, COST_BLOCK * COST_ITEMS);
  if (nb) {
    hipLaunchKernelGGL(k_cost, dim3(nb), dim3(COST_BLOCK), 0, s, L.n_obs, L.obs_img, L.obs_pt,
                       L.obs_xy, L.cam_of_img, L.cam_model, poses, cams, pts, loss, scale, mode,
                       part);
    HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_cost_final, dim3(1), dim3(1024), 0, s, nb, part, out);
  HIP_CHECK(hipGetLastError());
}

void launch_residual_norms(uint32_t n_obs, const sfmba_observation* obs, const uint64_t* cam_of_img,
                           const int32_t* cam_model, const double* poses, const double* cams,
                           const double* pts, double* norm, uint8_t* in_front, hipStream_t s) {
  if (n_obs == 0) return;
  hipLaunchKernelGGL(k_residual_norms, dim3(blocks_for(n_obs, 256)), dim3(256), 0, s, n_obs, obs, cam_of_img,
                     cam_model, poses, cams, pts, norm, in_front);
  HIP_CHECK(hipGetLastError());
}

void launch_jacobians_debug(const Layout& L, const double* poses, const double* cams,
                            const double* pts, int fp32, double* r, double* jp, double* jx,
                            double* jc, uint8_t* valid, hipStream_t s) {
  if (L.n_obs == 0) return;
  if (fp32)
    hipLaunchKernelGGL(k_jacobians_debug<float>, dim3(blocks_for(L.n_obs, 256)), dim3(256), 0, s,
                       L.n_obs, L.obs_img, L.obs_pt, L.obs_xy, L.cam_of_img, L.cam_model, poses,
                       cams, pts, r, jp, jx, jc, valid);
  else
    hipLaunchKernelGGL(k_jacobians_debug<double>, dim3(blocks_for(L.n_obs, 256)), dim3(256), 0, s,
                       L.n_obs, L.obs_img, L.obs_pt, L.obs_xy, L.cam_of_img, L.cam_model, poses,
                       cams, pts, r, jp, jx, jc, valid);
  HIP_CHECK(hipGetLastError());
}

template void launch_linearize<float>(const Layout&, Work<float>&, int, int, double, hipStream_t);
template void launch_linearize<double>(const Layout&, Work<double>&, int, int, double, hipStream_t);
template void launch_point_accumulate<float>(const Layout&, Work<float>&, hipStream_t);
template void launch_point_accumulate<double>(const Layout&, Work<double>&, hipStream_t);
template void launch_point_prepare<float>(const Layout&, Work<float>&, double, hipStream_t);
template void launch_point_prepare<double>(const Layout&, Work<double>&, double, hipStream_t);
template void launch_schur<float>(const Layout&, Work<float>&, double, hipStream_t);
template void launch_schur<double>(const Layout&, Work<double>&, double, hipStream_t);
template void launch_backsub_update<float>(const Layout&, Work<float>&, int, hipStream_t);
template void launch_backsub_update<double>(const Layout&, Work<double>&, int, hipStream_t);

}  // namespace sfmba
