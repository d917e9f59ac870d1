// Per-observation, per-point and Schur-reduction kernels of the engine.
//
// Reference correspondence (crates/sfm-ba/src/lib.rs):
//   k_linearize          hot loop A, 832-853 (analytic_jacobians + robust_weight)
//                        and the per-observation terms of hot loop B, 880-909
//   k_point_accumulate   c_diag / c_rhs sums, 885-886; the non-empty
//                        points_to_obs test, 960
//   k_point_prepare      damped point-block inverse, 958-965
//   k_schur / k_schur_reduce
//                        damped diagonal and cross blocks, 923-956, plus the
//                        point elimination S -= E C^-1 E^T, rhs += E C^-1 g
//                        of accumulate_point_schur, 112-211 / 967-1043
//   k_backsub            delta_point = -C^-1 (g + E^T delta), 1096-1120
//   k_update_*           trial state, 1122-1164 (perturb_pose 263-269)
//   k_cost               total_cost 727-745 / mean_reprojection_error 708-725
#include <hip/hip_runtime.h>

#include "camera_models.hpp"
#include "engine.hpp"

namespace sfmba {

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

// ---------------------------------------------------------------- linearize
// One thread per observation (point-major order).  Writes, for each of the
// observation's incidences, one J4 and one E4 row per unknown, and the
// observation's contribution to its point's 3x3 block and gradient.
template <typename T>
__global__ __launch_bounds__(256) void k_linearize(
    uint32_t n_obs, const uint32_t* __restrict__ obs_img, const uint32_t* __restrict__ obs_pt,
    const double2* __restrict__ obs_xy, const uint32_t* __restrict__ obs_inc,
    const uint32_t* __restrict__ inc_roff, const uint32_t* __restrict__ cam_of_img,
    const int32_t* __restrict__ cam_model, const uint32_t* __restrict__ img_slot,
    const uint32_t* __restrict__ cam_slot, const uint8_t* __restrict__ cam_free,
    const uint32_t* __restrict__ cam_nfree, const double* __restrict__ poses,
    const double* __restrict__ cams, const double* __restrict__ pts, int loss, double scale,
    typename Vec4<T>::type* __restrict__ J4, typename Vec4<T>::type* __restrict__ E4,
    T* __restrict__ obs_pc, uint8_t* __restrict__ obs_valid) {
  using V4 = typename Vec4<T>::type;
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_obs) return;
  const uint32_t img = obs_img[o];
  const uint32_t pt = obs_pt[o];
  const uint32_t cam = cam_of_img[img];
  const int model = cam_model[cam];
  const double* pose = poses + 7 * (size_t)img;
  const double* P = cams + 8 * (size_t)cam;
  const double X0 = pts[3 * (size_t)pt], X1 = pts[3 * (size_t)pt + 1], X2 = pts[3 * (size_t)pt + 2];
  double c0, c1, c2;
  transform_point(pose, X0, X1, X2, c0, c1, c2);

  const bool pose_free = img_slot[img] != NONE;
  const bool cam_is_free = cam_slot[cam] != NONE;
  const uint32_t kf = cam_is_free ? cam_nfree[cam] : 0u;
  const uint32_t ns = (pose_free ? 6u : 0u) + kf;
  const uint32_t inc0 = obs_inc[2 * o], inc1 = obs_inc[2 * o + 1];
  const uint32_t roff0 = inc0 != NONE ? inc_roff[inc0] : 0u;
  const bool two = inc1 != NONE;
  const uint32_t roff1 = two ? inc_roff[inc1] : 0u;
  T* pc_out = obs_pc + 9 * (size_t)o;

  if (c2 <= 1e-9) {
    // Behind the camera: dropped for this iteration (lib.rs:339-341).
    obs_valid[o] = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) pc_out[k] = T(0);
    const V4 z = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (uint32_t s = 0; s < 14; ++s) {
      if (s >= ns) break;
      const uint32_t row = (two && s >= 6) ? roff1 + (s - 6) : roff0 + s;
      J4[row] = z;
      E4[row] = z;
    }
    return;
  }
  obs_valid[o] = 1;
  const double2 xy = obs_xy[o];
  double r0, r1;
  linearization_residual(model, P, c0, c1, c2, xy.x, xy.y, r0, r1);
  const double wd = robust_weight(sqrt(r0 * r0 + r1 * r1), loss, scale);
  const T w = (T)wd;

  T du[3], dv[3], jc0[8], jc1[8];
  model_jacobian<T>(model, P, (T)c0, (T)c1, (T)c2, du, dv, jc0, jc1);
  T R[9];
  rotation_matrix<T>(pose, R);
  T jp[12], jx[6];
  pose_point_jacobians<T>(R, (T)(c0 - pose[4]), (T)(c1 - pose[5]), (T)(c2 - pose[6]), du, dv, jp,
                          jx);
  const T tr0 = (T)r0, tr1 = (T)r1;

  // Point terms: w Jx^T Jx (upper 6), w Jx^T r.
  const T wx0 = w * jx[0], wx1 = w * jx[1], wx2 = w * jx[2];
  const T wy0 = w * jx[3], wy1 = w * jx[4], wy2 = w * jx[5];
  pc_out[0] = wx0 * jx[0] + wy0 * jx[3];
  pc_out[1] = wx0 * jx[1] + wy0 * jx[4];
  pc_out[2] = wx0 * jx[2] + wy0 * jx[5];
  pc_out[3] = wx1 * jx[1] + wy1 * jx[4];
  pc_out[4] = wx1 * jx[2] + wy1 * jx[5];
  pc_out[5] = wx2 * jx[2] + wy2 * jx[5];
  pc_out[6] = wx0 * tr0 + wy0 * tr1;
  pc_out[7] = wx1 * tr0 + wy1 * tr1;
  pc_out[8] = wx2 * tr0 + wy2 * tr1;

  // Side vector: free pose columns, then the camera's free parameters.
  T s0[14], s1[14];
#pragma unroll
  for (int s = 0; s < 14; ++s) {
    s0[s] = T(0);
    s1[s] = T(0);
  }
  if (pose_free) {
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      s0[s] = jp[s];
      s1[s] = jp[6 + s];
    }
  }
  const uint8_t* fi = cam_free + 8 * (size_t)cam;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if ((uint32_t)j < kf) {
      const int idx = fi[j];
      T a0 = T(0), a1 = T(0);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q == idx) {
          a0 = jc0[q];
          a1 = jc1[q];
        }
      if (pose_free) {
#pragma unroll
        for (int s = 6; s < 14; ++s)
          if (s == 6 + j) {
            s0[s] = a0;
            s1[s] = a1;
          }
      } else {
#pragma unroll
        for (int s = 0; s < 8; ++s)
          if (s == j) {
            s0[s] = a0;
            s1[s] = a1;
          }
      }
    }
  }
#pragma unroll
  for (uint32_t s = 0; s < 14; ++s) {
    if (s >= ns) break;
    const uint32_t row = (two && s >= 6) ? roff1 + (s - 6) : roff0 + s;
    const T w0 = w * s0[s], w1 = w * s1[s];
    J4[row] = V4{s0[s], s1[s], w0, w1};
    E4[row] = V4{w0 * jx[0] + w1 * jx[3], w0 * jx[1] + w1 * jx[4], w0 * jx[2] + w1 * jx[5],
                 w0 * tr0 + w1 * tr1};
  }
}

// ------------------------------------------------------- point accumulation
template <typename T>
__global__ __launch_bounds__(256) void k_point_accumulate(
    uint32_t n_pt, const uint32_t* __restrict__ pt_ptr, const T* __restrict__ obs_pc,
    const uint8_t* __restrict__ obs_valid, const uint32_t* __restrict__ obs_inc,
    double* __restrict__ pt_C, double* __restrict__ pt_g, uint8_t* __restrict__ pt_active) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pt) return;
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) a[k] = 0.0;
  uint8_t active = 0;
  for (uint32_t o = pt_ptr[p]; o < pt_ptr[p + 1]; ++o) {
    if (!obs_valid[o]) continue;
    if (obs_inc[2 * o] != NONE) active = 1;
    const T* c = obs_pc + 9 * (size_t)o;
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] += (double)c[k];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) pt_C[6 * (size_t)p + k] = a[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) pt_g[3 * (size_t)p + k] = a[6 + k];
  pt_active[p] = active;
}

// Damped point block (C + lambda diag C)^-1 by the adjugate; a point without
// free couplings or with det == 0 is left out of the Schur complement and
// keeps its position (lib.rs:958-965, 1001, 1102).
__global__ __launch_bounds__(256) void k_point_prepare(uint32_t n_pt, const double* __restrict__ pt_C,
                                                       const uint8_t* __restrict__ pt_active,
                                                       double lambda, double* __restrict__ pt_cinv,
                                                       uint8_t* __restrict__ pt_ok) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pt) return;
  double* o = pt_cinv + 9 * (size_t)p;
  const double* C = pt_C + 6 * (size_t)p;
  bool ok = false;
  if (pt_active[p]) {
    const double m11 = C[0] + C[0] * lambda, m12 = C[1], m13 = C[2];
    const double m22 = C[3] + C[3] * lambda, m23 = C[4];
    const double m33 = C[5] + C[5] * lambda;
    const double m21 = m12, m31 = m13, m32 = m23;
    const double a = m22 * m33 - m32 * m23;
    const double b = m21 * m33 - m31 * m23;
    const double c = m21 * m32 - m31 * m22;
    const double det = m11 * a - m12 * b + m13 * c;
    if (det != 0.0) {
      ok = true;
      o[0] = a / det;
      o[1] = (m13 * m32 - m33 * m12) / det;
      o[2] = (m12 * m23 - m22 * m13) / det;
      o[3] = -b / det;
      o[4] = (m11 * m33 - m31 * m13) / det;
      o[5] = (m13 * m21 - m23 * m11) / det;
      o[6] = c / det;
      o[7] = (m12 * m31 - m32 * m11) / det;
      o[8] = (m11 * m22 - m21 * m12) / det;
    }
  }
  if (!ok)
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = 0.0;
  pt_ok[p] = ok ? 1 : 0;
}

// F = E Cinv and F . g (the incidence's rhs share), one thread per incidence.
template <typename T>
__global__ __launch_bounds__(256) void k_compute_F(uint32_t n_inc, const uint32_t* __restrict__ inc_pt,
                                                   const uint32_t* __restrict__ inc_roff,
                                                   const uint32_t* __restrict__ inc_froff,
                                                   const uint32_t* __restrict__ inc_grp,
                                                   const uint32_t* __restrict__ grp_dim,
                                                   const double* __restrict__ pt_cinv,
                                                   const double* __restrict__ pt_g,
                                                   const typename Vec4<T>::type* __restrict__ E4,
                                                   typename Vec4<T>::type* __restrict__ F4) {
  using V4 = typename Vec4<T>::type;
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_inc) return;
  const uint32_t pt = inc_pt[a];
  const double* ci = pt_cinv + 9 * (size_t)pt;
  T c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) c[k] = (T)ci[k];
  const T g0 = (T)pt_g[3 * (size_t)pt], g1 = (T)pt_g[3 * (size_t)pt + 1], g2 = (T)pt_g[3 * (size_t)pt + 2];
  const uint32_t r0 = inc_roff[a], f0ff = inc_froff[a], n = grp_dim[inc_grp[a]];
  for (uint32_t r = 0; r < n; ++r) {
    const V4 e = E4[r0 + r];
    const T f0 = e.x * c[0] + e.y * c[3] + e.z * c[6];
    const T f1 = e.x * c[1] + e.y * c[4] + e.z * c[7];
    const T f2 = e.x * c[2] + e.y * c[5] + e.z * c[8];
    // w: this row's share of rhs, F . g_point = E Cinv g_point
    F4[f0ff + r] = V4{f0, f1, f2, f0 * g0 + f1 * g1 + f2 * g2};
  }
}

// Compensated (Kahan) accumulator for the FP32 path; plain sum in FP64.
template <typename T>
struct Acc {
  T s = T(0), c = T(0);
  __device__ inline void add(T x) {
    const T y = x - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ inline double value() const { return (double)s - (double)c; }
};
template <>
struct Acc<double> {
  double s = 0.0;
  __device__ inline void add(double x) { s += x; }
  __device__ inline double value() const { return s; }
};

constexpr int PART_STRIDE = 2 * 196 + 16;

// One wavefront per chunk of a destination block (ga >= gb) of the reduced
// system; lane l owns entries l, l + 64, ... (NQ of them) of the da x db
// block.  Sums, over the chunk's point pairs (a, b), B = w J_a^T J_b for
// pairs of the same observation and Schur = F_a E_b^T = E_a Cinv E_b^T; on
// diagonal blocks also the right-hand side F_a g_point - w J_a^T r over
// pairs a == b (F4.w - E4.w).  A pair is two words: a's first F row (F is
// group-major) with bit 31 "same observation" and bit 30 "a's rows precede
// b's", and b's first row with bit 31 "a == b".  U pairs are processed per
// step with all their loads issued together.  Chunks run in destination
// order, so the workgroups in flight on all XCDs share one block row's F
// rows (a compact group-major range, held by L2 and the MALL).  Single-chunk
// destinations are written straight into S (with the Marquardt damping of
// the diagonal); others leave partials.
template <typename T, int NQ>
__global__ __launch_bounds__(256) void k_schur(
    uint32_t n_chunks, const uint4* __restrict__ chunks, const uint2* __restrict__ pairs,
    const uint32_t* __restrict__ grp_off, const uint32_t* __restrict__ grp_dim,
    const typename Vec4<T>::type* __restrict__ J4, const typename Vec4<T>::type* __restrict__ E4,
    const typename Vec4<T>::type* __restrict__ F4, double lambda, double* __restrict__ S, uint32_t ld,
    uint32_t Dp, double* __restrict__ parts) {
  using V4 = typename Vec4<T>::type;
  constexpr uint32_t HI = 0x80000000u, BEFORE = 0x40000000u, OFF = 0x3FFFFFFFu;
  constexpr int UMAX = sizeof(T) == 4 ? 4 : 2;  // measured best on gfx950
  constexpr int U = UMAX / NQ > 1 ? UMAX / NQ : 1;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t ci = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (ci >= n_chunks) return;
  const uint4 ch = chunks[ci];
  const uint32_t ga = ch.x >> 16, gb = ch.x & 0xFFFFu;
  const uint32_t da = grp_dim[ga], db = grp_dim[gb];
  const uint32_t ne = da * db;
  const bool diag = ga == gb;
  Acc<T> accS[NQ], accB[NQ], accR;
  uint32_t er[NQ], ec[NQ];
  bool on[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const uint32_t e = lane + 64 * q;
    on[q] = e < ne;
    er[q] = on[q] ? e % da : 0;
    ec[q] = on[q] ? e / da : 0;
  }
  const bool rlane = diag && lane < da;
  for (uint32_t p = ch.y; p < ch.z; p += U) {
    uint2 pr[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = p + u < ch.z;
      pr[u] = ok[u] ? pairs[p + u] : make_uint2(0u, 0u);
    }
    V4 f[U][NQ], e[U][NQ];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t ra = pr[u].x & OFF, rb = pr[u].y & OFF;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (ok[u] && on[q]) {
          f[u][q] = F4[ra + er[q]];
          e[u][q] = E4[rb + ec[q]];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      const uint32_t ra = pr[u].x & OFF, rb = pr[u].y & OFF;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (on[q]) accS[q].add(f[u][q].x * e[u][q].x + f[u][q].y * e[u][q].y + f[u][q].z * e[u][q].z);
      if (pr[u].x & HI) {
        // same observation: a's point-major first row from b's
        const bool self = (pr[u].y & HI) != 0u;
        const uint32_t raj = self ? rb : ((pr[u].x & BEFORE) ? rb - da : rb + db);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (on[q]) {
            const V4 ja = J4[raj + er[q]];
            const V4 jb = J4[rb + ec[q]];
            accB[q].add(ja.z * jb.x + ja.w * jb.y);
          }
        }
      }
      if (rlane && (pr[u].y & HI)) accR.add(F4[ra + lane].w - E4[rb + lane].w);
    }
  }
  const uint32_t oa = grp_off[ga], ob = grp_off[gb];
  if (ch.w == NONE) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (on[q]) {
        const uint32_t r = er[q], c = ec[q];
        if (diag && r < c) continue;
        const double bv = accB[q].value();
        double v = bv - accS[q].value();
        if (diag && r == c) v = (bv + bv * lambda) - accS[q].value();
        S[(size_t)(ob + c) * ld + oa + r] = v;
      }
    }
    if (rlane) S[(size_t)(oa + lane) * ld + Dp] = accR.value();
  } else {
    double* pp = parts + (size_t)ch.w * PART_STRIDE;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (on[q]) {
        pp[lane + 64 * q] = accB[q].value();
        pp[196 + lane + 64 * q] = accS[q].value();
      }
    }
    if (rlane) pp[392 + lane] = accR.value();
  }
}

// Destinations split over several chunks: sum the partials in chunk order.
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
    double bv = 0.0, sv = 0.0;
    for (uint32_t k = 0; k < np; ++k) {
      const double* pp = parts + (size_t)(p0 + k) * PART_STRIDE;
      bv += pp[e];
      sv += pp[196 + e];
    }
    double v = bv - sv;
    if (diag && r == c) v = (bv + bv * lambda) - sv;
    S[(size_t)(ob + c) * ld + oa + r] = v;
  }
  if (diag && lane < da) {
    double rv = 0.0;
    for (uint32_t k = 0; k < np; ++k) rv += parts[(size_t)(p0 + k) * PART_STRIDE + 392 + lane];
    S[(size_t)(oa + lane) * ld + Dp] = rv;
  }
}

// ------------------------------------------------- back-substitution + update
template <typename T>
__global__ __launch_bounds__(256) void k_backsub(
    uint32_t n_pt, const uint32_t* __restrict__ pt_inc_ptr, const uint32_t* __restrict__ inc_grp,
    const uint32_t* __restrict__ inc_roff, const uint32_t* __restrict__ grp_off,
    const uint32_t* __restrict__ grp_dim, const typename Vec4<T>::type* __restrict__ E4,
    const double* __restrict__ pt_g, const double* __restrict__ pt_cinv,
    const uint8_t* __restrict__ pt_ok, const double* __restrict__ delta,
    const double* __restrict__ pts_cur, double* __restrict__ pts_trial) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pt) return;
  const double* X = pts_cur + 3 * (size_t)p;
  double* Y = pts_trial + 3 * (size_t)p;
  if (!pt_ok[p]) {
    Y[0] = X[0];
    Y[1] = X[1];
    Y[2] = X[2];
    return;
  }
  double acc0 = pt_g[3 * (size_t)p], acc1 = pt_g[3 * (size_t)p + 1], acc2 = pt_g[3 * (size_t)p + 2];
  for (uint32_t a = pt_inc_ptr[p]; a < pt_inc_ptr[p + 1]; ++a) {
    const uint32_t g = inc_grp[a];
    const uint32_t r0 = inc_roff[a], n = grp_dim[g], off = grp_off[g];
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
      const auto e = E4[r0 + r];
      const double d = delta[off + r];
      c0 += (double)e.x * d;
      c1 += (double)e.y * d;
      c2 += (double)e.z * d;
    }
    acc0 += c0;
    acc1 += c1;
    acc2 += c2;
  }
  const double* ci = pt_cinv + 9 * (size_t)p;
  Y[0] = X[0] + -(ci[0] * acc0 + ci[1] * acc1 + ci[2] * acc2);
  Y[1] = X[1] + -(ci[3] * acc0 + ci[4] * acc1 + ci[5] * acc2);
  Y[2] = X[2] + -(ci[6] * acc0 + ci[7] * acc1 + ci[8] * acc2);
}

// perturb_pose: R' = exp(omega) R (nalgebra from_scaled_axis: identity when
// |omega/2|^2 <= eps^2, no renormalization), t' = t + dt.
__global__ __launch_bounds__(256) void k_update_poses(uint32_t n_img, const uint32_t* __restrict__ img_slot,
                                                      const double* __restrict__ delta,
                                                      const double* __restrict__ cur,
                                                      double* __restrict__ trial) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_img) return;
  const double* p = cur + 7 * (size_t)i;
  double* o = trial + 7 * (size_t)i;
  const uint32_t s = img_slot[i];
  double d[6] = {0, 0, 0, 0, 0, 0};
  if (s != NONE)
    for (int k = 0; k < 6; ++k) d[k] = delta[s + k];
  const double hx = d[0] / 2.0, hy = d[1] / 2.0, hz = d[2] / 2.0;
  const double nn = hx * hx + hy * hy + hz * hz;
  const double eps = 2.220446049250313e-16;
  double dw = 1.0, dx = 0.0, dy = 0.0, dz = 0.0;
  if (!(nn <= eps * eps)) {
    const double n = sqrt(nn);
    const double sn = sin(n) / n;
    dx = hx * sn;
    dy = hy * sn;
    dz = hz * sn;
    dw = cos(n);
  }
  o[0] = dw * p[0] - dx * p[1] - dy * p[2] - dz * p[3];
  o[1] = dw * p[1] + dx * p[0] + dy * p[3] - dz * p[2];
  o[2] = dw * p[2] - dx * p[3] + dy * p[0] + dz * p[1];
  o[3] = dw * p[3] + dx * p[2] - dy * p[1] + dz * p[0];
  o[4] = p[4] + d[3];
  o[5] = p[5] + d[4];
  o[6] = p[6] + d[5];
}

__global__ __launch_bounds__(256) void k_update_cams(uint32_t n_cam, const uint32_t* __restrict__ cam_slot,
                                                     const uint8_t* __restrict__ cam_free,
                                                     const uint32_t* __restrict__ cam_nfree,
                                                     const double* __restrict__ delta,
                                                     const double* __restrict__ cur,
                                                     double* __restrict__ trial) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cam) return;
  double p[8];
  for (int k = 0; k < 8; ++k) p[k] = cur[8 * (size_t)c + k];
  const uint32_t s = cam_slot[c];
  if (s != NONE) {
    const uint32_t nf = cam_nfree[c];
    for (uint32_t j = 0; j < nf; ++j) p[cam_free[8 * (size_t)c + j]] += delta[s + j];
  }
  for (int k = 0; k < 8; ++k) trial[8 * (size_t)c + k] = p[k];
}

// ------------------------------------------------------------------- cost
// mode 0: sum w(|r|) |r|^2 (total_cost); mode 1: sum |r| (the numerator of
// mean_reprojection_error).  Both also count the observations in front of
// their camera.  Deterministic two-level reduction.
constexpr int COST_BLOCK = 256;
constexpr int COST_ITEMS = 8;

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
#pragma unroll
  for (int it = 0; it < COST_ITEMS; ++it) {
    const uint32_t o = base + it * COST_BLOCK;
    if (o < n_obs) {
      const uint32_t img = obs_img[o], pt = obs_pt[o], cam = cam_of_img[img];
      double c0, c1, c2;
      transform_point(poses + 7 * (size_t)img, pts[3 * (size_t)pt], pts[3 * (size_t)pt + 1],
                      pts[3 * (size_t)pt + 2], c0, c1, c2);
      if (!(c2 <= 1e-9)) {
        const double2 xy = obs_xy[o];
        double r0, r1;
        project_residual(cam_model[cam], cams + 8 * (size_t)cam, c0, c1, c2, xy.x, xy.y, r0, r1);
        const double n = sqrt(r0 * r0 + r1 * r1);
        if (mode == 0) {
          const double w = robust_weight(n, loss, scale);
          s0 += w * n * n;
        } else {
          s0 += n;
        }
        s1 += 1.0;
      }
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

// ------------------------------------------------------------- launchers
static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

template <typename T>
void launch_linearize(const Layout& L, Work<T>& W, int cur, int loss, double scale, hipStream_t s) {
  using V4 = typename Vec4<T>::type;
  if (L.n_obs == 0) return;
  hipLaunchKernelGGL(k_linearize<T>, dim3(blocks_for(L.n_obs, 256)), dim3(256), 0, s, L.n_obs,
                     L.obs_img, L.obs_pt, L.obs_xy, L.obs_inc, L.inc_roff, L.cam_of_img,
                     L.cam_model, L.img_slot, L.cam_slot, L.cam_free, L.cam_nfree, W.poses[cur],
                     W.cams[cur], W.pts[cur], loss, scale, reinterpret_cast<V4*>(W.J4),
                     reinterpret_cast<V4*>(W.E4), W.obs_pc, W.obs_valid);
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_point_accumulate(const Layout& L, Work<T>& W, hipStream_t s) {
  if (L.n_pt == 0) return;
  hipLaunchKernelGGL(k_point_accumulate<T>, dim3(blocks_for(L.n_pt, 256)), dim3(256), 0, s, L.n_pt,
                     L.pt_ptr, W.obs_pc, W.obs_valid, L.obs_inc, W.pt_C, W.pt_g, W.pt_active);
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_point_prepare(const Layout& L, Work<T>& W, double lambda, hipStream_t s) {
  using V4 = typename Vec4<T>::type;
  if (L.n_pt == 0) return;
  hipLaunchKernelGGL(k_point_prepare, dim3(blocks_for(L.n_pt, 256)), dim3(256), 0, s, L.n_pt,
                     W.pt_C, W.pt_active, lambda, W.pt_cinv, W.pt_ok);
  HIP_CHECK(hipGetLastError());
  if (L.n_inc == 0) return;
  hipLaunchKernelGGL(k_compute_F<T>, dim3(blocks_for(L.n_inc, 256)), dim3(256), 0, s, L.n_inc,
                     L.inc_pt, L.inc_roff, L.inc_froff, L.inc_grp, L.grp_dim, W.pt_cinv, W.pt_g,
                     reinterpret_cast<const V4*>(W.E4), reinterpret_cast<V4*>(W.F4));
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_schur(const Layout& L, Work<T>& W, double lambda, hipStream_t s) {
  using V4 = typename Vec4<T>::type;
  if (L.n_chunks) {
    const unsigned grid = blocks_for(L.n_chunks, 4);
    const uint32_t ne = L.max_grp_dim * L.max_grp_dim;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, L.n_chunks, L.chunks,
                         reinterpret_cast<const uint2*>(L.pairs), L.grp_off, L.grp_dim,
                         reinterpret_cast<const V4*>(W.J4), reinterpret_cast<const V4*>(W.E4),
                         reinterpret_cast<const V4*>(W.F4), lambda, W.S, L.ld, L.Dp, W.parts);
    };
    if (ne <= 64) go(k_schur<T, 1>);
    else if (ne <= 128) go(k_schur<T, 2>);
    else if (ne <= 192) go(k_schur<T, 3>);
    else go(k_schur<T, 4>);
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
  using V4 = typename Vec4<T>::type;
  const int nxt = 1 - cur;
  if (L.n_pt) {
    hipLaunchKernelGGL(k_backsub<T>, dim3(blocks_for(L.n_pt, 256)), dim3(256), 0, s, L.n_pt,
                       L.pt_inc_ptr, L.inc_grp, L.inc_roff, L.grp_off, L.grp_dim,
                       reinterpret_cast<const V4*>(W.E4), W.pt_g, W.pt_cinv, W.pt_ok, W.delta,
                       W.pts[cur], W.pts[nxt]);
    HIP_CHECK(hipGetLastError());
  }
  if (L.n_img) {
    hipLaunchKernelGGL(k_update_poses, dim3(blocks_for(L.n_img, 256)), dim3(256), 0, s, L.n_img,
                       L.img_slot, W.delta, W.poses[cur], W.poses[nxt]);
    HIP_CHECK(hipGetLastError());
  }
  if (L.n_cam) {
    hipLaunchKernelGGL(k_update_cams, dim3(blocks_for(L.n_cam, 256)), dim3(256), 0, s, L.n_cam,
                       L.cam_slot, L.cam_free, L.cam_nfree, W.delta, W.cams[cur], W.cams[nxt]);
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
