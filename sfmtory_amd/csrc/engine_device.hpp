// Device-side view of one problem and the per-element work of an LM try.
//
// The grid-wide kernels (engine_kernels.hip) call these functions with one
// thread per element; the persistent window solver (window_kernels.hip)
// calls the same functions from one workgroup per problem.  Reference
// correspondence (crates/sfm-ba/src/lib.rs):
//   linearize_obs      hot loop A, 832-853 (analytic_jacobians + robust_weight)
//                      and the per-observation terms of hot loop B, 880-909
//   point_accumulate   c_diag / c_rhs sums, 885-886; the points_to_obs test, 960
//   point_prepare      damped point-block inverse, 958-965
//   compute_F / schur_chunk (schur_chunk_mfma8: the same reduction on the
//   matrix cores), for the window solver's row streams; compute_Y /
//   schur_chunk_rec, from incidence records, for the single-problem solve
//                      damped diagonal and cross blocks, 923-956, and the
//                      point elimination S -= E C^-1 E^T, rhs += E C^-1 g of
//                      accumulate_point_schur, 112-211 / 967-1043
//   backsub_point      delta_point = -C^-1 (g + E^T delta), 1096-1120
//   update_pose / update_cam
//                      trial state, 1122-1164 (perturb_pose 263-269)
//   obs_cost           one term of total_cost 727-745 / mean_reprojection_error 708-725
#pragma once

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

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// Pointers and sizes of one problem on the device (layout + work arrays).
template <typename T>
struct View {
  using V4 = typename Vec4<T>::type;
  using V2 = typename Vec2<T>::type;
  uint32_t n_img, n_cam, n_pt, n_obs, n_inc, n_chunks;
  uint32_t D, Dp, ld;
  const uint32_t* obs_img;
  const uint32_t* obs_pt;
  const double2* obs_xy;
  const uint32_t* obs_inc;
  const uint32_t* pt_ptr;
  const uint32_t* inc_grp;
  const uint32_t* inc_pt;
  const uint32_t* inc_roff;
  const uint32_t* inc_froff;
  const uint32_t* pt_inc_ptr;
  const uint32_t* cam_of_img;
  const int32_t* cam_model;
  const uint32_t* img_slot;
  const uint32_t* cam_slot;
  const uint8_t* cam_free;
  const uint32_t* cam_nfree;
  const uint32_t* grp_off;
  const uint32_t* grp_dim;
  const uint2* pairs;
  const uint4* chunks;
  V2* J2;
  V4* E4;
  V4* F4;
  T* rec;
  T* obs_pc;
  uint8_t* obs_valid;
  double* pt_C;
  double* pt_g;
  uint8_t* pt_active;
  double* pt_cinv;
  uint8_t* pt_ok;
  double* S;
  double* delta;
  double* poses[2];
  double* cams[2];
  double* pts[2];
};

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

// ---------------------------------------------------------------- linearize
// Observation o (point-major order): for each of its incidences one J2 and
// one E4 row per unknown (DB = 0), or its incidence records (DB > 0: X2 and
// R2, and one J2 row per unknown; rec_values), and its contribution to its
// point's 3x3 block and gradient (obs_pc).  With LPO > 1, LPO lanes share
// observation o: each computes the same values and lane r writes rows r,
// r + LPO, ... and obs_pc entries r, r + LPO, ..., so that one store
// instruction writes whole rows of consecutive observations (a full cache
// line per observation) instead of one 16- or 32-byte piece per lane.
// pc_reg (LPO = 1): the point terms go there (registers of the caller)
// instead of obs_pc.
template <typename T, int LPO = 1, int DB = 0>
__device__ inline void linearize_obs(const View<T>& v, uint32_t o, int cur, int loss, double scale,
                                     uint32_t r = 0, typename Vec2<T>::type* stage = nullptr,
                                     T* pc_reg = nullptr) {
  using V4 = typename Vec4<T>::type;
  using V2 = typename Vec2<T>::type;
  const uint32_t img = v.obs_img[o];
  const uint32_t pt = v.obs_pt[o];
  const uint32_t cam = v.cam_of_img[img];
  const int model = v.cam_model[cam];
  const double* pose = v.poses[cur] + 7 * (size_t)img;
  const double* P = v.cams[cur] + 8 * (size_t)cam;
  const double* Xp = v.pts[cur] + 3 * (size_t)pt;
  double c0, c1, c2;
  transform_point(pose, Xp[0], Xp[1], Xp[2], c0, c1, c2);

  const bool pose_free = v.img_slot[img] != NONE;
  const bool cam_is_free = v.cam_slot[cam] != NONE;
  const uint32_t kf = cam_is_free ? v.cam_nfree[cam] : 0u;
  const uint32_t ns = (pose_free ? 6u : 0u) + kf;
  const uint32_t inc0 = v.obs_inc[2 * o], inc1 = v.obs_inc[2 * o + 1];
  const uint32_t roff0 = inc0 != NONE ? v.inc_roff[inc0] : 0u;
  const bool two = inc1 != NONE;
  const uint32_t roff1 = two ? v.inc_roff[inc1] : 0u;
  T* pc_out = v.obs_pc + 9 * (size_t)o;
  // DB > 0: the incidences' records, in V2 units (an observation without a
  // free unknown has none and writes no record)
  V2* rec0 = nullptr;
  V2* rec1 = nullptr;
  if constexpr (DB > 0) {
    constexpr size_t RS2 = rec_values(DB) / 2;
    if (inc0 != NONE) rec0 = reinterpret_cast<V2*>(v.rec) + (size_t)inc0 * RS2;
    if (two) rec1 = reinterpret_cast<V2*>(v.rec) + (size_t)inc1 * RS2;
  }
  // With a stage (LPO = 1), an observation of one incidence leaves its
  // record pieces there (V2 k of X2 R2 at k, J2 row s at 4 + s) for the
  // caller to store; one of two incidences writes its records directly.
  const bool staged = DB > 0 && stage != nullptr && !two;
  auto put_xr = [&](uint32_t k, V2 val) {
    if (staged) {
      stage[k] = val;
    } else {
      if (rec0) rec0[k] = val;
      if (rec1) rec1[k] = val;
    }
  };
  auto put_row = [&](uint32_t s, V2 val) {
    if (staged)
      stage[4 + s] = val;
    else
      *((two && s >= 6) ? rec1 + 4 + (s - 6) : rec0 + 4 + s) = val;
  };

  if (c2 <= 1e-9) {
    // Behind the camera: dropped for this iteration (lib.rs:339-341).
    if (r == 0) v.obs_valid[o] = 0;
#pragma unroll
    for (uint32_t k0 = 0; k0 < 9; k0 += LPO)
      if (k0 + r < 9) {
        if (pc_reg)
          pc_reg[k0 + r] = T(0);
        else
          pc_out[k0 + r] = T(0);
      }
    const V4 z = {T(0), T(0), T(0), T(0)};
    const V2 z2 = {T(0), T(0)};
    if constexpr (DB > 0) {
#pragma unroll
      for (uint32_t k0 = 0; k0 < 4; k0 += LPO)
        if (k0 + r < 4 && rec0) put_xr(k0 + r, z2);
    }
#pragma unroll
    for (uint32_t s0 = 0; s0 < 14; s0 += LPO) {
      const uint32_t s = s0 + r;
      if (s >= ns) break;
      if constexpr (DB > 0) {
        put_row(s, z2);
      } else {
        const uint32_t row = (two && s >= 6) ? roff1 + (s - 6) : roff0 + s;
        v.J2[row] = z2;
        v.E4[row] = z;
      }
    }
    return;
  }
  if (r == 0) v.obs_valid[o] = 1;
  const double2 xy = v.obs_xy[o];
  double r0, r1;
  linearization_residual(model, P, c0, c1, c2, xy.x, xy.y, r0, r1);
  // Every normal-equation term of the observation is built from one scaled
  // Jacobian [sqrt(w) J | sqrt(w) Jx] and residual sqrt(w) r, rounded once:
  // B = J2^T J2, E = J2^T X2, g = J2^T R2, C = X2^T X2, g_p = X2^T R2.  The
  // terms are then exactly the Gram matrix of one rounded Jacobian, whose
  // Schur complement stays positive semi-definite up to the rounding of
  // the sums (with w and sqrt(w) rounded separately it does not, and an
  // FP32 assembly loses definiteness at a larger damping).
  const double wd = robust_weight(sqrt(r0 * r0 + r1 * r1), loss, scale);
  const T sw = (T)sqrt(wd);

  T du[3], dv[3], jc0[8], jc1[8];
  model_jacobian<T>(model, P, (T)c0, (T)c1, (T)c2, du, dv, jc0, jc1);
  T R[9];
  rotation_matrix<T>(pose, R);
  T jp[12], jx[6];
  pose_point_jacobians<T>(R, (T)(c0 - pose[4]), (T)(c1 - pose[5]), (T)(c2 - pose[6]), du, dv, jp,
                          jx);
  const T sr0 = (T)(sqrt(wd) * r0), sr1 = (T)(sqrt(wd) * r1);
  T sx[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) sx[q] = sw * jx[q];

  // Point terms: X2^T X2 (upper 6), X2^T R2.
  T pc[9];
  pc[0] = sx[0] * sx[0] + sx[3] * sx[3];
  pc[1] = sx[0] * sx[1] + sx[3] * sx[4];
  pc[2] = sx[0] * sx[2] + sx[3] * sx[5];
  pc[3] = sx[1] * sx[1] + sx[4] * sx[4];
  pc[4] = sx[1] * sx[2] + sx[4] * sx[5];
  pc[5] = sx[2] * sx[2] + sx[5] * sx[5];
  pc[6] = sx[0] * sr0 + sx[3] * sr1;
  pc[7] = sx[1] * sr0 + sx[4] * sr1;
  pc[8] = sx[2] * sr0 + sx[5] * sr1;
#pragma unroll
  for (uint32_t k0 = 0; k0 < 9; k0 += LPO) {
    T val = pc[k0];
#pragma unroll
    for (uint32_t k = k0 + 1; k < k0 + LPO && k < 9; ++k)
      if (k == k0 + r) val = pc[k];
    if (k0 + r < 9) {
      if (pc_reg)
        pc_reg[k0 + r] = val;
      else
        pc_out[k0 + r] = val;
    }
  }
  if constexpr (DB > 0) {
    // X2 (three V2) and R2 into every record of the observation
#pragma unroll
    for (uint32_t k0 = 0; k0 < 4; k0 += LPO) {
      const uint32_t k = k0 + r;
      if (k < 4) {
        const T x0 = k == 0 ? sx[0] : k == 1 ? sx[2] : k == 2 ? sx[4] : sr0;
        const T x1 = k == 0 ? sx[1] : k == 1 ? sx[3] : k == 2 ? sx[5] : sr1;
        if (rec0) put_xr(k, V2{x0, x1});
      }
    }
  }

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
  const uint8_t* fi = v.cam_free + 8 * (size_t)cam;
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
  for (uint32_t q0 = 0; q0 < 14; q0 += LPO) {
    const uint32_t s = q0 + r;
    if (s >= ns) break;
    T a0 = s0[q0], a1 = s1[q0];
#pragma unroll
    for (uint32_t q = q0 + 1; q < q0 + LPO && q < 14; ++q)
      if (q == s) {
        a0 = s0[q];
        a1 = s1[q];
      }
    const T j0 = sw * a0, j1 = sw * a1;
    if constexpr (DB > 0) {
      put_row(s, V2{j0, j1});
    } else {
      const uint32_t row = (two && s >= 6) ? roff1 + (s - 6) : roff0 + s;
      v.J2[row] = V2{j0, j1};
      v.E4[row] = V4{j0 * sx[0] + j1 * sx[3], j0 * sx[1] + j1 * sx[4], j0 * sx[2] + j1 * sx[5], j0 * sr0 + j1 * sr1};
    }
  }
}

// ------------------------------------------------------- point accumulation
// Point p's 3x3 block and gradient from its observations' terms.  With
// LPP > 1, LPP consecutive lanes share the point: lane r sums observations
// r, r + LPP, ... (so neighbouring lanes read neighbouring observations'
// terms), then the lanes' sums are combined by a fixed butterfly and lane 0
// writes (LPP a power of two <= 64, the lanes of one point in one
// wavefront).
template <typename T, int LPP = 1>
__device__ inline void point_accumulate(const View<T>& v, uint32_t p, uint32_t r = 0) {
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) a[k] = 0.0;
  int active = 0;
  for (uint32_t o = v.pt_ptr[p] + r; o < v.pt_ptr[p + 1]; o += LPP) {
    if (!v.obs_valid[o]) continue;
    if (v.obs_inc[2 * o] != NONE) active = 1;
    const T* c = v.obs_pc + 9 * (size_t)o;
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] += (double)c[k];
  }
#pragma unroll
  for (int m = 1; m < LPP; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] += __shfl_xor(a[k], m);
    active |= __shfl_xor(active, m);
  }
  if (r != 0) return;
#pragma unroll
  for (int k = 0; k < 6; ++k) v.pt_C[6 * (size_t)p + k] = a[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) v.pt_g[3 * (size_t)p + k] = a[6 + k];
  v.pt_active[p] = (uint8_t)active;
}

// Damped point block (C + lambda diag C)^-1 by the adjugate; a point without
// free couplings or with det == 0 is left out of the Schur complement and
// keeps its position (lib.rs:958-965, 1001, 1102).
template <typename T>
__device__ inline void point_prepare(const View<T>& v, uint32_t p, double lambda) {
  double* o = v.pt_cinv + 9 * (size_t)p;
  const double* C = v.pt_C + 6 * (size_t)p;
  bool ok = false;
  if (v.pt_active[p]) {
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
  v.pt_ok[p] = ok ? 1 : 0;
}

// F = E Cinv and F . g (the incidence's rhs share) for incidence a; F is
// stored group-major.
template <typename T>
__device__ inline void compute_F(const View<T>& v, uint32_t a) {
  using V4 = typename Vec4<T>::type;
  const uint32_t pt = v.inc_pt[a];
  const double* ci = v.pt_cinv + 9 * (size_t)pt;
  T c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) c[k] = (T)ci[k];
  const T g0 = (T)v.pt_g[3 * (size_t)pt], g1 = (T)v.pt_g[3 * (size_t)pt + 1],
          g2 = (T)v.pt_g[3 * (size_t)pt + 2];
  const uint32_t r0 = v.inc_roff[a], f0ff = v.inc_froff[a], n = v.grp_dim[v.inc_grp[a]];
  for (uint32_t r = 0; r < n; ++r) {
    const V4 e = v.E4[r0 + r];
    const T f0 = e.x * c[0] + e.y * c[3] + e.z * c[6];
    const T f1 = e.x * c[1] + e.y * c[4] + e.z * c[7];
    const T f2 = e.x * c[2] + e.y * c[5] + e.z * c[8];
    // w: this row's share of rhs, F . g_point = E Cinv g_point
    v.F4[f0ff + r] = V4{f0, f1, f2, f0 * g0 + f1 * g1 + f2 * g2};
  }
}

constexpr int PART_STRIDE = 2 * 196 + 16;

// One wavefront reduces one chunk of a destination block (ga >= gb) of the
// reduced system; lane l owns entries l, l + 64, ... (NQ of them) of the
// da x db block.  Sums, over the chunk's point pairs (a, b), B = w J_a^T J_b
// for pairs of the same observation and Schur = F_a E_b^T = E_a Cinv E_b^T;
// on diagonal blocks also the right-hand side F_a g_point - w J_a^T r over
// pairs a == b (F4.w - E4.w).  A pair is two words: a's first row with bit
// 31 "same observation", and b's first row with bit 31 "a == b" (J, E and F
// rows share one group-major row numbering).  U pairs are processed per
// step with all their loads issued together.  A chunk (ch.w == NONE) that
// covers its whole destination is written straight into S (with the
// Marquardt damping of the diagonal); others leave partials.
// UMAX: pairs in flight per step (the grid kernel's default measured best
// on gfx950 at full occupancy).
template <typename T, int NQ, int UMAX = (sizeof(T) == 4 ? 4 : 2)>
__device__ inline void schur_chunk(const View<T>& v, const uint4 ch, double lambda, double* parts) {
  using V4 = typename Vec4<T>::type;
  constexpr uint32_t HI = 0x80000000u, OFF = 0x7FFFFFFFu;
  constexpr int U = UMAX / NQ > 1 ? UMAX / NQ : 1;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t ga = ch.x >> 16, gb = ch.x & 0xFFFFu;
  const uint32_t da = v.grp_dim[ga], db = v.grp_dim[gb];
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
      pr[u] = ok[u] ? v.pairs[p + u] : make_uint2(0u, 0u);
    }
    V4 f[U][NQ], e[U][NQ];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t ra = pr[u].x & OFF, rb = pr[u].y & OFF;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (ok[u] && on[q]) {
          f[u][q] = v.F4[ra + er[q]];
          e[u][q] = v.E4[rb + ec[q]];
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
        // same observation: w J_a^T J_b from the J rows of a and b
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (on[q]) {
            const auto ja = v.J2[ra + er[q]];
            const auto jb = v.J2[rb + ec[q]];
            accB[q].add(ja.x * jb.x + ja.y * jb.y);
          }
        }
      }
      if (rlane && (pr[u].y & HI)) accR.add(v.F4[ra + lane].w - v.E4[rb + lane].w);
    }
  }
  const uint32_t oa = v.grp_off[ga], ob = v.grp_off[gb];
  if (ch.w == NONE) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (on[q]) {
        const uint32_t r = er[q], c = ec[q];
        if (diag && r < c) continue;
        const double bv = accB[q].value();
        double val = bv - accS[q].value();
        if (diag && r == c) val = (bv + bv * lambda) - accS[q].value();
        v.S[(size_t)(ob + c) * v.ld + oa + r] = val;
      }
    }
    if (rlane) v.S[(size_t)(oa + lane) * v.ld + v.Dp] = accR.value();
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

// ----------------------------------------------------- MFMA Schur chunk
// 16x16x4 matrix-core product for T: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; register v of lane l holds D[row(l, v)][l % 16], with
// row 4 (l / 16) + v in FP32 and l / 16 + 4 v in FP64 (checked on gfx950
// with a one-wavefront probe of both builtins).
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __attribute__((always_inline)) inline acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __attribute__((always_inline)) inline uint32_t row(uint32_t lane, int v) {
    return 4u * (lane >> 4) + (uint32_t)v;
  }
};
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __attribute__((always_inline)) inline acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __attribute__((always_inline)) inline uint32_t row(uint32_t lane, int v) {
    return (lane >> 4) + 4u * (uint32_t)v;
  }
};

// The destination-block reduction of schur_chunk on the matrix cores, for
// blocks of at most 8x8 (da, db <= 8; the window solver's row-stream
// layout).  Per pair, lane l supplies one component, F_a[row][k] and
// E_b[row][k] for the three point coordinates (K slot 3 zero), so one
// 16x16x4 MFMA adds F_a E_b^T; a pair of the same observation adds
// w J_a^T J_b with a second MFMA over the two residual components.  Two
// pairs per MFMA: rows 0-7 of both operands carry the even pair of a step, rows 8-15
// the odd pair, so the 16x16 tile's two diagonal quadrants sum the even and
// the odd pairs of the chunk and its off-diagonal quadrants (cross terms)
// are discarded.  Lane l loads row l % 8 of pair 2u + (l / 8) % 2,
// component l / 16: per load instruction twice the pairs of the one-pair
// form, at the same registers.  The quadrants are added at the end (even +
// odd); the FP32 instantiation flushes after every U steps, so no more than
// U terms are summed in FP32 per quadrant.
// LONG: chunks of any length (the window solver's one chunk per
// destination), taken 128 pairs at a time; otherwise at most PAIR_CHUNK.
template <typename T, int U, bool LONG = false>
__device__ inline void schur_chunk_mfma8(const View<T>& v, const uint4 ch, double lambda, double* parts) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr uint32_t HI = 0x80000000u, OFF = 0x7FFFFFFFu;
  static_assert(PAIR_CHUNK <= 128, "a chunk's pair words are held in two registers per lane");
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t i = lane & 7, h = (lane >> 3) & 1, k = lane >> 4;
  const uint32_t ga = ch.x >> 16, gb = ch.x & 0xFFFFu;
  const uint32_t da = v.grp_dim[ga], db = v.grp_dim[gb];
  const bool diag = ga == gb;
  const bool fa_on = i < da && k < 3, eb_on = i < db && k < 3;
  const bool ja_on = i < da && k < 2, jb_on = i < db && k < 2;
  const T* F = reinterpret_cast<const T*>(v.F4);
  const T* E = reinterpret_cast<const T*>(v.E4);
  const T* J = reinterpret_cast<const T*>(v.J2);
  const uint32_t n_all = ch.z - ch.y;
  acc_t s = {T(0), T(0), T(0), T(0)}, b = {T(0), T(0), T(0), T(0)};
  double ss[4] = {0.0, 0.0, 0.0, 0.0}, bs[4] = {0.0, 0.0, 0.0, 0.0};
  Acc<T> accR;
  const bool rlane = diag && lane < da;
  for (uint32_t base = 0; base < (LONG ? n_all : 1u); base += 128) {
    const uint32_t n = LONG ? min(128u, n_all - base) : n_all;
    const uint32_t p0 = ch.y + base;
    const uint2 w0 = lane < n ? v.pairs[p0 + lane] : make_uint2(0u, 0u);
    const uint2 w1 = lane + 64 < n ? v.pairs[p0 + 64 + lane] : make_uint2(0u, 0u);
    // Pair q's words (zero for q >= n).
    auto word = [&](uint32_t q, uint32_t& x, uint32_t& y) {
      x = 0u;
      y = 0u;
      if (q < n) {
        const uint32_t sl = q & 63;
        x = __builtin_amdgcn_readlane(q < 64 ? w0.x : w1.x, sl);
        y = __builtin_amdgcn_readlane(q < 64 ? w0.y : w1.y, sl);
      }
    };
    struct Stage {
      T fa[U], eb[U], ja[U], jb[U];
    };
    // Operands of U steps (2U pairs) starting at pair q.
    auto load = [&](uint32_t q, Stage& st) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        uint32_t x0, y0, x1, y1;
        word(q + 2 * u, x0, y0);
        word(q + 2 * u + 1, x1, y1);
        const bool ok = q + 2 * u + h < n;
        const uint32_t px = h ? x1 : x0, py = h ? y1 : y0;
        const uint32_t ra = px & OFF, rb = py & OFF;
        st.fa[u] = (ok && fa_on) ? F[4 * (size_t)(ra + i) + k] : T(0);
        st.eb[u] = (ok && eb_on) ? E[4 * (size_t)(rb + i) + k] : T(0);
        const bool same = ok && (px & HI);
        st.ja[u] = (same && ja_on) ? J[2 * (size_t)(ra + i) + k] : T(0);
        st.jb[u] = (same && jb_on) ? J[2 * (size_t)(rb + i) + k] : T(0);
      }
    };
    Stage cur, nxt;
    load(0, cur);
    for (uint32_t q = 0; q < n; q += 2 * U) {
      if (q + 2 * U < n) load(q + 2 * U, nxt);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (q + 2 * u >= n) break;
        uint32_t x0, y0, x1, y1;
        word(q + 2 * u, x0, y0);
        word(q + 2 * u + 1, x1, y1);
        s = M::step(cur.fa[u], cur.eb[u], s);
        if ((x0 | x1) & HI) b = M::step(cur.ja[u], cur.jb[u], b);
        // right-hand side in pair order
        if (rlane && (y0 & HI)) accR.add(v.F4[(x0 & OFF) + lane].w - v.E4[(y0 & OFF) + lane].w);
        if (rlane && (y1 & HI)) accR.add(v.F4[(x1 & OFF) + lane].w - v.E4[(y1 & OFF) + lane].w);
      }
      if (sizeof(T) == 4) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          ss[c] += (double)s[c];
          bs[c] += (double)b[c];
          s[c] = T(0);
          b[c] = T(0);
        }
      }
      cur = nxt;
    }
  }  // 128-pair blocks
  if (sizeof(T) == 8) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ss[c] = (double)s[c];
      bs[c] = (double)b[c];
    }
  }
  // Quadrant (8 + r, 8 + c) sits 40 lanes up in the same register (FP32
  // layout, row 4 (l / 16) + v) or 8 lanes up two registers on (FP64
  // layout, row l / 16 + 4 v).
  // Only registers holding rows 0-7 are combined (FP64: registers 0 and 1).
#pragma unroll
  for (int q = 0; q < (sizeof(T) == 4 ? 4 : 2); ++q) {
    const int src = sizeof(T) == 4 ? q : q + 2;
    const int dl = sizeof(T) == 4 ? 40 : 8;
    ss[q] += __shfl(ss[src], (int)((lane + dl) & 63));
    bs[q] += __shfl(bs[src], (int)((lane + dl) & 63));
  }
  const uint32_t oa = v.grp_off[ga], ob = v.grp_off[gb];
  const uint32_t c = lane & 15;
  if (ch.w == NONE) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t r = M::row(lane, q);
      if (r >= da || c >= db || (diag && r < c)) continue;
      double val = bs[q] - ss[q];
      if (diag && r == c) val = (bs[q] + bs[q] * lambda) - ss[q];
      v.S[(size_t)(ob + c) * v.ld + oa + r] = val;
    }
    if (rlane) v.S[(size_t)(oa + lane) * v.ld + v.Dp] = accR.value();
  } else {
    double* pp = parts + (size_t)ch.w * PART_STRIDE;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t r = M::row(lane, q);
      if (r >= da || c >= db) continue;
      pp[r + c * da] = bs[q];
      pp[196 + r + c * da] = ss[q];
    }
    if (rlane) pp[392 + lane] = accR.value();
  }
}

// ------------------------------------------------ incidence records
typedef __attribute__((address_space(3))) void lds_void_t;

// Y = X2 Cinv and z = Y g_point - R2 into incidence a's record (rec_values),
// after the damped point inverse.  A point left out of the Schur complement
// has Cinv = 0, so Y = 0 and z = -R2: its pairs add only the camera block
// J2_a^T J2_b of their observation and the gradient.  Rows of F = E Cinv
// and the rhs share F g_point are J2 Y and J2 (Y g_point), so the assembly
// needs no F (or E) rows: E = J2 X2.
template <typename T, int DB>
__device__ inline void compute_Y(const View<T>& v, uint32_t a, uint32_t q) {
  // Four lanes per incidence (q = 0..3): lane q reads and writes 16 bytes
  // (FP32) or 32 bytes (FP64) of the record, so a wavefront's loads and
  // stores cover whole pieces of 16 consecutive records.  The lanes of one
  // incidence share X2 and R2 by shuffles within the group.
  using V2 = typename Vec2<T>::type;
  V2* rc = reinterpret_cast<V2*>(v.rec + (size_t)a * rec_values(DB));
  const V2 w = rc[q];  // q = 0: X2[0][0..1], 1: X2[0][2], X2[1][0], 2: X2[1][1..2], 3: R2
  const int g4 = (int)(threadIdx.x & 63u) & ~3;
  T x[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[2 * k] = __shfl(w.x, g4 + k);
    x[2 * k + 1] = __shfl(w.y, g4 + k);
  }
  const uint32_t pt = v.inc_pt[a];
  const double* ci = v.pt_cinv + 9 * (size_t)pt;
  T c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) c[k] = (T)ci[k];
  const T g0 = (T)v.pt_g[3 * (size_t)pt], g1 = (T)v.pt_g[3 * (size_t)pt + 1],
          g2 = (T)v.pt_g[3 * (size_t)pt + 2];
  T y[6];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) y[3 * r + j] = x[3 * r] * c[j] + x[3 * r + 1] * c[3 + j] + x[3 * r + 2] * c[6 + j];
  V2 o;
  if (q == 3) {
    o.x = (y[0] * g0 + y[1] * g1 + y[2] * g2) - x[6];
    o.y = (y[3] * g0 + y[4] * g1 + y[5] * g2) - x[7];
  } else {
    o.x = q == 0 ? y[0] : q == 1 ? y[2] : y[4];
    o.y = q == 0 ? y[1] : q == 1 ? y[3] : y[5];
  }
  rc[4 + DB + q] = o;
}

// Shapes of schur_chunk_rec for records of DB rows.  A pair's a side (J2
// rows, Y, z) and b side (X2, R2, J2 rows) are SIDE values each, moved by
// LPS lanes of 16 bytes; one global_load_lds instruction (64 lanes, 1 KB of
// LDS) carries NPI sides, and a batch is INS instructions per side.  Blocks
// of at most 8 unknowns take four pairs per MFMA (two per 8x8 diagonal
// quadrant of the 16x16 tile), larger ones two (one per K-slot pair); the
// last step of a batch may be partly empty.  INS, measured on the
// 1,000-camera scene (FP64 / FP32 assembly per try): 1 -> 1.36 / 0.89 ms,
// 2 -> 1.26 / 0.85 ms, 4 -> 1.45 / 1.07 ms, 8 (FP64) -> 1.97 ms.  Two keeps
// 17 KB of LDS per four-wavefront workgroup, so eight wavefronts per SIMD
// (the register limit) are resident; four halve that.
template <typename T, int DB>
struct RecShape {
  static constexpr uint32_t RS = rec_values(DB);
  static constexpr uint32_t SIDE = 8 + 2 * DB;
  static constexpr uint32_t LPS = SIDE * sizeof(T) / 16;
  static constexpr uint32_t NPI = 64 / LPS;
  static constexpr uint32_t INS = 2;
  static constexpr uint32_t BP = INS * NPI;            // pairs per batch
  static constexpr uint32_t IE = 1024 / sizeof(T);     // LDS elements per load instruction
  static constexpr bool QUAD = DB <= 8;
  static constexpr uint32_t PPM = QUAD ? 4 : 2;        // pairs per MFMA
  static constexpr uint32_t LDS = 2 * INS * IE + 4 * BP;  // elements per wavefront
  static_assert(SIDE * sizeof(T) % 16 == 0, "sides of whole 16-byte pieces");
  static_assert(BP <= 64, "one pair per lane in the M pass");
  static_assert(LDS * sizeof(T) % 16 == 0, "16-byte aligned wavefront stages");
  // LDS offset of batch pair p's side (within its side's area)
  static __device__ inline uint32_t at(uint32_t p) { return (p / NPI) * IE + (p % NPI) * SIDE; }
};

// One wavefront reduces one chunk of a destination block (ga >= gb) from
// incidence records.  For a pair (a, b) of one point, with Y_a, X2_b of
// their observations,
//   F_a E_b^T = J2_a (Y_a X2_b^T) J2_b^T,
// and a pair of the same observation also adds J2_a J2_b^T, so the pair
// adds J2_a M J2_b^T with the 2 x 2 M = [same] I - Y_a X2_b^T.  Per batch:
// the pairs' a and b sides move from global memory straight into LDS
// (global_load_lds), every lane with a pair computes its M into LDS, then
// each MFMA step takes PPM pairs: lane l supplies A = (J2_a M)[r][c] and
// B = J2_b[r][c] for its pair, row r and residual component c.  On
// diagonal blocks, the lanes below the block size also sum, in pair order
// over pairs a == b, the right-hand side J2_a z_a (= F_a g_point - J2_a R2)
// and the damping diagonal |J2_a row|^2.  FP32 MFMA sums are flushed into
// FP64 after every step (two products per entry).  A chunk covering its
// whole destination writes S (with the Marquardt damping of the diagonal,
// lib.rs:923-956); others leave partials (sums, damping diagonal, rhs) for
// k_schur_reduce_rec.  lds: RecShape<T, DB>::LDS elements of this wavefront.
template <typename T, int DB>
__device__ inline void schur_chunk_rec(const View<T>& v, const uint4 ch, double lambda, double* parts, T* lds) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using Sh = RecShape<T, DB>;
  constexpr uint32_t HI = 0x80000000u, OFF = 0x7FFFFFFFu;
  static_assert(PAIR_CHUNK <= 128, "a chunk's pair words are held in two registers per lane");
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t ga = ch.x >> 16, gb = ch.x & 0xFFFFu;
  const uint32_t da = v.grp_dim[ga], db = v.grp_dim[gb];
  const bool diag = ga == gb;
  const bool rlane = diag && lane < da;
  // MFMA roles of this lane: row r (a block) / column r (b block), residual
  // component c (its K slot within the pair), pair pi of the step
  const uint32_t c = (lane >> 4) & 1;
  const uint32_t r = Sh::QUAD ? (lane & 7) : (lane & 15);
  const uint32_t pi = Sh::QUAD ? 2 * ((lane >> 3) & 1) + (lane >> 5) : (lane >> 5);
  const uint32_t rr = r < DB ? r : 0u;  // in-record row of the address (masked below)
  const bool a_on = r < da, b_on = r < db;
  T* As = lds;
  T* Bs = lds + Sh::INS * Sh::IE;
  T* Ms = lds + 2 * Sh::INS * Sh::IE;
  const char* recb = reinterpret_cast<const char*>(v.rec);
  constexpr size_t RSB = Sh::RS * sizeof(T);
  const uint32_t n = ch.z - ch.y;
  const uint2 w0 = lane < n ? v.pairs[ch.y + lane] : make_uint2(0u, 0u);
  const uint2 w1 = lane + 64 < n ? v.pairs[ch.y + 64 + lane] : make_uint2(0u, 0u);
  // per-pair flags of the chunk (bit q % 64 of half q / 64): same
  // observation, and a == b
  const uint64_t same_m[2] = {__ballot((w0.x & HI) != 0), __ballot((w1.x & HI) != 0)};
  const uint64_t eq_m[2] = {__ballot((w0.y & HI) != 0), __ballot((w1.y & HI) != 0)};
  // this lane's side slot and 16-byte piece in each load instruction
  const uint32_t slot = lane / Sh::LPS, piece = (lane % Sh::LPS) * 16;
  acc_t s = {T(0), T(0), T(0), T(0)};
  double ss[4] = {0.0, 0.0, 0.0, 0.0};
  Acc<T> accR;
  double bdg = 0.0;
  for (uint32_t q0 = 0; q0 < n; q0 += Sh::BP) {
    const uint32_t nb = min(Sh::BP, n - q0);
    // sides of pairs q0 .. q0 + nb - 1 into LDS; spare lanes and pairs past
    // the chunk's end load pair q0's sides again (never read back)
#pragma unroll
    for (uint32_t j = 0; j < Sh::INS; ++j) {
      if (j * Sh::NPI >= nb) break;
      uint32_t idx = q0 + j * Sh::NPI + slot;
      if (slot >= Sh::NPI || idx >= n) idx = q0;
      const int src = (int)(idx & 63);
      const uint32_t xl = (uint32_t)__shfl((int)w0.x, src), xh = (uint32_t)__shfl((int)w1.x, src);
      const uint32_t yl = (uint32_t)__shfl((int)w0.y, src), yh = (uint32_t)__shfl((int)w1.y, src);
      const uint32_t ra = (idx < 64 ? xl : xh) & OFF, rb = (idx < 64 ? yl : yh) & OFF;
      __builtin_amdgcn_global_load_lds((const void*)(recb + (size_t)ra * RSB + 8 * sizeof(T) + piece),
                                       (lds_void_t*)(As + j * Sh::IE), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(recb + (size_t)rb * RSB + piece),
                                       (lds_void_t*)(Bs + j * Sh::IE), 16, 0, 0);
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // M of every pair of the batch, column-major: Ms[4 p + 2 c + c'] = M[c'][c]
    if (lane < nb) {
      const uint32_t idx = q0 + lane;
      const bool same = (((idx < 64 ? same_m[0] : same_m[1]) >> (idx & 63)) & 1u) != 0;
      const T* ya = As + Sh::at(lane) + 2 * DB;
      const T* xb = Bs + Sh::at(lane);
      const T one = same ? T(1) : T(0);
      Ms[4 * lane + 0] = one - (ya[0] * xb[0] + ya[1] * xb[1] + ya[2] * xb[2]);
      Ms[4 * lane + 1] = -(ya[3] * xb[0] + ya[4] * xb[1] + ya[5] * xb[2]);
      Ms[4 * lane + 2] = -(ya[0] * xb[3] + ya[1] * xb[4] + ya[2] * xb[5]);
      Ms[4 * lane + 3] = one - (ya[3] * xb[3] + ya[4] * xb[4] + ya[5] * xb[5]);
    }
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    for (uint32_t u = 0; u < nb; u += Sh::PPM) {
      const uint32_t p = u + pi;
      const bool ok = p < nb;
      const uint32_t pq = ok ? p : 0u;
      const T* a = As + Sh::at(pq);
      const T* b = Bs + Sh::at(pq);
      const T m0 = Ms[4 * pq + 2 * c], m1 = Ms[4 * pq + 2 * c + 1];
      const T j0 = a[2 * rr], j1 = a[2 * rr + 1];
      const T av = (ok && a_on) ? j0 * m0 + j1 * m1 : T(0);
      const T bv = (ok && b_on) ? b[8 + 2 * rr + c] : T(0);
      s = M::step(av, bv, s);
      if (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ss[q] += (double)s[q];
          s[q] = T(0);
        }
      }
      if (diag) {
        // right-hand side and damping diagonal, in pair order
#pragma unroll
        for (uint32_t t = 0; t < Sh::PPM; ++t) {
          const uint32_t q = u + t, idx = q0 + q;
          if (q < nb && (((idx < 64 ? eq_m[0] : eq_m[1]) >> (idx & 63)) & 1u) && rlane) {
            const T* aq = As + Sh::at(q);
            const T g0 = aq[2 * lane], g1 = aq[2 * lane + 1];
            accR.add(g0 * aq[2 * DB + 6] + g1 * aq[2 * DB + 7]);
            bdg += (double)g0 * (double)g0 + (double)g1 * (double)g1;
          }
        }
      }
    }
    // the next batch's loads overwrite this one: every lane's reads first
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  }
  if (sizeof(T) == 8) {
#pragma unroll
    for (int q = 0; q < 4; ++q) ss[q] = (double)s[q];
  }
  if (Sh::QUAD) {
    // Quadrant (8 + r, 8 + c) sits 40 lanes up in the same register (FP32
    // layout, row 4 (l / 16) + v) or 8 lanes up two registers on (FP64
    // layout, row l / 16 + 4 v); only the registers of rows 0-7 combine.
#pragma unroll
    for (int q = 0; q < (sizeof(T) == 4 ? 4 : 2); ++q) {
      const int src = sizeof(T) == 4 ? q : q + 2;
      const int dl = sizeof(T) == 4 ? 40 : 8;
      ss[q] += __shfl(ss[src], (int)((lane + dl) & 63));
    }
  }
  const uint32_t oa = v.grp_off[ga], ob = v.grp_off[gb];
  const uint32_t col = lane & 15;
  const double bdc = __shfl(bdg, (int)col);  // damping diagonal of row col (diagonal blocks)
  if (ch.w == NONE) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t row = M::row(lane, q);
      if (row >= da || col >= db || (diag && row < col)) continue;
      double val = ss[q];
      if (diag && row == col) val += lambda * bdc;
      v.S[(size_t)(ob + col) * v.ld + oa + row] = val;
    }
    if (rlane) v.S[(size_t)(oa + lane) * v.ld + v.Dp] = accR.value();
  } else {
    double* pp = parts + (size_t)ch.w * PART_STRIDE;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t row = M::row(lane, q);
      if (row >= da || col >= db) continue;
      pp[row + col * da] = ss[q];
    }
    if (rlane) {
      pp[196 + lane] = bdg;
      pp[392 + lane] = accR.value();
    }
  }
}

// ------------------------------------------------- back-substitution + update
template <typename T>
__device__ inline void backsub_point(const View<T>& v, uint32_t p, int cur) {
  const double* X = v.pts[cur] + 3 * (size_t)p;
  double* Y = v.pts[1 - cur] + 3 * (size_t)p;
  if (!v.pt_ok[p]) {
    Y[0] = X[0];
    Y[1] = X[1];
    Y[2] = X[2];
    return;
  }
  double acc0 = v.pt_g[3 * (size_t)p], acc1 = v.pt_g[3 * (size_t)p + 1], acc2 = v.pt_g[3 * (size_t)p + 2];
  for (uint32_t a = v.pt_inc_ptr[p]; a < v.pt_inc_ptr[p + 1]; ++a) {
    const uint32_t g = v.inc_grp[a];
    const uint32_t r0 = v.inc_roff[a], n = v.grp_dim[g], off = v.grp_off[g];
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
      const auto e = v.E4[r0 + r];
      const double d = v.delta[off + r];
      c0 += (double)e.x * d;
      c1 += (double)e.y * d;
      c2 += (double)e.z * d;
    }
    acc0 += c0;
    acc1 += c1;
    acc2 += c2;
  }
  const double* ci = v.pt_cinv + 9 * (size_t)p;
  Y[0] = X[0] + -(ci[0] * acc0 + ci[1] * acc1 + ci[2] * acc2);
  Y[1] = X[1] + -(ci[3] * acc0 + ci[4] * acc1 + ci[5] * acc2);
  Y[2] = X[2] + -(ci[6] * acc0 + ci[7] * acc1 + ci[8] * acc2);
}

// perturb_pose: R' = exp(omega) R (nalgebra from_scaled_axis: identity when
// |omega/2|^2 <= eps^2, no renormalization), t' = t + dt.
template <typename T>
__device__ inline void update_pose(const View<T>& v, uint32_t i, int cur) {
  const double* p = v.poses[cur] + 7 * (size_t)i;
  double* o = v.poses[1 - cur] + 7 * (size_t)i;
  const uint32_t s = v.img_slot[i];
  double d[6] = {0, 0, 0, 0, 0, 0};
  if (s != NONE)
    for (int k = 0; k < 6; ++k) d[k] = v.delta[s + k];
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

template <typename T>
__device__ inline void update_cam(const View<T>& v, uint32_t c, int cur) {
  double p[8];
  for (int k = 0; k < 8; ++k) p[k] = v.cams[cur][8 * (size_t)c + k];
  const uint32_t s = v.cam_slot[c];
  if (s != NONE) {
    const uint32_t nf = v.cam_nfree[c];
    for (uint32_t j = 0; j < nf; ++j) p[v.cam_free[8 * (size_t)c + j]] += v.delta[s + j];
  }
  for (int k = 0; k < 8; ++k) v.cams[1 - cur][8 * (size_t)c + k] = p[k];
}

// ------------------------------------------------------------------- cost
// Observation o's term: mode 0 w(|r|) |r|^2 (total_cost), mode 1 |r| (the
// numerator of mean_reprojection_error); in_front is 1 for an observation
// in front of its camera.
__device__ inline void obs_cost(const uint32_t* obs_img, const uint32_t* obs_pt, const double2* obs_xy,
                                const uint32_t* cam_of_img, const int32_t* cam_model,
                                const double* poses, const double* cams, const double* pts,
                                uint32_t o, int loss, double scale, int mode, double& term,
                                double& in_front) {
  term = 0.0;
  in_front = 0.0;
  const uint32_t img = obs_img[o], pt = obs_pt[o], cam = cam_of_img[img];
  double c0, c1, c2;
  transform_point(poses + 7 * (size_t)img, pts[3 * (size_t)pt], pts[3 * (size_t)pt + 1],
                  pts[3 * (size_t)pt + 2], c0, c1, c2);
  if (c2 <= 1e-9) return;
  const double2 xy = obs_xy[o];
  double r0, r1;
  project_residual(cam_model[cam], cams + 8 * (size_t)cam, c0, c1, c2, xy.x, xy.y, r0, r1);
  const double n = sqrt(r0 * r0 + r1 * r1);
  if (mode == 0) {
    const double w = robust_weight(n, loss, scale);
    term = w * n * n;
  } else {
    term = n;
  }
  in_front = 1.0;
}

}  // namespace sfmba
