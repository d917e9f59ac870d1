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
//   compute_F / schur_chunk (schur_chunk_mfma, schur_chunk_mfma8: the same
//   reduction on the matrix cores)
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
// one E4 row per unknown, and its contribution to its point's 3x3 block and
// gradient (obs_pc).  With LPO > 1, LPO lanes share observation o: each
// computes the same values and lane r writes rows r, r + LPO, ... and
// obs_pc entries r, r + LPO, ..., so that one store instruction writes
// whole rows of consecutive observations (a full cache line per
// observation) instead of one 16- or 32-byte piece per lane.
template <typename T, int LPO = 1>
__device__ inline void linearize_obs(const View<T>& v, uint32_t o, int cur, int loss, double scale,
                                     uint32_t r = 0) {
  using V4 = typename Vec4<T>::type;
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

  if (c2 <= 1e-9) {
    // Behind the camera: dropped for this iteration (lib.rs:339-341).
    if (r == 0) v.obs_valid[o] = 0;
#pragma unroll
    for (uint32_t k0 = 0; k0 < 9; k0 += LPO)
      if (k0 + r < 9) pc_out[k0 + r] = T(0);
    const V4 z = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (uint32_t s0 = 0; s0 < 14; s0 += LPO) {
      const uint32_t s = s0 + r;
      if (s >= ns) break;
      const uint32_t row = (two && s >= 6) ? roff1 + (s - 6) : roff0 + s;
      v.J2[row] = typename Vec2<T>::type{T(0), T(0)};
      v.E4[row] = z;
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
    if (k0 + r < 9) pc_out[k0 + r] = val;
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
    const uint32_t row = (two && s >= 6) ? roff1 + (s - 6) : roff0 + s;
    const T j0 = sw * a0, j1 = sw * a1;
    v.J2[row] = typename Vec2<T>::type{j0, j1};
    v.E4[row] = V4{j0 * sx[0] + j1 * sx[3], j0 * sx[1] + j1 * sx[4], j0 * sx[2] + j1 * sx[5], j0 * sr0 + j1 * sr1};
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

// Row r of incidence a, for the grid kernel: the lanes of one incidence
// read and write its consecutive rows in one instruction (whole cache
// lines), and share the point's Cinv and g by broadcast.
template <typename T>
__device__ inline void compute_F_row(const View<T>& v, uint32_t a, uint32_t r) {
  using V4 = typename Vec4<T>::type;
  const uint32_t n = v.grp_dim[v.inc_grp[a]];
  if (r >= n) return;
  const uint32_t pt = v.inc_pt[a];
  const double* ci = v.pt_cinv + 9 * (size_t)pt;
  T c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) c[k] = (T)ci[k];
  const T g0 = (T)v.pt_g[3 * (size_t)pt], g1 = (T)v.pt_g[3 * (size_t)pt + 1],
          g2 = (T)v.pt_g[3 * (size_t)pt + 2];
  const V4 e = v.E4[v.inc_roff[a] + r];
  const T f0 = e.x * c[0] + e.y * c[3] + e.z * c[6];
  const T f1 = e.x * c[1] + e.y * c[4] + e.z * c[7];
  const T f2 = e.x * c[2] + e.y * c[5] + e.z * c[8];
  v.F4[v.inc_froff[a] + r] = V4{f0, f1, f2, f0 * g0 + f1 * g1 + f2 * g2};
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

// The same destination-block reduction as schur_chunk, on the matrix cores:
// the block is one 16x16 MFMA tile (da, db <= 16).  Per pair, lane l loads
// one component, F_a[l % 16][l / 16] and E_b[l % 16][l / 16] for the three
// point coordinates (K slot 3 stays zero), and one MFMA adds F_a E_b^T; a
// pair of the same observation adds w J_a^T J_b with a second MFMA over the
// two residual components.  The FP32 instantiation flushes its MFMA sums
// into FP64 after every U pairs, so no more than U terms are summed in
// FP32.  Output entries, partials and the right-hand side follow
// schur_chunk.
template <typename T, int U>
__device__ inline void schur_chunk_mfma(const View<T>& v, const uint4 ch, double lambda, double* parts) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr uint32_t HI = 0x80000000u, OFF = 0x7FFFFFFFu;
  static_assert(PAIR_CHUNK <= 128, "a chunk's pair words are held in two registers per lane");
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t i = lane & 15, k = lane >> 4;
  const uint32_t ga = ch.x >> 16, gb = ch.x & 0xFFFFu;
  const uint32_t da = v.grp_dim[ga], db = v.grp_dim[gb];
  const bool diag = ga == gb;
  const bool fa_on = i < da && k < 3, eb_on = i < db && k < 3;
  const bool ja_on = i < da && k < 2, jb_on = i < db && k < 2;
  const T* F = reinterpret_cast<const T*>(v.F4);
  const T* E = reinterpret_cast<const T*>(v.E4);
  const T* J = reinterpret_cast<const T*>(v.J2);
  // All pair words of the chunk in one coalesced load: lane l holds pairs
  // l and 64 + l; pair q reaches every lane by v_readlane.
  const uint32_t n = ch.z - ch.y;
  const uint2 w0 = lane < n ? v.pairs[ch.y + lane] : make_uint2(0u, 0u);
  const uint2 w1 = lane + 64 < n ? v.pairs[ch.y + 64 + lane] : make_uint2(0u, 0u);
  auto word = [&](uint32_t q, uint32_t& x, uint32_t& y) {
    const uint32_t sl = q & 63;
    x = __builtin_amdgcn_readlane(q < 64 ? w0.x : w1.x, sl);
    y = __builtin_amdgcn_readlane(q < 64 ? w0.y : w1.y, sl);
  };
  // Operands of U pairs starting at q (zero past the chunk's end).
  struct Stage {
    T fa[U], eb[U], ja[U], jb[U];
  };
  auto load = [&](uint32_t q, Stage& st) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = q + u < n;
      uint32_t px = 0u, py = 0u;
      if (ok) word(q + u, px, py);
      const uint32_t ra = px & OFF, rb = py & OFF;
      st.fa[u] = (ok && fa_on) ? F[4 * (size_t)(ra + i) + k] : T(0);
      st.eb[u] = (ok && eb_on) ? E[4 * (size_t)(rb + i) + k] : T(0);
      const bool same = ok && (px & HI);
      st.ja[u] = (same && ja_on) ? J[2 * (size_t)(ra + i) + k] : T(0);
      st.jb[u] = (same && jb_on) ? J[2 * (size_t)(rb + i) + k] : T(0);
    }
  };
  acc_t s = {T(0), T(0), T(0), T(0)}, b = {T(0), T(0), T(0), T(0)};
  double ss[4] = {0.0, 0.0, 0.0, 0.0}, bs[4] = {0.0, 0.0, 0.0, 0.0};
  Acc<T> accR;
  const bool rlane = diag && lane < da;
  Stage cur, nxt;
  load(0, cur);
  for (uint32_t q = 0; q < n; q += U) {
    if (q + U < n) load(q + U, nxt);  // next group's loads in flight during this group's MFMAs
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (q + u >= n) break;
      uint32_t px, py;
      word(q + u, px, py);
      s = M::step(cur.fa[u], cur.eb[u], s);
      if (px & HI) b = M::step(cur.ja[u], cur.jb[u], b);
      if (rlane && (py & HI)) accR.add(v.F4[(px & OFF) + lane].w - v.E4[(py & OFF) + lane].w);
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
  if (sizeof(T) == 8) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ss[c] = (double)s[c];
      bs[c] = (double)b[c];
    }
  }
  const uint32_t oa = v.grp_off[ga], ob = v.grp_off[gb];
  const uint32_t c = i;
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

// schur_chunk_mfma for blocks of at most 8x8 (da, db <= 8), two pairs per
// MFMA: rows 0-7 of both operands carry the even pair of a step, rows 8-15
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

// ------------------------------------------------ LDS-staged Schur chunk
typedef __attribute__((address_space(3))) void lds_void_t;

// Pairs of one LDS batch of schur_chunk_lds: 8 KB of F and E row blocks
// (and 4 KB of J2 blocks of both sides when the chunk has pairs of the same
// observation).
template <typename T>
constexpr int lds_batch_pairs() {
  return sizeof(T) == 4 ? 32 : 16;
}
template <typename T, bool SAME>
constexpr int lds_chunk_elems() {
  return lds_batch_pairs<T>() * (2 * 32 + (SAME ? 2 * 16 : 0));
}

// schur_chunk_mfma8 with its operands staged through LDS (blocks of at most
// 8x8): the 8-row blocks of a batch of pairs move from global memory
// straight into LDS (global_load_lds, 16 bytes per lane, one instruction per
// 8 FP32 or 4 FP64 blocks, no registers held while they are in flight),
// then the two-pairs-per-MFMA products read them from LDS.  The register
// form is bound by its gathers' latency; here the bytes in flight per CU
// are set by LDS instead of registers.  SAME: the chunk has pairs of the
// same observation, so the sqrt(w) J blocks (two components per row) of
// both sides are staged too, for w J_a^T J_b (a second MFMA, K = the two
// residual components; b's block is a's when a == b, and is not loaded when
// every pair of the chunk has a == b) and, on diagonal blocks, the
// right-hand side F_a g_point - w J_a^T r over pairs a == b, in pair order.
// lds: this wavefront's lds_chunk_elems<T, SAME>() elements.  Sums, FP32
// flushing every two steps, the quadrant merge and the outputs are
// schur_chunk_mfma8's.
template <typename T, bool SAME>
__device__ inline void schur_chunk_lds(const View<T>& v, const uint4 ch, double lambda, double* parts, T* lds) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr uint32_t HI = 0x80000000u, OFF = 0x7FFFFFFFu;
  constexpr int BP = lds_batch_pairs<T>();
  constexpr uint32_t ROWB = 4 * sizeof(T);  // bytes per row (4 components)
  constexpr uint32_t LPB = 8 * ROWB / 16;   // lanes per 8-row block
  constexpr uint32_t PPI = 64 / LPB;        // blocks per load instruction
  static_assert(BP % PPI == 0 && 64 % BP == 0, "batches of whole load instructions within a 64-pair half");
  static_assert(PAIR_CHUNK <= 128, "a chunk's pair words are held in two registers per lane");
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t i = lane & 7, h = (lane >> 3) & 1, k = lane >> 4;
  const uint32_t ga = ch.x >> 16, gb = ch.x & 0xFFFFu;
  const uint32_t da = v.grp_dim[ga], db = v.grp_dim[gb];
  const bool diag = ga == gb;
  const bool fa_on = i < da && k < 3, eb_on = i < db && k < 3;
  const bool ja_on = i < da && k < 2, jb_on = i < db && k < 2;
  const bool rlane = diag && lane < da;
  const char* Fg = reinterpret_cast<const char*>(v.F4);
  const char* Eg = reinterpret_cast<const char*>(v.E4);
  const char* Jg = reinterpret_cast<const char*>(v.J2);
  constexpr uint32_t JROWB = 2 * sizeof(T);  // bytes per J2 row
  constexpr uint32_t JLPB = 8 * JROWB / 16;  // lanes per 8-row J2 block
  constexpr uint32_t JPPI = 64 / JLPB;       // J2 blocks per load instruction
  static_assert(BP % JPPI == 0, "whole J2 load instructions per batch");
  T* Fs = lds;
  T* Es = lds + BP * 32;
  T* JAs = lds + 2 * BP * 32;
  T* JBs = lds + 2 * BP * 32 + BP * 16;
  const uint32_t n = ch.z - ch.y;
  const uint2 w0 = lane < n ? v.pairs[ch.y + lane] : make_uint2(0u, 0u);
  const uint2 w1 = lane + 64 < n ? v.pairs[ch.y + 64 + lane] : make_uint2(0u, 0u);
  // per-pair flags of the chunk (bit q % 64 of half q / 64): same
  // observation, and a == b
  const uint64_t same_m[2] = {__ballot((w0.x & HI) != 0), __ballot((w1.x & HI) != 0)};
  const uint64_t eq_m[2] = {__ballot((w0.y & HI) != 0), __ballot((w1.y & HI) != 0)};
  // byte offset of this lane's 16 bytes in a block; rows past the block's
  // dimension re-read row 0 of the same block (in bounds, masked below)
  const uint32_t off = (lane % LPB) * 16;
  const uint32_t offa = off / ROWB < da ? off : off % ROWB;
  const uint32_t offb = off / ROWB < db ? off : off % ROWB;
  // J2 blocks: FP32 rows are 8 bytes, so a lane's 16 bytes may end one row
  // past the block (the row array is padded by a block for that)
  const uint32_t joff = (lane % JLPB) * 16;
  const uint32_t joffa = joff / JROWB < da ? joff : 0u;
  const uint32_t joffb = joff / JROWB < db ? joff : 0u;
  const bool all_eq = SAME && (uint32_t)(__popcll(eq_m[0]) + __popcll(eq_m[1])) == n;
  acc_t s = {T(0), T(0), T(0), T(0)}, b = {T(0), T(0), T(0), T(0)};
  double ss[4] = {0.0, 0.0, 0.0, 0.0}, bs[4] = {0.0, 0.0, 0.0, 0.0};
  Acc<T> accR;
  for (uint32_t q0 = 0; q0 < n; q0 += BP) {
    const bool upper = q0 >= 64;
    // pairs q0 .. q0 + BP - 1 into LDS; a lane past the chunk's end loads
    // pair q0's block again (never read back)
#pragma unroll
    for (int j = 0; j < BP / (int)PPI; ++j) {
      const uint32_t q = q0 + j * PPI;
      uint32_t idx = q + lane / LPB;
      if (idx >= n) idx = q0;
      const uint32_t src = idx & 63;
      const uint32_t ra = (uint32_t)__shfl((int)(upper ? w1.x : w0.x), (int)src) & OFF;
      const uint32_t rb = (uint32_t)__shfl((int)(upper ? w1.y : w0.y), (int)src) & OFF;
      __builtin_amdgcn_global_load_lds((const void*)(Fg + (size_t)ra * ROWB + offa),
                                       (lds_void_t*)(Fs + j * PPI * 32), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(Eg + (size_t)rb * ROWB + offb),
                                       (lds_void_t*)(Es + j * PPI * 32), 16, 0, 0);
    }
    if (SAME) {
#pragma unroll
      for (int j = 0; j < BP / (int)JPPI; ++j) {
        const uint32_t q = q0 + j * JPPI;
        if (q >= n) break;
        uint32_t idx = q + lane / JLPB;
        if (idx >= n) idx = q0;
        const uint32_t src = idx & 63;
        const uint32_t ra = (uint32_t)__shfl((int)(upper ? w1.x : w0.x), (int)src) & OFF;
        __builtin_amdgcn_global_load_lds((const void*)(Jg + (size_t)ra * JROWB + joffa),
                                         (lds_void_t*)(JAs + j * JPPI * 16), 16, 0, 0);
        if (!all_eq) {
          const uint32_t rb = (uint32_t)__shfl((int)(upper ? w1.y : w0.y), (int)src) & OFF;
          __builtin_amdgcn_global_load_lds((const void*)(Jg + (size_t)rb * JROWB + joffb),
                                           (lds_void_t*)(JBs + j * JPPI * 16), 16, 0, 0);
        }
      }
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const uint32_t nb = min((uint32_t)BP, n - q0);
    const uint64_t sm = same_m[upper ? 1 : 0] >> (q0 & 63), em = eq_m[upper ? 1 : 0] >> (q0 & 63);
    for (uint32_t u = 0; u < nb; u += 2) {
      const uint32_t p = u + h;
      const bool ok = p < nb;
      const T fa = (ok && fa_on) ? Fs[p * 32 + i * 4 + k] : T(0);
      const T eb = (ok && eb_on) ? Es[p * 32 + i * 4 + k] : T(0);
      s = M::step(fa, eb, s);
      if (SAME) {
        if ((sm >> u) & 3u) {
          const bool same = ok && ((sm >> p) & 1u);
          const T* jbs = ((em >> p) & 1u) ? JAs : JBs;  // a == b: b's block is a's
          const T ja = (same && ja_on) ? JAs[p * 16 + i * 2 + k] : T(0);
          const T jb = (same && jb_on) ? jbs[p * 16 + i * 2 + k] : T(0);
          b = M::step(ja, jb, b);
        }
        // right-hand side in pair order
        if (rlane && ((em >> u) & 1u)) accR.add(Fs[u * 32 + lane * 4 + 3] - Es[u * 32 + lane * 4 + 3]);
        if (rlane && u + 1 < nb && ((em >> (u + 1)) & 1u))
          accR.add(Fs[(u + 1) * 32 + lane * 4 + 3] - Es[(u + 1) * 32 + lane * 4 + 3]);
      }
      if (sizeof(T) == 4 && (u & 2)) {
        // FP32: at most two products per quadrant summed before FP64
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          ss[c] += (double)s[c];
          bs[c] += (double)b[c];
          s[c] = T(0);
          b[c] = T(0);
        }
      }
    }
    // the next batch's loads overwrite this one: every lane's reads first
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    ss[c] += (double)s[c];
    bs[c] += (double)b[c];
  }
  // Quadrant (8 + r, 8 + c) sits 40 lanes up in the same register (FP32
  // layout) or 8 lanes up two registers on (FP64 layout).
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
