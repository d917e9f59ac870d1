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
//   compute_Y / schur_chunk_rec (the reduction from incidence records on
//   the matrix cores)
//                      damped diagonal and cross blocks, 923-956, and the
//                      point elimination S -= E C^-1 E^T, rhs += E C^-1 g of
//                      accumulate_point_schur, 112-211 / 967-1043
//   backsub_point_rec  delta_point = -C^-1 (g + E^T delta), 1096-1120
//   update_pose / update_cam
//                      trial state, 1122-1164 (perturb_pose 263-269)
//   obs_cost           one term of total_cost 727-745 / mean_reprojection_error 708-725
#pragma once

#include <hip/hip_runtime.h>

#include "camera_models.hpp"
#include "engine.hpp"

namespace sfmba {

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
// Observation o (point-major order): the records of its incidences
// (rec_values: X2 and R2, and one J2 row per unknown of the block) and its
// contribution to its point's 3x3 block and gradient (obs_pc).  With a
// stage, an observation of one incidence leaves its record pieces there
// (V2 k of X2 R2 at k, J2 row s at 4 + s, zero rows up to DB) for the
// caller to store; one of two incidences writes its records directly.
// With pc_reg, the point terms go there (the caller's registers) instead
// of obs_pc.  Returns whether the point is in front of the camera
// (obs_valid).
template <typename T, int DB>
__device__ inline bool linearize_obs(const View<T>& v, uint32_t o, int cur, int loss, double scale,
                                     typename Vec2<T>::type* stage = nullptr, T* pc_reg = nullptr) {
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
  const bool two = inc1 != NONE;
  T* pc_out = v.obs_pc + 9 * (size_t)o;
  // the incidences' records, in V2 units (an observation without a free
  // unknown has none and writes no record)
  constexpr size_t RS2 = rec_values(DB) / 2;
  V2* rec0 = inc0 != NONE ? reinterpret_cast<V2*>(v.rec) + (size_t)inc0 * RS2 : nullptr;
  V2* rec1 = two ? reinterpret_cast<V2*>(v.rec) + (size_t)inc1 * RS2 : nullptr;
  const bool staged = stage != nullptr && !two;
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
  auto put_pc = [&](int k, T val) {
    if (pc_reg)
      pc_reg[k] = val;
    else
      pc_out[k] = val;
  };

  if (c2 <= 1e-9) {
    // Behind the camera: dropped for this iteration (lib.rs:339-341).
    v.obs_valid[o] = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) put_pc(k, T(0));
    const V2 z2 = {T(0), T(0)};
    if (rec0)
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) put_xr(k, z2);
#pragma unroll
    for (uint32_t s = 0; s < 14; ++s) {
      if (s >= ns) break;
      put_row(s, z2);
    }
    if (staged)
#pragma unroll
      for (uint32_t s = 0; s < DB; ++s)
        if (s >= ns) stage[4 + s] = z2;
    return false;
  }
  v.obs_valid[o] = 1;
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
  put_pc(0, sx[0] * sx[0] + sx[3] * sx[3]);
  put_pc(1, sx[0] * sx[1] + sx[3] * sx[4]);
  put_pc(2, sx[0] * sx[2] + sx[3] * sx[5]);
  put_pc(3, sx[1] * sx[1] + sx[4] * sx[4]);
  put_pc(4, sx[1] * sx[2] + sx[4] * sx[5]);
  put_pc(5, sx[2] * sx[2] + sx[5] * sx[5]);
  put_pc(6, sx[0] * sr0 + sx[3] * sr1);
  put_pc(7, sx[1] * sr0 + sx[4] * sr1);
  put_pc(8, sx[2] * sr0 + sx[5] * sr1);
  // X2 (three V2) and R2 into every record of the observation
  if (rec0) {
    put_xr(0, V2{sx[0], sx[1]});
    put_xr(1, V2{sx[2], sx[3]});
    put_xr(2, V2{sx[4], sx[5]});
    put_xr(3, V2{sr0, sr1});
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
  for (uint32_t s = 0; s < 14; ++s) {
    if (s >= ns) break;
    put_row(s, V2{sw * s0[s], sw * s1[s]});
  }
  if (staged)
#pragma unroll
    for (uint32_t s = 0; s < DB; ++s)
      if (s >= ns) stage[4 + s] = V2{T(0), T(0)};
  return true;
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

constexpr int PART_STRIDE = 2 * 196 + 16;

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
template <typename T, int DB, int INS_ = 2>
struct RecShape {
  static constexpr uint32_t RS = rec_values(DB);
  static constexpr uint32_t SIDE = 8 + 2 * DB;
  static constexpr uint32_t LPS = SIDE * sizeof(T) / 16;
  static constexpr uint32_t NPI = 64 / LPS;
  static constexpr uint32_t INS = INS_;
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
// k_schur_reduce.  lds: RecShape<T, DB, INS>::LDS elements of this wavefront.
template <typename T, int DB, bool LONG = false, int INS = 2>
__device__ inline void schur_chunk_rec(const View<T>& v, const uint4 ch, double lambda, double* parts, T* lds) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using Sh = RecShape<T, DB, INS>;
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
  // this lane's side slot and 16-byte piece in each load instruction
  const uint32_t slot = lane / Sh::LPS, piece = (lane % Sh::LPS) * 16;
  acc_t s = {T(0), T(0), T(0), T(0)};
  double ss[4] = {0.0, 0.0, 0.0, 0.0};
  Acc<T> accR;
  double bdg = 0.0;
  const uint32_t n_all = ch.z - ch.y;
  // LONG: a chunk of any length (the window solver's one chunk per
  // destination), 128 pairs at a time, in pair order
  for (uint32_t base = 0; base < (LONG ? n_all : 1u); base += 128) {
  const uint32_t n = LONG ? min(128u, n_all - base) : n_all;
  const uint32_t p0 = ch.y + base;
  const uint2 w0 = lane < n ? v.pairs[p0 + lane] : make_uint2(0u, 0u);
  const uint2 w1 = lane + 64 < n ? v.pairs[p0 + 64 + lane] : make_uint2(0u, 0u);
  // per-pair flags of the block (bit q % 64 of half q / 64): same
  // observation, and a == b
  const uint64_t same_m[2] = {__ballot((w0.x & HI) != 0), __ballot((w1.x & HI) != 0)};
  const uint64_t eq_m[2] = {__ballot((w0.y & HI) != 0), __ballot((w1.y & HI) != 0)};
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
  }  // 128-pair blocks
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
// The point back-substitution from incidence records: E_a^T delta = X2^T (J2_a^T delta)
// per incidence, in incidence order.
template <typename T, int DB>
__device__ inline void backsub_point_rec(const View<T>& v, uint32_t p, int cur) {
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
    const uint32_t n = v.grp_dim[g], off = v.grp_off[g];
    const T* rc = v.rec + (size_t)a * rec_values(DB);
    double t0 = 0.0, t1 = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
      const double d = v.delta[off + r];
      t0 += (double)rc[8 + 2 * r] * d;
      t1 += (double)rc[8 + 2 * r + 1] * d;
    }
    acc0 += (double)rc[0] * t0 + (double)rc[3] * t1;
    acc1 += (double)rc[1] * t0 + (double)rc[4] * t1;
    acc2 += (double)rc[2] * t0 + (double)rc[5] * t1;
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
// The term of an observation of point pt in image img at xy: mode 0
// w(|r|) |r|^2 (total_cost), mode 1 |r| (the numerator of
// mean_reprojection_error); in_front is 1 for an observation in front of
// its camera.
__device__ inline void obs_cost_at(uint32_t img, uint32_t pt, double2 xy, const uint32_t* cam_of_img,
                                   const int32_t* cam_model, const double* poses, const double* cams,
                                   const double* pts, int loss, double scale, int mode, double& term,
                                   double& in_front) {
  term = 0.0;
  in_front = 0.0;
  const uint32_t cam = cam_of_img[img];
  double c0, c1, c2;
  transform_point(poses + 7 * (size_t)img, pts[3 * (size_t)pt], pts[3 * (size_t)pt + 1],
                  pts[3 * (size_t)pt + 2], c0, c1, c2);
  if (c2 <= 1e-9) return;
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

// Observation o's term (obs_cost_at).
__device__ inline void obs_cost(const uint32_t* obs_img, const uint32_t* obs_pt, const double2* obs_xy,
                                const uint32_t* cam_of_img, const int32_t* cam_model,
                                const double* poses, const double* cams, const double* pts,
                                uint32_t o, int loss, double scale, int mode, double& term,
                                double& in_front) {
  obs_cost_at(obs_img[o], obs_pt[o], obs_xy[o], cam_of_img, cam_model, poses, cams, pts, loss, scale, mode, term,
              in_front);
}

}  // namespace sfmba
