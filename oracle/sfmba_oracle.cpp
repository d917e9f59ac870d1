// ORACLE - test infrastructure only.
//
// CPU FP64 restatement of the reference's bundle-adjustment path, used by
// tests/, the smoke check and bench.py's cpu_baseline leg as the checker.
// The product library (sfmtory_amd/csrc) never links or calls this file.
//
// Followed, function by function (paths relative to the reference repo):
//   crates/sfm-core/src/pose.rs:46-48        Pose::transform_point
//   crates/sfm-core/src/camera.rs:204-244    CameraModel::project
//   crates/sfm-ba/src/lib.rs:249-261         reprojection_residual
//   crates/sfm-ba/src/lib.rs:263-269         perturb_pose
//   crates/sfm-ba/src/lib.rs:282-310         pose_point_jacobians
//   crates/sfm-ba/src/lib.rs:327-590         analytic_jacobians
//   crates/sfm-ba/src/lib.rs:599-670         observation_jacobians / camera_jacobian (FD)
//   crates/sfm-ba/src/lib.rs:677-694         robust_weight
//   crates/sfm-ba/src/lib.rs:708-745         mean_reprojection_error / total_cost
//   crates/sfm-ba/src/lib.rs:747-1232        bundle_adjust (+ accumulate_point_schur 112-211)
//   crates/sfm-ba/src/intrinsics.rs:33-238   default_fixed_params_mask / refine_intrinsics
//
// The reference links nalgebra 0.33.3 for small-matrix algebra, Cholesky, LU
// and quaternions, and rayon 1.12.0 for threading; neither is vendored.  Those
// pieces are restated from their published algorithms (noted where used).
// Summation order is the reference's serial order; the threaded variant
// (nthreads > 1) follows the reference's parallel structure for CPU timing.
//
// Pinning (tests/test_oracle_*.py): central FD Jacobians at the reference's
// own test cameras and tolerances, sympy-derived Jacobians, scipy rotations,
// closed-form residuals, numpy/LAPACK solves of the full damped normal
// equations (Schur + back-substitution), LAPACK Cholesky/LU, and the
// reference's convergence tests.  Parity unpinned: bit-level agreement with
// nalgebra's rounding (quaternion product, gemm order, Cholesky/LU pivot
// tests) and rayon's reduction grouping.

#include "../include/sfmba.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <chrono>
#include <thread>
#include <vector>

namespace {

constexpr size_t FIXED = std::numeric_limits<size_t>::max();

int param_count(int model) {
  switch (model) {
    case SFMBA_SIMPLE_PINHOLE: return 3;
    case SFMBA_PINHOLE: return 4;
    case SFMBA_SIMPLE_RADIAL: return 4;
    case SFMBA_RADIAL: return 5;
    case SFMBA_OPENCV: return 8;
    case SFMBA_OPENCV_FISHEYE: return 8;
    default: return -1;
  }
}

struct V3 {
  double x, y, z;
};

// nalgebra UnitQuaternion * Vector3: t = 2 (q_v x v); v' = v + w t + q_v x t.
V3 rotate(const sfmba_pose& p, const V3& v) {
  const double tx = 2.0 * (p.qy * v.z - p.qz * v.y);
  const double ty = 2.0 * (p.qz * v.x - p.qx * v.z);
  const double tz = 2.0 * (p.qx * v.y - p.qy * v.x);
  const double cx = p.qy * tz - p.qz * ty;
  const double cy = p.qz * tx - p.qx * tz;
  const double cz = p.qx * ty - p.qy * tx;
  return V3{tx * p.qw + cx + v.x, ty * p.qw + cy + v.y, tz * p.qw + cz + v.z};
}

// pose.rs:46-48
V3 transform_point(const sfmba_pose& p, const V3& X) {
  V3 r = rotate(p, X);
  return V3{r.x + p.tx, r.y + p.ty, r.z + p.tz};
}

// Rotation matrix of a (unit) quaternion, homogeneous form (row-major).
void rotation_matrix(const sfmba_pose& p, double R[9]) {
  const double w = p.qw, i = p.qx, j = p.qy, k = p.qz;
  const double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  const double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0;
  const double ik = i * k * 2.0, jk = j * k * 2.0, wi = w * i * 2.0;
  R[0] = ww + ii - jj - kk; R[1] = ij - wk;          R[2] = wj + ik;
  R[3] = wk + ij;          R[4] = ww - ii + jj - kk; R[5] = jk - wi;
  R[6] = ik - wj;          R[7] = wi + jk;           R[8] = ww - ii - jj + kk;
}

// camera.rs:204-244
void project(const sfmba_camera& cam, const V3& pc, double* u, double* v) {
  const double x = pc.x / pc.z;
  const double y = pc.y / pc.z;
  const double* P = cam.params;
  switch (cam.model_id) {
    case SFMBA_SIMPLE_PINHOLE:
      *u = P[0] * x + P[1];
      *v = P[0] * y + P[2];
      return;
    case SFMBA_PINHOLE:
      *u = P[0] * x + P[2];
      *v = P[1] * y + P[3];
      return;
    case SFMBA_SIMPLE_RADIAL: {
      const double r2 = x * x + y * y;
      const double d = 1.0 + P[3] * r2;
      *u = P[0] * (x * d) + P[1];
      *v = P[0] * (y * d) + P[2];
      return;
    }
    case SFMBA_RADIAL: {
      const double r2 = x * x + y * y;
      const double d = 1.0 + P[3] * r2 + P[4] * r2 * r2;
      *u = P[0] * (x * d) + P[1];
      *v = P[0] * (y * d) + P[2];
      return;
    }
    case SFMBA_OPENCV: {
      const double k1 = P[4], k2 = P[5], p1 = P[6], p2 = P[7];
      const double r2 = x * x + y * y;
      const double radial = 1.0 + k1 * r2 + k2 * r2 * r2;
      const double xd = x * radial + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
      const double yd = y * radial + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
      *u = P[0] * xd + P[2];
      *v = P[1] * yd + P[3];
      return;
    }
    case SFMBA_OPENCV_FISHEYE: {
      const double k1 = P[4], k2 = P[5], k3 = P[6], k4 = P[7];
      const double r = std::max(std::sqrt(x * x + y * y), 1e-12);
      const double theta = std::atan(r);
      const double t2 = theta * theta;
      const double t4 = t2 * t2, t6 = t2 * t2 * t2, t8 = t4 * t4;
      const double theta_d = theta * (1.0 + k1 * t2 + k2 * t4 + k3 * t6 + k4 * t8);
      const double scale = theta_d / r;
      *u = P[0] * (x * scale) + P[2];
      *v = P[1] * (y * scale) + P[3];
      return;
    }
  }
  *u = std::nan("");
  *v = std::nan("");
}

// lib.rs:249-261
bool reprojection_residual(const sfmba_pose& pose, const sfmba_camera& cam,
                           const V3& X, double ox, double oy, double r[2]) {
  V3 pc = transform_point(pose, X);
  if (pc.z <= 1e-9) return false;
  double u, v;
  project(cam, pc, &u, &v);
  r[0] = u - ox;
  r[1] = v - oy;
  return true;
}

// lib.rs:263-269.  UnitQuaternion::from_scaled_axis(w) = exp(w/2) with
// nalgebra's Quaternion::exp (identity when |w/2|^2 <= eps^2), then the
// Hamilton product dq * q with no renormalization; translation additive.
sfmba_pose perturb_pose(const sfmba_pose& p, const double d[6]) {
  const double hx = d[0] / 2.0, hy = d[1] / 2.0, hz = d[2] / 2.0;
  const double nn = hx * hx + hy * hy + hz * hz;
  const double eps = std::numeric_limits<double>::epsilon();
  double dw = 1.0, dx = 0.0, dy = 0.0, dz = 0.0;
  if (!(nn <= eps * eps)) {
    const double n = std::sqrt(nn);
    const double s = 1.0 * std::sin(n) / n;
    dx = hx * s;
    dy = hy * s;
    dz = hz * s;
    dw = 1.0 * std::cos(n);
  }
  sfmba_pose o;
  o.qw = dw * p.qw - dx * p.qx - dy * p.qy - dz * p.qz;
  o.qx = dw * p.qx + dx * p.qw + dy * p.qz - dz * p.qy;
  o.qy = dw * p.qy - dx * p.qz + dy * p.qw + dz * p.qx;
  o.qz = dw * p.qz + dx * p.qy - dy * p.qx + dz * p.qw;
  o.tx = p.tx + d[3];
  o.ty = p.ty + d[4];
  o.tz = p.tz + d[5];
  return o;
}

// lib.rs:677-694
double robust_weight(double n, int loss, double scale) {
  const double s = std::max(scale, 1e-9);
  switch (loss) {
    case SFMBA_LOSS_NONE: return 1.0;
    case SFMBA_LOSS_HUBER: return n <= s ? 1.0 : s / n;
    case SFMBA_LOSS_CAUCHY: {
      const double q = n / s;
      return 1.0 / (1.0 + q * q);
    }
  }
  return 1.0;
}

double norm2(const double r[2]) { return std::sqrt(r[0] * r[0] + r[1] * r[1]); }

// lib.rs:282-310 + 327-590.  jp: 2x6 row-major, jx: 2x3, jc: 2x8 (k used).
bool analytic_jacobians(const sfmba_pose& pose, const sfmba_camera& cam, const V3& X,
                        double ox, double oy, double r[2], double jp[12], double jx[6],
                        double jc[16]) {
  V3 pc = transform_point(pose, X);
  if (pc.z <= 1e-9) return false;
  const double inv_z = 1.0 / pc.z;
  const double xp = pc.x * inv_z;
  const double yp = pc.y * inv_z;
  const double* P = cam.params;
  double du[3], dv[3];
  for (int i = 0; i < 16; ++i) jc[i] = 0.0;
  double* jc0 = jc;      // row 0
  double* jc1 = jc + 8;  // row 1

  // Shared tail of every distorted model: d(u,v)/d(pc) from d(u,v)/d(xp,yp).
  auto chain = [&](double du_dxp, double du_dyp, double dv_dxp, double dv_dyp) {
    du[0] = du_dxp * inv_z;
    du[1] = du_dyp * inv_z;
    du[2] = -(du_dxp * pc.x + du_dyp * pc.y) * inv_z * inv_z;
    dv[0] = dv_dxp * inv_z;
    dv[1] = dv_dyp * inv_z;
    dv[2] = -(dv_dxp * pc.x + dv_dyp * pc.y) * inv_z * inv_z;
  };

  switch (cam.model_id) {
    case SFMBA_SIMPLE_PINHOLE: {
      const double f = P[0], cx = P[1], cy = P[2];
      r[0] = f * xp + cx - ox;
      r[1] = f * yp + cy - oy;
      du[0] = f * inv_z; du[1] = 0.0; du[2] = -f * xp * inv_z;
      dv[0] = 0.0; dv[1] = f * inv_z; dv[2] = -f * yp * inv_z;
      jc0[0] = xp; jc0[1] = 1.0;
      jc1[0] = yp; jc1[2] = 1.0;
      break;
    }
    case SFMBA_PINHOLE: {
      const double fx = P[0], fy = P[1], cx = P[2], cy = P[3];
      r[0] = fx * xp + cx - ox;
      r[1] = fy * yp + cy - oy;
      du[0] = fx * inv_z; du[1] = 0.0; du[2] = -fx * xp * inv_z;
      dv[0] = 0.0; dv[1] = fy * inv_z; dv[2] = -fy * yp * inv_z;
      jc0[0] = xp; jc0[2] = 1.0;
      jc1[1] = yp; jc1[3] = 1.0;
      break;
    }
    case SFMBA_SIMPLE_RADIAL: {
      const double f = P[0], cx = P[1], cy = P[2], k = P[3];
      const double r2 = xp * xp + yp * yp;
      const double d = 1.0 + k * r2;
      r[0] = f * xp * d + cx - ox;
      r[1] = f * yp * d + cy - oy;
      const double du_dxp = f * (d + 2.0 * k * xp * xp);
      const double du_dyp = 2.0 * f * k * xp * yp;
      const double dv_dyp = f * (d + 2.0 * k * yp * yp);
      chain(du_dxp, du_dyp, du_dyp, dv_dyp);
      jc0[0] = xp * d; jc0[1] = 1.0; jc0[3] = f * xp * r2;
      jc1[0] = yp * d; jc1[2] = 1.0; jc1[3] = f * yp * r2;
      break;
    }
    case SFMBA_RADIAL: {
      const double f = P[0], cx = P[1], cy = P[2], k1 = P[3], k2 = P[4];
      const double r2 = xp * xp + yp * yp;
      const double d = 1.0 + k1 * r2 + k2 * r2 * r2;
      r[0] = f * xp * d + cx - ox;
      r[1] = f * yp * d + cy - oy;
      const double kr = k1 + 2.0 * k2 * r2;
      const double du_dxp = f * (d + 2.0 * kr * xp * xp);
      const double du_dyp = 2.0 * f * kr * xp * yp;
      const double dv_dyp = f * (d + 2.0 * kr * yp * yp);
      chain(du_dxp, du_dyp, du_dyp, dv_dyp);
      jc0[0] = xp * d; jc0[1] = 1.0; jc0[3] = f * xp * r2; jc0[4] = f * xp * r2 * r2;
      jc1[0] = yp * d; jc1[2] = 1.0; jc1[3] = f * yp * r2; jc1[4] = f * yp * r2 * r2;
      break;
    }
    case SFMBA_OPENCV: {
      const double fx = P[0], fy = P[1], cx = P[2], cy = P[3];
      const double k1 = P[4], k2 = P[5], p1 = P[6], p2 = P[7];
      const double r2 = xp * xp + yp * yp;
      const double radial = 1.0 + k1 * r2 + k2 * r2 * r2;
      const double xd = xp * radial + 2.0 * p1 * xp * yp + p2 * (r2 + 2.0 * xp * xp);
      const double yd = yp * radial + p1 * (r2 + 2.0 * yp * yp) + 2.0 * p2 * xp * yp;
      r[0] = fx * xd + cx - ox;
      r[1] = fy * yd + cy - oy;
      const double kr = k1 + 2.0 * k2 * r2;
      const double dxd_dxp = radial + 2.0 * kr * xp * xp + 2.0 * p1 * yp + 6.0 * p2 * xp;
      const double dxd_dyp = 2.0 * kr * xp * yp + 2.0 * p1 * xp + 2.0 * p2 * yp;
      const double dyd_dxp = dxd_dyp;
      const double dyd_dyp = radial + 2.0 * kr * yp * yp + 6.0 * p1 * yp + 2.0 * p2 * xp;
      chain(fx * dxd_dxp, fx * dxd_dyp, fy * dyd_dxp, fy * dyd_dyp);
      jc0[0] = xd; jc0[2] = 1.0;
      jc0[4] = fx * xp * r2; jc0[5] = fx * xp * r2 * r2;
      jc0[6] = fx * 2.0 * xp * yp; jc0[7] = fx * (r2 + 2.0 * xp * xp);
      jc1[1] = yd; jc1[3] = 1.0;
      jc1[4] = fy * yp * r2; jc1[5] = fy * yp * r2 * r2;
      jc1[6] = fy * (r2 + 2.0 * yp * yp); jc1[7] = fy * 2.0 * xp * yp;
      break;
    }
    case SFMBA_OPENCV_FISHEYE: {
      const double fx = P[0], fy = P[1], cx = P[2], cy = P[3];
      const double k1 = P[4], k2 = P[5], k3 = P[6], k4 = P[7];
      const double rr = std::max(std::sqrt(xp * xp + yp * yp), 1e-12);
      const double theta = std::atan(rr);
      const double t2 = theta * theta;
      const double t4 = t2 * t2, t6 = t2 * t2 * t2, t8 = t4 * t4;
      const double poly = 1.0 + k1 * t2 + k2 * t4 + k3 * t6 + k4 * t8;
      const double theta_d = theta * poly;
      const double scale = theta_d / rr;
      r[0] = fx * (xp * scale) + cx - ox;
      r[1] = fy * (yp * scale) + cy - oy;
      const double poly_deriv = k1 + 2.0 * k2 * t2 + 3.0 * k3 * t4 + 4.0 * k4 * t6;
      const double dtd_dt = poly + 2.0 * t2 * poly_deriv;
      const double dt_dr = 1.0 / (1.0 + rr * rr);
      const double ds_dr = (dtd_dt * dt_dr * rr - theta_d) / (rr * rr);
      const double ds_dxp = ds_dr * (xp / rr);
      const double ds_dyp = ds_dr * (yp / rr);
      chain(fx * (scale + xp * ds_dxp), fx * xp * ds_dyp, fy * yp * ds_dxp,
            fy * (scale + yp * ds_dyp));
      const double d1 = theta * t2, d2 = theta * t4, d3 = theta * t6, d4 = theta * t8;
      jc0[0] = xp * scale; jc0[2] = 1.0;
      jc0[4] = fx * xp * d1 / rr; jc0[5] = fx * xp * d2 / rr;
      jc0[6] = fx * xp * d3 / rr; jc0[7] = fx * xp * d4 / rr;
      jc1[1] = yp * scale; jc1[3] = 1.0;
      jc1[4] = fy * yp * d1 / rr; jc1[5] = fy * yp * d2 / rr;
      jc1[6] = fy * yp * d3 / rr; jc1[7] = fy * yp * d4 / rr;
      break;
    }
    default:
      return false;
  }

  // pose_point_jacobians (lib.rs:282-310)
  double R[9];
  rotation_matrix(pose, R);
  for (int c = 0; c < 3; ++c) {
    jx[c] = R[0 * 3 + c] * du[0] + R[1 * 3 + c] * du[1] + R[2 * 3 + c] * du[2];
    jx[3 + c] = R[0 * 3 + c] * dv[0] + R[1 * 3 + c] * dv[1] + R[2 * 3 + c] * dv[2];
  }
  const double qx = pc.x - pose.tx, qy = pc.y - pose.ty, qz = pc.z - pose.tz;
  // skew(q) * d = q x d
  jp[0] = -qz * du[1] + qy * du[2];
  jp[1] = qz * du[0] - qx * du[2];
  jp[2] = -qy * du[0] + qx * du[1];
  jp[6] = -qz * dv[1] + qy * dv[2];
  jp[7] = qz * dv[0] - qx * dv[2];
  jp[8] = -qy * dv[0] + qx * dv[1];
  for (int c = 0; c < 3; ++c) {
    jp[3 + c] = du[c];
    jp[9 + c] = dv[c];
  }
  return true;
}

bool camera_from_params(int model, const double* p, sfmba_camera* out) {
  const int k = param_count(model);
  if (k < 0) return false;
  out->model_id = model;
  out->reserved = 0;
  for (int i = 0; i < 8; ++i) out->params[i] = i < k ? p[i] : 0.0;
  return true;
}

struct Problem {
  const sfmba_input* in;
  std::vector<sfmba_pose> poses;
  std::vector<sfmba_camera> cams;
  std::vector<double> pts;
};

inline V3 point_at(const std::vector<double>& pts, size_t i) {
  return V3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
}

// lib.rs:727-745 (serial sum; rayon's sum grouping is not reproduced).
double total_cost(const sfmba_input* in, const std::vector<sfmba_pose>& poses,
                  const std::vector<sfmba_camera>& cams, const std::vector<double>& pts,
                  const sfmba_params* prm, int nthreads) {
  const size_t n = in->num_observations;
  auto range_cost = [&](size_t lo, size_t hi) {
    double s = 0.0;
    for (size_t o = lo; o < hi; ++o) {
      const sfmba_observation& ob = in->observations[o];
      double r[2];
      if (reprojection_residual(poses[ob.image_idx], cams[in->camera_of_image[ob.image_idx]],
                                point_at(pts, ob.point_idx), ob.x, ob.y, r)) {
        const double nr = norm2(r);
        const double w = robust_weight(nr, prm->robust_loss, prm->loss_scale_px);
        s += w * nr * nr;
      }
    }
    return s;
  };
  if (nthreads <= 1 || n < 4096) return range_cost(0, n);
  std::vector<double> part(nthreads, 0.0);
  std::vector<std::thread> th;
  const size_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      part[t] = range_cost(std::min(n, t * chunk), std::min(n, (t + 1) * chunk));
    });
  for (auto& x : th) x.join();
  double s = 0.0;
  for (double v : part) s += v;
  return s;
}

// Cholesky (lower, column-major, left-looking column by column), the
// published algorithm of nalgebra's Cholesky::new: fails when a pivot is
// zero, negative or NaN.  Then L y = b, L^T x = y.
bool cholesky_solve(std::vector<double>& A, size_t n, std::vector<double>& b) {
  for (size_t j = 0; j < n; ++j) {
    double* colj = &A[j * n];
    for (size_t k = 0; k < j; ++k) {
      const double f = -A[k * n + j];
      const double* colk = &A[k * n];
      for (size_t i = j; i < n; ++i) colj[i] += f * colk[i];
    }
    const double diag = colj[j];
    if (diag != 0.0 && diag >= 0.0) {
      const double den = std::sqrt(diag);
      colj[j] = den;
      for (size_t i = j + 1; i < n; ++i) colj[i] /= den;
      continue;
    }
    return false;
  }
  for (size_t i = 0; i < n; ++i) {
    const double c = b[i] / A[i * n + i];
    b[i] = c;
    const double* col = &A[i * n];
    for (size_t r = i + 1; r < n; ++r) b[r] += -c * col[r];
  }
  for (size_t ii = n; ii-- > 0;) {
    const double* col = &A[ii * n];
    double dot = 0.0;
    for (size_t r = ii + 1; r < n; ++r) dot += col[r] * b[r];
    b[ii] = (b[ii] - dot) / col[ii];
  }
  return true;
}

// LU with partial pivoting (first maximal |a|; a zero column is skipped),
// the published algorithm of nalgebra's LU::new, then solve, which fails
// when a diagonal entry of U is exactly zero.  Column-major.
bool lu_solve(std::vector<double>& A, size_t n, std::vector<double>& b) {
  std::vector<size_t> perm_a, perm_b;
  for (size_t i = 0; i < n; ++i) {
    size_t piv = i;
    double best = std::fabs(A[i * n + i]);
    for (size_t r = i + 1; r < n; ++r) {
      const double v = std::fabs(A[i * n + r]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    const double diag = A[i * n + piv];
    if (diag == 0.0) continue;
    if (piv != i) {
      perm_a.push_back(i);
      perm_b.push_back(piv);
      for (size_t c = 0; c < n; ++c) std::swap(A[c * n + i], A[c * n + piv]);
    }
    const double inv = 1.0 / diag;
    for (size_t r = i + 1; r < n; ++r) A[i * n + r] *= inv;
    for (size_t c = i + 1; c < n; ++c) {
      const double f = A[c * n + i];
      if (f == 0.0) continue;
      double* colc = &A[c * n];
      const double* coli = &A[i * n];
      for (size_t r = i + 1; r < n; ++r) colc[r] += -f * coli[r];
    }
  }
  for (size_t t = 0; t < perm_a.size(); ++t) std::swap(b[perm_a[t]], b[perm_b[t]]);
  // unit lower
  for (size_t i = 0; i < n; ++i) {
    const double c = b[i];
    for (size_t r = i + 1; r < n; ++r) b[r] += -c * A[i * n + r];
  }
  for (size_t ii = n; ii-- > 0;) {
    const double diag = A[ii * n + ii];
    if (diag == 0.0) return false;
    const double c = b[ii] / diag;
    b[ii] = c;
    for (size_t r = 0; r < ii; ++r) b[r] += -c * A[ii * n + r];
  }
  return true;
}

// 3x3 inverse by the adjugate; singular exactly when det == 0 (the published
// criterion of nalgebra's Matrix3::try_inverse).  Row-major.
bool inverse3(const double m[9], double o[9]) {
  const double m11 = m[0], m12 = m[1], m13 = m[2];
  const double m21 = m[3], m22 = m[4], m23 = m[5];
  const double m31 = m[6], m32 = m[7], m33 = m[8];
  const double a = m22 * m33 - m32 * m23;
  const double b = m21 * m33 - m31 * m23;
  const double c = m21 * m32 - m31 * m22;
  const double det = m11 * a - m12 * b + m13 * c;
  if (det == 0.0) return false;
  o[0] = a / det;
  o[1] = (m13 * m32 - m33 * m12) / det;
  o[2] = (m12 * m23 - m22 * m13) / det;
  o[3] = -b / det;
  o[4] = (m11 * m33 - m31 * m13) / det;
  o[5] = (m13 * m21 - m23 * m11) / det;
  o[6] = c / det;
  o[7] = (m12 * m31 - m32 * m11) / det;
  o[8] = (m11 * m22 - m21 * m12) / det;
  return true;
}

struct Prepared {
  size_t image, camera, point;
  double r[2], jp[12], jx[6], jc[16];
  double w;
};

// One coupling block of a point: rows = 6 (pose) or k (camera), 3 columns.
struct EBlock {
  bool is_pose;
  size_t idx;
  int rows;
  double e[8 * 3];  // row-major rows x 3
};

// accumulate_point_schur, lib.rs:112-211.  S column-major D x D.
void accumulate_point_schur(const double cinv[9], const std::vector<EBlock>& list,
                            const double g[3], const std::vector<size_t>& pose_slot,
                            const std::vector<size_t>& cam_slot, double* S, double* rhs,
                            size_t D) {
  auto off_of = [&](const EBlock& b) { return b.is_pose ? pose_slot[b.idx] : cam_slot[b.idx]; };
  for (size_t i = 0; i < list.size(); ++i) {
    const EBlock& bi = list[i];
    const size_t oi = off_of(bi);
    double ec[8 * 3];  // E_i C^-1
    for (int r = 0; r < bi.rows; ++r)
      for (int c = 0; c < 3; ++c)
        ec[r * 3 + c] = bi.e[r * 3 + 0] * cinv[0 * 3 + c] + bi.e[r * 3 + 1] * cinv[1 * 3 + c] +
                        bi.e[r * 3 + 2] * cinv[2 * 3 + c];
    for (size_t j = i; j < list.size(); ++j) {
      const EBlock& bj = list[j];
      const size_t oj = off_of(bj);
      // Each unordered pair once, mirrored unless it is the same entry
      // (lib.rs:144-147, :156-157, :175-176, :184-187): the block is
      // subtracted whole, then its transpose.
      double blk[8 * 8];
      for (int r = 0; r < bi.rows; ++r)
        for (int c = 0; c < bj.rows; ++c)
          blk[r * 8 + c] = ec[r * 3 + 0] * bj.e[c * 3 + 0] + ec[r * 3 + 1] * bj.e[c * 3 + 1] +
                           ec[r * 3 + 2] * bj.e[c * 3 + 2];
      for (int r = 0; r < bi.rows; ++r)
        for (int c = 0; c < bj.rows; ++c) S[(oj + c) * D + (oi + r)] -= blk[r * 8 + c];
      if (j != i)
        for (int r = 0; r < bi.rows; ++r)
          for (int c = 0; c < bj.rows; ++c) S[(oi + r) * D + (oj + c)] -= blk[r * 8 + c];
    }
  }
  double cg[3];
  for (int r = 0; r < 3; ++r) cg[r] = cinv[r * 3 + 0] * g[0] + cinv[r * 3 + 1] * g[1] + cinv[r * 3 + 2] * g[2];
  for (const EBlock& b : list) {
    const size_t o = off_of(b);
    for (int r = 0; r < b.rows; ++r)
      rhs[o + r] += b.e[r * 3 + 0] * cg[0] + b.e[r * 3 + 1] * cg[1] + b.e[r * 3 + 2] * cg[2];
  }
}

bool flag_at(const uint8_t* v, size_t n, size_t i) { return i < n && v[i] != 0; }

bool validate(const sfmba_input* in) {
  if (!in) return false;
  if (in->num_camera_of_image < in->num_images) return false;
  for (size_t i = 0; i < in->num_images; ++i)
    if (in->camera_of_image[i] >= in->num_cameras) return false;
  for (size_t c = 0; c < in->num_cameras; ++c)
    if (param_count(in->cameras[c].model_id) < 0) return false;
  for (size_t o = 0; o < in->num_observations; ++o) {
    if (in->observations[o].image_idx >= in->num_images) return false;
    if (in->observations[o].point_idx >= in->num_points) return false;
  }
  return true;
}

struct Trace {
  double* outer;
  size_t cap, *len;
  double* trial;
  size_t tcap, *tlen;
  void push_outer(double v) {
    if (outer && *len < cap) outer[*len] = v;
    if (len) ++*len;
  }
  void push_trial(double v) {
    if (trial && *tlen < tcap) trial[*tlen] = v;
    if (tlen) ++*tlen;
  }
};

// Stage timing of one outer iteration (CPU baseline): when non-null,
// bundle_adjust records wall-clock seconds of [0] the initial cost,
// [1] hot loop A (Jacobians), [2] hot loop B (accumulation), [3] the first
// reduced-system assembly (damped blocks, point inverses, Schur, pinning,
// symmetrization), and returns right after that assembly.
struct StageTimes {
  double t[4];
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// lib.rs:747-1232
int bundle_adjust(const sfmba_input* in, const sfmba_params* prm, sfmba_output* out,
                  Trace tr, int nthreads, StageTimes* timing = nullptr) {
  const size_t num_images = in->num_images, num_points = in->num_points;
  const size_t num_cameras = in->num_cameras;
  std::vector<sfmba_pose> poses(in->poses, in->poses + num_images);
  std::vector<sfmba_camera> cams(in->cameras, in->cameras + num_cameras);
  std::vector<double> pts(in->points, in->points + 3 * num_points);
  std::vector<int> cam_dof(num_cameras);
  for (size_t c = 0; c < num_cameras; ++c) cam_dof[c] = param_count(cams[c].model_id);

  std::vector<size_t> pose_slot(num_images, FIXED), cam_slot(num_cameras, FIXED);
  size_t D = 0;
  for (size_t i = 0; i < num_images; ++i)
    if (!flag_at(in->fixed_poses, in->num_fixed_poses, i)) {
      pose_slot[i] = D;
      D += 6;
    }
  for (size_t c = 0; c < num_cameras; ++c)
    if (cam_dof[c] > 0 && !flag_at(in->fixed_cameras, in->num_fixed_cameras, c)) {
      cam_slot[c] = D;
      D += cam_dof[c];
    }
  auto finish = [&](double initial_cost, double cost, size_t iters) {
    std::memcpy(out->poses, poses.data(), num_images * sizeof(sfmba_pose));
    std::memcpy(out->cameras, cams.data(), num_cameras * sizeof(sfmba_camera));
    std::memcpy(out->points, pts.data(), 3 * num_points * sizeof(double));
    out->initial_cost = initial_cost;
    out->final_cost = cost;
    out->iterations_run = iters;
    return 0;
  };
  if (D == 0) {
    const double c0 = total_cost(in, poses, cams, pts, prm, nthreads);
    return finish(c0, c0, 0);
  }
  // Per-camera mask offsets into the flattened fixed_camera_params.
  std::vector<size_t> mask_off(in->num_fixed_camera_params + 1, 0);
  for (size_t c = 0; c < in->num_fixed_camera_params; ++c)
    mask_off[c + 1] = mask_off[c] + in->fixed_camera_params_len[c];

  std::vector<double> S(D * D), rhs(D);
  double tmark = now_s();
  const double initial_cost = total_cost(in, poses, cams, pts, prm, nthreads);
  if (timing) {
    timing->t[0] = now_s() - tmark;
    tmark = now_s();
  }
  double lambda = prm->initial_lambda;
  double cost = initial_cost;
  double nu = 2.0;
  size_t iterations_run = 0;
  const double FUNCTION_TOLERANCE = 1e-6;
  const size_t nobs = in->num_observations;

  std::vector<Prepared> prep_all(nobs);
  std::vector<uint8_t> prep_ok(nobs);
  for (size_t outer = 0; outer < prm->max_iterations; ++outer) {
    ++iterations_run;
    const double cost_before_iter = cost;

    // Hot loop A (lib.rs:832-853)
    auto lin = [&](size_t lo, size_t hi) {
      for (size_t o = lo; o < hi; ++o) {
        const sfmba_observation& ob = in->observations[o];
        Prepared& p = prep_all[o];
        p.image = ob.image_idx;
        p.camera = in->camera_of_image[ob.image_idx];
        p.point = ob.point_idx;
        prep_ok[o] = analytic_jacobians(poses[p.image], cams[p.camera], point_at(pts, p.point),
                                        ob.x, ob.y, p.r, p.jp, p.jx, p.jc);
        if (prep_ok[o]) p.w = robust_weight(norm2(p.r), prm->robust_loss, prm->loss_scale_px);
      }
    };
    if (nthreads <= 1) {
      lin(0, nobs);
    } else {
      std::vector<std::thread> th;
      const size_t chunk = (nobs + nthreads - 1) / nthreads;
      for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t] { lin(std::min(nobs, t * chunk), std::min(nobs, (t + 1) * chunk)); });
      for (auto& x : th) x.join();
    }
    size_t nprep = 0;
    for (size_t o = 0; o < nobs; ++o) nprep += prep_ok[o];
    if (timing) {
      timing->t[1] = now_s() - tmark;
      tmark = now_s();
    }
    if (nprep == 0) break;

    // Hot loop B (lib.rs:859-909)
    std::vector<double> b_pose(num_images * 36, 0.0), g_pose(num_images * 6, 0.0);
    std::vector<std::vector<double>> b_cam(num_cameras), g_cam(num_cameras), cross(num_images);
    for (size_t c = 0; c < num_cameras; ++c) {
      b_cam[c].assign(cam_dof[c] * cam_dof[c], 0.0);
      g_cam[c].assign(cam_dof[c], 0.0);
    }
    for (size_t i = 0; i < num_images; ++i) cross[i].assign(6 * cam_dof[in->camera_of_image[i]], 0.0);
    std::vector<double> c_diag(num_points * 9, 0.0), c_rhs(num_points * 3, 0.0);
    std::vector<std::vector<EBlock>> points_to_obs(num_points);

    for (size_t o = 0; o < nobs; ++o) {
      if (!prep_ok[o]) continue;
      const Prepared& p = prep_all[o];
      const double w = p.w;
      const int k = cam_dof[p.camera];
      // w * J^T * J / w * J^T * r, the scalar applied to J^T first.
      double* C = &c_diag[p.point * 9];
      for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b)
          C[a * 3 + b] += (w * p.jx[a]) * p.jx[b] + (w * p.jx[3 + a]) * p.jx[3 + b];
        c_rhs[p.point * 3 + a] += (w * p.jx[a]) * p.r[0] + (w * p.jx[3 + a]) * p.r[1];
      }
      const bool pose_free = pose_slot[p.image] != FIXED;
      const bool cam_free = cam_slot[p.camera] != FIXED;
      if (pose_free) {
        double* B = &b_pose[p.image * 36];
        EBlock e;
        e.is_pose = true;
        e.idx = p.image;
        e.rows = 6;
        for (int a = 0; a < 6; ++a) {
          for (int b = 0; b < 6; ++b)
            B[a * 6 + b] += (w * p.jp[a]) * p.jp[b] + (w * p.jp[6 + a]) * p.jp[6 + b];
          g_pose[p.image * 6 + a] += (w * p.jp[a]) * p.r[0] + (w * p.jp[6 + a]) * p.r[1];
          for (int b = 0; b < 3; ++b)
            e.e[a * 3 + b] = (w * p.jp[a]) * p.jx[b] + (w * p.jp[6 + a]) * p.jx[3 + b];
        }
        points_to_obs[p.point].push_back(e);
      }
      if (cam_free) {
        double* B = b_cam[p.camera].data();
        EBlock e;
        e.is_pose = false;
        e.idx = p.camera;
        e.rows = k;
        for (int a = 0; a < k; ++a) {
          for (int b = 0; b < k; ++b)
            B[a * k + b] += (w * p.jc[a]) * p.jc[b] + (w * p.jc[8 + a]) * p.jc[8 + b];
          g_cam[p.camera][a] += (w * p.jc[a]) * p.r[0] + (w * p.jc[8 + a]) * p.r[1];
          for (int b = 0; b < 3; ++b)
            e.e[a * 3 + b] = (w * p.jc[a]) * p.jx[b] + (w * p.jc[8 + a]) * p.jx[3 + b];
        }
        points_to_obs[p.point].push_back(e);
      }
      if (pose_free && cam_free) {
        double* X = cross[p.image].data();
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < k; ++b)
            X[a * k + b] += (w * p.jp[a]) * p.jc[b] + (w * p.jp[6 + a]) * p.jc[8 + b];
      }
    }

    if (timing) {
      timing->t[2] = now_s() - tmark;
      tmark = now_s();
    }
    bool accepted = false;
    std::vector<double> cp_inv(num_points * 9);
    std::vector<uint8_t> cp_ok(num_points);
    for (int inner = 0; inner < 8; ++inner) {
      std::fill(S.begin(), S.end(), 0.0);
      std::fill(rhs.begin(), rhs.end(), 0.0);
      for (size_t i = 0; i < num_images; ++i) {
        const size_t off = pose_slot[i];
        if (off == FIXED) continue;
        const double* B = &b_pose[i * 36];
        for (int a = 0; a < 6; ++a) {
          for (int b = 0; b < 6; ++b) {
            double v = B[a * 6 + b];
            if (a == b) v = v + B[a * 6 + a] * lambda;
            S[(off + b) * D + off + a] = v;
          }
          rhs[off + a] = -g_pose[i * 6 + a];
        }
      }
      for (size_t c = 0; c < num_cameras; ++c) {
        const int k = cam_dof[c];
        const size_t off = cam_slot[c];
        if (k == 0 || off == FIXED) continue;
        const double* B = b_cam[c].data();
        for (int a = 0; a < k; ++a) {
          for (int b = 0; b < k; ++b) {
            double v = B[a * k + b];
            if (a == b) v = v + B[a * k + a] * lambda;
            S[(off + b) * D + off + a] = v;
          }
          rhs[off + a] = -g_cam[c][a];
        }
      }
      for (size_t i = 0; i < num_images; ++i) {
        const size_t c = in->camera_of_image[i];
        const size_t po = pose_slot[i], co = cam_slot[c];
        const int k = cam_dof[c];
        if (po == FIXED || co == FIXED || k == 0) continue;
        const double* X = cross[i].data();
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < k; ++b) {
            S[(co + b) * D + po + a] = X[a * k + b];
            S[(po + a) * D + co + b] = X[a * k + b];
          }
      }
      for (size_t p = 0; p < num_points; ++p) {
        cp_ok[p] = 0;
        if (points_to_obs[p].empty()) continue;
        const double* C = &c_diag[p * 9];
        double m[9];
        for (int a = 0; a < 9; ++a) m[a] = C[a];
        for (int a = 0; a < 3; ++a) m[a * 4] = C[a * 4] + C[a * 4] * lambda;
        cp_ok[p] = inverse3(m, &cp_inv[p * 9]);
      }

      // Hot loop C (lib.rs:967-1043)
      if (nthreads > 1 && num_points >= 1500) {
        const size_t chunk = std::max<size_t>(1, (num_points + nthreads - 1) / nthreads);
        std::vector<std::vector<double>> ps(nthreads), pr(nthreads);
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t)
          th.emplace_back([&, t] {
            ps[t].assign(D * D, 0.0);
            pr[t].assign(D, 0.0);
            const size_t lo = std::min(num_points, t * chunk), hi = std::min(num_points, (t + 1) * chunk);
            for (size_t p = lo; p < hi; ++p)
              if (cp_ok[p])
                accumulate_point_schur(&cp_inv[p * 9], points_to_obs[p], &c_rhs[p * 3], pose_slot,
                                       cam_slot, ps[t].data(), pr[t].data(), D);
          });
        for (auto& x : th) x.join();
        for (int t = 0; t < nthreads; ++t) {
          for (size_t e = 0; e < D * D; ++e) S[e] += ps[t][e];
          for (size_t e = 0; e < D; ++e) rhs[e] += pr[t][e];
        }
      } else {
        for (size_t p = 0; p < num_points; ++p)
          if (cp_ok[p])
            accumulate_point_schur(&cp_inv[p * 9], points_to_obs[p], &c_rhs[p * 3], pose_slot,
                                   cam_slot, S.data(), rhs.data(), D);
      }

      // Pinned parameters (lib.rs:1051-1070)
      for (size_t c = 0; c < num_cameras && c < in->num_fixed_camera_params; ++c) {
        const size_t len = in->fixed_camera_params_len[c];
        for (size_t k = 0; k < len; ++k) {
          if (!in->fixed_camera_params[mask_off[c] + k] || k >= (size_t)cam_dof[c]) continue;
          if (cam_slot[c] == FIXED) continue;
          const size_t row = cam_slot[c] + k;
          for (size_t col = 0; col < D; ++col) {
            S[col * D + row] = 0.0;
            S[row * D + col] = 0.0;
          }
          S[row * D + row] = 1.0;
          rhs[row] = 0.0;
        }
      }
      // Symmetrize (lib.rs:1074)
      std::vector<double> Ssym(D * D);
      for (size_t c = 0; c < D; ++c)
        for (size_t r = 0; r < D; ++r) Ssym[c * D + r] = 0.5 * (S[c * D + r] + S[r * D + c]);

      if (timing) {
        timing->t[3] = now_s() - tmark;
        return finish(initial_cost, cost, iterations_run);
      }
      std::vector<double> delta = rhs;
      if (!cholesky_solve(Ssym, D, delta)) {
        for (size_t c = 0; c < D; ++c)
          for (size_t r = 0; r < D; ++r) Ssym[c * D + r] = 0.5 * (S[c * D + r] + S[r * D + c]);
        delta = rhs;
        if (!lu_solve(Ssym, D, delta)) {
          lambda *= nu;
          nu *= 2.0;
          continue;
        }
      }

      // Back-substitution (lib.rs:1100-1120)
      std::vector<double> dpt(num_points * 3, 0.0);
      for (size_t p = 0; p < num_points; ++p) {
        if (!cp_ok[p]) continue;
        double acc[3] = {c_rhs[p * 3], c_rhs[p * 3 + 1], c_rhs[p * 3 + 2]};
        for (const EBlock& b : points_to_obs[p]) {
          const size_t off = b.is_pose ? pose_slot[b.idx] : cam_slot[b.idx];
          double con[3] = {0, 0, 0};
          for (int r = 0; r < b.rows; ++r)
            for (int c = 0; c < 3; ++c) con[c] += b.e[r * 3 + c] * delta[off + r];
          for (int c = 0; c < 3; ++c) acc[c] += con[c];
        }
        const double* ci = &cp_inv[p * 9];
        for (int r = 0; r < 3; ++r)
          dpt[p * 3 + r] = -(ci[r * 3 + 0] * acc[0] + ci[r * 3 + 1] * acc[1] + ci[r * 3 + 2] * acc[2]);
      }

      // Trial state (lib.rs:1122-1175)
      std::vector<sfmba_pose> tposes(num_images);
      for (size_t i = 0; i < num_images; ++i) {
        double d[6] = {0, 0, 0, 0, 0, 0};
        if (pose_slot[i] != FIXED)
          for (int a = 0; a < 6; ++a) d[a] = delta[pose_slot[i] + a];
        tposes[i] = perturb_pose(poses[i], d);
      }
      std::vector<double> tpts(pts);
      for (size_t e = 0; e < 3 * num_points; ++e) tpts[e] += dpt[e];
      std::vector<sfmba_camera> tcams(cams);
      for (size_t c = 0; c < num_cameras; ++c) {
        const int k = cam_dof[c];
        if (k == 0 || cam_slot[c] == FIXED) continue;
        for (int j = 0; j < k; ++j) tcams[c].params[j] += delta[cam_slot[c] + j];
      }
      const double trial_cost = total_cost(in, tposes, tcams, tpts, prm, nthreads);
      tr.push_trial(trial_cost);
      if (std::isfinite(trial_cost) && trial_cost < cost) {
        poses.swap(tposes);
        pts.swap(tpts);
        cams.swap(tcams);
        cost = trial_cost;
        lambda = std::max(lambda / 3.0, 1e-12);
        nu = 2.0;
        accepted = true;
        break;
      } else {
        lambda *= nu;
        nu *= 2.0;
      }
    }
    tr.push_outer(cost);
    if (!accepted) break;
    const double rel = (cost_before_iter - cost) / std::max(cost_before_iter, 1e-12);
    if (rel < FUNCTION_TOLERANCE) break;
  }
  return finish(initial_cost, cost, iterations_run);
}

// intrinsics.rs:92-111
double camera_cost(const sfmba_camera& cam, const std::vector<const sfmba_pose*>& poses,
                   const std::vector<V3>& pts, const std::vector<double>& obs,
                   const sfmba_intrinsics_params* prm) {
  double s = 0.0;
  for (size_t i = 0; i < poses.size(); ++i) {
    double r[2];
    if (reprojection_residual(*poses[i], cam, pts[i], obs[2 * i], obs[2 * i + 1], r)) {
      const double n = norm2(r);
      const double w = robust_weight(n, prm->robust_loss, prm->loss_scale_px);
      s += w * n * n;
    }
  }
  return s;
}

// intrinsics.rs:148-238
sfmba_camera refine_one_camera(const sfmba_camera& cam0, const std::vector<const sfmba_observation*>& ol,
                               const sfmba_pose* poses, const double* points,
                               const sfmba_intrinsics_params* prm) {
  const int n = param_count(cam0.model_id);
  std::vector<const sfmba_pose*> ip;
  std::vector<V3> pts;
  std::vector<double> px;
  for (auto* o : ol) {
    ip.push_back(&poses[o->image_idx]);
    pts.push_back(V3{points[3 * o->point_idx], points[3 * o->point_idx + 1], points[3 * o->point_idx + 2]});
    px.push_back(o->x);
    px.push_back(o->y);
  }
  const double EPS = 1e-6;
  sfmba_camera cur = cam0;
  double cost = camera_cost(cur, ip, pts, px, prm);
  double lambda = prm->initial_lambda;
  for (size_t it = 0; it < prm->max_iterations; ++it) {
    std::vector<double> jtj(n * n, 0.0), jtr(n, 0.0);
    for (size_t i = 0; i < ol.size(); ++i) {
      double r0[2];
      if (!reprojection_residual(*ip[i], cur, pts[i], px[2 * i], px[2 * i + 1], r0)) continue;
      const double w = robust_weight(norm2(r0), prm->robust_loss, prm->loss_scale_px);
      double jac[8][2];
      for (int k = 0; k < n; ++k) {
        sfmba_camera cp = cur, cm = cur;
        cp.params[k] += EPS;
        cm.params[k] += -EPS;
        double rp[2], rm[2];
        if (!reprojection_residual(*ip[i], cp, pts[i], px[2 * i], px[2 * i + 1], rp)) { rp[0] = r0[0]; rp[1] = r0[1]; }
        if (!reprojection_residual(*ip[i], cm, pts[i], px[2 * i], px[2 * i + 1], rm)) { rm[0] = r0[0]; rm[1] = r0[1]; }
        jac[k][0] = (rp[0] - rm[0]) / (2.0 * EPS);
        jac[k][1] = (rp[1] - rm[1]) / (2.0 * EPS);
      }
      for (int a = 0; a < n; ++a) {
        jtr[a] += w * (jac[a][0] * r0[0] + jac[a][1] * r0[1]);
        for (int b = 0; b < n; ++b) jtj[a * n + b] += w * (jac[a][0] * jac[b][0] + jac[a][1] * jac[b][1]);
      }
    }
    bool accepted = false;
    for (int t = 0; t < 8; ++t) {
      std::vector<double> damped(n * n);  // column-major == row-major (symmetric build)
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) damped[b * n + a] = jtj[a * n + b];
      for (int a = 0; a < n; ++a) damped[a * n + a] += lambda * std::max(jtj[a * n + a], 1e-12);
      std::vector<double> delta(n);
      for (int a = 0; a < n; ++a) delta[a] = -jtr[a];
      std::vector<double> A = damped;
      if (!cholesky_solve(A, n, delta)) {
        A = damped;
        for (int a = 0; a < n; ++a) delta[a] = -jtr[a];
        if (!lu_solve(A, n, delta)) {
          lambda *= 3.0;
          continue;
        }
      }
      sfmba_camera trial = cur;
      for (int k = 0; k < n; ++k) trial.params[k] += delta[k];
      const double tc = camera_cost(trial, ip, pts, px, prm);
      if (std::isfinite(tc) && tc < cost) {
        cur = trial;
        cost = tc;
        lambda = std::max(lambda / 3.0, 1e-12);
        accepted = true;
        break;
      }
      lambda *= 3.0;
    }
    if (!accepted) break;
  }
  return cur;
}

}  // namespace

extern "C" {

int oracle_param_count(int model) { return param_count(model); }

int oracle_project(const sfmba_camera* cam, const double* pc, double* uv) {
  project(*cam, V3{pc[0], pc[1], pc[2]}, &uv[0], &uv[1]);
  return 0;
}

int oracle_transform_point(const sfmba_pose* pose, const double* X, double* out) {
  V3 r = transform_point(*pose, V3{X[0], X[1], X[2]});
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  return 0;
}

int oracle_rotation_matrix(const sfmba_pose* pose, double* R) {
  rotation_matrix(*pose, R);
  return 0;
}

int oracle_perturb_pose(const sfmba_pose* pose, const double* d, sfmba_pose* out) {
  *out = perturb_pose(*pose, d);
  return 0;
}

double oracle_robust_weight(double n, int loss, double scale) { return robust_weight(n, loss, scale); }

// 1 = valid, 0 = behind the camera.
int oracle_reprojection_residual(const sfmba_pose* pose, const sfmba_camera* cam, const double* X,
                                 double ox, double oy, double* r) {
  return reprojection_residual(*pose, *cam, V3{X[0], X[1], X[2]}, ox, oy, r) ? 1 : 0;
}

int oracle_analytic_jacobians(const sfmba_pose* pose, const sfmba_camera* cam, const double* X,
                              double ox, double oy, double* r, double* jp, double* jx, double* jc) {
  return analytic_jacobians(*pose, *cam, V3{X[0], X[1], X[2]}, ox, oy, r, jp, jx, jc) ? 1 : 0;
}

// lib.rs:599-670: central differences (h = eps) of the residual w.r.t. the
// 6 pose parameters (through perturb_pose), the point and the camera's own
// parameters; a perturbation that fails falls back to r0.
int oracle_numeric_jacobians(const sfmba_pose* pose, const sfmba_camera* cam, const double* Xp,
                             double ox, double oy, double eps, double* r0, double* jp, double* jx,
                             double* jc) {
  const V3 X{Xp[0], Xp[1], Xp[2]};
  if (!reprojection_residual(*pose, *cam, X, ox, oy, r0)) return 0;
  for (int k = 0; k < 6; ++k) {
    double dp[6] = {0, 0, 0, 0, 0, 0}, dm[6] = {0, 0, 0, 0, 0, 0};
    dp[k] = eps;
    dm[k] = -eps;
    sfmba_pose pp = perturb_pose(*pose, dp), pm = perturb_pose(*pose, dm);
    double rp[2], rm[2];
    if (!reprojection_residual(pp, *cam, X, ox, oy, rp)) { rp[0] = r0[0]; rp[1] = r0[1]; }
    if (!reprojection_residual(pm, *cam, X, ox, oy, rm)) { rm[0] = r0[0]; rm[1] = r0[1]; }
    jp[k] = (rp[0] - rm[0]) / (2.0 * eps);
    jp[6 + k] = (rp[1] - rm[1]) / (2.0 * eps);
  }
  for (int k = 0; k < 3; ++k) {
    V3 xp = X, xm = X;
    (&xp.x)[k] += eps;
    (&xm.x)[k] -= eps;
    double rp[2], rm[2];
    if (!reprojection_residual(*pose, *cam, xp, ox, oy, rp)) { rp[0] = r0[0]; rp[1] = r0[1]; }
    if (!reprojection_residual(*pose, *cam, xm, ox, oy, rm)) { rm[0] = r0[0]; rm[1] = r0[1]; }
    jx[k] = (rp[0] - rm[0]) / (2.0 * eps);
    jx[3 + k] = (rp[1] - rm[1]) / (2.0 * eps);
  }
  for (int i = 0; i < 16; ++i) jc[i] = 0.0;
  const int n = param_count(cam->model_id);
  for (int k = 0; k < n; ++k) {
    sfmba_camera cp = *cam, cm = *cam;
    cp.params[k] += eps;
    cm.params[k] += -eps;
    double rp[2], rm[2];
    if (!reprojection_residual(*pose, cp, X, ox, oy, rp)) { rp[0] = r0[0]; rp[1] = r0[1]; }
    if (!reprojection_residual(*pose, cm, X, ox, oy, rm)) { rm[0] = r0[0]; rm[1] = r0[1]; }
    jc[k] = (rp[0] - rm[0]) / (2.0 * eps);
    jc[8 + k] = (rp[1] - rm[1]) / (2.0 * eps);
  }
  return 1;
}

int oracle_total_cost(const sfmba_input* in, const sfmba_params* prm, double* out) {
  if (!validate(in)) return SFMBA_ERR_INVALID_ARGUMENT;
  std::vector<sfmba_pose> poses(in->poses, in->poses + in->num_images);
  std::vector<sfmba_camera> cams(in->cameras, in->cameras + in->num_cameras);
  std::vector<double> pts(in->points, in->points + 3 * in->num_points);
  *out = total_cost(in, poses, cams, pts, prm, 1);
  return 0;
}

// |reprojection_residual| of every observation (lib.rs:249-261) and its
// in-front flag (0: behind the camera, norm 0): the plain RMS and mean
// reprojection statistics of a state.
int oracle_residual_norms(const sfmba_input* in, double* norm, uint8_t* in_front) {
  if (!validate(in)) return SFMBA_ERR_INVALID_ARGUMENT;
  for (size_t o = 0; o < in->num_observations; ++o) {
    const sfmba_observation& ob = in->observations[o];
    double r[2];
    const double* X = &in->points[3 * ob.point_idx];
    const bool ok = reprojection_residual(in->poses[ob.image_idx], in->cameras[in->camera_of_image[ob.image_idx]],
                                          V3{X[0], X[1], X[2]}, ob.x, ob.y, r);
    in_front[o] = ok ? 1 : 0;
    norm[o] = ok ? norm2(r) : 0.0;
  }
  return 0;
}

// lib.rs:708-725
int oracle_mean_reprojection_error(const sfmba_input* in, double* out) {
  if (!validate(in)) return SFMBA_ERR_INVALID_ARGUMENT;
  double s = 0.0;
  size_t count = 0;
  for (size_t o = 0; o < in->num_observations; ++o) {
    const sfmba_observation& ob = in->observations[o];
    double r[2];
    const double* X = &in->points[3 * ob.point_idx];
    if (reprojection_residual(in->poses[ob.image_idx], in->cameras[in->camera_of_image[ob.image_idx]],
                              V3{X[0], X[1], X[2]}, ob.x, ob.y, r)) {
      s += norm2(r);
      ++count;
    }
  }
  *out = count == 0 ? 0.0 : s / (double)count;
  return 0;
}

int oracle_bundle_adjust(const sfmba_input* in, const sfmba_params* prm, sfmba_output* out,
                         int nthreads, double* outer_costs, size_t outer_cap, size_t* outer_len,
                         double* trial_costs, size_t trial_cap, size_t* trial_len) {
  if (!validate(in) || !prm || !out) return SFMBA_ERR_INVALID_ARGUMENT;
  size_t dummy_len = 0, dummy_tlen = 0;
  if (outer_len) *outer_len = 0;
  if (trial_len) *trial_len = 0;
  Trace tr{outer_costs, outer_cap, outer_len ? outer_len : &dummy_len,
           trial_costs, trial_cap, trial_len ? trial_len : &dummy_tlen};
  return bundle_adjust(in, prm, out, tr, nthreads);
}

// intrinsics.rs:49-61
// CPU-baseline stage timing on the reference's threading structure
// (parallel Jacobians, serial accumulation, per-thread D x D Schur
// accumulators when num_points >= 1500); times[0..3] as in StageTimes,
// times[4] = D of the reference layout.  The state is not advanced.
int oracle_stage_times(const sfmba_input* in, const sfmba_params* prm, int nthreads, double* times) {
  if (!validate(in) || !prm || !times) return SFMBA_ERR_INVALID_ARGUMENT;
  std::vector<sfmba_pose> p(in->num_images);
  std::vector<sfmba_camera> c(in->num_cameras);
  std::vector<double> x(3 * in->num_points);
  sfmba_output out{p.data(), x.data(), c.data(), 0, 0, 0};
  size_t a = 0, b = 0;
  Trace tr{nullptr, 0, &a, nullptr, 0, &b};
  StageTimes st{{0, 0, 0, 0}};
  sfmba_params one = *prm;
  one.max_iterations = 1;
  bundle_adjust(in, &one, &out, tr, nthreads, &st);
  for (int i = 0; i < 4; ++i) times[i] = st.t[i];
  size_t D = 0;
  for (size_t i = 0; i < in->num_images; ++i) D += flag_at(in->fixed_poses, in->num_fixed_poses, i) ? 0 : 6;
  for (size_t k = 0; k < in->num_cameras; ++k)
    D += flag_at(in->fixed_cameras, in->num_fixed_cameras, k) ? 0 : param_count(in->cameras[k].model_id);
  times[4] = (double)D;
  return 0;
}

// Seconds of one single-threaded Cholesky factorization + solve (the
// reference's dense solve, lib.rs:1081-1082) of an n x n SPD matrix.
double oracle_cholesky_seconds(size_t n) {
  std::vector<double> A(n * n), b(n, 1.0);
  for (size_t c = 0; c < n; ++c)
    for (size_t r = 0; r < n; ++r) A[c * n + r] = (r == c ? (double)n : 0.0) + 1.0 / (1.0 + r + c);
  const double t0 = now_s();
  cholesky_solve(A, n, b);
  return now_s() - t0;
}

int oracle_default_fixed_params_mask(const sfmba_camera* cams, size_t n, uint8_t* mask, size_t* lens) {
  for (size_t c = 0; c < n; ++c) {
    const int k = param_count(cams[c].model_id);
    if (k < 0) return SFMBA_ERR_INVALID_ARGUMENT;
    lens[c] = (size_t)k;
    for (int i = 0; i < 8; ++i) mask[c * 8 + i] = 0;
    const int m = cams[c].model_id;
    const bool simple = m == SFMBA_SIMPLE_PINHOLE || m == SFMBA_SIMPLE_RADIAL || m == SFMBA_RADIAL;
    mask[c * 8 + (simple ? 1 : 2)] = 1;
    mask[c * 8 + (simple ? 2 : 3)] = 1;
  }
  return 0;
}

// intrinsics.rs:116-146
int oracle_refine_intrinsics(sfmba_camera* cameras, size_t num_cameras, const uint64_t* camera_of_image,
                             size_t num_images, const sfmba_pose* poses, const double* points,
                             size_t num_points, const sfmba_observation* obs, size_t nobs,
                             const sfmba_intrinsics_params* prm) {
  std::vector<std::vector<const sfmba_observation*>> by_cam(num_cameras);
  for (size_t o = 0; o < nobs; ++o) {
    if (obs[o].image_idx >= num_images || obs[o].point_idx >= num_points) return SFMBA_ERR_INVALID_ARGUMENT;
    const uint64_t c = camera_of_image[obs[o].image_idx];
    if (c >= num_cameras) return SFMBA_ERR_INVALID_ARGUMENT;
    by_cam[c].push_back(&obs[o]);
  }
  std::vector<sfmba_camera> refined(cameras, cameras + num_cameras);
  for (size_t c = 0; c < num_cameras; ++c) {
    if (by_cam[c].size() < 8) continue;
    refined[c] = refine_one_camera(cameras[c], by_cam[c], poses, points, prm);
  }
  std::memcpy(cameras, refined.data(), num_cameras * sizeof(sfmba_camera));
  return 0;
}

// The reduced camera system of one linearization at damping lambda, in the
// reference layout (D_ref, pinned rows set to identity, symmetrized), so
// tests can check the Schur complement against a brute-force solve of the
// full damped normal equations.  Returns D (or a negative error); S is
// column-major and must hold D*D doubles.
long oracle_reduced_system(const sfmba_input* in, const sfmba_params* prm, double lambda, double* S_out,
                           double* rhs_out, size_t cap) {
  if (!validate(in)) return SFMBA_ERR_INVALID_ARGUMENT;
  // Reuse bundle_adjust's first linearization by running it with a capture
  // hook would complicate the restatement; the assembly is repeated here in
  // the same order instead.
  const size_t num_images = in->num_images, num_points = in->num_points, num_cameras = in->num_cameras;
  std::vector<int> cam_dof(num_cameras);
  for (size_t c = 0; c < num_cameras; ++c) cam_dof[c] = param_count(in->cameras[c].model_id);
  std::vector<size_t> pose_slot(num_images, FIXED), cam_slot(num_cameras, FIXED);
  size_t D = 0;
  for (size_t i = 0; i < num_images; ++i)
    if (!flag_at(in->fixed_poses, in->num_fixed_poses, i)) { pose_slot[i] = D; D += 6; }
  for (size_t c = 0; c < num_cameras; ++c)
    if (!flag_at(in->fixed_cameras, in->num_fixed_cameras, c)) { cam_slot[c] = D; D += cam_dof[c]; }
  if (D * D > cap) return (long)D;
  std::vector<double> S(D * D, 0.0), rhs(D, 0.0);
  std::vector<double> c_diag(num_points * 9, 0.0), c_rhs(num_points * 3, 0.0);
  std::vector<std::vector<EBlock>> pto(num_points);
  for (size_t o = 0; o < in->num_observations; ++o) {
    const sfmba_observation& ob = in->observations[o];
    const size_t cam = in->camera_of_image[ob.image_idx];
    Prepared p;
    const double* X = &in->points[3 * ob.point_idx];
    if (!analytic_jacobians(in->poses[ob.image_idx], in->cameras[cam], V3{X[0], X[1], X[2]}, ob.x, ob.y,
                            p.r, p.jp, p.jx, p.jc))
      continue;
    const double w = robust_weight(norm2(p.r), prm->robust_loss, prm->loss_scale_px);
    const int k = cam_dof[cam];
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b)
        c_diag[ob.point_idx * 9 + a * 3 + b] += (w * p.jx[a]) * p.jx[b] + (w * p.jx[3 + a]) * p.jx[3 + b];
      c_rhs[ob.point_idx * 3 + a] += (w * p.jx[a]) * p.r[0] + (w * p.jx[3 + a]) * p.r[1];
    }
    // Rows of [Jp | Jc] for the free sides, and their slots.
    std::vector<std::pair<size_t, const double*>> rows;  // (slot, row ptr pair via index)
    double J[2][14];
    size_t slot[14];
    int nr = 0;
    bool pf = pose_slot[ob.image_idx] != FIXED, cf = cam_slot[cam] != FIXED;
    if (pf) {
      EBlock e{true, ob.image_idx, 6, {}};
      for (int a = 0; a < 6; ++a) {
        J[0][nr] = p.jp[a]; J[1][nr] = p.jp[6 + a]; slot[nr++] = pose_slot[ob.image_idx] + a;
        for (int b = 0; b < 3; ++b) e.e[a * 3 + b] = (w * p.jp[a]) * p.jx[b] + (w * p.jp[6 + a]) * p.jx[3 + b];
      }
      pto[ob.point_idx].push_back(e);
    }
    if (cf) {
      EBlock e{false, cam, k, {}};
      for (int a = 0; a < k; ++a) {
        J[0][nr] = p.jc[a]; J[1][nr] = p.jc[8 + a]; slot[nr++] = cam_slot[cam] + a;
        for (int b = 0; b < 3; ++b) e.e[a * 3 + b] = (w * p.jc[a]) * p.jx[b] + (w * p.jc[8 + a]) * p.jx[3 + b];
      }
      pto[ob.point_idx].push_back(e);
    }
    for (int a = 0; a < nr; ++a) {
      for (int b = 0; b < nr; ++b) S[slot[b] * D + slot[a]] += (w * J[0][a]) * J[0][b] + (w * J[1][a]) * J[1][b];
      rhs[slot[a]] -= (w * J[0][a]) * p.r[0] + (w * J[1][a]) * p.r[1];
    }
  }
  // Damping on the diagonal blocks' diagonal.
  for (size_t e = 0; e < D; ++e) S[e * D + e] += S[e * D + e] * lambda;
  for (size_t p = 0; p < num_points; ++p) {
    if (pto[p].empty()) continue;
    double m[9], ci[9];
    for (int a = 0; a < 9; ++a) m[a] = c_diag[p * 9 + a];
    for (int a = 0; a < 3; ++a) m[a * 4] += c_diag[p * 9 + a * 4] * lambda;
    if (!inverse3(m, ci)) continue;
    accumulate_point_schur(ci, pto[p], &c_rhs[p * 3], pose_slot, cam_slot, S.data(), rhs.data(), D);
  }
  std::vector<size_t> moff(in->num_fixed_camera_params + 1, 0);
  for (size_t c = 0; c < in->num_fixed_camera_params; ++c) moff[c + 1] = moff[c] + in->fixed_camera_params_len[c];
  for (size_t c = 0; c < num_cameras && c < in->num_fixed_camera_params; ++c)
    for (size_t k = 0; k < in->fixed_camera_params_len[c]; ++k) {
      if (!in->fixed_camera_params[moff[c] + k] || k >= (size_t)cam_dof[c] || cam_slot[c] == FIXED) continue;
      const size_t row = cam_slot[c] + k;
      for (size_t col = 0; col < D; ++col) { S[col * D + row] = 0.0; S[row * D + col] = 0.0; }
      S[row * D + row] = 1.0;
      rhs[row] = 0.0;
    }
  for (size_t c = 0; c < D; ++c)
    for (size_t r = 0; r < D; ++r) S_out[c * D + r] = 0.5 * (S[c * D + r] + S[r * D + c]);
  std::memcpy(rhs_out, rhs.data(), D * sizeof(double));
  return (long)D;
}

int oracle_cholesky_solve(double* A, size_t n, double* b) {
  std::vector<double> a(A, A + n * n), x(b, b + n);
  if (!cholesky_solve(a, n, x)) return 0;
  std::memcpy(A, a.data(), n * n * sizeof(double));
  std::memcpy(b, x.data(), n * sizeof(double));
  return 1;
}

int oracle_lu_solve(double* A, size_t n, double* b) {
  std::vector<double> a(A, A + n * n), x(b, b + n);
  if (!lu_solve(a, n, x)) return 0;
  std::memcpy(b, x.data(), n * sizeof(double));
  return 1;
}

int oracle_inverse3(const double* m, double* o) { return inverse3(m, o) ? 1 : 0; }

}  // extern "C"
