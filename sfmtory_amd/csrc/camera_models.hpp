// Residual and analytic Jacobian of one observation, for every COLMAP model
// the reference supports, as device functions.
//
// Residual: crates/sfm-ba/src/lib.rs:249-261 (via CameraModel::project,
// crates/sfm-core/src/camera.rs:204-244) for the cost, and the residual
// expression inside analytic_jacobians (lib.rs:353-585) for linearization.
// Jacobians: analytic_jacobians (lib.rs:327-590) with the pose/point chain
// rule of pose_point_jacobians (lib.rs:282-310): left-multiplicative
// rotation perturbation, d(pc)/d(omega) = -[R X]x, d(pc)/d(t) = I.
//
// The residual is always evaluated in FP64 (pixel coordinates near 2000 px
// leave FP32 only ~1e-4 px of resolution); the Jacobian is templated on the
// arithmetic type so the mixed-precision path builds it in FP32.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sfmba.h"

namespace sfmba {

__host__ __device__ inline int model_param_count(int model) {
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

// X_cam = q X q* + t with the two-cross-product form of the quaternion
// rotation (t = 2 q_v x X, X' = X + w t + q_v x t), pose = (w x y z tx ty tz).
__device__ inline void transform_point(const double* pose, double X0, double X1, double X2,
                                       double& c0, double& c1, double& c2) {
  const double w = pose[0], qx = pose[1], qy = pose[2], qz = pose[3];
  const double t0 = 2.0 * (qy * X2 - qz * X1);
  const double t1 = 2.0 * (qz * X0 - qx * X2);
  const double t2 = 2.0 * (qx * X1 - qy * X0);
  const double u0 = qy * t2 - qz * t1;
  const double u1 = qz * t0 - qx * t2;
  const double u2 = qx * t1 - qy * t0;
  c0 = (t0 * w + u0 + X0) + pose[4];
  c1 = (t1 * w + u1 + X1) + pose[5];
  c2 = (t2 * w + u2 + X2) + pose[6];
}

// Row-major rotation matrix of the pose quaternion (homogeneous form).
template <typename T>
__device__ inline void rotation_matrix(const double* pose, T R[9]) {
  const double w = pose[0], i = pose[1], j = pose[2], k = pose[3];
  const double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  const double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0;
  const double ik = i * k * 2.0, jk = j * k * 2.0, wi = w * i * 2.0;
  R[0] = (T)(ww + ii - jj - kk); R[1] = (T)(ij - wk);           R[2] = (T)(wj + ik);
  R[3] = (T)(wk + ij);           R[4] = (T)(ww - ii + jj - kk); R[5] = (T)(jk - wi);
  R[6] = (T)(ik - wj);           R[7] = (T)(wi + jk);           R[8] = (T)(ww - ii - jj + kk);
}

// CameraModel::project on a camera-space point (division form), then the
// residual against the observation.  Used for the cost and the mean error.
__device__ inline void project_residual(int model, const double* P, double c0, double c1,
                                        double c2, double ox, double oy, double& r0,
                                        double& r1) {
  const double x = c0 / c2, y = c1 / c2;
  double u, v;
  switch (model) {
    case SFMBA_SIMPLE_PINHOLE: u = P[0] * x + P[1]; v = P[0] * y + P[2]; break;
    case SFMBA_PINHOLE: u = P[0] * x + P[2]; v = P[1] * y + P[3]; break;
    case SFMBA_SIMPLE_RADIAL: {
      const double r2 = x * x + y * y;
      const double d = 1.0 + P[3] * r2;
      u = P[0] * (x * d) + P[1];
      v = P[0] * (y * d) + P[2];
      break;
    }
    case SFMBA_RADIAL: {
      const double r2 = x * x + y * y;
      const double d = 1.0 + P[3] * r2 + P[4] * r2 * r2;
      u = P[0] * (x * d) + P[1];
      v = P[0] * (y * d) + P[2];
      break;
    }
    case SFMBA_OPENCV: {
      const double k1 = P[4], k2 = P[5], p1 = P[6], p2 = P[7];
      const double r2 = x * x + y * y;
      const double rad = 1.0 + k1 * r2 + k2 * r2 * r2;
      const double xd = x * rad + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
      const double yd = y * rad + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
      u = P[0] * xd + P[2];
      v = P[1] * yd + P[3];
      break;
    }
    default: {  // SFMBA_OPENCV_FISHEYE
      const double rr = fmax(sqrt(x * x + y * y), 1e-12);
      const double th = atan(rr);
      const double t2 = th * th, t4 = t2 * t2, t6 = t4 * t2, t8 = t4 * t4;
      const double td = th * (1.0 + P[4] * t2 + P[5] * t4 + P[6] * t6 + P[7] * t8);
      const double sc = td / rr;
      u = P[0] * (x * sc) + P[2];
      v = P[1] * (y * sc) + P[3];
      break;
    }
  }
  r0 = u - ox;
  r1 = v - oy;
}

// Residual in the form analytic_jacobians evaluates it (xp = pc.x * (1/z)).
__device__ inline void linearization_residual(int model, const double* P, double c0, double c1,
                                              double c2, double ox, double oy, double& r0,
                                              double& r1) {
  const double inv_z = 1.0 / c2;
  const double xp = c0 * inv_z, yp = c1 * inv_z;
  switch (model) {
    case SFMBA_SIMPLE_PINHOLE: r0 = P[0] * xp + P[1] - ox; r1 = P[0] * yp + P[2] - oy; return;
    case SFMBA_PINHOLE: r0 = P[0] * xp + P[2] - ox; r1 = P[1] * yp + P[3] - oy; return;
    case SFMBA_SIMPLE_RADIAL: {
      const double d = 1.0 + P[3] * (xp * xp + yp * yp);
      r0 = P[0] * xp * d + P[1] - ox;
      r1 = P[0] * yp * d + P[2] - oy;
      return;
    }
    case SFMBA_RADIAL: {
      const double r2 = xp * xp + yp * yp;
      const double d = 1.0 + P[3] * r2 + P[4] * r2 * r2;
      r0 = P[0] * xp * d + P[1] - ox;
      r1 = P[0] * yp * d + P[2] - oy;
      return;
    }
    case SFMBA_OPENCV: {
      const double k1 = P[4], k2 = P[5], p1 = P[6], p2 = P[7];
      const double r2 = xp * xp + yp * yp;
      const double rad = 1.0 + k1 * r2 + k2 * r2 * r2;
      const double xd = xp * rad + 2.0 * p1 * xp * yp + p2 * (r2 + 2.0 * xp * xp);
      const double yd = yp * rad + p1 * (r2 + 2.0 * yp * yp) + 2.0 * p2 * xp * yp;
      r0 = P[0] * xd + P[2] - ox;
      r1 = P[1] * yd + P[3] - oy;
      return;
    }
    default: {
      const double rr = fmax(sqrt(xp * xp + yp * yp), 1e-12);
      const double th = atan(rr);
      const double t2 = th * th, t4 = t2 * t2, t6 = t4 * t2, t8 = t4 * t4;
      const double sc = th * (1.0 + P[4] * t2 + P[5] * t4 + P[6] * t6 + P[7] * t8) / rr;
      r0 = P[0] * (xp * sc) + P[2] - ox;
      r1 = P[1] * (yp * sc) + P[3] - oy;
      return;
    }
  }
}

// Analytic Jacobian: du/dv = d(u,v)/d(pc) (3 each), jc = d(u,v)/d(params)
// as two rows of 8 in CameraModel::params() order.
template <typename T>
__device__ inline void model_jacobian(int model, const double* Pd, T c0, T c1, T c2, T du[3],
                                      T dv[3], T jc0[8], T jc1[8]) {
  const T inv_z = T(1) / c2;
  const T xp = c0 * inv_z, yp = c1 * inv_z;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    jc0[i] = T(0);
    jc1[i] = T(0);
  }
  T du_dxp, du_dyp, dv_dxp, dv_dyp;
  switch (model) {
    case SFMBA_SIMPLE_PINHOLE: {
      const T f = (T)Pd[0];
      du[0] = f * inv_z; du[1] = T(0); du[2] = -f * xp * inv_z;
      dv[0] = T(0); dv[1] = f * inv_z; dv[2] = -f * yp * inv_z;
      jc0[0] = xp; jc0[1] = T(1);
      jc1[0] = yp; jc1[2] = T(1);
      return;
    }
    case SFMBA_PINHOLE: {
      const T fx = (T)Pd[0], fy = (T)Pd[1];
      du[0] = fx * inv_z; du[1] = T(0); du[2] = -fx * xp * inv_z;
      dv[0] = T(0); dv[1] = fy * inv_z; dv[2] = -fy * yp * inv_z;
      jc0[0] = xp; jc0[2] = T(1);
      jc1[1] = yp; jc1[3] = T(1);
      return;
    }
    case SFMBA_SIMPLE_RADIAL: {
      const T f = (T)Pd[0], k = (T)Pd[3];
      const T r2 = xp * xp + yp * yp;
      const T d = T(1) + k * r2;
      du_dxp = f * (d + T(2) * k * xp * xp);
      du_dyp = T(2) * f * k * xp * yp;
      dv_dxp = du_dyp;
      dv_dyp = f * (d + T(2) * k * yp * yp);
      jc0[0] = xp * d; jc0[1] = T(1); jc0[3] = f * xp * r2;
      jc1[0] = yp * d; jc1[2] = T(1); jc1[3] = f * yp * r2;
      break;
    }
    case SFMBA_RADIAL: {
      const T f = (T)Pd[0], k1 = (T)Pd[3], k2 = (T)Pd[4];
      const T r2 = xp * xp + yp * yp;
      const T d = T(1) + k1 * r2 + k2 * r2 * r2;
      const T kr = k1 + T(2) * k2 * r2;
      du_dxp = f * (d + T(2) * kr * xp * xp);
      du_dyp = T(2) * f * kr * xp * yp;
      dv_dxp = du_dyp;
      dv_dyp = f * (d + T(2) * kr * yp * yp);
      jc0[0] = xp * d; jc0[1] = T(1); jc0[3] = f * xp * r2; jc0[4] = f * xp * r2 * r2;
      jc1[0] = yp * d; jc1[2] = T(1); jc1[3] = f * yp * r2; jc1[4] = f * yp * r2 * r2;
      break;
    }
    case SFMBA_OPENCV: {
      const T fx = (T)Pd[0], fy = (T)Pd[1];
      const T k1 = (T)Pd[4], k2 = (T)Pd[5], p1 = (T)Pd[6], p2 = (T)Pd[7];
      const T r2 = xp * xp + yp * yp;
      const T rad = T(1) + k1 * r2 + k2 * r2 * r2;
      const T xd = xp * rad + T(2) * p1 * xp * yp + p2 * (r2 + T(2) * xp * xp);
      const T yd = yp * rad + p1 * (r2 + T(2) * yp * yp) + T(2) * p2 * xp * yp;
      const T kr = k1 + T(2) * k2 * r2;
      const T dxd_dxp = rad + T(2) * kr * xp * xp + T(2) * p1 * yp + T(6) * p2 * xp;
      const T dxd_dyp = T(2) * kr * xp * yp + T(2) * p1 * xp + T(2) * p2 * yp;
      const T dyd_dyp = rad + T(2) * kr * yp * yp + T(6) * p1 * yp + T(2) * p2 * xp;
      du_dxp = fx * dxd_dxp;
      du_dyp = fx * dxd_dyp;
      dv_dxp = fy * dxd_dyp;
      dv_dyp = fy * dyd_dyp;
      jc0[0] = xd; jc0[2] = T(1);
      jc0[4] = fx * xp * r2; jc0[5] = fx * xp * r2 * r2;
      jc0[6] = fx * T(2) * xp * yp; jc0[7] = fx * (r2 + T(2) * xp * xp);
      jc1[1] = yd; jc1[3] = T(1);
      jc1[4] = fy * yp * r2; jc1[5] = fy * yp * r2 * r2;
      jc1[6] = fy * (r2 + T(2) * yp * yp); jc1[7] = fy * T(2) * xp * yp;
      break;
    }
    default: {  // SFMBA_OPENCV_FISHEYE
      const T fx = (T)Pd[0], fy = (T)Pd[1];
      const T k1 = (T)Pd[4], k2 = (T)Pd[5], k3 = (T)Pd[6], k4 = (T)Pd[7];
      const T rr = fmax(sqrt(xp * xp + yp * yp), T(1e-12));
      const T th = atan(rr);
      const T t2 = th * th, t4 = t2 * t2, t6 = t4 * t2, t8 = t4 * t4;
      const T poly = T(1) + k1 * t2 + k2 * t4 + k3 * t6 + k4 * t8;
      const T td = th * poly;
      const T sc = td / rr;
      const T pd = k1 + T(2) * k2 * t2 + T(3) * k3 * t4 + T(4) * k4 * t6;
      const T dtd = poly + T(2) * t2 * pd;
      const T dth = T(1) / (T(1) + rr * rr);
      const T dsc = (dtd * dth * rr - td) / (rr * rr);
      const T dsx = dsc * (xp / rr), dsy = dsc * (yp / rr);
      du_dxp = fx * (sc + xp * dsx);
      du_dyp = fx * xp * dsy;
      dv_dxp = fy * yp * dsx;
      dv_dyp = fy * (sc + yp * dsy);
      const T d1 = th * t2, d2 = th * t4, d3 = th * t6, d4 = th * t8;
      jc0[0] = xp * sc; jc0[2] = T(1);
      jc0[4] = fx * xp * d1 / rr; jc0[5] = fx * xp * d2 / rr;
      jc0[6] = fx * xp * d3 / rr; jc0[7] = fx * xp * d4 / rr;
      jc1[1] = yp * sc; jc1[3] = T(1);
      jc1[4] = fy * yp * d1 / rr; jc1[5] = fy * yp * d2 / rr;
      jc1[6] = fy * yp * d3 / rr; jc1[7] = fy * yp * d4 / rr;
      break;
    }
  }
  du[0] = du_dxp * inv_z;
  du[1] = du_dyp * inv_z;
  du[2] = -(du_dxp * c0 + du_dyp * c1) * inv_z * inv_z;
  dv[0] = dv_dxp * inv_z;
  dv[1] = dv_dyp * inv_z;
  dv[2] = -(dv_dxp * c0 + dv_dyp * c1) * inv_z * inv_z;
}

// pose_point_jacobians: jx (2x3) = (R^T du, R^T dv)^T, jp (2x6) =
// [ (pc - t) x du | du ; (pc - t) x dv | dv ].
template <typename T>
__device__ inline void pose_point_jacobians(const T R[9], T q0, T q1, T q2, const T du[3],
                                            const T dv[3], T jp[12], T jx[6]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    jx[c] = R[c] * du[0] + R[3 + c] * du[1] + R[6 + c] * du[2];
    jx[3 + c] = R[c] * dv[0] + R[3 + c] * dv[1] + R[6 + c] * dv[2];
  }
  jp[0] = -q2 * du[1] + q1 * du[2];
  jp[1] = q2 * du[0] - q0 * du[2];
  jp[2] = -q1 * du[0] + q0 * du[1];
  jp[6] = -q2 * dv[1] + q1 * dv[2];
  jp[7] = q2 * dv[0] - q0 * dv[2];
  jp[8] = -q1 * dv[0] + q0 * dv[1];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    jp[3 + c] = du[c];
    jp[9 + c] = dv[c];
  }
}

// robust_weight (lib.rs:677-694)
__device__ inline double robust_weight(double n, int loss, double scale) {
  const double s = fmax(scale, 1e-9);
  if (loss == SFMBA_LOSS_HUBER) return n <= s ? 1.0 : s / n;
  if (loss == SFMBA_LOSS_CAUCHY) {
    const double q = n / s;
    return 1.0 / (1.0 + q * q);
  }
  return 1.0;
}

}  // namespace sfmba
