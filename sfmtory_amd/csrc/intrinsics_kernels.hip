// Intrinsics-only refinement (crates/sfm-ba/src/intrinsics.rs:116-238).
//
// One workgroup per camera with at least 8 observations runs that camera's
// whole dense Levenberg-Marquardt loop on the device: central-difference
// Jacobians of the residual w.r.t. the camera's own parameters (h = 1e-6),
// J^T W J and J^T W r by block reduction, damping lambda * max(diag, 1e-12),
// Cholesky with LU fallback on thread 0, trial cost by block reduction, and
// the fixed x3 / /3 lambda schedule with at most 8 tries per iteration.
#include <hip/hip_runtime.h>

#include "camera_models.hpp"
#include "engine.hpp"

namespace sfmba {

constexpr int RI_BLOCK = 256;

__device__ __attribute__((always_inline)) inline double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int h = RI_BLOCK / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline bool ri_residual(const double* pose, const double* X, int model, const double* P,
                                   double2 xy, double& r0, double& r1) {
  double c0, c1, c2;
  transform_point(pose, X[0], X[1], X[2], c0, c1, c2);
  if (c2 <= 1e-9) return false;
  project_residual(model, P, c0, c1, c2, xy.x, xy.y, r0, r1);
  return true;
}

// Column-major n x n solves on one thread (n <= 8).
__device__ inline bool small_cholesky_solve(double* A, int n, double* b) {
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < j; ++k) {
      const double f = -A[k * n + j];
      for (int i = j; i < n; ++i) A[j * n + i] += f * A[k * n + i];
    }
    const double d = A[j * n + j];
    if (!(d > 0.0)) return false;
    const double den = sqrt(d);
    A[j * n + j] = den;
    for (int i = j + 1; i < n; ++i) A[j * n + i] /= den;
  }
  for (int i = 0; i < n; ++i) {
    const double c = b[i] / A[i * n + i];
    b[i] = c;
    for (int r = i + 1; r < n; ++r) b[r] += -c * A[i * n + r];
  }
  for (int i = n - 1; i >= 0; --i) {
    double dot = 0.0;
    for (int r = i + 1; r < n; ++r) dot += A[i * n + r] * b[r];
    b[i] = (b[i] - dot) / A[i * n + i];
  }
  return true;
}

__device__ inline bool small_lu_solve(double* A, int n, double* b) {
  int pa[8], pb[8], np = 0;
  for (int i = 0; i < n; ++i) {
    int piv = i;
    double best = fabs(A[i * n + i]);
    for (int r = i + 1; r < n; ++r)
      if (fabs(A[i * n + r]) > best) {
        best = fabs(A[i * n + r]);
        piv = r;
      }
    const double diag = A[i * n + piv];
    if (diag == 0.0) continue;
    if (piv != i) {
      pa[np] = i;
      pb[np++] = piv;
      for (int c = 0; c < n; ++c) {
        const double t = A[c * n + i];
        A[c * n + i] = A[c * n + piv];
        A[c * n + piv] = t;
      }
    }
    const double inv = 1.0 / diag;
    for (int r = i + 1; r < n; ++r) A[i * n + r] *= inv;
    for (int c = i + 1; c < n; ++c) {
      const double f = A[c * n + i];
      for (int r = i + 1; r < n; ++r) A[c * n + r] += -f * A[i * n + r];
    }
  }
  for (int t = 0; t < np; ++t) {
    const double v = b[pa[t]];
    b[pa[t]] = b[pb[t]];
    b[pb[t]] = v;
  }
  for (int i = 0; i < n; ++i)
    for (int r = i + 1; r < n; ++r) b[r] += -b[i] * A[i * n + r];
  for (int i = n - 1; i >= 0; --i) {
    const double d = A[i * n + i];
    if (d == 0.0) return false;
    const double c = b[i] / d;
    b[i] = c;
    for (int r = 0; r < i; ++r) b[r] += -c * A[i * n + r];
  }
  return true;
}

__global__ __launch_bounds__(RI_BLOCK) void k_refine_intrinsics(
    const uint32_t* __restrict__ cam_list, const uint32_t* __restrict__ cam_obs_ptr,
    const uint32_t* __restrict__ obs_img, const uint32_t* __restrict__ obs_pt,
    const double2* __restrict__ obs_xy, const double* __restrict__ poses,
    const double* __restrict__ pts, const int32_t* __restrict__ cam_model,
    double* __restrict__ cams, uint64_t max_iter, double lambda0, int loss, double scale) {
  __shared__ double red[RI_BLOCK];
  __shared__ double cur[8], trial[8], jtj[64], jtr[8];
  __shared__ double sh_lambda;
  __shared__ int sh_state;  // 0 evaluate trial, 1 solve failed
  const int t = threadIdx.x;
  const uint32_t cam = cam_list[blockIdx.x];
  const int model = cam_model[cam];
  const int n = model_param_count(model);
  const uint32_t o0 = cam_obs_ptr[blockIdx.x], o1 = cam_obs_ptr[blockIdx.x + 1];
  if (t < 8) cur[t] = cams[8 * (size_t)cam + t];
  if (t == 0) sh_lambda = lambda0;
  __syncthreads();
  const double EPS = 1e-6;

  auto cost_of = [&](const double* P) {
    double s = 0.0;
    for (uint32_t o = o0 + t; o < o1; o += RI_BLOCK) {
      double r0, r1;
      if (ri_residual(poses + 7 * (size_t)obs_img[o], pts + 3 * (size_t)obs_pt[o], model, P, obs_xy[o], r0, r1)) {
        const double nr = sqrt(r0 * r0 + r1 * r1);
        const double w = robust_weight(nr, loss, scale);
        s += w * nr * nr;
      }
    }
    return block_sum(s, red);
  };

  double cost = cost_of(cur);
  for (uint64_t it = 0; it < max_iter; ++it) {
    double ltt[64], ltr[8];
    for (int k = 0; k < 64; ++k) ltt[k] = 0.0;
    for (int k = 0; k < 8; ++k) ltr[k] = 0.0;
    for (uint32_t o = o0 + t; o < o1; o += RI_BLOCK) {
      const double* pose = poses + 7 * (size_t)obs_img[o];
      const double* X = pts + 3 * (size_t)obs_pt[o];
      const double2 xy = obs_xy[o];
      double r0a, r0b;
      if (!ri_residual(pose, X, model, cur, xy, r0a, r0b)) continue;
      const double w = robust_weight(sqrt(r0a * r0a + r0b * r0b), loss, scale);
      double ja[8], jb[8];
      for (int k = 0; k < 8; ++k) {
        ja[k] = 0.0;
        jb[k] = 0.0;
        if (k >= n) continue;
        double pp[8], pm[8];
        for (int q = 0; q < 8; ++q) pp[q] = pm[q] = cur[q];
        pp[k] += EPS;
        pm[k] += -EPS;
        double ra, rb, sa, sb;
        if (!ri_residual(pose, X, model, pp, xy, ra, rb)) { ra = r0a; rb = r0b; }
        if (!ri_residual(pose, X, model, pm, xy, sa, sb)) { sa = r0a; sb = r0b; }
        ja[k] = (ra - sa) / (2.0 * EPS);
        jb[k] = (rb - sb) / (2.0 * EPS);
      }
      for (int a = 0; a < n; ++a) {
        ltr[a] += w * (ja[a] * r0a + jb[a] * r0b);
        for (int b = 0; b < n; ++b) ltt[a * 8 + b] += w * (ja[a] * ja[b] + jb[a] * jb[b]);
      }
    }
    for (int a = 0; a < n; ++a) {
      const double v = block_sum(ltr[a], red);
      if (t == 0) jtr[a] = v;
      for (int b = 0; b < n; ++b) {
        const double u = block_sum(ltt[a * 8 + b], red);
        if (t == 0) jtj[a * 8 + b] = u;
      }
    }
    __syncthreads();
    bool accepted = false;
    for (int tr = 0; tr < 8; ++tr) {
      if (t == 0) {
        double A[64], A2[64], d[8];
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) A[b * n + a] = jtj[a * 8 + b];
        for (int a = 0; a < n; ++a) A[a * n + a] += sh_lambda * fmax(jtj[a * 8 + a], 1e-12);
        for (int q = 0; q < n * n; ++q) A2[q] = A[q];
        for (int a = 0; a < n; ++a) d[a] = -jtr[a];
        bool ok = small_cholesky_solve(A, n, d);
        if (!ok) {
          for (int a = 0; a < n; ++a) d[a] = -jtr[a];
          ok = small_lu_solve(A2, n, d);
        }
        if (!ok) {
          sh_lambda *= 3.0;
          sh_state = 1;
        } else {
          for (int q = 0; q < 8; ++q) trial[q] = cur[q];
          for (int a = 0; a < n; ++a) trial[a] += d[a];
          sh_state = 0;
        }
      }
      __syncthreads();
      const int state = sh_state;
      __syncthreads();  // every thread has read sh_state before thread 0 rewrites it
      if (state == 1) continue;
      const double tc = cost_of(trial);
      if (isfinite(tc) && tc < cost) {
        if (t == 0) sh_lambda = fmax(sh_lambda / 3.0, 1e-12);
        if (t < 8) cur[t] = trial[t];
        cost = tc;
        accepted = true;
        __syncthreads();
        break;
      }
      if (t == 0) sh_lambda *= 3.0;
      __syncthreads();
    }
    if (!accepted) break;
  }
  if (t < 8) cams[8 * (size_t)cam + t] = cur[t];
}

void launch_refine_intrinsics(uint32_t n_cams_refined, const uint32_t* cam_list,
                              const uint32_t* cam_obs_ptr, const uint32_t* obs_img,
                              const uint32_t* obs_pt, const double2* obs_xy, const double* poses,
                              const double* pts, const int32_t* cam_model, double* cams,
                              const sfmba_intrinsics_params* p, hipStream_t s) {
  if (!n_cams_refined) return;
  hipLaunchKernelGGL(k_refine_intrinsics, dim3(n_cams_refined), dim3(RI_BLOCK), 0, s, cam_list,
                     cam_obs_ptr, obs_img, obs_pt, obs_xy, poses, pts, cam_model, cams,
                     (uint64_t)p->max_iterations, p->initial_lambda, p->robust_loss,
                     p->loss_scale_px);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sfmba
