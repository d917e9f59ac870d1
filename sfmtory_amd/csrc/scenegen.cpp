// Seeded synthetic bundle-adjustment scenes (host code, FP64).
//
// Cameras sit on a circle of radius R in the world XZ plane, look at the
// origin with world +Y as image-down (OpenCV camera axes, as in
// crates/sfm-core/src/pose.rs:4-9), at angles 2*pi*i/n.  Points are uniform
// in a cube.  Each point draws a track length, then that many distinct
// cameras uniformly among those that see it (z > 0.1 and inside the image),
// by rejection sampling.  Observations get N(0, sigma) pixel noise and an
// optional fraction of uniform outliers.  The starting estimate perturbs
// every pose but pose 0 (the gauge), every point and every focal length.
//
// PRNG: xoshiro256** seeded through splitmix64; normals by Box-Muller.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/sfmba.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n) % n; }
  double normal() {
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  // failures before the first success, success probability p
  int64_t geometric(double p) {
    double u = uniform();
    while (u <= 0.0) u = uniform();
    return (int64_t)std::floor(std::log(u) / std::log(1.0 - p));
  }
};

struct V3 {
  double x, y, z;
};

void quat_from_matrix(const double R[9], double q[4]) {
  const double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    const double s = std::sqrt(tr + 1.0) * 2;
    q[0] = 0.25 * s;
    q[1] = (R[7] - R[5]) / s;
    q[2] = (R[2] - R[6]) / s;
    q[3] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    const double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    q[0] = (R[7] - R[5]) / s;
    q[1] = 0.25 * s;
    q[2] = (R[1] + R[3]) / s;
    q[3] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    const double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    q[0] = (R[2] - R[6]) / s;
    q[1] = (R[1] + R[3]) / s;
    q[2] = 0.25 * s;
    q[3] = (R[5] + R[7]) / s;
  } else {
    const double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    q[0] = (R[3] - R[1]) / s;
    q[1] = (R[2] + R[6]) / s;
    q[2] = (R[5] + R[7]) / s;
    q[3] = 0.25 * s;
  }
}

void rot(const sfmba_pose& p, const double v[3], double o[3]) {
  const double w = p.qw, i = p.qx, j = p.qy, k = p.qz;
  const double R[9] = {w * w + i * i - j * j - k * k, 2 * (i * j - w * k), 2 * (w * j + i * k),
                       2 * (w * k + i * j), w * w - i * i + j * j - k * k, 2 * (j * k - w * i),
                       2 * (i * k - w * j), 2 * (w * i + j * k), w * w - i * i - j * j + k * k};
  for (int r = 0; r < 3; ++r) o[r] = R[r * 3] * v[0] + R[r * 3 + 1] * v[1] + R[r * 3 + 2] * v[2];
}

bool project(const sfmba_camera& c, const double pc[3], double* u, double* v) {
  if (pc[2] <= 1e-9) return false;
  const double x = pc[0] / pc[2], y = pc[1] / pc[2];
  const double* P = c.params;
  switch (c.model_id) {
    case SFMBA_SIMPLE_PINHOLE: *u = P[0] * x + P[1]; *v = P[0] * y + P[2]; return true;
    case SFMBA_PINHOLE: *u = P[0] * x + P[2]; *v = P[1] * y + P[3]; return true;
    case SFMBA_SIMPLE_RADIAL: {
      const double d = 1 + P[3] * (x * x + y * y);
      *u = P[0] * x * d + P[1]; *v = P[0] * y * d + P[2]; return true;
    }
    case SFMBA_RADIAL: {
      const double r2 = x * x + y * y, d = 1 + P[3] * r2 + P[4] * r2 * r2;
      *u = P[0] * x * d + P[1]; *v = P[0] * y * d + P[2]; return true;
    }
    case SFMBA_OPENCV: {
      const double r2 = x * x + y * y, rad = 1 + P[4] * r2 + P[5] * r2 * r2;
      const double xd = x * rad + 2 * P[6] * x * y + P[7] * (r2 + 2 * x * x);
      const double yd = y * rad + P[6] * (r2 + 2 * y * y) + 2 * P[7] * x * y;
      *u = P[0] * xd + P[2]; *v = P[1] * yd + P[3]; return true;
    }
    case SFMBA_OPENCV_FISHEYE: {
      const double r = std::fmax(std::sqrt(x * x + y * y), 1e-12), th = std::atan(r), t2 = th * th;
      const double td = th * (1 + P[4] * t2 + P[5] * t2 * t2 + P[6] * t2 * t2 * t2 + P[7] * t2 * t2 * t2 * t2);
      *u = P[0] * x * td / r + P[2]; *v = P[1] * y * td / r + P[3]; return true;
    }
  }
  return false;
}

sfmba_pose perturb_rotation(const sfmba_pose& p, const double w[3]) {
  const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double dw = 1, dx = 0, dy = 0, dz = 0;
  if (th > 0) {
    const double s = std::sin(th / 2) / th;
    dw = std::cos(th / 2);
    dx = w[0] * s;
    dy = w[1] * s;
    dz = w[2] * s;
  }
  sfmba_pose o = p;
  o.qw = dw * p.qw - dx * p.qx - dy * p.qy - dz * p.qz;
  o.qx = dw * p.qx + dx * p.qw + dy * p.qz - dz * p.qy;
  o.qy = dw * p.qy - dx * p.qz + dy * p.qw + dz * p.qx;
  o.qz = dw * p.qz + dx * p.qy - dy * p.qx + dz * p.qw;
  return o;
}

}  // namespace

extern "C" {

// Scene description.  track_kind: 0 uniform in [track_min, track_max],
// 1 track_min + Geometric(track_p), 2 discrete Pareto floor(track_min *
// U^(-1/track_alpha)) capped at track_max.
typedef struct scene_spec {
  uint64_t seed;
  int32_t model_id;
  int32_t num_images;
  int64_t num_points;
  double width, height;
  double focal;          // nominal focal length (px)
  double focal_jitter;   // true focal = focal * (1 + U(-j, j))
  double k_fixed;        // SIMPLE_RADIAL / RADIAL first coefficient when dist_sigma == 0
  double dist_sigma;     // distortion coefficients ~ N(0, dist_sigma) when > 0
  double radius;         // camera circle radius (m)
  double cube_half;      // points uniform in [-h, h]^3
  int32_t track_kind;
  int32_t track_min;
  int32_t track_max;
  int32_t shared_camera; // 1: one camera for all images
  double track_p;
  double track_alpha;
  double pixel_sigma;
  double outlier_frac;   // fraction of observations shifted by U(-outlier_px, outlier_px)
  double outlier_px;
  double sigma_rot;      // rad
  double sigma_t;        // relative to |t|
  double sigma_x;        // relative to the cube edge
  double sigma_f;        // relative
} scene_spec;

typedef struct scene_data {
  uint64_t num_images, num_cameras, num_points, num_observations;
  uint64_t* camera_of_image;
  sfmba_camera* cameras_true;
  sfmba_camera* cameras_init;
  sfmba_pose* poses_true;
  sfmba_pose* poses_init;
  double* points_true;
  double* points_init;
  sfmba_observation* observations;
} scene_data;

void scene_free(scene_data* s) {
  if (!s) return;
  std::free(s->camera_of_image);
  std::free(s->cameras_true);
  std::free(s->cameras_init);
  std::free(s->poses_true);
  std::free(s->poses_init);
  std::free(s->points_true);
  std::free(s->points_init);
  std::free(s->observations);
  std::memset(s, 0, sizeof(*s));
}

int scene_generate(const scene_spec* sp, scene_data* out) {
  std::memset(out, 0, sizeof(*out));
  if (sp->num_images < 2 || sp->num_points < 1) return -1;
  Rng rng(sp->seed);
  const int n = sp->num_images;
  const int ncam = sp->shared_camera ? 1 : n;
  std::vector<sfmba_camera> cams(ncam);
  for (int c = 0; c < ncam; ++c) {
    sfmba_camera& k = cams[c];
    std::memset(&k, 0, sizeof(k));
    k.model_id = sp->model_id;
    const double f = sp->focal * (1.0 + rng.uniform(-sp->focal_jitter, sp->focal_jitter));
    const double cx = sp->width / 2, cy = sp->height / 2;
    auto dist = [&]() { return sp->dist_sigma > 0 ? sp->dist_sigma * rng.normal() : 0.0; };
    switch (sp->model_id) {
      case SFMBA_SIMPLE_PINHOLE: k.params[0] = f; k.params[1] = cx; k.params[2] = cy; break;
      case SFMBA_PINHOLE: k.params[0] = f; k.params[1] = f; k.params[2] = cx; k.params[3] = cy; break;
      case SFMBA_SIMPLE_RADIAL:
        k.params[0] = f; k.params[1] = cx; k.params[2] = cy;
        k.params[3] = sp->dist_sigma > 0 ? dist() : sp->k_fixed;
        break;
      case SFMBA_RADIAL:
        k.params[0] = f; k.params[1] = cx; k.params[2] = cy;
        k.params[3] = sp->dist_sigma > 0 ? dist() : sp->k_fixed;
        k.params[4] = dist();
        break;
      case SFMBA_OPENCV:
      case SFMBA_OPENCV_FISHEYE:
        k.params[0] = f; k.params[1] = f; k.params[2] = cx; k.params[3] = cy;
        for (int j = 4; j < 8; ++j) k.params[j] = dist();
        break;
      default: return -1;
    }
  }
  std::vector<sfmba_pose> poses(n);
  std::vector<uint64_t> coi(n);
  for (int i = 0; i < n; ++i) {
    coi[i] = sp->shared_camera ? 0 : (uint64_t)i;
    const double th = 2.0 * M_PI * i / n;
    const double C[3] = {sp->radius * std::cos(th), 0.0, sp->radius * std::sin(th)};
    const double zn = std::sqrt(C[0] * C[0] + C[2] * C[2]);
    const double z[3] = {-C[0] / zn, 0.0, -C[2] / zn};
    const double y[3] = {0.0, 1.0, 0.0};
    const double x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
    const double R[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
    double q[4];
    quat_from_matrix(R, q);
    sfmba_pose& p = poses[i];
    p.qw = q[0]; p.qx = q[1]; p.qy = q[2]; p.qz = q[3];
    p.tx = -(R[0] * C[0] + R[1] * C[1] + R[2] * C[2]);
    p.ty = -(R[3] * C[0] + R[4] * C[1] + R[5] * C[2]);
    p.tz = -(R[6] * C[0] + R[7] * C[1] + R[8] * C[2]);
  }
  const int64_t np = sp->num_points;
  std::vector<double> pts(3 * np);
  std::vector<sfmba_observation> obs;
  obs.reserve((size_t)np * 6);
  std::vector<int> chosen;
  std::vector<uint8_t> used(n, 0);
  for (int64_t p = 0; p < np; ++p) {
    double X[3];
    int t = 0;
    for (int attempt = 0;; ++attempt) {
      for (int a = 0; a < 3; ++a) X[a] = rng.uniform(-sp->cube_half, sp->cube_half);
      if (sp->track_kind == 0)
        t = sp->track_min + (int)rng.below((uint64_t)(sp->track_max - sp->track_min + 1));
      else if (sp->track_kind == 1)
        t = sp->track_min + (int)rng.geometric(sp->track_p);
      else {
        double u = rng.uniform();
        while (u <= 0.0) u = rng.uniform();
        t = (int)std::floor(sp->track_min * std::pow(u, -1.0 / sp->track_alpha));
      }
      if (t > sp->track_max) t = sp->track_max;
      if (t > n) t = n;
      if (t < 2) t = 2;
      // Rejection-sample distinct visible cameras.
      chosen.clear();
      int tries = 0;
      while ((int)chosen.size() < t && tries < 64 * t + 256) {
        ++tries;
        const int i = (int)rng.below((uint64_t)n);
        if (used[i]) continue;
        double pc[3], u, v;
        rot(poses[i], X, pc);
        pc[0] += poses[i].tx; pc[1] += poses[i].ty; pc[2] += poses[i].tz;
        if (pc[2] <= 0.1) continue;
        if (!project(cams[coi[i]], pc, &u, &v)) continue;
        if (u < 0 || v < 0 || u >= sp->width || v >= sp->height) continue;
        used[i] = 1;
        chosen.push_back(i);
      }
      for (int i : chosen) used[i] = 0;
      if ((int)chosen.size() >= 2 || attempt > 1000) break;
    }
    for (int a = 0; a < 3; ++a) pts[3 * p + a] = X[a];
    for (int i : chosen) {
      double pc[3], u, v;
      rot(poses[i], X, pc);
      pc[0] += poses[i].tx; pc[1] += poses[i].ty; pc[2] += poses[i].tz;
      project(cams[coi[i]], pc, &u, &v);
      u += sp->pixel_sigma * rng.normal();
      v += sp->pixel_sigma * rng.normal();
      if (sp->outlier_frac > 0 && rng.uniform() < sp->outlier_frac) {
        u += rng.uniform(-sp->outlier_px, sp->outlier_px);
        v += rng.uniform(-sp->outlier_px, sp->outlier_px);
      }
      obs.push_back(sfmba_observation{(uint64_t)i, (uint64_t)p, u, v});
    }
  }
  // Starting estimate.
  std::vector<sfmba_pose> pinit(poses);
  for (int i = 1; i < n; ++i) {
    double ax[3] = {rng.normal(), rng.normal(), rng.normal()};
    const double an = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    const double ang = sp->sigma_rot * rng.normal();
    double w[3] = {ax[0] / an * ang, ax[1] / an * ang, ax[2] / an * ang};
    sfmba_pose q = perturb_rotation(poses[i], w);
    const double tn = std::sqrt(poses[i].tx * poses[i].tx + poses[i].ty * poses[i].ty + poses[i].tz * poses[i].tz);
    q.tx += sp->sigma_t * tn * rng.normal();
    q.ty += sp->sigma_t * tn * rng.normal();
    q.tz += sp->sigma_t * tn * rng.normal();
    pinit[i] = q;
  }
  std::vector<double> xinit(pts);
  const double edge = 2.0 * sp->cube_half;
  for (int64_t e = 0; e < 3 * np; ++e) xinit[e] += sp->sigma_x * edge * rng.normal();
  std::vector<sfmba_camera> cinit(cams);
  for (int c = 0; c < ncam; ++c) {
    const double s = 1.0 + sp->sigma_f * rng.normal();
    cinit[c].params[0] *= s;
    if (sp->model_id == SFMBA_PINHOLE || sp->model_id == SFMBA_OPENCV || sp->model_id == SFMBA_OPENCV_FISHEYE)
      cinit[c].params[1] *= s;
  }

  out->num_images = n;
  out->num_cameras = ncam;
  out->num_points = np;
  out->num_observations = obs.size();
  auto dup = [](const void* src, size_t bytes) {
    void* d = std::malloc(bytes ? bytes : 1);
    if (bytes) std::memcpy(d, src, bytes);
    return d;
  };
  out->camera_of_image = (uint64_t*)dup(coi.data(), n * sizeof(uint64_t));
  out->cameras_true = (sfmba_camera*)dup(cams.data(), ncam * sizeof(sfmba_camera));
  out->cameras_init = (sfmba_camera*)dup(cinit.data(), ncam * sizeof(sfmba_camera));
  out->poses_true = (sfmba_pose*)dup(poses.data(), n * sizeof(sfmba_pose));
  out->poses_init = (sfmba_pose*)dup(pinit.data(), n * sizeof(sfmba_pose));
  out->points_true = (double*)dup(pts.data(), 3 * np * sizeof(double));
  out->points_init = (double*)dup(xinit.data(), 3 * np * sizeof(double));
  out->observations = (sfmba_observation*)dup(obs.data(), obs.size() * sizeof(sfmba_observation));
  return 0;
}

}  // extern "C"
