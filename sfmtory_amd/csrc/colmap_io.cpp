// COLMAP plain-text sparse model (cameras.txt / images.txt / points3D.txt)
// and the standalone refine pass over it.
//
// Reader and writer follow crates/sfm-io/src/colmap.rs:25-300: the same
// fields and formatting (10 decimals for poses, parameters and points, 3 for
// keypoints, which are parsed as single precision, 6 for point errors),
// blank and comment lines skipped before header and keypoint lines are
// paired, ids kept in ascending order.  The refine pass is the reference's
// cmd_refine (crates/sfm-cli/src/main.rs:946-1130): compact indices, image 0
// (in id order) fixed, intrinsics fixed or free with pinned principal
// points, 200 LM iterations, per-point errors recomputed against the refined
// model, the model written back.  The solve and the residuals run on the GPU
// engine through its C ABI; this file is host bookkeeping only.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdarg>
#include <limits>
#include <type_traits>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sfmba.h"
#include "engine.hpp"

namespace sfmba {
namespace {

struct Camera {
  uint32_t id = 0;
  int32_t model = 0;
  uint32_t width = 0, height = 0;
  std::vector<double> params;
};
struct Image {
  uint32_t id = 0, camera_id = 0;
  double q[4] = {1, 0, 0, 0}, t[3] = {0, 0, 0};
  std::string name;
  std::vector<float> kx, ky;
  std::vector<int64_t> pid;
};
struct Point {
  uint64_t id = 0;
  double xyz[3] = {0, 0, 0};
  unsigned rgb[3] = {0, 0, 0};
  double error = 0;
  std::vector<std::pair<uint32_t, uint32_t>> track;
};
struct Model {
  std::map<uint32_t, Camera> cameras;
  std::map<uint32_t, Image> images;
  std::map<uint64_t, Point> points;
};

const char* const MODEL_NAMES[6] = {"SIMPLE_PINHOLE", "PINHOLE", "SIMPLE_RADIAL",
                                    "RADIAL", "OPENCV", "OPENCV_FISHEYE"};

[[noreturn]] void parse_fail(const std::string& file, const std::string& msg) {
  throw Error(SFMBA_ERR_INVALID_ARGUMENT, file + ": " + msg);
}

std::vector<std::string> tokens(const std::string& line) {
  std::vector<std::string> t;
  std::istringstream is(line);
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

bool data_line(std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n\v\f");
  if (a == std::string::npos) return false;
  size_t b = s.find_last_not_of(" \t\r\n\v\f");
  s = s.substr(a, b - a + 1);
  return s[0] != '#';
}

template <typename T>
T parse_num(const std::string& tok, const std::string& file, const char* what) {
  const char* s = tok.c_str();
  char* end = nullptr;
  errno = 0;
  T v{};
  if constexpr (std::is_same<T, double>::value) {
    v = std::strtod(s, &end);
  } else if constexpr (std::is_same<T, float>::value) {
    v = std::strtof(s, &end);
  } else if constexpr (std::is_signed<T>::value) {
    const long long x = std::strtoll(s, &end, 10);
    v = (T)x;
    if (x < (long long)std::numeric_limits<T>::min() || x > (long long)std::numeric_limits<T>::max()) errno = ERANGE;
  } else {
    if (tok[0] == '-') parse_fail(file, std::string("bad ") + what);
    const unsigned long long x = std::strtoull(s, &end, 10);
    v = (T)x;
    if (x > (unsigned long long)std::numeric_limits<T>::max()) errno = ERANGE;
  }
  // Floats: Rust's grammar (no hex, no nan(...)); overflow reads as inf and
  // underflow as the nearest subnormal or zero, as Rust's parser rounds.
  const bool is_float = std::is_floating_point<T>::value;
  if (is_float && tok.find_first_of("xX(") != std::string::npos) parse_fail(file, std::string("bad ") + what);
  if (end == s || *end != '\0' || (!is_float && errno == ERANGE)) parse_fail(file, std::string("bad ") + what);
  return v;
}

std::vector<std::string> read_lines(const std::filesystem::path& p) {
  std::ifstream f(p);
  if (!f) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "cannot read " + p.string());
  std::vector<std::string> out;
  std::string line;
  while (std::getline(f, line))
    if (data_line(line)) out.push_back(line);
  return out;
}

Model read_model(const std::filesystem::path& dir) {
  Model m;
  for (const std::string& line : read_lines(dir / "cameras.txt")) {
    const auto t = tokens(line);
    if (t.size() < 4) parse_fail("cameras.txt", "malformed line: " + line);
    Camera c;
    c.id = parse_num<uint32_t>(t[0], "cameras.txt", "CAMERA_ID");
    c.model = -1;
    for (int k = 0; k < 6; ++k)
      if (t[1] == MODEL_NAMES[k]) c.model = k;
    c.width = parse_num<uint32_t>(t[2], "cameras.txt", "WIDTH");
    c.height = parse_num<uint32_t>(t[3], "cameras.txt", "HEIGHT");
    for (size_t k = 4; k < t.size(); ++k) c.params.push_back(parse_num<double>(t[k], "cameras.txt", "PARAMS"));
    if (c.model < 0) parse_fail("cameras.txt", "unknown camera model " + t[1]);
    if ((int)c.params.size() < model_param_count_host(c.model)) parse_fail("cameras.txt", "too few PARAMS");
    c.params.resize(model_param_count_host(c.model));
    m.cameras[c.id] = c;
  }
  const auto il = read_lines(dir / "images.txt");
  for (size_t k = 0; k < il.size(); ++k) {
    const auto t = tokens(il[k]);
    if (t.size() < 10) parse_fail("images.txt", "malformed header: " + il[k]);
    Image im;
    im.id = parse_num<uint32_t>(t[0], "images.txt", "IMAGE_ID");
    for (int j = 0; j < 4; ++j) im.q[j] = parse_num<double>(t[1 + j], "images.txt", "quaternion");
    for (int j = 0; j < 3; ++j) im.t[j] = parse_num<double>(t[5 + j], "images.txt", "translation");
    im.camera_id = parse_num<uint32_t>(t[8], "images.txt", "CAMERA_ID");
    im.name = t[9];
    // Pose::from_quaternion_wxyz_translation (crates/sfm-core/src/pose.rs:74-80)
    // keeps q / |q|, |q|^2 summed in storage order x, y, z, w.
    const double qn = std::sqrt(((im.q[1] * im.q[1] + im.q[2] * im.q[2]) + im.q[3] * im.q[3]) + im.q[0] * im.q[0]);
    for (double& c : im.q) c /= qn;
    if (++k >= il.size()) parse_fail("images.txt", "missing POINTS2D line");
    const auto p = tokens(il[k]);
    for (size_t j = 0; j + 2 < p.size(); j += 3) {
      im.kx.push_back(parse_num<float>(p[j], "images.txt", "point x"));
      im.ky.push_back(parse_num<float>(p[j + 1], "images.txt", "point y"));
      im.pid.push_back(parse_num<int64_t>(p[j + 2], "images.txt", "POINT3D_ID"));
    }
    m.images[im.id] = im;
  }
  for (const std::string& line : read_lines(dir / "points3D.txt")) {
    const auto t = tokens(line);
    if (t.size() < 8) parse_fail("points3D.txt", "malformed line: " + line);
    Point p;
    p.id = parse_num<uint64_t>(t[0], "points3D.txt", "POINT3D_ID");
    static const char* const XYZ[3] = {"X", "Y", "Z"}, *const RGB[3] = {"R", "G", "B"};
    for (int j = 0; j < 3; ++j) p.xyz[j] = parse_num<double>(t[1 + j], "points3D.txt", XYZ[j]);
    for (int j = 0; j < 3; ++j) p.rgb[j] = parse_num<uint8_t>(t[4 + j], "points3D.txt", RGB[j]);
    p.error = parse_num<double>(t[7], "points3D.txt", "ERROR");
    for (size_t j = 8; j + 1 < t.size(); j += 2)
      p.track.emplace_back(parse_num<uint32_t>(t[j], "points3D.txt", "TRACK IMAGE_ID"),
                           parse_num<uint32_t>(t[j + 1], "points3D.txt", "TRACK POINT2D_IDX"));
    m.points[p.id] = p;
  }
  return m;
}

void write_file(const std::filesystem::path& p, const std::string& s) {
  std::ofstream f(p, std::ios::binary);
  if (!f || !(f << s)) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "cannot write " + p.string());
}

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...) {
  va_list a, b;
  va_start(a, f);
  va_copy(b, a);
  const int n = std::vsnprintf(nullptr, 0, f, a);
  va_end(a);
  std::string s((size_t)n, '\0');
  std::vsnprintf(&s[0], (size_t)n + 1, f, b);
  va_end(b);
  return s;
}

// A float with `prec` decimals; non-finite values spelled as Rust's Display
// spells them (NaN, inf, -inf).
std::string num(double v, int prec) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  return fmt("%.*f", prec, v);
}

void write_model(const Model& m, const std::filesystem::path& dir) {
  std::filesystem::create_directories(dir);
  std::string c = "# Camera list with one line of data per camera:\n"
                  "#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n";
  c += fmt("# Number of cameras: %zu\n", m.cameras.size());
  for (const auto& [id, cam] : m.cameras) {
    c += fmt("%u %s %u %u ", cam.id, MODEL_NAMES[cam.model], cam.width, cam.height);
    for (size_t k = 0; k < cam.params.size(); ++k) c += (k ? " " : "") + num(cam.params[k], 10);
    c += "\n";
  }
  write_file(dir / "cameras.txt", c);

  size_t n2d = 0;
  for (const auto& [id, im] : m.images) n2d += im.kx.size();
  const double mean_obs = m.images.empty() ? 0.0 : (double)n2d / (double)m.images.size();
  std::string s = "# Image list with two lines of data per image:\n"
                  "#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n"
                  "#   POINTS2D[] as (X, Y, POINT3D_ID)\n";
  s += fmt("# Number of images: %zu, mean observations per image: %.1f\n", m.images.size(), mean_obs);
  for (const auto& [id, im] : m.images) {
    s += fmt("%u", im.id);
    for (double v : im.q) s += " " + num(v, 10);
    for (double v : im.t) s += " " + num(v, 10);
    s += fmt(" %u %s\n", im.camera_id, im.name.c_str());
    for (size_t k = 0; k < im.kx.size(); ++k)
      s += (k ? " " : "") + num(im.kx[k], 3) + " " + num(im.ky[k], 3) +
           fmt(" %" PRId64, im.pid[k] < 0 ? (int64_t)-1 : im.pid[k]);
    s += "\n";
  }
  write_file(dir / "images.txt", s);

  std::string p = "# 3D point list with one line of data per point:\n"
                  "#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)\n";
  p += fmt("# Number of points: %zu\n", m.points.size());
  for (const auto& [id, pt] : m.points) {
    p += fmt("%" PRIu64, pt.id);
    for (double v : pt.xyz) p += " " + num(v, 10);
    p += fmt(" %u %u %u ", pt.rgb[0], pt.rgb[1], pt.rgb[2]) + num(pt.error, 6) + " ";
    for (size_t k = 0; k < pt.track.size(); ++k) p += fmt(k ? " %u %u" : "%u %u", pt.track[k].first, pt.track[k].second);
    p += "\n";
  }
  write_file(dir / "points3D.txt", p);
}

// Reconstruction::mean_reprojection_error (crates/sfm-core/src/model.rs:84-97):
// stored point errors weighted by track length.
double model_mean_error(const Model& m) {
  double total = 0.0, count = 0.0;
  for (const auto& [id, p] : m.points) {
    total += p.error * (double)p.track.size();
    count += (double)p.track.size();
  }
  return count > 0.0 ? total / count : 0.0;
}

void check(int rc) {
  if (rc != SFMBA_OK) throw Error(rc, sfmba_last_error());
}

void refine(const char* in_dir, const char* out_dir, const sfmba_colmap_refine_options* ro,
            const sfmba_options* opt, sfmba_colmap_refine_report* rep) {
  Model m = read_model(in_dir);
  if (m.images.size() < 2) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "need at least 2 registered images to refine");
  if (m.points.empty()) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "no 3D points to refine");
  std::vector<uint32_t> image_ids;
  std::map<uint32_t, size_t> image_pos;
  for (const auto& [id, im] : m.images) {
    image_pos[id] = image_ids.size();
    image_ids.push_back(id);
  }
  std::vector<sfmba_camera> cams;
  std::vector<uint32_t> cam_ids;
  std::map<uint32_t, size_t> cam_index;
  std::vector<uint64_t> coi;
  std::vector<sfmba_pose> poses;
  for (uint32_t id : image_ids) {
    const Image& im = m.images[id];
    auto it = cam_index.find(im.camera_id);
    if (it == cam_index.end()) {
      auto c = m.cameras.find(im.camera_id);
      if (c == m.cameras.end()) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "image references an unknown camera");
      sfmba_camera sc{};
      sc.model_id = c->second.model;
      for (size_t k = 0; k < c->second.params.size(); ++k) sc.params[k] = c->second.params[k];
      it = cam_index.emplace(im.camera_id, cams.size()).first;
      cams.push_back(sc);
      cam_ids.push_back(im.camera_id);
    }
    coi.push_back(it->second);
    poses.push_back(sfmba_pose{im.q[0], im.q[1], im.q[2], im.q[3], im.t[0], im.t[1], im.t[2]});
  }
  std::vector<uint64_t> point_ids;
  std::vector<double> pts;
  std::vector<sfmba_observation> obs;
  std::vector<size_t> obs_point;  // which point (compact) each observation belongs to
  for (const auto& [id, p] : m.points) {
    const size_t cp = point_ids.size();
    point_ids.push_back(id);
    pts.insert(pts.end(), p.xyz, p.xyz + 3);
    for (const auto& [img, k] : p.track) {
      auto ip = image_pos.find(img);
      if (ip == image_pos.end()) continue;
      const Image& im = m.images[img];
      if (k >= im.kx.size()) continue;
      obs.push_back(sfmba_observation{ip->second, cp, (double)im.kx[k], (double)im.ky[k]});
      obs_point.push_back(cp);
    }
  }
  if (obs.empty())
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "no observations linking tracked points to registered images");

  std::vector<uint8_t> fixed_poses(image_ids.size(), 0);
  fixed_poses[0] = 1;
  std::vector<uint8_t> fixed_cams(cams.size(), ro->refine_intrinsics ? 0 : 1);
  std::vector<uint8_t> mask_flat;
  std::vector<size_t> mask_len;
  if (ro->refine_intrinsics) {
    std::vector<uint8_t> mask(8 * cams.size());
    mask_len.resize(cams.size());
    check(sfmba_default_fixed_params_mask(cams.data(), cams.size(), mask.data(), mask_len.data()));
    for (size_t c = 0; c < cams.size(); ++c) mask_flat.insert(mask_flat.end(), &mask[8 * c], &mask[8 * c] + mask_len[c]);
  }
  sfmba_input in{};
  in.num_images = image_ids.size();
  in.num_cameras = cams.size();
  in.num_points = point_ids.size();
  in.num_observations = obs.size();
  in.camera_of_image = coi.data();
  in.num_camera_of_image = coi.size();
  in.cameras = cams.data();
  in.poses = poses.data();
  in.points = pts.data();
  in.observations = obs.data();
  in.fixed_poses = fixed_poses.data();
  in.num_fixed_poses = fixed_poses.size();
  in.fixed_cameras = fixed_cams.data();
  in.num_fixed_cameras = fixed_cams.size();
  in.fixed_camera_params = mask_flat.data();
  in.fixed_camera_params_len = mask_len.data();
  in.num_fixed_camera_params = mask_len.size();
  sfmba_params p;
  sfmba_params_default(&p);
  p.robust_loss = ro->robust_loss;
  p.max_iterations = ro->max_iterations;

  const double before = model_mean_error(m);
  std::vector<sfmba_pose> out_poses(poses.size());
  std::vector<double> out_pts(pts.size());
  std::vector<sfmba_camera> out_cams(cams.size());
  sfmba_output out{out_poses.data(), out_pts.data(), out_cams.data(), 0.0, 0.0, 0};
  check(sfmba_bundle_adjust_ex(&in, &p, opt, &out, nullptr));

  for (size_t i = 0; i < image_ids.size(); ++i) {
    Image& im = m.images[image_ids[i]];
    const sfmba_pose& q = out_poses[i];
    const double v[7] = {q.qw, q.qx, q.qy, q.qz, q.tx, q.ty, q.tz};
    std::copy(v, v + 4, im.q);
    std::copy(v + 4, v + 7, im.t);
  }
  for (size_t k = 0; k < point_ids.size(); ++k) std::copy(&out_pts[3 * k], &out_pts[3 * k] + 3, m.points[point_ids[k]].xyz);
  if (ro->refine_intrinsics)
    for (size_t c = 0; c < cam_ids.size(); ++c) {
      Camera& cam = m.cameras[cam_ids[c]];
      for (size_t k = 0; k < cam.params.size(); ++k) cam.params[k] = out_cams[c].params[k];
    }
  // Per-point error: mean pixel distance over the track elements in front of
  // their camera, against the refined model (kept when none is).
  std::vector<double> norm;
  std::vector<uint8_t> front;
  residual_norms(obs, coi, ro->refine_intrinsics ? out_cams : cams, out_poses, out_pts, opt->device, norm, front);
  std::vector<double> tot(point_ids.size(), 0.0);
  std::vector<size_t> cnt(point_ids.size(), 0);
  for (size_t o = 0; o < obs.size(); ++o)
    if (front[o]) {
      tot[obs_point[o]] += norm[o];
      ++cnt[obs_point[o]];
    }
  for (size_t k = 0; k < point_ids.size(); ++k)
    if (cnt[k]) m.points[point_ids[k]].error = tot[k] / (double)cnt[k];
  const double after = model_mean_error(m);
  write_model(m, out_dir);
  if (rep) {
    rep->num_images = image_ids.size();
    rep->num_points3d = point_ids.size();
    rep->num_observations = obs.size();
    rep->initial_cost = out.initial_cost;
    rep->final_cost = out.final_cost;
    rep->iterations_run = out.iterations_run;
    rep->mean_reprojection_error_px_before = before;
    rep->mean_reprojection_error_px_after = after;
  }
}

}  // namespace
}  // namespace sfmba

using sfmba::Error;

template <typename F>
static int colmap_guarded(F&& f) {
  sfmba::set_error("");
  try {
    f();
    return SFMBA_OK;
  } catch (const Error& e) {
    sfmba::set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    sfmba::set_error(e.what());
    return SFMBA_ERR_INVALID_ARGUMENT;
  }
}

extern "C" {

void sfmba_colmap_refine_options_default(sfmba_colmap_refine_options* o) {
  if (!o) return;
  *o = sfmba_colmap_refine_options{};
  o->refine_intrinsics = 0;
  o->robust_loss = SFMBA_LOSS_HUBER;
  o->max_iterations = 200;
}

int sfmba_colmap_refine(const char* model_dir, const char* output_dir, const sfmba_colmap_refine_options* options,
                        const sfmba_options* engine_options, sfmba_colmap_refine_report* report) {
  return colmap_guarded([&] {
    if (!model_dir || !output_dir) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL directory");
    sfmba_colmap_refine_options ro;
    sfmba_colmap_refine_options_default(&ro);
    if (options) ro = *options;
    sfmba_options o;
    sfmba_options_default(&o);
    if (engine_options) o = *engine_options;
    sfmba::refine(model_dir, output_dir, &ro, &o, report);
  });
}

int sfmba_colmap_rewrite(const char* model_dir, const char* output_dir, uint64_t* counts) {
  return colmap_guarded([&] {
    if (!model_dir || !output_dir) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL directory");
    const sfmba::Model m = sfmba::read_model(model_dir);
    sfmba::write_model(m, output_dir);
    if (counts) {
      counts[0] = m.cameras.size();
      counts[1] = m.images.size();
      counts[2] = m.points.size();
    }
  });
}

}  // extern "C"
