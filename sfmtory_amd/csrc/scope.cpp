// Problem assembly for one bundle adjustment of a growing reconstruction:
// global or local scope, the co-visibility window, the held poses, the
// compact cameras / points / observations, and the write-back of a solve.
// Follows run_bundle_adjustment's scope handling
// (crates/sfm-reconstruction/src/lib.rs:1339-1474, :1697-1707).  Host-side
// index bookkeeping only; the solves run on the GPU through the engine's
// C ABI.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/sfmba.h"
#include "engine.hpp"

struct sfmba_window {
  std::vector<uint64_t> image_idx, point_idx;  // original index of each compact one
  std::vector<size_t> camera_entry;            // rec->cameras entry of each compact camera
  std::vector<uint64_t> coi;
  std::vector<sfmba_camera> cams;
  std::vector<sfmba_pose> poses;
  std::vector<double> pts;
  std::vector<sfmba_observation> obs;
  std::vector<uint8_t> fixed_poses, fixed_cams, mask_flat, caller_fixed;
  std::vector<size_t> mask_len;
  uint64_t num_free_images = 0;
  bool empty() const { return image_idx.empty(); }
};

namespace sfmba {
namespace {

void check(int rc) {
  if (rc != SFMBA_OK) throw Error(rc, sfmba_last_error());
}

void validate(const sfmba_reconstruction* r) {
  if (!r) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL reconstruction");
  if (r->num_images && (!r->registered || !r->poses || !r->image_camera_id || !r->keypoint_offsets))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL image array");
  if (r->num_points && (!r->points || !r->track_offsets))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL point array");
  if (r->num_cameras && (!r->camera_ids || !r->cameras))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL camera array");
  if (r->num_caller_fixed_camera_ids && !r->caller_fixed_camera_ids)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL caller_fixed_camera_ids");
  for (size_t i = 0; i < r->num_images; ++i)
    if (r->keypoint_offsets[i + 1] < r->keypoint_offsets[i])
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "keypoint_offsets must be non-decreasing");
  if (r->num_images && r->keypoint_offsets[r->num_images] && !r->keypoints)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL keypoints");
  for (size_t p = 0; p < r->num_points; ++p)
    if (r->track_offsets[p + 1] < r->track_offsets[p])
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "track_offsets must be non-decreasing");
  const uint64_t nt = r->num_points ? r->track_offsets[r->num_points] : 0;
  if (nt && (!r->track_images || !r->track_keypoints))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL track arrays");
  for (uint64_t k = 0; k < nt; ++k)
    if (r->track_images[k] >= r->num_images)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "track image index out of range");
}

sfmba_window* build(const sfmba_reconstruction* r, int32_t scope, uint64_t center) {
  validate(r);
  if (scope != SFMBA_SCOPE_GLOBAL && scope != SFMBA_SCOPE_LOCAL)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown scope");
  auto w = std::make_unique<sfmba_window>();
  const size_t ni = r->num_images;
  auto track = [&](size_t p) { return std::make_pair(r->track_offsets[p], r->track_offsets[p + 1]); };

  // Points (lib.rs:1339-1356) and, for a local window, the free images
  // (lib.rs:1358-1378).
  std::vector<uint8_t> free_img;  // empty: every image free but the seed
  for (size_t p = 0; p < r->num_points; ++p) {
    const auto [b, e] = track(p);
    if (e - b < 2) continue;
    if (scope == SFMBA_SCOPE_LOCAL) {
      bool sees = false;
      for (uint64_t k = b; k < e && !sees; ++k) sees = r->track_images[k] == center;
      if (!sees) continue;
    }
    w->point_idx.push_back(p);
  }
  if (scope == SFMBA_SCOPE_LOCAL) {
    std::vector<uint64_t> covis(ni, 0);
    for (uint64_t p : w->point_idx) {
      const auto [b, e] = track(p);
      for (uint64_t k = b; k < e; ++k) {
        const uint32_t img = r->track_images[k];
        if (img != center && r->registered[img]) ++covis[img];
      }
    }
    std::vector<uint32_t> ranked;
    for (size_t i = 0; i < ni; ++i)
      if (covis[i]) ranked.push_back((uint32_t)i);
    std::stable_sort(ranked.begin(), ranked.end(), [&](uint32_t a, uint32_t b) { return covis[a] > covis[b]; });
    free_img.assign(ni, 0);
    for (size_t k = 0; k < ranked.size() && k < (size_t)SFMBA_LOCAL_BA_NUM_IMAGES; ++k) free_img[ranked[k]] = 1;
    if (center < ni) free_img[center] = 1;
  }

  // Images (lib.rs:1386-1407): every registered one, or for a local window
  // the registered images observing one of its points.
  std::vector<int64_t> image_pos(ni, -1);
  std::vector<uint8_t> seen(ni, scope == SFMBA_SCOPE_GLOBAL ? 1 : 0);
  if (scope == SFMBA_SCOPE_LOCAL)
    for (uint64_t p : w->point_idx) {
      const auto [b, e] = track(p);
      for (uint64_t k = b; k < e; ++k) seen[r->track_images[k]] = 1;
    }
  for (size_t i = 0; i < ni; ++i)
    if (r->registered[i] && seen[i]) {
      image_pos[i] = (int64_t)w->image_idx.size();
      w->image_idx.push_back(i);
    }
  if (w->image_idx.size() < 2) {
    w->image_idx.clear();
    w->point_idx.clear();
    return w.release();
  }

  // Cameras in order of first use (lib.rs:1409-1421).
  std::unordered_map<uint32_t, size_t> entry_of_id, compact_of_id;
  for (size_t c = 0; c < r->num_cameras; ++c) entry_of_id.emplace(r->camera_ids[c], c);
  std::unordered_set<uint32_t> caller_fixed(r->caller_fixed_camera_ids,
                                            r->caller_fixed_camera_ids + r->num_caller_fixed_camera_ids);
  for (uint64_t orig : w->image_idx) {
    const uint32_t id = r->image_camera_id[orig];
    auto it = compact_of_id.find(id);
    if (it == compact_of_id.end()) {
      auto e = entry_of_id.find(id);
      if (e == entry_of_id.end()) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "image references an unknown camera id");
      it = compact_of_id.emplace(id, w->cams.size()).first;
      w->camera_entry.push_back(e->second);
      w->cams.push_back(r->cameras[e->second]);
      w->caller_fixed.push_back(caller_fixed.count(id) ? 1 : 0);
    }
    w->coi.push_back(it->second);
    w->poses.push_back(r->poses[orig]);
  }
  for (uint64_t p : w->point_idx) w->pts.insert(w->pts.end(), r->points + 3 * p, r->points + 3 * p + 3);

  // Held poses (lib.rs:1433-1455).
  for (uint64_t orig : w->image_idx) {
    bool held = orig == r->seed_image || (!free_img.empty() && !free_img[orig]);
    if (r->pose_pinned && r->pose_pinned[orig]) held = true;
    w->fixed_poses.push_back(held ? 1 : 0);
  }
  if (std::none_of(w->fixed_poses.begin(), w->fixed_poses.end(), [](uint8_t f) { return f != 0; }))
    w->fixed_poses[0] = 1;
  for (uint8_t f : w->fixed_poses) w->num_free_images += f ? 0 : 1;

  // Observations (lib.rs:1457-1474): track elements whose image is in the problem.
  for (size_t cp = 0; cp < w->point_idx.size(); ++cp) {
    const auto [b, e] = track(w->point_idx[cp]);
    for (uint64_t k = b; k < e; ++k) {
      const uint32_t img = r->track_images[k];
      if (image_pos[img] < 0) continue;
      const uint64_t kp = r->keypoint_offsets[img] + r->track_keypoints[k];
      if (r->track_keypoints[k] >= r->keypoint_offsets[img + 1] - r->keypoint_offsets[img])
        throw Error(SFMBA_ERR_INVALID_ARGUMENT, "track keypoint index out of range");
      w->obs.push_back(sfmba_observation{(uint64_t)image_pos[img], cp, (double)r->keypoints[2 * kp],
                                         (double)r->keypoints[2 * kp + 1]});
    }
  }
  if (w->obs.empty()) {
    *w = sfmba_window{};
    return w.release();
  }

  // Intrinsics held, principal points masked (lib.rs:1489, :717).
  const size_t nc = w->cams.size();
  w->fixed_cams.assign(nc, 1);
  std::vector<uint8_t> mask(8 * nc);
  w->mask_len.resize(nc);
  check(sfmba_default_fixed_params_mask(w->cams.data(), nc, mask.data(), w->mask_len.data()));
  for (size_t c = 0; c < nc; ++c) w->mask_flat.insert(w->mask_flat.end(), &mask[8 * c], &mask[8 * c] + w->mask_len[c]);
  return w.release();
}

sfmba_input view(const sfmba_window& w) {
  sfmba_input in{};
  if (w.empty()) return in;
  in.num_images = w.image_idx.size();
  in.num_cameras = w.cams.size();
  in.num_points = w.point_idx.size();
  in.num_observations = w.obs.size();
  in.camera_of_image = w.coi.data();
  in.num_camera_of_image = w.coi.size();
  in.cameras = w.cams.data();
  in.poses = w.poses.data();
  in.points = w.pts.data();
  in.observations = w.obs.data();
  in.fixed_poses = w.fixed_poses.data();
  in.num_fixed_poses = w.fixed_poses.size();
  in.fixed_cameras = w.fixed_cams.data();
  in.num_fixed_cameras = w.fixed_cams.size();
  in.fixed_camera_params = w.mask_flat.data();
  in.fixed_camera_params_len = w.mask_len.data();
  in.num_fixed_camera_params = w.mask_len.size();
  return in;
}

void scatter(const sfmba_window& w, const sfmba_output* out, sfmba_reconstruction* r) {
  if (w.empty()) return;
  if (!out || !out->poses || !out->points || !out->cameras) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL output");
  for (size_t i = 0; i < w.image_idx.size(); ++i) r->poses[w.image_idx[i]] = out->poses[i];
  for (size_t p = 0; p < w.point_idx.size(); ++p)
    std::memcpy(r->points + 3 * w.point_idx[p], out->points + 3 * p, 3 * sizeof(double));
  for (size_t c = 0; c < w.cams.size(); ++c) r->cameras[w.camera_entry[c]] = out->cameras[c];
}

struct Result {
  std::vector<sfmba_pose> poses;
  std::vector<double> pts;
  std::vector<sfmba_camera> cams;
  sfmba_output o{};
  explicit Result(const sfmba_window& w) : poses(w.poses.size()), pts(w.pts.size()), cams(w.cams.size()) {
    o.poses = poses.data();
    o.points = pts.data();
    o.cameras = cams.data();
  }
};

void summarize(const sfmba_window& w, const sfmba_output* o, sfmba_window_summary* s) {
  if (!s) return;
  *s = sfmba_window_summary{};
  s->num_images = w.image_idx.size();
  s->num_free_images = w.num_free_images;
  s->num_cameras = w.cams.size();
  s->num_points = w.point_idx.size();
  s->num_observations = w.obs.size();
  if (o) {
    s->initial_cost = o->initial_cost;
    s->final_cost = o->final_cost;
    s->iterations_run = o->iterations_run;
  }
}

}  // namespace
}  // namespace sfmba

using sfmba::Error;

template <typename F>
static int scope_guarded(F&& f) {
  sfmba::set_error("");
  try {
    f();
    return SFMBA_OK;
  } catch (const Error& e) {
    sfmba::set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    sfmba::set_error("host allocation failed");
    return SFMBA_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    sfmba::set_error(e.what());
    return SFMBA_ERR_INVALID_ARGUMENT;
  }
}

extern "C" {

int sfmba_window_create(const sfmba_reconstruction* rec, int32_t scope, uint64_t center, sfmba_window** out) {
  return scope_guarded([&] {
    if (!out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL out");
    *out = nullptr;
    *out = sfmba::build(rec, scope, center);
  });
}

int sfmba_window_input(const sfmba_window* w, sfmba_input* input) {
  return scope_guarded([&] {
    if (!w || !input) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    *input = sfmba::view(*w);
  });
}

int sfmba_window_indices(const sfmba_window* w, const uint64_t** image_indices, size_t* num_images,
                         const uint64_t** point_indices, size_t* num_points) {
  return scope_guarded([&] {
    if (!w) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL window");
    if (image_indices) *image_indices = w->image_idx.data();
    if (num_images) *num_images = w->image_idx.size();
    if (point_indices) *point_indices = w->point_idx.data();
    if (num_points) *num_points = w->point_idx.size();
  });
}

int sfmba_window_scatter(const sfmba_window* w, const sfmba_output* out, sfmba_reconstruction* rec) {
  return scope_guarded([&] {
    if (!w || !rec) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    sfmba::scatter(*w, out, rec);
  });
}

void sfmba_window_destroy(sfmba_window* w) { delete w; }

int sfmba_reconstruction_bundle_adjust(sfmba_reconstruction* rec, int32_t scope, uint64_t center,
                                       const sfmba_policy* policy, const sfmba_options* options,
                                       sfmba_window_summary* summary, sfmba_policy_report* report) {
  return scope_guarded([&] {
    if (!policy) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL policy");
    std::unique_ptr<sfmba_window> w(sfmba::build(rec, scope, center));
    if (report) *report = sfmba_policy_report{};
    if (w->empty()) {
      sfmba::summarize(*w, nullptr, summary);
      return;
    }
    sfmba_policy pol = *policy;
    pol.camera_caller_fixed = w->caller_fixed.data();
    pol.num_camera_caller_fixed = w->caller_fixed.size();
    const sfmba_input in = sfmba::view(*w);
    sfmba::Result res(*w);
    sfmba::check(sfmba_run_bundle_adjustment(&in, &pol, options, &res.o, report));
    sfmba::scatter(*w, &res.o, rec);
    sfmba::summarize(*w, &res.o, summary);
  });
}

int sfmba_reconstruction_bundle_adjust_batch(sfmba_reconstruction* recs, size_t count, const int32_t* scopes,
                                             const uint64_t* centers, const sfmba_params* params,
                                             const sfmba_options* options, sfmba_window_summary* summaries) {
  return scope_guarded([&] {
    if (count && (!recs || !scopes || !centers)) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    sfmba_params p;
    sfmba_params_default(&p);
    if (params) p = *params;
    std::vector<std::unique_ptr<sfmba_window>> ws;
    for (size_t k = 0; k < count; ++k) ws.emplace_back(sfmba::build(&recs[k], scopes[k], centers[k]));
    std::vector<sfmba_input> ins;
    std::vector<size_t> which;
    std::vector<std::unique_ptr<sfmba::Result>> res;
    for (size_t k = 0; k < count; ++k)
      if (!ws[k]->empty()) {
        which.push_back(k);
        ins.push_back(sfmba::view(*ws[k]));
        res.emplace_back(new sfmba::Result(*ws[k]));
      }
    std::vector<sfmba_output> outs;
    for (auto& r : res) outs.push_back(r->o);
    if (!ins.empty())
      sfmba::check(sfmba_bundle_adjust_batch(ins.data(), ins.size(), &p, options, outs.data(), nullptr));
    for (size_t k = 0; k < count; ++k) sfmba::summarize(*ws[k], nullptr, summaries ? &summaries[k] : nullptr);
    for (size_t j = 0; j < which.size(); ++j) {
      sfmba::scatter(*ws[which[j]], &outs[j], &recs[which[j]]);
      sfmba::summarize(*ws[which[j]], &outs[j], summaries ? &summaries[which[j]] : nullptr);
    }
  });
}

}  // extern "C"
