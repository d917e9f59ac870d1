// Host side of the engine: C ABI entry points, problem packing and the
// Levenberg-Marquardt driver.  All arithmetic on the problem runs in the
// kernels of engine_kernels.hip and dense_kernels.hip; the host only builds
// index structures, launches, and reads back two scalars per damping try
// (the trial cost and the factorization status), exactly the values the
// reference's control flow branches on (crates/sfm-ba/src/lib.rs:1081-1221).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sfmba.h"
#include "camera_models.hpp"
#include "engine.hpp"
#include "engine_device.hpp"
#include "pcg.hpp"
#include "window.hpp"

namespace sfmba {

static thread_local std::string g_last_error;

void set_error(const std::string& m) { g_last_error = m; }

// Device allocations owned by one call; freed in the destructor.
struct DevArena {
  std::vector<void*> ptrs;
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename X>
  X* alloc(size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, n * sizeof(X)));
    ptrs.push_back(p);
    return static_cast<X*>(p);
  }
  template <typename X>
  X* upload(const std::vector<X>& v, hipStream_t s) {
    X* d = alloc<X>(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, s));
    return d;
  }
};

struct PinnedHost {
  void* p = nullptr;
  explicit PinnedHost(size_t bytes) { HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
  ~PinnedHost() {
    if (p) (void)hipHostFree(p);
  }
};

struct StreamGuard {
  hipStream_t s = nullptr;
  explicit StreamGuard(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      throw Error(SFMBA_ERR_DEVICE, "no HIP device available (this engine has no CPU path)");
    if (device < 0 || device >= n) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};

static inline bool flag_at(const uint8_t* v, size_t n, size_t i) { return v && i < n && v[i] != 0; }

static void validate(const sfmba_input* in) {
  if (!in) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "input is NULL");
  auto need = [](bool ok, const char* m) {
    if (!ok) throw Error(SFMBA_ERR_INVALID_ARGUMENT, m);
  };
  need(in->num_images == 0 || in->poses, "poses is NULL");
  need(in->num_cameras == 0 || in->cameras, "cameras is NULL");
  need(in->num_points == 0 || in->points, "points is NULL");
  need(in->num_observations == 0 || in->observations, "observations is NULL");
  need(in->num_camera_of_image >= in->num_images && (in->num_images == 0 || in->camera_of_image),
       "camera_of_image shorter than poses");
  need(in->num_observations < 0xFFFFFFF0ull && in->num_points < 0xFFFFFFF0ull &&
           in->num_images < 0xFFFFFFF0ull && in->num_cameras < 0xFFFFFFF0ull,
       "problem too large for 32-bit indices");
  need(in->num_fixed_camera_params == 0 || in->fixed_camera_params_len,
       "fixed_camera_params_len is NULL");
  for (size_t i = 0; i < in->num_images; ++i)
    if (in->camera_of_image[i] >= in->num_cameras)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "camera_of_image[" + std::to_string(i) + "] out of range");
  for (size_t c = 0; c < in->num_cameras; ++c)
    if (model_param_count_host(in->cameras[c].model_id) < 0)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "camera " + std::to_string(c) + " has an unknown model id");
  for (size_t o = 0; o < in->num_observations; ++o) {
    if (in->observations[o].image_idx >= in->num_images)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "observation " + std::to_string(o) + ": image_idx out of range");
    if (in->observations[o].point_idx >= in->num_points)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "observation " + std::to_string(o) + ": point_idx out of range");
  }
}

// Host-built structure of one problem (see engine.hpp for the layout).
struct Pack {
  uint32_t n_img = 0, n_cam = 0, n_pt = 0, n_obs = 0;
  uint32_t D = 0, Dp = 0, ld = 0, n_grp = 0;
  std::vector<uint32_t> obs_img, obs_pt, obs_inc, pt_ptr;
  std::vector<uint32_t> pt_fix;  // points left to k_point_accumulate
  std::vector<double2> obs_xy;
  std::vector<uint32_t> inc_grp, inc_pt, inc_obs, pt_inc_ptr;
  std::vector<uint32_t> cam_of_img, img_slot, cam_slot, cam_nfree, grp_off, grp_dim;
  std::vector<int32_t> cam_model;
  std::vector<uint8_t> cam_free;
  std::vector<uint32_t> pairs;
  std::vector<uint4> chunks;
  std::vector<uint32_t> multi_dest, dest_part0, dest_nparts;
  uint32_t n_parts = 0;
  std::vector<double> poses, cams, pts;  // packed FP64 state
};

// pair_chunk: most pairs one Schur chunk reduces (0: no limit, one chunk
// per destination).
static void build_pack(const sfmba_input* in, Pack& P, bool need_pairs,
                       uint32_t pair_chunk = PAIR_CHUNK) {
  P.n_img = (uint32_t)in->num_images;
  P.n_cam = (uint32_t)in->num_cameras;
  P.n_pt = (uint32_t)in->num_points;
  P.n_obs = (uint32_t)in->num_observations;
  P.cam_of_img.resize(P.n_img);
  for (uint32_t i = 0; i < P.n_img; ++i) P.cam_of_img[i] = (uint32_t)in->camera_of_image[i];
  P.cam_model.resize(P.n_cam);
  P.cams.assign(8 * (size_t)P.n_cam, 0.0);
  std::vector<int> dof(P.n_cam);
  for (uint32_t c = 0; c < P.n_cam; ++c) {
    P.cam_model[c] = in->cameras[c].model_id;
    dof[c] = model_param_count_host(in->cameras[c].model_id);
    for (int k = 0; k < dof[c]; ++k) P.cams[8 * (size_t)c + k] = in->cameras[c].params[k];
  }
  P.poses.resize(7 * (size_t)P.n_img);
  for (uint32_t i = 0; i < P.n_img; ++i) {
    const sfmba_pose& q = in->poses[i];
    const double v[7] = {q.qw, q.qx, q.qy, q.qz, q.tx, q.ty, q.tz};
    std::memcpy(&P.poses[7 * (size_t)i], v, sizeof(v));
  }
  P.pts.assign(in->points, in->points + 3 * (size_t)P.n_pt);

  // Free unknowns: free poses (6) and the unpinned parameters of free
  // cameras (lib.rs:753-776, 1045-1070; pinned parameters are left out of
  // the system, which is the exact decoupling the reference builds).
  std::vector<uint8_t> pose_free(P.n_img), cam_is_free(P.n_cam);
  for (uint32_t i = 0; i < P.n_img; ++i) pose_free[i] = !flag_at(in->fixed_poses, in->num_fixed_poses, i);
  P.cam_free.assign(8 * (size_t)P.n_cam, 0);
  P.cam_nfree.assign(P.n_cam, 0);
  size_t moff = 0;
  std::vector<size_t> mask_off(in->num_fixed_camera_params + 1, 0);
  for (size_t c = 0; c < in->num_fixed_camera_params; ++c) {
    mask_off[c] = moff;
    moff += in->fixed_camera_params_len[c];
  }
  for (uint32_t c = 0; c < P.n_cam; ++c) {
    cam_is_free[c] = dof[c] > 0 && !flag_at(in->fixed_cameras, in->num_fixed_cameras, c);
    if (!cam_is_free[c]) continue;
    uint32_t nf = 0;
    for (int k = 0; k < dof[c]; ++k) {
      bool pinned = false;
      if (c < in->num_fixed_camera_params && (size_t)k < in->fixed_camera_params_len[c])
        pinned = in->fixed_camera_params[mask_off[c] + k] != 0;
      if (!pinned) P.cam_free[8 * (size_t)c + nf++] = (uint8_t)k;
    }
    P.cam_nfree[c] = nf;
    if (nf == 0) cam_is_free[c] = 0;
  }
  std::vector<uint32_t> imgs_of_cam(P.n_cam, 0);
  for (uint32_t i = 0; i < P.n_img; ++i) ++imgs_of_cam[P.cam_of_img[i]];

  // Blocks of the reduced system: a free pose (merged with its camera when
  // that camera is free and used by this image alone), then the remaining
  // free cameras.
  P.img_slot.assign(P.n_img, NONE);
  P.cam_slot.assign(P.n_cam, NONE);
  std::vector<uint32_t> grp_of_pose(P.n_img, NONE), grp_of_cam(P.n_cam, NONE);
  uint32_t D = 0;
  for (uint32_t i = 0; i < P.n_img; ++i) {
    if (!pose_free[i]) continue;
    const uint32_t g = (uint32_t)P.grp_off.size();
    P.grp_off.push_back(D);
    uint32_t dim = 6;
    P.img_slot[i] = D;
    grp_of_pose[i] = g;
    const uint32_t c = P.cam_of_img[i];
    if (cam_is_free[c] && imgs_of_cam[c] == 1) {
      P.cam_slot[c] = D + 6;
      grp_of_cam[c] = g;
      dim += P.cam_nfree[c];
    }
    P.grp_dim.push_back(dim);
    D += dim;
  }
  for (uint32_t c = 0; c < P.n_cam; ++c) {
    if (!cam_is_free[c] || grp_of_cam[c] != NONE) continue;
    const uint32_t g = (uint32_t)P.grp_off.size();
    P.grp_off.push_back(D);
    P.grp_dim.push_back(P.cam_nfree[c]);
    P.cam_slot[c] = D;
    grp_of_cam[c] = g;
    D += P.cam_nfree[c];
  }
  P.n_grp = (uint32_t)P.grp_off.size();
  P.D = D;
  P.Dp = ((D + NB - 1) / NB) * NB;
  P.ld = P.Dp + NB;

  // Point-major observation order (stable counting sort).
  P.pt_ptr.assign((size_t)P.n_pt + 1, 0);
  for (uint32_t o = 0; o < P.n_obs; ++o) ++P.pt_ptr[in->observations[o].point_idx + 1];
  for (uint32_t p = 0; p < P.n_pt; ++p) P.pt_ptr[p + 1] += P.pt_ptr[p];
  // The linearization sums the point terms of a point whose observations
  // lie in one aligned span of LINEARIZE_SPAN; the others (and points
  // without observations) are summed by the accumulation pass.
  P.pt_fix.clear();
  for (uint32_t p = 0; p < P.n_pt; ++p) {
    const uint32_t b = P.pt_ptr[p], e = P.pt_ptr[p + 1];
    if (b == e || b / LINEARIZE_SPAN != (e - 1) / LINEARIZE_SPAN) P.pt_fix.push_back(p);
  }
  std::vector<uint32_t> fill(P.pt_ptr.begin(), P.pt_ptr.end() - 1), order(P.n_obs);
  for (uint32_t o = 0; o < P.n_obs; ++o) order[fill[in->observations[o].point_idx]++] = o;
  P.obs_img.resize(P.n_obs);
  P.obs_pt.resize(P.n_obs);
  P.obs_xy.resize(P.n_obs);
  P.obs_inc.assign(2 * (size_t)P.n_obs, NONE);
  for (uint32_t s = 0; s < P.n_obs; ++s) {
    const sfmba_observation& ob = in->observations[order[s]];
    P.obs_img[s] = (uint32_t)ob.image_idx;
    P.obs_pt[s] = (uint32_t)ob.point_idx;
    P.obs_xy[s] = double2{ob.x, ob.y};
  }

  // Incidences, in observation (hence point) order.
  P.pt_inc_ptr.assign((size_t)P.n_pt + 1, 0);
  auto add_inc = [&](uint32_t s, uint32_t g, int slot) {
    const uint32_t a = (uint32_t)P.inc_grp.size();
    P.inc_grp.push_back(g);
    P.inc_pt.push_back(P.obs_pt[s]);
    P.inc_obs.push_back(s);
    P.obs_inc[2 * (size_t)s + slot] = a;
  };
  for (uint32_t s = 0; s < P.n_obs; ++s) {
    const uint32_t img = P.obs_img[s], c = P.cam_of_img[img];
    const bool pf = pose_free[img], cf = cam_is_free[c];
    if (pf && cf && grp_of_cam[c] == grp_of_pose[img]) {
      add_inc(s, grp_of_pose[img], 0);
    } else {
      int slot = 0;
      if (pf) add_inc(s, grp_of_pose[img], slot++);
      if (cf) add_inc(s, grp_of_cam[c], slot++);
    }
    P.pt_inc_ptr[P.obs_pt[s] + 1] = (uint32_t)P.inc_grp.size();
  }
  for (uint32_t p = 0; p < P.n_pt; ++p) P.pt_inc_ptr[p + 1] = std::max(P.pt_inc_ptr[p + 1], P.pt_inc_ptr[p]);
  if (P.inc_grp.size() >= 0x80000000u)
    throw Error(SFMBA_ERR_UNSUPPORTED, "too many incidences for 31-bit pair words");
  if (!need_pairs || D == 0) return;
  if (P.n_grp > 0xFFFFu) throw Error(SFMBA_ERR_UNSUPPORTED, "more than 65535 parameter blocks");

  // Destination-sorted point pairs.  Every lower-triangular block (ga >= gb)
  // is a destination, so each try rewrites all of S.
  const uint64_t G = P.n_grp;
  const uint64_t n_dests = G * (G + 1) / 2;
  if (n_dests > 400000000ull) throw Error(SFMBA_ERR_UNSUPPORTED, "reduced system has too many blocks");
  std::vector<uint32_t> cnt(n_dests + 1, 0);
  for (uint32_t p = 0; p < P.n_pt; ++p) {
    const uint32_t a0 = P.pt_inc_ptr[p], a1 = P.pt_inc_ptr[p + 1];
    for (uint32_t a = a0; a < a1; ++a)
      for (uint32_t b = a0; b < a1; ++b) {
        const uint64_t ga = P.inc_grp[a], gb = P.inc_grp[b];
        if (ga < gb) continue;
        ++cnt[ga * (ga + 1) / 2 + gb + 1];
      }
  }
  for (uint64_t d = 0; d < n_dests; ++d) cnt[d + 1] += cnt[d];
  const uint64_t n_pairs = cnt[n_dests];
  if (n_pairs > 0x7FFFFFFFull) throw Error(SFMBA_ERR_UNSUPPORTED, "too many point pairs");
  P.pairs.resize(2 * n_pairs);
  {
    std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
    for (uint32_t p = 0; p < P.n_pt; ++p) {
      const uint32_t a0 = P.pt_inc_ptr[p], a1 = P.pt_inc_ptr[p + 1];
      for (uint32_t a = a0; a < a1; ++a)
        for (uint32_t b = a0; b < a1; ++b) {
          const uint64_t ga = P.inc_grp[a], gb = P.inc_grp[b];
          if (ga < gb) continue;
          const uint32_t q = pos[ga * (ga + 1) / 2 + gb]++;
          // First word: incidence a, bit 31 "same observation".  Second
          // word: incidence b, bit 31 "a == b".
          const bool same = P.inc_obs[a] == P.inc_obs[b];
          P.pairs[2 * (size_t)q] = a | (same ? 0x80000000u : 0u);
          P.pairs[2 * (size_t)q + 1] = b | (a == b ? 0x80000000u : 0u);
        }
    }
  }
  // Chunks name their destination as (ga << 16) | gb.  They are emitted in
  // tiles of SCHUR_TILE x SCHUR_TILE destination blocks, so the wavefronts in
  // flight read the F and E rows of about 2 SCHUR_TILE groups (a working set
  // that stays in the MALL) rather than of every group of a block row.
  auto emit = [&](uint64_t ga, uint64_t gb) {
    const uint64_t d = ga * (ga + 1) / 2 + gb;
    const uint32_t code = (uint32_t)((ga << 16) | gb);
    const uint32_t b0 = cnt[d], b1 = cnt[d + 1];
    const uint32_t nch =
        pair_chunk == 0 ? 1u : std::max<uint32_t>(1, (b1 - b0 + pair_chunk - 1) / pair_chunk);
    if (nch == 1) {
      P.chunks.push_back(uint4{code, b0, b1, NONE});
      return;
    }
    P.multi_dest.push_back(code);
    P.dest_part0.push_back(P.n_parts);
    P.dest_nparts.push_back(nch);
    for (uint32_t k = 0; k < nch; ++k) {
      const uint32_t lo = b0 + k * pair_chunk, hi = std::min(b1, lo + pair_chunk);
      P.chunks.push_back(uint4{code, lo, hi, P.n_parts++});
    }
  };
  for (uint64_t ta = 0; ta < G; ta += SCHUR_TILE)
    for (uint64_t tb = 0; tb <= ta; tb += SCHUR_TILE)
      for (uint64_t ga = ta; ga < std::min<uint64_t>(G, ta + SCHUR_TILE); ++ga)
        for (uint64_t gb = tb; gb < std::min<uint64_t>(ga + 1, tb + SCHUR_TILE); ++gb) emit(ga, gb);
}

struct Problem {
  DevArena arena;
  Layout L;
};

static void upload_layout(const Pack& P, Problem& pr, hipStream_t s) {
  Layout& L = pr.L;
  DevArena& A = pr.arena;
  L.n_img = P.n_img;
  L.n_cam = P.n_cam;
  L.n_pt = P.n_pt;
  L.n_obs = P.n_obs;
  L.n_inc = (uint32_t)P.inc_grp.size();
  L.n_grp = P.n_grp;
  L.D = P.D;
  L.Dp = P.Dp;
  L.ld = P.ld;
  L.n_pairs = P.pairs.size() / 2;
  L.n_chunks = (uint32_t)P.chunks.size();
  L.n_multi = (uint32_t)P.multi_dest.size();
  L.n_parts = P.n_parts;
  for (uint32_t g : P.grp_dim) L.max_grp_dim = std::max(L.max_grp_dim, g);
  L.rec_db = rec_db_for(L.max_grp_dim);
  L.obs_img = A.upload(P.obs_img, s);
  L.obs_pt = A.upload(P.obs_pt, s);
  L.obs_xy = A.upload(P.obs_xy, s);
  L.obs_inc = A.upload(P.obs_inc, s);
  L.pt_ptr = A.upload(P.pt_ptr, s);
  L.n_pt_fix = (uint32_t)P.pt_fix.size();
  L.pt_fix = A.upload(P.pt_fix, s);
  L.inc_grp = A.upload(P.inc_grp, s);
  L.inc_pt = A.upload(P.inc_pt, s);
  L.inc_obs = A.upload(P.inc_obs, s);
  L.pt_inc_ptr = A.upload(P.pt_inc_ptr, s);
  L.cam_of_img = A.upload(P.cam_of_img, s);
  L.cam_model = A.upload(P.cam_model, s);
  L.img_slot = A.upload(P.img_slot, s);
  L.cam_slot = A.upload(P.cam_slot, s);
  L.cam_free = A.upload(P.cam_free, s);
  L.cam_nfree = A.upload(P.cam_nfree, s);
  L.grp_off = A.upload(P.grp_off, s);
  L.grp_dim = A.upload(P.grp_dim, s);
  L.pairs = A.upload(P.pairs, s);
  {
    // Chunks without same-observation pairs first (the LDS-staged
    // assembly), then those with them; tile order kept within each class
    // (partials are indexed by chunk, so the order does not change sums).
    std::vector<uint4> ordered;
    ordered.reserve(P.chunks.size());
    std::vector<uint8_t> same(P.chunks.size(), 0);
    for (size_t c = 0; c < P.chunks.size(); ++c)
      for (uint32_t q = P.chunks[c].y; q < P.chunks[c].z && !same[c]; ++q)
        same[c] = (P.pairs[2 * (size_t)q] & 0x80000000u) != 0;
    for (size_t c = 0; c < P.chunks.size(); ++c)
      if (!same[c]) ordered.push_back(P.chunks[c]);
    L.n_chunks_plain = (uint32_t)ordered.size();
    for (size_t c = 0; c < P.chunks.size(); ++c)
      if (same[c]) ordered.push_back(P.chunks[c]);
    L.chunks = A.upload(ordered, s);
  }
  L.multi_dest = A.upload(P.multi_dest, s);
  L.dest_part0 = A.upload(P.dest_part0, s);
  L.dest_nparts = A.upload(P.dest_nparts, s);
  L.h_img_slot = P.img_slot;
  L.h_cam_slot = P.cam_slot;
  L.h_cam_nfree = P.cam_nfree;
  L.h_grp_off = P.grp_off;
  L.h_grp_dim = P.grp_dim;
  L.h_cam_free = P.cam_free;
}

// Phase timing with HIP events on the solve stream.
struct Profiler {
  bool on = false;
  hipStream_t s = nullptr;
  struct Rec {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  int open_phase = -1;
  hipEvent_t open_ev = nullptr;
  KernelTimer trailing;  // the factorization's trailing-update launches
  ~Profiler() {
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    if (open_ev) (void)hipEventDestroy(open_ev);
  }
  void begin(int phase) {
    if (!on) return;
    HIP_CHECK(hipEventCreate(&open_ev));
    HIP_CHECK(hipEventRecord(open_ev, s));
    open_phase = phase;
  }
  void end() {
    if (!on) return;
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, s));
    recs.push_back(Rec{open_phase, open_ev, e});
    open_ev = nullptr;
  }
  void collect(sfmba_stats* st) {
    if (!on || !st) return;
    HIP_CHECK(hipStreamSynchronize(s));
    for (auto& r : recs) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      const double sec = ms * 1e-3;
      // phases 5 and 6: the FP32 linearization and assembly of the mixed form
      if (r.phase == 5) {
        st->time_linearize_fp32_s += sec;
        st->count_linearize_fp32++;
      } else if (r.phase == 6) {
        st->time_schur_fp32_s += sec;
        st->count_schur_fp32++;
      }
      switch (r.phase % 5) {
        case 0: st->time_linearize_s += sec; st->count_linearize++; break;
        case 1: st->time_schur_s += sec; st->count_schur++; break;
        case 2: st->time_factor_s += sec; st->count_factor++; break;
        case 3: st->time_solve_s += sec; break;
        case 4: st->time_update_s += sec; break;
      }
    }
    trailing.collect(st->time_trailing_update_s, st->count_trailing_update, st->flop_trailing_update);
  }
};

static void push_stat(double* arr, uint64_t cap, uint64_t* len, double v) {
  if (arr && *len < cap) arr[*len] = v;
  ++*len;
}

// Zero a caller's stats, keeping its array pointers and capacities.
static void reset_stats(sfmba_stats* st) {
  sfmba_stats z{};
  z.outer_costs = st->outer_costs;
  z.outer_costs_capacity = st->outer_costs_capacity;
  z.trial_costs = st->trial_costs;
  z.trial_costs_capacity = st->trial_costs_capacity;
  z.try_lambdas = st->try_lambdas;
  z.try_outcomes = st->try_outcomes;
  z.tries_capacity = st->tries_capacity;
  *st = z;
}

// One damping try's damping and outcome (SFMBA_TRY_*) into the trace.
static void push_try(sfmba_stats* st, double lambda, int outcome) {
  if (!st) return;
  if (st->tries_len < st->tries_capacity) {
    if (st->try_lambdas) st->try_lambdas[st->tries_len] = lambda;
    if (st->try_outcomes) st->try_outcomes[st->tries_len] = (uint8_t)outcome;
  }
  ++st->tries_len;
  if (outcome & SFMBA_TRY_FP32) ++st->num_fp32_tries;
}


}  // namespace sfmba

// A problem resident on one device: packed structure, initial state and the
// per-precision work buffers, reused by every solve of the problem.
struct sfmba_problem {
  int device = 0;
  hipStream_t s = nullptr;
  sfmba::Problem pr;
  double* init_poses = nullptr;
  double* init_cams = nullptr;
  double* init_pts = nullptr;
  std::vector<sfmba_camera> cams_in;
  sfmba::Work<float> w32;
  sfmba::Work<double> w64;
  bool have32 = false, have64 = false;
  double* vec2 = nullptr;
  double* vec3 = nullptr;
  int* piv = nullptr;
  double* pinned = nullptr;
  sfmba::PcgBuffers pcg;
  bool have_pcg = false;
  double pack_s = 0.0, upload_s = 0.0;
  ~sfmba_problem() {
    // device buffers (pr.arena) are freed after this body, by member destruction
    if (s) (void)hipStreamSynchronize(s);
    if (pinned) (void)hipHostFree(pinned);
    if (pcg.host) (void)hipHostFree(pcg.host);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace sfmba {

static sfmba_problem* create_problem(const sfmba_input* in, int device) {
  using clock = std::chrono::steady_clock;
  validate(in);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw Error(SFMBA_ERR_DEVICE, "no HIP device available (this engine has no CPU path)");
  if (device < 0 || device >= n) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  HIP_CHECK(hipSetDevice(device));
  std::unique_ptr<sfmba_problem> H(new sfmba_problem());
  H->device = device;
  HIP_CHECK(hipStreamCreateWithFlags(&H->s, hipStreamNonBlocking));
  HIP_CHECK(hipHostMalloc((void**)&H->pinned, 64, hipHostMallocDefault));
  const auto t0 = clock::now();
  Pack P;
  build_pack(in, P, true);
  const auto t1 = clock::now();
  upload_layout(P, H->pr, H->s);
  H->init_poses = H->pr.arena.upload(P.poses, H->s);
  H->init_cams = H->pr.arena.upload(P.cams, H->s);
  H->init_pts = H->pr.arena.upload(P.pts, H->s);
  const Layout& L = H->pr.L;
  H->vec2 = H->pr.arena.alloc<double>(std::max<uint32_t>(L.Dp, 1));
  H->vec3 = H->pr.arena.alloc<double>(std::max<uint32_t>(L.Dp, 1));
  H->piv = H->pr.arena.alloc<int>(std::max<uint32_t>(L.Dp, 1));
  H->cams_in.assign(in->cameras, in->cameras + in->num_cameras);
  HIP_CHECK(hipStreamSynchronize(H->s));
  H->pack_s = std::chrono::duration<double>(t1 - t0).count();
  H->upload_s = std::chrono::duration<double>(clock::now() - t1).count();
  return H.release();
}

// PCG work buffers, allocated on the first PCG solve of a problem (the FP32
// copy of the reduced system on the first mixed-storage one).
static void alloc_pcg(sfmba_problem& H, bool mixed_storage) {
  if (H.have_pcg) {
    if (mixed_storage && !H.pcg.S32) {
      const Layout& L = H.pr.L;
      H.pcg.S32 = reinterpret_cast<float*>(H.pr.arena.alloc<double>(((size_t)L.ld * L.Dp + 1) / 2));
      H.pcg.x_best = H.pr.arena.alloc<double>(std::max<uint32_t>(L.Dp, 1));
    }
    return;
  }
  const Layout& L = H.pr.L;
  DevArena& A = H.pr.arena;
  PcgBuffers& B = H.pcg;
  std::vector<uint32_t> moff(L.n_grp), row_grp(std::max<uint32_t>(L.D, 1), 0);
  uint32_t m = 0;
  for (uint32_t g = 0; g < L.n_grp; ++g) {
    moff[g] = m;
    m += L.h_grp_dim[g] * L.h_grp_dim[g];
    for (uint32_t r = 0; r < L.h_grp_dim[g]; ++r) row_grp[L.h_grp_off[g] + r] = g;
  }
  B.Minv = A.alloc<double>(std::max<uint32_t>(m, 1));
  B.moff = A.upload(moff, H.s);
  B.row_grp = A.upload(row_grp, H.s);
  const size_t n = std::max<uint32_t>(L.Dp, 1);
  B.r = A.alloc<double>(n);
  B.r2 = A.alloc<double>(n);
  B.z = A.alloc<double>(n);
  B.p = A.alloc<double>(n);
  B.q = A.alloc<double>(n);
  B.symv_part = A.alloc<double>(std::max<size_t>(pcg_symv_part_doubles(L.D), 1));
  B.dot_part = A.alloc<double>(pcg_dot_part_doubles(L.D));
  B.scal = A.alloc<double>(16);
  if (mixed_storage) {
    B.S32 = reinterpret_cast<float*>(A.alloc<double>(((size_t)L.ld * L.Dp + 1) / 2));
    B.x_best = A.alloc<double>(n);
  }
  HIP_CHECK(hipHostMalloc((void**)&B.host, 8 * sizeof(double), hipHostMallocDefault));
  HIP_CHECK(hipStreamSynchronize(H.s));
  H.have_pcg = true;
}

// Work buffers of one precision.  The reduced system S and the factor
// workspace are FP64 in both precisions and never used by both at once, so
// a second precision's work shares them (shared_S / shared_Linv).
template <typename T>
static void alloc_work(sfmba_problem& H, Work<T>& W, double* shared_S = nullptr, double* shared_Linv = nullptr) {
  const Layout& L = H.pr.L;
  DevArena& A = H.pr.arena;
  hipStream_t s = H.s;
  for (int k = 0; k < 2; ++k) {
    W.poses[k] = A.alloc<double>(7 * (size_t)L.n_img);
    W.cams[k] = A.alloc<double>(8 * (size_t)L.n_cam);
    W.pts[k] = A.alloc<double>(3 * (size_t)L.n_pt);
  }
  const uint32_t n_cost_part = (L.n_obs + COST_BLOCK * COST_ITEMS - 1) / (COST_BLOCK * COST_ITEMS) + 1;
  W.cost_part = A.alloc<double>(2 * (size_t)n_cost_part);
  W.cost = A.alloc<double>(8);
  W.flags = A.alloc<int>(4);
  if (L.D == 0) return;
  W.chain = A.alloc<uint32_t>(L.Dp / NB);
  HIP_CHECK(hipMemsetAsync(W.chain, 0, (L.Dp / NB) * sizeof(uint32_t), s));
  // records: zeroed once (rows past a block's dimension are never written)
  const size_t rec_elems = (size_t)std::max<uint32_t>(L.n_inc, 1) * rec_values(L.rec_db);
  W.rec = A.alloc<T>(rec_elems);
  HIP_CHECK(hipMemsetAsync(W.rec, 0, rec_elems * sizeof(T), s));
  W.obs_pc = A.alloc<T>(9 * (size_t)L.n_obs);
  W.obs_valid = A.alloc<uint8_t>(L.n_obs);
  W.pt_C = A.alloc<double>(6 * (size_t)L.n_pt);
  W.pt_g = A.alloc<double>(3 * (size_t)L.n_pt);
  W.pt_active = A.alloc<uint8_t>(L.n_pt);
  W.pt_cinv = A.alloc<double>(9 * (size_t)L.n_pt);
  W.pt_ok = A.alloc<uint8_t>(L.n_pt);
  const size_t s_elems = (size_t)L.ld * L.Dp;
  W.delta = A.alloc<double>(L.Dp);
  W.vec = A.alloc<double>(L.Dp);
  W.rhs = A.alloc<double>(L.Dp);
  W.parts = A.alloc<double>((size_t)std::max<uint32_t>(L.n_parts, 1) * (2 * 196 + 16));
  W.inc_part = A.alloc<double>(3 * (size_t)std::max<uint32_t>(L.n_inc, 1));
  if (shared_S) {
    W.S = shared_S;
    W.Linv = shared_Linv;
    HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  W.S = A.alloc<double>(s_elems);
  W.Linv = A.alloc<double>(dense_workspace_doubles(L.Dp));
  // S: zero once, identity on the padding diagonal; the kernels rewrite the
  // D x D lower triangle and the rhs row on every try and never touch the
  // padding.
  HIP_CHECK(hipMemsetAsync(W.S, 0, s_elems * sizeof(double), s));
  std::vector<double> one(1, 1.0);
  for (uint32_t d = L.D; d < L.Dp; ++d)
    HIP_CHECK(hipMemcpyAsync(W.S + (size_t)d * L.ld + d, one.data(), sizeof(double), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// The stages of an LM solve in one precision T (the precision of the
// residual Jacobians and the Schur assembly; the point inverses, the
// factorization and the cost are FP64 in both).  The current state lives in
// W.poses/cams/pts[cur]; a trial step writes the other buffer.
template <typename T>
struct Stages {
  sfmba_problem& H;
  Work<T>& W;
  const sfmba_options* opt;
  sfmba_stats* st;
  Profiler& prof;
  int loss;
  double scale;
  bool refine, pcg;

  Stages(sfmba_problem& h, Work<T>& w, const sfmba_params* prm, const sfmba_options* o, sfmba_stats* s_,
         Profiler& p)
      : H(h), W(w), opt(o), st(s_), prof(p), loss(prm->robust_loss), scale(prm->loss_scale_px) {
    const Layout& L = H.pr.L;
    refine = sizeof(T) == 4 && opt->refine_steps > 0 && L.D;
    if (refine && !W.S_copy) W.S_copy = H.pr.arena.alloc<double>((size_t)L.ld * L.Dp);
    pcg = opt->solver == SFMBA_SOLVER_PCG || opt->solver == SFMBA_SOLVER_PCG_MIXED;
  }
  // State buffer 0 <- the given device arrays.
  void load(const double* poses, const double* cams, const double* pts) {
    const Layout& L = H.pr.L;
    HIP_CHECK(hipMemcpyAsync(W.poses[0], poses, 7 * (size_t)L.n_img * sizeof(double), hipMemcpyDeviceToDevice, H.s));
    HIP_CHECK(hipMemcpyAsync(W.cams[0], cams, 8 * (size_t)L.n_cam * sizeof(double), hipMemcpyDeviceToDevice, H.s));
    HIP_CHECK(hipMemcpyAsync(W.pts[0], pts, 3 * (size_t)L.n_pt * sizeof(double), hipMemcpyDeviceToDevice, H.s));
  }
  // total_cost and the in-front count of state buffer b (synchronizes).
  void cost(int b, double& c, double& front) {
    const Layout& L = H.pr.L;
    launch_cost(L, W.poses[b], W.cams[b], W.pts[b], loss, scale, W.cost_part, W.cost, 0, H.s);
    HIP_CHECK(hipMemcpyAsync(H.pinned, W.cost, 2 * sizeof(double), hipMemcpyDeviceToHost, H.s));
    HIP_CHECK(hipStreamSynchronize(H.s));
    c = H.pinned[0];
    front = H.pinned[1];
  }
  // Hot loops A and B (lib.rs:816-909) at state cur.
  void linearize(int cur) {
    prof.begin(sizeof(T) == 4 ? 5 : 0);
    launch_linearize<T>(H.pr.L, W, cur, loss, scale, H.s);
    launch_point_accumulate<T>(H.pr.L, W, H.s);
    prof.end();
    if (st) st->num_linearizations++;
  }
  // Trial state from W.delta and its cost; one host synchronization.
  void step_and_cost(int cur, double& trial, double& front) {
    const Layout& L = H.pr.L;
    double* h = H.pinned;
    prof.begin(4);
    launch_backsub_update<T>(L, W, cur, H.s);
    launch_cost(L, W.poses[1 - cur], W.cams[1 - cur], W.pts[1 - cur], loss, scale, W.cost_part, W.cost, 0, H.s);
    prof.end();
    HIP_CHECK(hipMemcpyAsync(h, W.cost, 2 * sizeof(double), hipMemcpyDeviceToHost, H.s));
    HIP_CHECK(hipMemcpyAsync(h + 4, W.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, H.s));
    HIP_CHECK(hipStreamSynchronize(H.s));
    trial = h[0];
    front = h[1];
    if (reinterpret_cast<const int*>(h + 4)[2] != 0)
      throw Error(SFMBA_ERR_DEVICE, "back substitution hand-over timed out");
  }
  // Reduced system at damping lam and its factorization, enqueued.  PCG
  // (which synchronizes) returns false on a breakdown; the Cholesky's
  // failure flag (flags[0], nalgebra's test) is read with the trial cost.
  bool assemble_factor(double lam) {
    const Layout& L = H.pr.L;
    hipStream_t s = H.s;
    prof.begin(sizeof(T) == 4 ? 6 : 1);
    launch_point_prepare<T>(L, W, lam, s);
    launch_schur<T>(L, W, lam, s);
    prof.end();
    HIP_CHECK(hipMemsetAsync(W.flags, 0, 4 * sizeof(int), s));
    if (pcg) {
      // Iterative solve of the damped reduced system (no factorization).
      prof.begin(3);
      const PcgResult pr = pcg_solve(H.pcg, W.S, L.ld, L.D, L.Dp, L.n_grp, L.grp_off, L.grp_dim,
                                     opt->pcg_max_iterations, opt->pcg_tolerance, W.delta, W.flags, H.pcg.host, s,
                                     opt->solver == SFMBA_SOLVER_PCG_MIXED);
      prof.end();
      if (st) st->num_pcg_iterations += pr.iterations;
      return pr.ok;
    }
    if (refine) {
      dense_mirror_copy(W.S, W.S_copy, L.Dp, L.ld, s);
      dense_extract_row(W.S, L.Dp, L.ld, L.Dp, W.rhs, s);
    }
    prof.begin(2);
    dense_cholesky(W.S, W.Linv, L.Dp, L.ld, L.D, W.flags, s, prof.on ? &prof.trailing : nullptr);
    prof.end();
    return true;
  }
  // The step from the factor (L^T delta = L^-1 rhs), with optional FP64
  // iterative refinement against the assembled system.
  void solve_factored() {
    const Layout& L = H.pr.L;
    hipStream_t s = H.s;
    prof.begin(3);
    dense_extract_row(W.S, L.Dp, L.ld, L.Dp, W.vec, s);
    dense_backsolve(W.S, W.Linv, L.Dp, L.ld, W.vec, W.delta, W.chain, next_epoch(W.chain_epoch), W.flags + 2, s);
    for (int it = 0; refine && it < opt->refine_steps; ++it) {
      // r = rhs - S delta, delta += S^-1 r with the same factor
      dense_symv_residual(W.S_copy, L.Dp, L.ld, W.delta, W.rhs, W.vec, s);
      dense_forwardsolve(W.S, W.Linv, L.Dp, L.ld, W.vec, H.vec2, s);
      dense_backsolve(W.S, W.Linv, L.Dp, L.ld, H.vec2, H.vec3, W.chain, next_epoch(W.chain_epoch), W.flags + 2, s);
      dense_vec_add(W.delta, H.vec3, L.Dp, s);
    }
    prof.end();
  }
  // One damping try (lib.rs:1076-1176): Cholesky, then (FP64 only) LU on
  // the rebuilt system.  Returns false when no step was formed; bits
  // receives SFMBA_TRY_LU when the LU fallback gave the step.
  bool try_step(int cur, double lam, double& trial, double& front, int& bits) {
    const Layout& L = H.pr.L;
    if (assemble_factor(lam)) {
      // step, trial state and cost are enqueued behind the factorization
      // and read back with its failure flag in one synchronization (a
      // failed factorization's trial is discarded)
      if (!pcg) solve_factored();
      step_and_cost(cur, trial, front);
      if (pcg || reinterpret_cast<const int*>(H.pinned + 4)[0] == 0) return true;
    }
    if (st) st->num_cholesky_failures++;
    if (pcg || sizeof(T) == 4) return false;
    // Not positive definite: rebuild S and solve by LU (lib.rs:1083-1093).
    bits |= SFMBA_TRY_LU;
    launch_point_prepare<T>(L, W, lam, H.s);
    launch_schur<T>(L, W, lam, H.s);
    dense_lu_solve(W.S, L.Dp, L.ld, L.D, W.delta, H.piv, W.flags, H.s);
    HIP_CHECK(hipMemcpyAsync(H.pinned + 4, W.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, H.s));
    HIP_CHECK(hipStreamSynchronize(H.s));
    if (st) st->num_lu_solves++;
    if (reinterpret_cast<const int*>(H.pinned + 4)[1] != 0) return false;
    step_and_cost(cur, trial, front);
    return true;
  }
};


// Bundle adjustment on a resident problem (lib.rs:747-1232).  FP64: every
// stage in FP64, the reference's arithmetic.  Mixed: Jacobians and the
// Schur assembly in FP32 while the Marquardt damping is at least
// MIXED_MIN_LAMBDA; once it falls below, or an FP32-assembled system fails
// to factor, the solve continues from its current state with FP64
// assembly.  The LM schedule (lambda, nu, accept/reject, stopping rule) is
// the same in both forms.
static void solve(sfmba_problem& H, const sfmba_params* prm, const sfmba_options* opt, sfmba_output* out,
                  sfmba_stats* st) {
  using clock = std::chrono::steady_clock;
  HIP_CHECK(hipSetDevice(H.device));
  const Layout& L = H.pr.L;
  hipStream_t s = H.s;
  const bool mixed = opt->precision == SFMBA_PRECISION_MIXED;
  if (opt->solver != SFMBA_SOLVER_CHOLESKY && L.D) alloc_pcg(H, opt->solver == SFMBA_SOLVER_PCG_MIXED);
  // FP64 work always (the mixed form finishes in FP64); FP32 work for the
  // mixed form, sharing S and the factor workspace.
  if (!H.have64) {
    alloc_work<double>(H, H.w64);
    H.have64 = true;
  }
  if (mixed && !H.have32) {
    alloc_work<float>(H, H.w32, H.w64.S, H.w64.Linv);
    H.have32 = true;
  }
  Profiler prof;
  prof.on = opt->profile != 0;
  prof.s = s;
  if (st) {
    st->dim = L.D;
    st->num_schur_pairs = L.n_pairs;
    st->time_upload_s = H.pack_s + H.upload_s;
  }
  Stages<double> s64(H, H.w64, prm, opt, st, prof);
  Stages<float> s32(H, H.w32, prm, opt, st, prof);
  bool fp32 = mixed;
  // Every solve starts from the problem's initial state.
  if (fp32) s32.load(H.init_poses, H.init_cams, H.init_pts);
  else s64.load(H.init_poses, H.init_cams, H.init_pts);
  HIP_CHECK(hipStreamSynchronize(s));
  const auto t1 = clock::now();

  int cur = 0;
  double initial_cost, n_front;
  if (fp32) s32.cost(0, initial_cost, n_front);
  else s64.cost(0, initial_cost, n_front);
  double cost = initial_cost;
  size_t iterations_run = 0;
  // The mixed form's switch to FP64 assembly: the current state moves to
  // the FP64 buffers and is linearized again there.
  auto escalate = [&]() {
    s64.load(H.w32.poses[cur], H.w32.cams[cur], H.w32.pts[cur]);
    cur = 0;
    fp32 = false;
    s64.linearize(cur);
  };

  if (L.D != 0) {
    double lambda = prm->initial_lambda;
    double nu = 2.0;
    const double FUNCTION_TOLERANCE = 1e-6;
    for (uint64_t outer = 0; outer < prm->max_iterations; ++outer) {
      ++iterations_run;
      const double cost_before = cost;
      if (n_front == 0.0) break;  // no observation in front of its camera (lib.rs:855-857)
      if (fp32 && lambda < MIXED_MIN_LAMBDA) escalate();
      else if (fp32) s32.linearize(cur);
      else s64.linearize(cur);
      bool accepted = false;
      for (int inner = 0; inner < 8; ++inner) {
        if (fp32 && lambda < MIXED_MIN_LAMBDA) escalate();
        double trial = 0.0, front = 0.0;
        int bits = 0;
        bool ok;
        if (fp32) {
          bits = SFMBA_TRY_FP32;
          ok = s32.try_step(cur, lambda, trial, front, bits);
          if (st) st->num_factorizations++;
          if (!ok && opt->solver == SFMBA_SOLVER_CHOLESKY) {
            // the FP32-assembled system did not factor: this damping again
            // with FP64 assembly
            push_try(st, lambda, SFMBA_TRY_NO_STEP | bits);
            escalate();
            bits = 0;
            ok = s64.try_step(cur, lambda, trial, front, bits);
            if (st) st->num_factorizations++;
          }
        } else {
          ok = s64.try_step(cur, lambda, trial, front, bits);
          if (st) st->num_factorizations++;
        }
        if (!ok) {
          push_try(st, lambda, SFMBA_TRY_NO_STEP | bits);
          lambda *= nu;
          nu *= 2.0;
          continue;
        }
        if (st) push_stat(st->trial_costs, st->trial_costs_capacity, &st->trial_costs_len, trial);
        if (std::isfinite(trial) && trial < cost) {
          push_try(st, lambda, SFMBA_TRY_ACCEPTED | bits);
          cur = 1 - cur;
          cost = trial;
          n_front = front;
          lambda = std::max(lambda / 3.0, 1e-12);
          nu = 2.0;
          accepted = true;
          break;
        }
        push_try(st, lambda, SFMBA_TRY_REJECTED | bits);
        if (st) st->num_rejected_steps++;
        lambda *= nu;
        nu *= 2.0;
      }
      if (st) push_stat(st->outer_costs, st->outer_costs_capacity, &st->outer_costs_len, cost);
      if (!accepted) break;
      const double rel = (cost_before - cost) / std::max(cost_before, 1e-12);
      if (rel < FUNCTION_TOLERANCE) break;
    }
  }
  HIP_CHECK(hipStreamSynchronize(s));
  if (st) st->time_total_s = std::chrono::duration<double>(clock::now() - t1).count();
  prof.collect(st);

  // Results back to the caller's arrays.
  const double* rp = fp32 ? H.w32.poses[cur] : H.w64.poses[cur];
  const double* rc = fp32 ? H.w32.cams[cur] : H.w64.cams[cur];
  const double* rx = fp32 ? H.w32.pts[cur] : H.w64.pts[cur];
  std::vector<double> hp(7 * (size_t)L.n_img), hc(8 * (size_t)L.n_cam);
  HIP_CHECK(hipMemcpyAsync(hp.data(), rp, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(hc.data(), rc, hc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  if (L.n_pt)
    HIP_CHECK(hipMemcpyAsync(out->points, rx, 3 * (size_t)L.n_pt * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < L.n_img; ++i) {
    sfmba_pose& q = out->poses[i];
    const double* v = &hp[7 * (size_t)i];
    q.qw = v[0]; q.qx = v[1]; q.qy = v[2]; q.qz = v[3];
    q.tx = v[4]; q.ty = v[5]; q.tz = v[6];
  }
  for (uint32_t c = 0; c < L.n_cam; ++c) {
    out->cameras[c] = H.cams_in[c];
    const int k = model_param_count_host(H.cams_in[c].model_id);
    for (int j = 0; j < k; ++j) out->cameras[c].params[j] = hc[8 * (size_t)c + j];
  }
  out->initial_cost = initial_cost;
  out->final_cost = cost;
  out->iterations_run = iterations_run;
}

static void solve_problem(sfmba_problem* H, const sfmba_params* params, const sfmba_options* options,
                          sfmba_output* out, sfmba_stats* stats) {
  if (!H || !params || !out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
  const Layout& L = H->pr.L;
  if ((L.n_img && !out->poses) || (L.n_pt && !out->points) || (L.n_cam && !out->cameras))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "output arrays are NULL");
  sfmba_options o;
  sfmba_options_default(&o);
  if (options) o = *options;
  if (stats) reset_stats(stats);  // keeps the caller's array pointers and capacities
  if (o.solver != SFMBA_SOLVER_CHOLESKY && o.solver != SFMBA_SOLVER_PCG && o.solver != SFMBA_SOLVER_PCG_MIXED)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown solver");
  if (o.solver != SFMBA_SOLVER_CHOLESKY && !(o.pcg_max_iterations > 0 && o.pcg_tolerance >= 0.0))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "PCG needs pcg_max_iterations > 0 and pcg_tolerance >= 0");
  if (o.precision != SFMBA_PRECISION_MIXED && o.precision != SFMBA_PRECISION_FP64)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown precision");
  solve(*H, params, &o, out, stats);
}

// Layout without pair lists for the cost / Jacobian entry points.
static void upload_state(const sfmba_input* in, Problem& pr, Pack& P, hipStream_t s, double** poses,
                         double** cams, double** pts) {
  build_pack(in, P, false);
  upload_layout(P, pr, s);
  *poses = pr.arena.upload(P.poses, s);
  *cams = pr.arena.upload(P.cams, s);
  *pts = pr.arena.upload(P.pts, s);
}

}  // namespace sfmba

// ------------------------------------------------------------ batched solve
// Many small problems (local-BA windows) in one persistent launch
// (window_kernels.hip): host packing on worker threads, every window's
// arrays in one device allocation uploaded by one copy, one kernel launch,
// one copy of all final states back.  Windows whose reduced system exceeds
// WINDOW_MAX_DIM go through the single-problem path afterwards.
namespace sfmba {

// Offsets of one window's arrays inside the batch allocation.
struct WinPlan {
  size_t obs_img, obs_pt, obs_xy, obs_inc, pt_ptr, inc_grp, inc_pt, pt_inc_ptr;
  size_t cam_of_img, cam_model, img_slot, cam_slot, cam_free, cam_nfree, grp_off, grp_dim, pairs, chunks;
  size_t state;  // poses[0], cams[0], pts[0], poses[1], cams[1], pts[1] (doubles)
  size_t rec, obs_pc, obs_valid, pt_C, pt_g, pt_active, pt_cinv, pt_ok, S, delta;
};

struct Planner {
  size_t n = 0;
  size_t take(size_t bytes) {
    const size_t o = n;
    n += (bytes + 255) & ~size_t(255);
    return o;
  }
};

// A batch packed and resident on the device: the windows the kernel
// solves (win) and those too large for it (big, solved by the
// single-problem path).
template <typename T>
struct WindowBatch {
  std::vector<Pack> packs;
  std::vector<std::vector<sfmba_camera>> cams_in;
  std::vector<size_t> win, big;
  uint32_t max_dim = 1;
  std::vector<WinPlan> plan;
  std::vector<size_t> state_dbl;
  size_t up_struct = 0, state_end = 0;
  DevArena arena;
  uint8_t* base = nullptr;
  View<T>* d_views = nullptr;
  WindowResult* d_res = nullptr;
  double* d_oc = nullptr;
  uint64_t max_it = 1;
  WindowParams wp{};
  std::vector<uint8_t> host;  // structure + initial state, as uploaded
  uint64_t total_obs = 0;
};

template <typename T>
static void batch_prepare(WindowBatch<T>& B, const sfmba_input* inputs, size_t count, const sfmba_params* prm,
                          hipStream_t s) {
  for (size_t w = 0; w < count; ++w) validate(&inputs[w]);
  B.packs.resize(count);
  B.cams_in.resize(count);
  for (size_t w = 0; w < count; ++w)
    B.cams_in[w].assign(inputs[w].cameras, inputs[w].cameras + inputs[w].num_cameras);
  // Host packing on worker threads.
  {
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex err_mu;
    const unsigned nth = std::max(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned k = 0; k < std::min<size_t>(nth, count); ++k)
      th.emplace_back([&] {
        for (size_t w; (w = next++) < count;) {
          try {
            build_pack(&inputs[w], B.packs[w], true, 0);
          } catch (...) {
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
          }
        }
      });
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
  }
  for (size_t w = 0; w < count; ++w) {
    if (B.packs[w].D > WINDOW_MAX_DIM) {
      B.big.push_back(w);
      continue;
    }
    B.win.push_back(w);
    B.total_obs += B.packs[w].n_obs;
    for (uint32_t g : B.packs[w].grp_dim) B.max_dim = std::max(B.max_dim, g);
  }

  // One allocation: uploaded structure and initial state first, then the
  // device-only work arrays.
  const size_t nw = B.win.size();
  B.plan.resize(nw);
  Planner up;
  for (size_t k = 0; k < nw; ++k) {
    const Pack& P = B.packs[B.win[k]];
    WinPlan& q = B.plan[k];
    q.obs_img = up.take(4 * (size_t)P.n_obs);
    q.obs_pt = up.take(4 * (size_t)P.n_obs);
    q.obs_xy = up.take(16 * (size_t)P.n_obs);
    q.obs_inc = up.take(8 * (size_t)P.n_obs);
    q.pt_ptr = up.take(4 * ((size_t)P.n_pt + 1));
    q.inc_grp = up.take(4 * P.inc_grp.size());
    q.inc_pt = up.take(4 * P.inc_pt.size());
    q.pt_inc_ptr = up.take(4 * P.pt_inc_ptr.size());
    q.cam_of_img = up.take(4 * (size_t)P.n_img);
    q.cam_model = up.take(4 * (size_t)P.n_cam);
    q.img_slot = up.take(4 * (size_t)P.n_img);
    q.cam_slot = up.take(4 * (size_t)P.n_cam);
    q.cam_free = up.take(8 * (size_t)P.n_cam);
    q.cam_nfree = up.take(4 * (size_t)P.n_cam);
    q.grp_off = up.take(4 * P.grp_off.size());
    q.grp_dim = up.take(4 * P.grp_dim.size());
    q.pairs = up.take(4 * P.pairs.size());
    q.chunks = up.take(16 * P.chunks.size());
  }
  B.up_struct = up.n;
  B.state_dbl.resize(nw);
  for (size_t k = 0; k < nw; ++k) {
    const Pack& P = B.packs[B.win[k]];
    B.state_dbl[k] = 7 * (size_t)P.n_img + 8 * (size_t)P.n_cam + 3 * (size_t)P.n_pt;
    B.plan[k].state = up.take(2 * 8 * B.state_dbl[k]);
  }
  B.state_end = up.n;
  Planner& wk = up;  // work arrays follow in the same allocation
  for (size_t k = 0; k < nw; ++k) {
    const Pack& P = B.packs[B.win[k]];
    WinPlan& q = B.plan[k];
    q.rec = wk.take(sizeof(T) * rec_values(rec_db_for(B.max_dim)) * std::max<size_t>(P.inc_grp.size(), 1));
    q.obs_pc = wk.take(9 * sizeof(T) * (size_t)P.n_obs);
    q.obs_valid = wk.take(P.n_obs);
    q.pt_C = wk.take(6 * 8 * (size_t)P.n_pt);
    q.pt_g = wk.take(3 * 8 * (size_t)P.n_pt);
    q.pt_active = wk.take(P.n_pt);
    q.pt_cinv = wk.take(9 * 8 * (size_t)P.n_pt);
    q.pt_ok = wk.take(P.n_pt);
    q.S = wk.take(8 * ((size_t)P.D + 1) * P.D);
    q.delta = wk.take(8 * (size_t)P.D);
  }
  B.base = B.arena.template alloc<uint8_t>(std::max<size_t>(up.n, 1));
  B.host.assign(B.state_end, 0);
  auto put = [&](size_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(B.host.data() + off, src, bytes);
  };
  std::vector<View<T>> views(nw);
  for (size_t k = 0; k < nw; ++k) {
    const Pack& P = B.packs[B.win[k]];
    const WinPlan& q = B.plan[k];
    put(q.obs_img, P.obs_img.data(), 4 * (size_t)P.n_obs);
    put(q.obs_pt, P.obs_pt.data(), 4 * (size_t)P.n_obs);
    put(q.obs_xy, P.obs_xy.data(), 16 * (size_t)P.n_obs);
    put(q.obs_inc, P.obs_inc.data(), 8 * (size_t)P.n_obs);
    put(q.pt_ptr, P.pt_ptr.data(), 4 * P.pt_ptr.size());
    put(q.inc_grp, P.inc_grp.data(), 4 * P.inc_grp.size());
    put(q.inc_pt, P.inc_pt.data(), 4 * P.inc_pt.size());
    put(q.pt_inc_ptr, P.pt_inc_ptr.data(), 4 * P.pt_inc_ptr.size());
    put(q.cam_of_img, P.cam_of_img.data(), 4 * (size_t)P.n_img);
    put(q.cam_model, P.cam_model.data(), 4 * (size_t)P.n_cam);
    put(q.img_slot, P.img_slot.data(), 4 * (size_t)P.n_img);
    put(q.cam_slot, P.cam_slot.data(), 4 * (size_t)P.n_cam);
    put(q.cam_free, P.cam_free.data(), 8 * (size_t)P.n_cam);
    put(q.cam_nfree, P.cam_nfree.data(), 4 * (size_t)P.n_cam);
    put(q.grp_off, P.grp_off.data(), 4 * P.grp_off.size());
    put(q.grp_dim, P.grp_dim.data(), 4 * P.grp_dim.size());
    put(q.pairs, P.pairs.data(), 4 * P.pairs.size());
    put(q.chunks, P.chunks.data(), 16 * P.chunks.size());
    put(q.state, P.poses.data(), 8 * P.poses.size());
    put(q.state + 8 * P.poses.size(), P.cams.data(), 8 * P.cams.size());
    put(q.state + 8 * (P.poses.size() + P.cams.size()), P.pts.data(), 8 * P.pts.size());

    View<T>& v = views[k];
    v = View<T>{};
    v.n_img = P.n_img;
    v.n_cam = P.n_cam;
    v.n_pt = P.n_pt;
    v.n_obs = P.n_obs;
    v.n_inc = (uint32_t)P.inc_grp.size();
    v.n_chunks = (uint32_t)P.chunks.size();
    v.D = P.D;
    v.Dp = P.D;  // rhs row right below the system
    v.ld = P.D + 1;
    auto at = [&](size_t off) { return B.base + off; };
    v.obs_img = reinterpret_cast<const uint32_t*>(at(q.obs_img));
    v.obs_pt = reinterpret_cast<const uint32_t*>(at(q.obs_pt));
    v.obs_xy = reinterpret_cast<const double2*>(at(q.obs_xy));
    v.obs_inc = reinterpret_cast<const uint32_t*>(at(q.obs_inc));
    v.pt_ptr = reinterpret_cast<const uint32_t*>(at(q.pt_ptr));
    v.inc_grp = reinterpret_cast<const uint32_t*>(at(q.inc_grp));
    v.inc_pt = reinterpret_cast<const uint32_t*>(at(q.inc_pt));
    v.pt_inc_ptr = reinterpret_cast<const uint32_t*>(at(q.pt_inc_ptr));
    v.cam_of_img = reinterpret_cast<const uint32_t*>(at(q.cam_of_img));
    v.cam_model = reinterpret_cast<const int32_t*>(at(q.cam_model));
    v.img_slot = reinterpret_cast<const uint32_t*>(at(q.img_slot));
    v.cam_slot = reinterpret_cast<const uint32_t*>(at(q.cam_slot));
    v.cam_free = reinterpret_cast<const uint8_t*>(at(q.cam_free));
    v.cam_nfree = reinterpret_cast<const uint32_t*>(at(q.cam_nfree));
    v.grp_off = reinterpret_cast<const uint32_t*>(at(q.grp_off));
    v.grp_dim = reinterpret_cast<const uint32_t*>(at(q.grp_dim));
    v.pairs = reinterpret_cast<const uint2*>(at(q.pairs));
    v.chunks = reinterpret_cast<const uint4*>(at(q.chunks));
    v.rec = reinterpret_cast<T*>(at(q.rec));
    v.obs_pc = reinterpret_cast<T*>(at(q.obs_pc));
    v.obs_valid = at(q.obs_valid);
    v.pt_C = reinterpret_cast<double*>(at(q.pt_C));
    v.pt_g = reinterpret_cast<double*>(at(q.pt_g));
    v.pt_active = at(q.pt_active);
    v.pt_cinv = reinterpret_cast<double*>(at(q.pt_cinv));
    v.pt_ok = at(q.pt_ok);
    v.S = reinterpret_cast<double*>(at(q.S));
    v.delta = reinterpret_cast<double*>(at(q.delta));
    double* st = reinterpret_cast<double*>(at(q.state));
    for (int c = 0; c < 2; ++c) {
      double* sb = st + (size_t)c * B.state_dbl[k];
      v.poses[c] = sb;
      v.cams[c] = sb + 7 * (size_t)P.n_img;
      v.pts[c] = sb + 7 * (size_t)P.n_img + 8 * (size_t)P.n_cam;
    }
  }
  B.d_views = B.arena.template alloc<View<T>>(std::max<size_t>(nw, 1));
  B.d_res = B.arena.template alloc<WindowResult>(std::max<size_t>(nw, 1));
  B.max_it = std::max<uint64_t>(prm->max_iterations, 1);
  B.d_oc = B.arena.template alloc<double>(std::max<size_t>(nw, 1) * B.max_it);
  if (nw) {
    HIP_CHECK(hipMemcpyAsync(B.base, B.host.data(), B.state_end, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(B.d_views, views.data(), nw * sizeof(View<T>), hipMemcpyHostToDevice, s));
  }
  HIP_CHECK(hipStreamSynchronize(s));
  B.wp = WindowParams{};
  B.wp.max_iterations = prm->max_iterations;
  B.wp.initial_lambda = prm->initial_lambda;
  B.wp.loss_scale_px = prm->loss_scale_px;
  B.wp.robust_loss = prm->robust_loss;
}

// Writes window k's solution (final state sb: poses, camera parameters,
// points) and result into the caller's output and statistics.
template <typename T>
static void batch_unpack(const WindowBatch<T>& B, size_t k, const double* sb, const WindowResult& r,
                         const double* oc, double kernel_s, double upload_s, sfmba_output* outs,
                         sfmba_stats* stats) {
  const size_t w = B.win[k];
  const Pack& P = B.packs[w];
  sfmba_output* out = &outs[w];
  for (uint32_t i = 0; i < P.n_img; ++i) {
    const double* q = sb + 7 * (size_t)i;
    out->poses[i] = sfmba_pose{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
  }
  const double* cs = sb + 7 * (size_t)P.n_img;
  for (uint32_t c = 0; c < P.n_cam; ++c) {
    out->cameras[c] = B.cams_in[w][c];
    const int kp = model_param_count_host(B.cams_in[w][c].model_id);
    for (int j = 0; j < kp; ++j) out->cameras[c].params[j] = cs[8 * (size_t)c + j];
  }
  if (P.n_pt) std::memcpy(out->points, cs + 8 * (size_t)P.n_cam, 24 * (size_t)P.n_pt);
  out->initial_cost = r.initial_cost;
  out->final_cost = r.final_cost;
  out->iterations_run = r.iterations_run;
  if (stats) {
    sfmba_stats& st = stats[w];
    st.dim = P.D;
    st.num_linearizations = r.num_linearizations;
    st.num_factorizations = r.num_factorizations;
    st.num_cholesky_failures = r.num_cholesky_failures;
    st.num_lu_solves = r.num_lu_solves;
    st.num_rejected_steps = r.num_rejected_steps;
    st.num_schur_pairs = P.pairs.size() / 2;
    for (uint64_t i = 0; i < r.num_outer; ++i) push_stat(st.outer_costs, st.outer_costs_capacity, &st.outer_costs_len, oc[i]);
    st.time_total_s = kernel_s;
    st.time_upload_s = upload_s;
  }
}

// ---- launch-shape autotuner
// Tuned shapes, per precision, destination block class and batch size
// class (log2 of the window count and of the mean observations per
// window), for the life of the process.
static std::mutex g_tune_mu;
static std::map<std::array<uint32_t, 4>, int> g_tuned_shape;

template <typename T>
static std::array<uint32_t, 4> tune_key(const WindowBatch<T>& B) {
  auto lg = [](uint64_t x) {
    uint32_t l = 0;
    while ((1ull << l) < x) ++l;
    return l;
  };
  const uint64_t nw = std::max<size_t>(B.win.size(), 1);
  return {(uint32_t)sizeof(T), rec_db_for(B.max_dim), lg(nw), lg(std::max<uint64_t>(B.total_obs / nw, 1))};
}

template <typename T>
static int lookup_shape(const WindowBatch<T>& B) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tuned_shape.find(tune_key(B));
  return it == g_tuned_shape.end() ? window_default_shape<T>() : it->second;
}

}  // namespace sfmba

// The batch plan of the C ABI: a batch resident on the device, its tuned
// launch shape and one solve captured in a HIP graph.
struct sfmba_batch {
  int device = 0;
  int precision = 0;
  hipStream_t s = nullptr;
  sfmba::WindowBatch<float> b32;
  sfmba::WindowBatch<double> b64;
  int shape = 0;
  sfmba_batch_info info{};
  sfmba::DevArena arena;
  double* d_pristine = nullptr;  // copy of the state region at creation
  double* d_out = nullptr;       // final states, compact
  std::vector<uint64_t> out_off;
  double* h_out = nullptr;       // pinned mirrors
  sfmba::WindowResult* h_res = nullptr;
  double* h_oc = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::unique_ptr<sfmba_problem>> big;  // windows above the kernel's limit
  sfmba_params prm{};
  sfmba_options opt{};
  ~sfmba_batch();
};

namespace sfmba {

// Restores the initial state of every window (both state buffers) from the
// pristine copy: a copy kernel, so that graph captures hold no memset or
// device-to-device copy node.
template <typename T>
static void batch_restore(const WindowBatch<T>& B, const double* pristine, hipStream_t s) {
  if (B.win.empty()) return;
  dense_copy(pristine, reinterpret_cast<double*>(B.base + B.up_struct), (B.state_end - B.up_struct) / 8, s);
}

// Times every launch shape on the batch with the iteration count capped at
// tune_iterations (the initial state restored before each run, one
// one-iteration warm-up launch per shape first) and records the fastest.
template <typename T>
static int tune_batch(WindowBatch<T>& B, const double* pristine, hipStream_t s, int tune_iterations,
                      double* seconds) {
  const uint32_t nw = (uint32_t)B.win.size();
  WindowParams wp = B.wp;
  wp.max_iterations = std::min<uint64_t>(B.wp.max_iterations, (uint64_t)std::max(tune_iterations, 1));
  WindowParams w1 = B.wp;
  w1.max_iterations = 1;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  int best = window_default_shape<T>();
  double best_t = 1e300;
  for (int sh = 0; sh < N_WINDOW_SHAPES; ++sh) {
    batch_restore(B, pristine, s);
    launch_window_lm<T>(B.d_views, std::min<uint32_t>(nw, 1), B.max_dim, w1, B.d_res, B.d_oc, sh, s);
    batch_restore(B, pristine, s);
    HIP_CHECK(hipEventRecord(e0, s));
    launch_window_lm<T>(B.d_views, nw, B.max_dim, wp, B.d_res, B.d_oc, sh, s);
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    seconds[sh] = ms * 1e-3;
    if (seconds[sh] < best_t) {
      best_t = seconds[sh];
      best = sh;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  batch_restore(B, pristine, s);
  HIP_CHECK(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tuned_shape[tune_key(B)] = best;
  return best;
}

template <typename T>
static void batch_solve(const sfmba_input* inputs, size_t count, const sfmba_params* prm,
                        const sfmba_options& opt, sfmba_output* outs, sfmba_stats* stats) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  StreamGuard sg(opt.device);
  WindowBatch<T> B;
  batch_prepare(B, inputs, count, prm, sg.s);
  const size_t nw = B.win.size();
  const auto t2 = clock::now();
  launch_window_lm<T>(B.d_views, (uint32_t)nw, B.max_dim, B.wp, B.d_res, B.d_oc, lookup_shape(B), sg.s);
  HIP_CHECK(hipStreamSynchronize(sg.s));
  const auto t3 = clock::now();
  std::vector<WindowResult> res(nw);
  std::vector<double> oc(nw * B.max_it);
  if (nw) {
    HIP_CHECK(hipMemcpyAsync(res.data(), B.d_res, nw * sizeof(WindowResult), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(oc.data(), B.d_oc, oc.size() * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(B.host.data() + B.up_struct, B.base + B.up_struct, B.state_end - B.up_struct,
                             hipMemcpyDeviceToHost, sg.s));
  }
  HIP_CHECK(hipStreamSynchronize(sg.s));
  const double kernel_s = std::chrono::duration<double>(t3 - t2).count();
  const double upload_s = std::chrono::duration<double>(t2 - t0).count();
  for (size_t k = 0; k < nw; ++k) {
    const double* sb =
        reinterpret_cast<const double*>(B.host.data() + B.plan[k].state) + (size_t)res[k].cur * B.state_dbl[k];
    batch_unpack(B, k, sb, res[k], oc.data() + k * B.max_it, kernel_s, upload_s, outs, stats);
  }
  // Windows too large for one workgroup: the single-problem path.
  for (size_t w : B.big) {
    std::unique_ptr<sfmba_problem> H(create_problem(&inputs[w], opt.device));
    solve_problem(H.get(), prm, &opt, &outs[w], stats ? &stats[w] : nullptr);
  }
}

// Plan creation: packing and upload, the pristine state copy, the
// autotuner (when asked), then one solve captured as a graph: restore the
// initial state, the window kernel, the gather of the final states, and
// the copies of results, outer costs and states into pinned host memory.
template <typename T>
static void batch_plan_build(sfmba_batch& Q, WindowBatch<T>& B, const sfmba_input* inputs, size_t count,
                             int autotune) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  hipStream_t s = Q.s;
  batch_prepare(B, inputs, count, &Q.prm, s);
  const size_t nw = B.win.size();
  const size_t n_state = (B.state_end - B.up_struct) / 8;
  Q.d_pristine = Q.arena.alloc<double>(std::max<size_t>(n_state, 1));
  if (nw) dense_copy(reinterpret_cast<const double*>(B.base + B.up_struct), Q.d_pristine, n_state, s);
  std::vector<uint64_t> st_off(nw), nd(nw);
  Q.out_off.resize(nw);
  uint64_t n_out = 0;
  for (size_t k = 0; k < nw; ++k) {
    st_off[k] = B.plan[k].state / 8;
    nd[k] = B.state_dbl[k];
    Q.out_off[k] = n_out;
    n_out += nd[k];
  }
  uint64_t* d_meta = Q.arena.alloc<uint64_t>(std::max<size_t>(3 * nw, 1));
  Q.d_out = Q.arena.alloc<double>(std::max<uint64_t>(n_out, 1));
  std::vector<uint64_t> meta;
  meta.insert(meta.end(), st_off.begin(), st_off.end());
  meta.insert(meta.end(), nd.begin(), nd.end());
  meta.insert(meta.end(), Q.out_off.begin(), Q.out_off.end());
  if (nw) HIP_CHECK(hipMemcpyAsync(d_meta, meta.data(), meta.size() * 8, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipHostMalloc((void**)&Q.h_out, std::max<uint64_t>(n_out, 1) * 8, hipHostMallocDefault));
  HIP_CHECK(hipHostMalloc((void**)&Q.h_res, std::max<size_t>(nw, 1) * sizeof(WindowResult), hipHostMallocDefault));
  HIP_CHECK(hipHostMalloc((void**)&Q.h_oc, std::max<size_t>(nw, 1) * B.max_it * 8, hipHostMallocDefault));
  HIP_CHECK(hipStreamSynchronize(s));
  Q.info.create_seconds = std::chrono::duration<double>(clock::now() - t0).count();
  Q.info.num_windows = nw;
  Q.info.num_large = B.big.size();
  Q.info.num_shapes = N_WINDOW_SHAPES;
  for (int k = 0; k < N_WINDOW_SHAPES; ++k) {
    Q.info.shape_threads[k] = WINDOW_SHAPES[k].threads;
    Q.info.shape_pairs[k] = WINDOW_SHAPES[k].batch;
  }
  if (autotune && nw) {
    Q.shape = tune_batch(B, Q.d_pristine, s, autotune, Q.info.shape_seconds);
    Q.info.tuned = 1;
  } else {
    Q.shape = lookup_shape(B);
  }
  Q.info.threads_per_window = WINDOW_SHAPES[Q.shape].threads;
  Q.info.pairs_in_flight = WINDOW_SHAPES[Q.shape].batch;
  // capture (a batch without windows for the kernel has no graph)
  const auto t1 = clock::now();
  if (nw) {
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      batch_restore(B, Q.d_pristine, s);
      launch_window_lm<T>(B.d_views, (uint32_t)nw, B.max_dim, B.wp, B.d_res, B.d_oc, Q.shape, s);
      launch_window_gather((uint32_t)nw, B.d_res, d_meta, d_meta + nw, d_meta + 2 * nw,
                           reinterpret_cast<const double*>(B.base), Q.d_out, s);
      HIP_CHECK(hipMemcpyAsync(Q.h_res, B.d_res, nw * sizeof(WindowResult), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(Q.h_oc, B.d_oc, nw * B.max_it * 8, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(Q.h_out, Q.d_out, n_out * 8, hipMemcpyDeviceToHost, s));
    } catch (...) {
      // leave capture mode before the error propagates
      hipGraph_t partial = nullptr;
      if (hipStreamEndCapture(s, &partial) == hipSuccess && partial) (void)hipGraphDestroy(partial);
      throw;
    }
    const hipError_t ce = hipStreamEndCapture(s, &Q.graph);
    if (ce != hipSuccess) throw Error(SFMBA_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
    HIP_CHECK(hipGraphInstantiate(&Q.exec, Q.graph, nullptr, nullptr, 0));
  }
  Q.info.capture_seconds = std::chrono::duration<double>(clock::now() - t1).count();
  for (size_t w : B.big) Q.big.emplace_back(create_problem(&inputs[w], Q.device));
}

template <typename T>
static void batch_plan_solve(sfmba_batch& Q, WindowBatch<T>& B, sfmba_output* outs, sfmba_stats* stats) {
  using clock = std::chrono::steady_clock;
  HIP_CHECK(hipSetDevice(Q.device));
  const auto t0 = clock::now();
  if (Q.exec) {
    HIP_CHECK(hipGraphLaunch(Q.exec, Q.s));
    HIP_CHECK(hipStreamSynchronize(Q.s));
  }
  const double graph_s = std::chrono::duration<double>(clock::now() - t0).count();
  for (size_t k = 0; k < B.win.size(); ++k)
    batch_unpack(B, k, Q.h_out + Q.out_off[k], Q.h_res[k], Q.h_oc + k * B.max_it, graph_s, 0.0, outs, stats);
  for (size_t j = 0; j < B.big.size(); ++j) {
    const size_t w = B.big[j];
    solve_problem(Q.big[j].get(), &Q.prm, &Q.opt, &outs[w], stats ? &stats[w] : nullptr);
  }
}

}  // namespace sfmba

// ------------------------------------------------------- caller policy
// The solve policy of the reference's caller (crates/sfm-reconstruction/
// src/lib.rs:1479-1768) on an assembled problem: the intrinsics modes, the
// fixed- vs free-intrinsics solves (run concurrently, each on its own
// stream), the residual filter of filter_and_reoptimize (residual norms on
// the device, compaction on the host), the distortion plausibility bound and
// the plain mean-error comparison.
namespace sfmba {

// An owned copy of a problem, from which filtered or re-seeded inputs are made.
struct OwnedInput {
  std::vector<uint64_t> coi;
  std::vector<sfmba_camera> cams;
  std::vector<sfmba_pose> poses;
  std::vector<double> pts;
  std::vector<sfmba_observation> obs;
  std::vector<uint8_t> fp, fc, fcp;
  std::vector<size_t> fcp_len;
  sfmba_input view() const {
    sfmba_input v{};
    v.num_images = poses.size();
    v.num_cameras = cams.size();
    v.num_points = pts.size() / 3;
    v.num_observations = obs.size();
    v.camera_of_image = coi.data();
    v.num_camera_of_image = coi.size();
    v.cameras = cams.data();
    v.poses = poses.data();
    v.points = pts.data();
    v.observations = obs.data();
    v.fixed_poses = fp.data();
    v.num_fixed_poses = fp.size();
    v.fixed_cameras = fc.data();
    v.num_fixed_cameras = fc.size();
    v.fixed_camera_params = fcp.data();
    v.fixed_camera_params_len = fcp_len.data();
    v.num_fixed_camera_params = fcp_len.size();
    return v;
  }
  void set_state(const sfmba_output& o) {
    std::copy(o.poses, o.poses + poses.size(), poses.begin());
    std::copy(o.points, o.points + pts.size(), pts.begin());
    std::copy(o.cameras, o.cameras + cams.size(), cams.begin());
  }
};

static OwnedInput own_input(const sfmba_input* in) {
  OwnedInput c;
  c.coi.assign(in->camera_of_image, in->camera_of_image + in->num_images);
  c.cams.assign(in->cameras, in->cameras + in->num_cameras);
  c.poses.assign(in->poses, in->poses + in->num_images);
  c.pts.assign(in->points, in->points + 3 * in->num_points);
  c.obs.assign(in->observations, in->observations + in->num_observations);
  if (in->num_fixed_poses) c.fp.assign(in->fixed_poses, in->fixed_poses + in->num_fixed_poses);
  if (in->num_fixed_cameras) c.fc.assign(in->fixed_cameras, in->fixed_cameras + in->num_fixed_cameras);
  size_t m = 0;
  for (size_t k = 0; k < in->num_fixed_camera_params; ++k) m += in->fixed_camera_params_len[k];
  if (m) c.fcp.assign(in->fixed_camera_params, in->fixed_camera_params + m);
  if (in->num_fixed_camera_params)
    c.fcp_len.assign(in->fixed_camera_params_len, in->fixed_camera_params_len + in->num_fixed_camera_params);
  return c;
}

// Output arrays owned by the caller of the policy helpers.
struct OwnedOutput {
  std::vector<sfmba_pose> poses;
  std::vector<double> pts;
  std::vector<sfmba_camera> cams;
  sfmba_output o{};
  explicit OwnedOutput(const OwnedInput& in) : poses(in.poses.size()), pts(in.pts.size()), cams(in.cams.size()) {
    o.poses = poses.data();
    o.points = pts.data();
    o.cameras = cams.data();
  }
};

static void solve_once(const sfmba_input* in, const sfmba_params* p, const sfmba_options* o, sfmba_output* out) {
  std::unique_ptr<sfmba_problem> H(create_problem(in, o->device));
  solve_problem(H.get(), p, o, out, nullptr);
}

static double mean_error(const sfmba_input* in, int device) {
  StreamGuard sg(device);
  Problem pr;
  Pack P;
  double *po, *ca, *pt;
  upload_state(in, pr, P, sg.s, &po, &ca, &pt);
  double* part = pr.arena.alloc<double>(2 * ((size_t)pr.L.n_obs / (COST_BLOCK * COST_ITEMS) + 2));
  double* res = pr.arena.alloc<double>(2);
  launch_cost(pr.L, po, ca, pt, SFMBA_LOSS_NONE, 1.0, part, res, 1, sg.s);
  double h[2];
  HIP_CHECK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, sg.s));
  HIP_CHECK(hipStreamSynchronize(sg.s));
  return h[1] == 0.0 ? 0.0 : h[0] / h[1];
}

// Observations that survive the filter (in front, |r| <= threshold) under
// the state held by `st`.
void residual_norms(const std::vector<sfmba_observation>& obs, const std::vector<uint64_t>& coi,
                    const std::vector<sfmba_camera>& cams, const std::vector<sfmba_pose>& pose_list,
                    const std::vector<double>& pts, int device, std::vector<double>& norm,
                    std::vector<uint8_t>& front) {
  const size_t n = obs.size();
  norm.assign(n, 0.0);
  front.assign(n, 0);
  if (n == 0) return;
  StreamGuard sg(device);
  DevArena A;
  std::vector<int32_t> model(cams.size());
  std::vector<double> params(8 * cams.size());
  for (size_t c = 0; c < cams.size(); ++c) {
    model[c] = cams[c].model_id;
    std::copy(cams[c].params, cams[c].params + 8, params.begin() + 8 * c);
  }
  std::vector<double> poses(7 * pose_list.size());
  for (size_t i = 0; i < pose_list.size(); ++i) {
    const sfmba_pose& q = pose_list[i];
    const double v[7] = {q.qw, q.qx, q.qy, q.qz, q.tx, q.ty, q.tz};
    std::copy(v, v + 7, poses.begin() + 7 * i);
  }
  const sfmba_observation* d_obs = A.upload(obs, sg.s);
  const uint64_t* d_coi = A.upload(coi, sg.s);
  const int32_t* d_model = A.upload(model, sg.s);
  const double* d_cams = A.upload(params, sg.s);
  const double* d_poses = A.upload(poses, sg.s);
  const double* d_pts = A.upload(pts, sg.s);
  double* d_norm = A.alloc<double>(n);
  uint8_t* d_front = A.alloc<uint8_t>(n);
  launch_residual_norms((uint32_t)n, d_obs, d_coi, d_model, d_poses, d_cams, d_pts, d_norm, d_front, sg.s);
  HIP_CHECK(hipMemcpyAsync(norm.data(), d_norm, n * sizeof(double), hipMemcpyDeviceToHost, sg.s));
  HIP_CHECK(hipMemcpyAsync(front.data(), d_front, n, hipMemcpyDeviceToHost, sg.s));
  HIP_CHECK(hipStreamSynchronize(sg.s));
}

static std::vector<uint8_t> filter_mask(const OwnedInput& st, double threshold, int device) {
  std::vector<double> norm;
  std::vector<uint8_t> keep;
  residual_norms(st.obs, st.coi, st.cams, st.poses, st.pts, device, norm, keep);
  for (size_t o = 0; o < keep.size(); ++o) keep[o] = keep[o] && norm[o] <= threshold;
  return keep;
}

// filter_and_reoptimize (lib.rs:1710-1768): solve, then up to two rounds of
// dropping observations behind their camera or with |r| > threshold and
// re-solving from the current state; stops when nothing was dropped or
// fewer than 12 observations remain.
static void filter_and_reoptimize(const sfmba_input* in, const sfmba_params* p, const sfmba_options* o,
                                  double threshold, sfmba_output* out, uint64_t* kept, uint64_t* rounds) {
  OwnedInput cur = own_input(in);
  solve_once(in, p, o, out);
  uint64_t nr = 0;
  for (int round = 0; round < 2; ++round) {
    OwnedInput st = cur;
    st.set_state(*out);
    const std::vector<uint8_t> keep = filter_mask(st, threshold, o->device);
    std::vector<sfmba_observation> surv;
    surv.reserve(st.obs.size());
    for (size_t k = 0; k < st.obs.size(); ++k)
      if (keep[k]) surv.push_back(st.obs[k]);
    if (surv.size() == st.obs.size() || surv.size() < 12) break;
    st.obs = std::move(surv);
    cur = std::move(st);
    const sfmba_input v = cur.view();
    solve_once(&v, p, o, out);
    ++nr;
  }
  if (kept) *kept = cur.obs.size();
  if (rounds) *rounds = nr;
}

// |distortion coefficient| < 0.5 for every camera (lib.rs:1674-1679, with
// CameraModel::opencv_distortion, camera.rs:192-200).
static bool distortion_plausible(const sfmba_camera* cams, size_t n) {
  for (size_t c = 0; c < n; ++c) {
    const double* q = cams[c].params;
    double d[4] = {0, 0, 0, 0};
    int nd = 0;
    switch (cams[c].model_id) {
      case SFMBA_SIMPLE_RADIAL: d[0] = q[3]; nd = 1; break;
      case SFMBA_RADIAL: d[0] = q[3]; d[1] = q[4]; nd = 2; break;
      case SFMBA_OPENCV:
      case SFMBA_OPENCV_FISHEYE: d[0] = q[4]; d[1] = q[5]; d[2] = q[6]; d[3] = q[7]; nd = 4; break;
      default: break;
    }
    for (int k = 0; k < nd; ++k)
      if (!(std::fabs(d[k]) < 0.5)) return false;
  }
  return true;
}

static void run_policy(const sfmba_input* in, const sfmba_policy* pol, const sfmba_options* o,
                       sfmba_output* out, sfmba_policy_report* rep) {
  validate(in);
  if (pol->intrinsics < SFMBA_INTRINSICS_FIXED || pol->intrinsics > SFMBA_INTRINSICS_FREE_GUARDED)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown intrinsics mode");
  sfmba_policy_report r{};
  OwnedInput base = own_input(in);
  OwnedOutput keep_state(base);
  // Nothing to adjust (lib.rs:1426-1428, :1472-1474): the state is returned unchanged.
  if (in->num_images < 2 || in->num_observations == 0) {
    std::copy(base.poses.begin(), base.poses.end(), out->poses);
    std::copy(base.pts.begin(), base.pts.end(), out->points);
    std::copy(base.cams.begin(), base.cams.end(), out->cameras);
    out->initial_cost = out->final_cost = 0.0;
    out->iterations_run = 0;
    if (rep) *rep = r;
    return;
  }
  const size_t nc = base.cams.size();
  // principal points pinned (lib.rs:1489)
  base.fcp.assign(8 * nc, 0);
  base.fcp_len.assign(nc, 0);
  {
    std::vector<uint8_t> mask(8 * nc);
    std::vector<size_t> lens(nc);
    for (size_t c = 0; c < nc; ++c) {
      const int m = base.cams[c].model_id;
      lens[c] = (size_t)model_param_count_host(m);
      const bool f_first = m == SFMBA_SIMPLE_PINHOLE || m == SFMBA_SIMPLE_RADIAL || m == SFMBA_RADIAL;
      mask[8 * c + (f_first ? 1 : 2)] = 1;
      mask[8 * c + (f_first ? 2 : 3)] = 1;
    }
    base.fcp.clear();
    for (size_t c = 0; c < nc; ++c) base.fcp.insert(base.fcp.end(), mask.begin() + 8 * c, mask.begin() + 8 * c + lens[c]);
    base.fcp_len = lens;
  }
  std::vector<size_t> per_cam(nc, 0);
  for (uint64_t c : base.coi) ++per_cam[c];
  auto caller_fixed = [&](size_t c) {
    return c < pol->num_camera_caller_fixed && pol->camera_caller_fixed[c] != 0;
  };
  sfmba_params p;
  sfmba_params_default(&p);
  p.robust_loss = pol->robust_loss;
  p.max_iterations = pol->intrinsics == SFMBA_INTRINSICS_FREE_GUARDED ? 200 : p.max_iterations;
  OwnedInput fixed_in = base, free_in = base;
  fixed_in.fc.assign(nc, 1);
  free_in.fc.assign(nc, 0);
  bool any_eligible = false;
  for (size_t c = 0; c < nc; ++c) {
    free_in.fc[c] = (per_cam[c] < 5 || caller_fixed(c)) ? 1 : 0;
    any_eligible = any_eligible || per_cam[c] >= 5;
  }
  if (pol->intrinsics == SFMBA_INTRINSICS_FIXED || pol->intrinsics == SFMBA_INTRINSICS_FREE) {
    const OwnedInput& use = pol->intrinsics == SFMBA_INTRINSICS_FIXED ? fixed_in : free_in;
    const sfmba_input v = use.view();
    solve_once(&v, &p, o, out);
    r.chose_free = pol->intrinsics == SFMBA_INTRINSICS_FREE;
    if (rep) *rep = r;
    return;
  }
  // FreeGuarded: both candidates, filtered, concurrently.
  OwnedOutput fixed_out(base), free_out(base);
  std::exception_ptr err;
  std::thread free_thread;
  if (any_eligible)
    free_thread = std::thread([&] {
      try {
        HIP_CHECK(hipSetDevice(o->device));
        const sfmba_input v = free_in.view();
        filter_and_reoptimize(&v, &p, o, pol->max_reprojection_error_px, &free_out.o, &r.observations_kept_free,
                              &r.filter_rounds_free);
      } catch (...) {
        err = std::current_exception();
      }
    });
  {
    const sfmba_input v = fixed_in.view();
    try {
      filter_and_reoptimize(&v, &p, o, pol->max_reprojection_error_px, &fixed_out.o, &r.observations_kept_fixed,
                            &r.filter_rounds_fixed);
    } catch (...) {
      if (free_thread.joinable()) free_thread.join();
      throw;
    }
  }
  if (free_thread.joinable()) free_thread.join();
  if (err) std::rethrow_exception(err);
  // Plain mean reprojection error over all of the problem's observations
  // (lib.rs:1542-1554).
  OwnedInput eval = base;
  eval.fp.clear();
  eval.fc.clear();
  eval.fcp.clear();
  eval.fcp_len.clear();
  eval.set_state(fixed_out.o);
  {
    const sfmba_input v = eval.view();
    r.fixed_mean_error_px = mean_error(&v, o->device);
  }
  const OwnedOutput* chosen = &fixed_out;
  if (any_eligible) {
    r.distortion_plausible = distortion_plausible(free_out.cams.data(), nc) ? 1 : 0;
    eval.set_state(free_out.o);
    const sfmba_input v = eval.view();
    r.free_mean_error_px = mean_error(&v, o->device);
    if (r.distortion_plausible && r.free_mean_error_px < r.fixed_mean_error_px) {
      chosen = &free_out;
      r.chose_free = 1;
    }
  }
  std::copy(chosen->poses.begin(), chosen->poses.end(), out->poses);
  std::copy(chosen->pts.begin(), chosen->pts.end(), out->points);
  std::copy(chosen->cams.begin(), chosen->cams.end(), out->cameras);
  out->initial_cost = chosen->o.initial_cost;
  out->final_cost = chosen->o.final_cost;
  out->iterations_run = chosen->o.iterations_run;
  if (rep) *rep = r;
}

}  // namespace sfmba

using namespace sfmba;

template <typename F>
static int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return SFMBA_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host out of memory");
    return SFMBA_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SFMBA_ERR_DEVICE;
  }
}

extern "C" {

void sfmba_params_default(sfmba_params* p) {
  if (!p) return;
  p->max_iterations = 50;
  p->initial_lambda = 1e-3;
  p->robust_loss = SFMBA_LOSS_HUBER;
  p->reserved = 0;
  p->loss_scale_px = 2.0;
}

void sfmba_options_default(sfmba_options* o) {
  if (!o) return;
  o->precision = SFMBA_PRECISION_FP64;
  o->solver = SFMBA_SOLVER_CHOLESKY;
  o->device = 0;
  o->profile = 0;
  o->refine_steps = 0;
  o->pcg_max_iterations = 500;
  o->pcg_tolerance = 1e-10;
}

void sfmba_intrinsics_params_default(sfmba_intrinsics_params* p) {
  if (!p) return;
  p->max_iterations = 30;
  p->initial_lambda = 1e-3;
  p->robust_loss = SFMBA_LOSS_HUBER;
  p->reserved = 0;
  p->loss_scale_px = 2.0;
}

const char* sfmba_last_error(void) { return g_last_error.c_str(); }

int sfmba_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int sfmba_bundle_adjust_ex(const sfmba_input* input, const sfmba_params* params,
                           const sfmba_options* options, sfmba_output* out, sfmba_stats* stats) {
  return guarded([&] {
    if (!params || !out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "params/output is NULL");
    sfmba_options o;
    sfmba_options_default(&o);
    if (options) o = *options;
    std::unique_ptr<sfmba_problem> H(create_problem(input, o.device));
    solve_problem(H.get(), params, &o, out, stats);
  });
}

int sfmba_problem_create(const sfmba_input* input, int32_t device, sfmba_problem** problem) {
  return guarded([&] {
    if (!problem) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "problem is NULL");
    *problem = nullptr;
    *problem = create_problem(input, device);
  });
}

int sfmba_problem_solve(sfmba_problem* problem, const sfmba_params* params, const sfmba_options* options,
                        sfmba_output* out, sfmba_stats* stats) {
  return guarded([&] { solve_problem(problem, params, options, out, stats); });
}

void sfmba_problem_destroy(sfmba_problem* problem) {
  if (!problem) return;
  (void)hipSetDevice(problem->device);
  delete problem;
}

int sfmba_bundle_adjust(const sfmba_input* input, const sfmba_params* params, sfmba_output* out) {
  return sfmba_bundle_adjust_ex(input, params, nullptr, out, nullptr);
}

int sfmba_total_cost(const sfmba_input* input, const sfmba_params* params, double* out) {
  return guarded([&] {
    validate(input);
    if (!params || !out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "params/out is NULL");
    StreamGuard sg(0);
    Problem pr;
    Pack P;
    double *po, *ca, *pt;
    upload_state(input, pr, P, sg.s, &po, &ca, &pt);
    double* part = pr.arena.alloc<double>(2 * ((size_t)pr.L.n_obs / (COST_BLOCK * COST_ITEMS) + 2));
    double* res = pr.arena.alloc<double>(2);
    launch_cost(pr.L, po, ca, pt, params->robust_loss, params->loss_scale_px, part, res, 0, sg.s);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipStreamSynchronize(sg.s));
    *out = h[0];
  });
}

int sfmba_mean_reprojection_error(const sfmba_input* input, double* out) {
  return guarded([&] {
    validate(input);
    if (!out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = mean_error(input, 0);
  });
}

int sfmba_filter_and_reoptimize(const sfmba_input* input, const sfmba_params* params,
                                const sfmba_options* options, double threshold_px, sfmba_output* out,
                                uint64_t* num_observations_kept) {
  return guarded([&] {
    if (!input || !params || !out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    validate(input);
    sfmba_options o;
    sfmba_options_default(&o);
    if (options) o = *options;
    uint64_t kept = 0, rounds = 0;
    filter_and_reoptimize(input, params, &o, threshold_px, out, &kept, &rounds);
    if (num_observations_kept) *num_observations_kept = kept;
  });
}

void sfmba_policy_default(sfmba_policy* p) {
  if (!p) return;
  *p = sfmba_policy{};
  p->intrinsics = SFMBA_INTRINSICS_FREE_GUARDED;
  p->robust_loss = SFMBA_LOSS_HUBER;
  p->max_reprojection_error_px = 4.0;
}

int sfmba_run_bundle_adjustment(const sfmba_input* input, const sfmba_policy* policy,
                                const sfmba_options* options, sfmba_output* out, sfmba_policy_report* report) {
  return guarded([&] {
    if (!input || !policy || !out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((input->num_images && !out->poses) || (input->num_points && !out->points) ||
        (input->num_cameras && !out->cameras))
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "output arrays are NULL");
    sfmba_options o;
    sfmba_options_default(&o);
    if (options) o = *options;
    run_policy(input, policy, &o, out, report);
  });
}

int sfmba_observation_jacobians(const sfmba_input* input, int32_t precision, int32_t device,
                                double* r, double* jp, double* jx, double* jc, uint8_t* valid) {
  return guarded([&] {
    validate(input);
    if (!r || !jp || !jx || !jc || !valid) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "output is NULL");
    StreamGuard sg(device);
    Problem pr;
    Pack P;
    // Keep the caller's observation order: pack without sorting by point.
    sfmba_input one = *input;
    double *po, *ca, *pt;
    upload_state(&one, pr, P, sg.s, &po, &ca, &pt);
    // Re-upload observations in input order.
    const size_t n = input->num_observations;
    std::vector<uint32_t> oi(n), op(n);
    std::vector<double2> oxy(n);
    for (size_t o = 0; o < n; ++o) {
      oi[o] = (uint32_t)input->observations[o].image_idx;
      op[o] = (uint32_t)input->observations[o].point_idx;
      oxy[o] = double2{input->observations[o].x, input->observations[o].y};
    }
    Layout L = pr.L;
    L.obs_img = pr.arena.upload(oi, sg.s);
    L.obs_pt = pr.arena.upload(op, sg.s);
    L.obs_xy = pr.arena.upload(oxy, sg.s);
    double* dr = pr.arena.alloc<double>(2 * n);
    double* djp = pr.arena.alloc<double>(12 * n);
    double* djx = pr.arena.alloc<double>(6 * n);
    double* djc = pr.arena.alloc<double>(16 * n);
    uint8_t* dv = pr.arena.alloc<uint8_t>(n);
    launch_jacobians_debug(L, po, ca, pt, precision == SFMBA_PRECISION_MIXED, dr, djp, djx, djc, dv, sg.s);
    HIP_CHECK(hipMemcpyAsync(r, dr, 2 * n * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(jp, djp, 12 * n * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(jx, djx, 6 * n * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(jc, djc, 16 * n * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(valid, dv, n, hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipStreamSynchronize(sg.s));
  });
}

int sfmba_dense_solve(const double* A, size_t n, const double* b, int32_t method, int32_t device,
                      double* x, int32_t* status) {
  return guarded([&] {
    if (!A || !b || !x || !status) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (method != 0 && method != 1) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown method");
    if (n == 0) {
      *status = 0;
      return;
    }
    if (n > 60000) throw Error(SFMBA_ERR_UNSUPPORTED, "matrix too large");
    StreamGuard sg(device);
    DevArena ar;
    const uint32_t D = (uint32_t)n, Dp = ((D + NB - 1) / NB) * NB, ld = Dp + NB;
    std::vector<double> h((size_t)ld * Dp, 0.0);
    for (uint32_t c = 0; c < D; ++c) {
      for (uint32_t r = c; r < D; ++r) h[(size_t)c * ld + r] = A[(size_t)c * n + r];
      h[(size_t)c * ld + Dp] = b[c];
    }
    for (uint32_t d = D; d < Dp; ++d) h[(size_t)d * ld + d] = 1.0;
    double* S = ar.upload(h, sg.s);
    double* Linv = ar.alloc<double>(dense_workspace_doubles(Dp));
    double* y = ar.alloc<double>(Dp);
    double* xd = ar.alloc<double>(Dp);
    int* flags = ar.alloc<int>(4);
    int* piv = ar.alloc<int>(Dp);
    uint32_t* chain = ar.alloc<uint32_t>(Dp / NB);
    HIP_CHECK(hipMemsetAsync(flags, 0, 4 * sizeof(int), sg.s));
    HIP_CHECK(hipMemsetAsync(chain, 0, (Dp / NB) * sizeof(uint32_t), sg.s));
    if (method == 0) {
      dense_cholesky(S, Linv, Dp, ld, D, flags, sg.s);
      dense_extract_row(S, Dp, ld, Dp, y, sg.s);
      dense_backsolve(S, Linv, Dp, ld, y, xd, chain, 1, flags + 2, sg.s);
    } else {
      dense_lu_solve(S, Dp, ld, D, xd, piv, flags, sg.s);
    }
    int hf[4];
    HIP_CHECK(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipMemcpyAsync(x, xd, n * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipStreamSynchronize(sg.s));
    if (hf[2] != 0) throw Error(SFMBA_ERR_DEVICE, "back substitution hand-over timed out");
    *status = method == 0 ? (hf[0] != 0) : (hf[1] != 0);
  });
}

int sfmba_default_fixed_params_mask(const sfmba_camera* cameras, size_t num_cameras, uint8_t* mask,
                                    size_t* lens) {
  return guarded([&] {
    if (num_cameras && (!cameras || !mask || !lens)) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (size_t c = 0; c < num_cameras; ++c) {
      const int m = cameras[c].model_id;
      const int k = model_param_count_host(m);
      if (k < 0) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown model id");
      lens[c] = (size_t)k;
      std::memset(mask + 8 * c, 0, 8);
      // principal-point indices (intrinsics.rs:33-42)
      const bool f_first = m == SFMBA_SIMPLE_PINHOLE || m == SFMBA_SIMPLE_RADIAL || m == SFMBA_RADIAL;
      mask[8 * c + (f_first ? 1 : 2)] = 1;
      mask[8 * c + (f_first ? 2 : 3)] = 1;
    }
  });
}

int sfmba_refine_intrinsics(sfmba_camera* cameras, size_t num_cameras,
                            const uint64_t* camera_of_image, size_t num_images,
                            const sfmba_pose* poses, const double* points, size_t num_points,
                            const sfmba_observation* observations, size_t num_observations,
                            const sfmba_intrinsics_params* params) {
  return guarded([&] {
    if (!params) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "params is NULL");
    if ((num_cameras && !cameras) || (num_images && (!camera_of_image || !poses)) ||
        (num_points && !points) || (num_observations && !observations))
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (size_t c = 0; c < num_cameras; ++c)
      if (model_param_count_host(cameras[c].model_id) < 0)
        throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown model id");
    // Group observations by camera, keeping their order (intrinsics.rs:125-128).
    std::vector<std::vector<uint32_t>> by_cam(num_cameras);
    for (size_t o = 0; o < num_observations; ++o) {
      const sfmba_observation& ob = observations[o];
      if (ob.image_idx >= num_images || ob.point_idx >= num_points)
        throw Error(SFMBA_ERR_INVALID_ARGUMENT, "observation " + std::to_string(o) + " out of range");
      const uint64_t c = camera_of_image[ob.image_idx];
      if (c >= num_cameras) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "camera_of_image out of range");
      by_cam[c].push_back((uint32_t)o);
    }
    std::vector<uint32_t> cam_list, cam_ptr(1, 0), oi, op;
    std::vector<double2> oxy;
    for (size_t c = 0; c < num_cameras; ++c) {
      if (by_cam[c].size() < 8) continue;  // left exactly as is (intrinsics.rs:135-140)
      cam_list.push_back((uint32_t)c);
      for (uint32_t o : by_cam[c]) {
        oi.push_back((uint32_t)observations[o].image_idx);
        op.push_back((uint32_t)observations[o].point_idx);
        oxy.push_back(double2{observations[o].x, observations[o].y});
      }
      cam_ptr.push_back((uint32_t)oi.size());
    }
    if (cam_list.empty()) return;
    StreamGuard sg(0);
    DevArena A;
    std::vector<double> hp(7 * num_images), hc(8 * num_cameras, 0.0), hx(points, points + 3 * num_points);
    std::vector<int32_t> hm(num_cameras);
    for (size_t i = 0; i < num_images; ++i) {
      const sfmba_pose& q = poses[i];
      const double v[7] = {q.qw, q.qx, q.qy, q.qz, q.tx, q.ty, q.tz};
      std::memcpy(&hp[7 * i], v, sizeof(v));
    }
    for (size_t c = 0; c < num_cameras; ++c) {
      hm[c] = cameras[c].model_id;
      const int k = model_param_count_host(hm[c]);
      for (int j = 0; j < k; ++j) hc[8 * c + j] = cameras[c].params[j];
    }
    double* dc = A.upload(hc, sg.s);
    launch_refine_intrinsics((uint32_t)cam_list.size(), A.upload(cam_list, sg.s), A.upload(cam_ptr, sg.s),
                             A.upload(oi, sg.s), A.upload(op, sg.s), A.upload(oxy, sg.s), A.upload(hp, sg.s),
                             A.upload(hx, sg.s), A.upload(hm, sg.s), dc, params, sg.s);
    HIP_CHECK(hipMemcpyAsync(hc.data(), dc, hc.size() * sizeof(double), hipMemcpyDeviceToHost, sg.s));
    HIP_CHECK(hipStreamSynchronize(sg.s));
    for (uint32_t c : cam_list) {
      const int k = model_param_count_host(cameras[c].model_id);
      for (int j = 0; j < k; ++j) cameras[c].params[j] = hc[8 * (size_t)c + j];
    }
  });
}

int sfmba_bundle_adjust_batch(const sfmba_input* inputs, size_t count, const sfmba_params* params,
                              const sfmba_options* options, sfmba_output* outputs,
                              sfmba_stats* stats) {
  return guarded([&] {
    if (count && (!inputs || !outputs || !params)) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    sfmba_options o;
    sfmba_options_default(&o);
    if (options) o = *options;
    if (o.precision != SFMBA_PRECISION_FP64 && o.precision != SFMBA_PRECISION_MIXED)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown precision");
    if (o.solver != SFMBA_SOLVER_CHOLESKY)
      throw Error(SFMBA_ERR_UNSUPPORTED, "batched windows are solved by Cholesky only");
    for (size_t i = 0; i < count; ++i) {
      const sfmba_output& out = outputs[i];
      if ((inputs[i].num_images && !out.poses) || (inputs[i].num_points && !out.points) ||
          (inputs[i].num_cameras && !out.cameras))
        throw Error(SFMBA_ERR_INVALID_ARGUMENT, "problem " + std::to_string(i) + ": output arrays are NULL");
      if (stats) {
        reset_stats(&stats[i]);
      }
    }
    if (o.precision == SFMBA_PRECISION_MIXED)
      batch_solve<float>(inputs, count, params, o, outputs, stats);
    else
      batch_solve<double>(inputs, count, params, o, outputs, stats);
  });
}

int sfmba_batch_create(const sfmba_input* inputs, size_t count, const sfmba_params* params,
                       const sfmba_options* options, int32_t autotune_iterations, sfmba_batch** batch) {
  return guarded([&] {
    if (!batch) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    *batch = nullptr;
    if ((count && !inputs) || !params) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (autotune_iterations < 0) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "autotune_iterations is negative");
    sfmba_options o;
    sfmba_options_default(&o);
    if (options) o = *options;
    if (o.precision != SFMBA_PRECISION_FP64 && o.precision != SFMBA_PRECISION_MIXED)
      throw Error(SFMBA_ERR_INVALID_ARGUMENT, "unknown precision");
    if (o.solver != SFMBA_SOLVER_CHOLESKY)
      throw Error(SFMBA_ERR_UNSUPPORTED, "batched windows are solved by Cholesky only");
    std::unique_ptr<sfmba_batch> Q(new sfmba_batch());
    {
      StreamGuard probe(o.device);  // device checks
    }
    Q->device = o.device;
    Q->precision = o.precision;
    Q->prm = *params;
    Q->opt = o;
    HIP_CHECK(hipSetDevice(o.device));
    HIP_CHECK(hipStreamCreateWithFlags(&Q->s, hipStreamNonBlocking));
    if (o.precision == SFMBA_PRECISION_MIXED)
      batch_plan_build(*Q, Q->b32, inputs, count, autotune_iterations);
    else
      batch_plan_build(*Q, Q->b64, inputs, count, autotune_iterations);
    *batch = Q.release();
  });
}

int sfmba_batch_solve(sfmba_batch* batch, sfmba_output* outputs, sfmba_stats* stats) {
  return guarded([&] {
    if (!batch) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t count = batch->precision == SFMBA_PRECISION_MIXED ? batch->b32.packs.size() : batch->b64.packs.size();
    if (count && !outputs) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (size_t i = 0; i < count; ++i) {
      const sfmba::Pack& P = batch->precision == SFMBA_PRECISION_MIXED ? batch->b32.packs[i] : batch->b64.packs[i];
      const sfmba_output& out = outputs[i];
      if ((P.n_img && !out.poses) || (P.n_pt && !out.points) || (P.n_cam && !out.cameras))
        throw Error(SFMBA_ERR_INVALID_ARGUMENT, "problem " + std::to_string(i) + ": output arrays are NULL");
      if (stats) {
        reset_stats(&stats[i]);
      }
    }
    if (batch->precision == SFMBA_PRECISION_MIXED)
      batch_plan_solve(*batch, batch->b32, outputs, stats);
    else
      batch_plan_solve(*batch, batch->b64, outputs, stats);
  });
}

int sfmba_batch_get_info(const sfmba_batch* batch, sfmba_batch_info* info) {
  return guarded([&] {
    if (!batch || !info) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
    *info = batch->info;
  });
}

void sfmba_batch_destroy(sfmba_batch* batch) { delete batch; }

}  // extern "C"

sfmba_batch::~sfmba_batch() {
  if (device >= 0) (void)hipSetDevice(device);
  if (s) (void)hipStreamSynchronize(s);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (h_out) (void)hipHostFree(h_out);
  if (h_res) (void)hipHostFree(h_res);
  if (h_oc) (void)hipHostFree(h_oc);
  big.clear();
  if (s) (void)hipStreamDestroy(s);
}

