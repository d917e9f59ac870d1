"""Problem assembly of a scoped bundle adjustment (sfmba_window_create,
host code, no GPU) against the oracle restatement of
crates/sfm-reconstruction/src/lib.rs:1339-1474: the same images, points,
cameras, held poses and observations, bit for bit, for global and local
scope, ties in the co-visibility ranking, the gauge fallback, pinned poses,
empty windows and the write-back."""
import numpy as np
import pytest

import recon_models as R
from oracle import scope as OS
from sfmtory_amd import ba, scenes
from sfmtory_amd.scope import Reconstruction, Window


def same_window(rec, scope, center):
    ref = OS.build_window(rec, scope, center)
    with Window(rec, scope, center) as w:
        if ref is None:
            assert w.empty and len(w.input.observations) == 0
            return None
        assert w.image_indices.tolist() == ref["image_ids"]
        assert w.point_indices.tolist() == ref["point_ids"]
        a, b = w.input, ref["input"]
        assert np.array_equal(a.camera_of_image, b.camera_of_image)
        assert a.cameras.tobytes() == b.cameras.tobytes()
        assert np.array_equal(a.poses, b.poses) and np.array_equal(a.points, b.points)
        assert a.observations.tobytes() == b.observations.tobytes()
        assert a.fixed_poses.tolist() == list(b.fixed_poses)
        assert a.fixed_cameras.tolist() == list(b.fixed_cameras)
        assert a.fixed_camera_params == b.fixed_camera_params
        return ref


@pytest.fixture(scope="module")
def rec():
    sc = scenes.generate("c0", num_images=24, num_points=900, seed=41, shared_camera=0)
    return R.from_input(sc.inp, seed=1, registered_frac=0.75, pinned=(5,))


def test_global_window_matches_oracle(rec):
    ref = same_window(rec, "global", 0)
    assert len(ref["image_ids"]) == int(rec.registered.sum())
    # only the seed (image 0) and the pinned image 5 (if registered) are held
    held = [orig for orig, f in zip(ref["image_ids"], ref["input"].fixed_poses) if f]
    assert held == [i for i in (0, 5) if rec.registered[i]]


def test_every_local_window_matches_oracle(rec):
    sizes = []
    for center in range(len(rec.registered)):
        ref = same_window(rec, "local", center)
        if ref is not None:
            free = [o for o, f in zip(ref["image_ids"], ref["input"].fixed_poses) if not f]
            assert len(free) <= 1 + 6
            sizes.append(len(ref["image_ids"]))
    assert len(sizes) >= 10 and max(sizes) > 7  # windows carry held neighbours


def test_covisibility_ties_break_by_lower_index():
    # Center 0 shares exactly one point with each of images 1..9: every
    # co-visibility count ties, so the free images are 1..6 (lib.rs:1371).
    ni = 10
    poses = np.tile([1.0, 0, 0, 0, 0, 0, 5.0], (ni, 1))
    tracks = [[(0, p), (p + 1, 0)] for p in range(ni - 1)]
    rec = Reconstruction(
        registered=np.ones(ni, np.uint8), poses=poses, image_camera_id=np.full(ni, 3, np.uint32),
        keypoint_offsets=np.array([0, 9] + [9 + k for k in range(1, ni)], np.uint64),
        keypoints=np.arange(2 * (9 + ni - 1), dtype=np.float32).reshape(-1, 2),
        points=np.zeros((ni - 1, 3)), track_offsets=np.arange(0, 2 * ni - 1, 2, dtype=np.uint64),
        track_images=np.array([e[0] for t in tracks for e in t], np.uint32),
        track_keypoints=np.array([e[1] for t in tracks for e in t], np.uint32),
        camera_ids=np.array([3], np.uint32), cameras=ba.cameras_to_array([ba.CameraModel.pinhole(500, 500, 0, 0)]),
        seed_image=0)
    ref = same_window(rec, "local", 0)
    free = [o for o, f in zip(ref["image_ids"], ref["input"].fixed_poses) if not f]
    assert free == [1, 2, 3, 4, 5, 6]  # 0 is the seed: held
    assert ref["image_ids"] == list(range(ni))


def test_gauge_fallback_holds_the_first_image():
    # The seed is not in the window and every image is free: image 0 of the
    # window is held (lib.rs:1453-1455).
    sc = scenes.generate("c0", num_images=6, num_points=200, seed=42)
    rec = R.from_input(sc.inp, seed=2, registered_frac=1.0, short_every=0, seed_image=99)
    ref = same_window(rec, "local", 3)
    assert list(ref["input"].fixed_poses) == [True] + [False] * (len(ref["image_ids"]) - 1)


def test_pinned_poses_are_held_in_both_scopes():
    sc = scenes.generate("c0", num_images=8, num_points=300, seed=43)
    rec = R.from_input(sc.inp, seed=3, registered_frac=1.0, pinned=(2, 4))
    for scope, center in (("global", 0), ("local", 2)):
        ref = same_window(rec, scope, center)
        fixed = dict(zip(ref["image_ids"], ref["input"].fixed_poses))
        assert fixed[2] and fixed[4]


@pytest.mark.parametrize("case", ["unregistered_world", "center_without_points", "one_image"])
def test_empty_windows(case):
    sc = scenes.generate("c0", num_images=6, num_points=120, seed=44)
    rec = R.from_input(sc.inp, seed=4, registered_frac=1.0)
    center = 2
    if case == "unregistered_world":
        rec.registered[:] = 0
    elif case == "center_without_points":
        center = 1_000  # no track contains it
    else:
        rec.registered[:] = 0
        rec.registered[2] = 1
    assert same_window(rec, "local", center) is None
    if case != "center_without_points":
        assert same_window(rec, "global", 0) is None


def test_scatter_writes_back_only_the_window(rec):
    r = R.copy(rec)
    with Window(r, "local", 7) as w:
        assert not w.empty
        n_i, n_p, n_c = len(w.image_indices), len(w.point_indices), len(w.input.cameras)
        poses = w.input.poses + 1.0
        pts = w.input.points * 2.0
        cams = w.input.cameras.copy()
        cams["params"] += 0.5
        w.scatter(ba.BaOutput(poses, pts, cams, 0.0, 0.0, 0))
        idx_i, idx_p = w.image_indices.astype(np.int64), w.point_indices.astype(np.int64)
    assert np.array_equal(r.poses[idx_i], poses) and np.array_equal(r.points[idx_p], pts)
    other_i = np.setdiff1d(np.arange(len(r.poses)), idx_i)
    other_p = np.setdiff1d(np.arange(len(r.points)), idx_p)
    assert np.array_equal(r.poses[other_i], rec.poses[other_i])
    assert np.array_equal(r.points[other_p], rec.points[other_p])
    changed = np.nonzero(np.any(r.cameras["params"] != rec.cameras["params"], axis=1))[0]
    assert len(changed) == n_c and n_i >= 2 and n_p > 0


def test_malformed_states_are_rejected(rec):
    r = R.copy(rec)
    r.track_images[3] = len(r.registered)
    with pytest.raises(IndexError, match="track image index out of range"):
        Window(r, "global", 0)
    # a keypoint index is only looked up for observations the window uses
    # (keypoint_px, lib.rs:1463): point 0's single-element track is not one
    r = R.copy(rec)
    r.track_keypoints[0] = 10_000_000
    assert not Window(r, "global", 0).empty
    r.track_keypoints[:] = 10_000_000
    with pytest.raises(IndexError, match="keypoint index out of range"):
        Window(r, "global", 0)
    r = R.copy(rec)
    r.image_camera_id[0] = 1
    with pytest.raises(ValueError, match="unknown camera id"):
        Window(r, "global", 0)
    with pytest.raises(KeyError):
        Window(rec, "regional", 0)
