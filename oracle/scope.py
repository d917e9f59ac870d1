"""ORACLE - test infrastructure only.  Pure-Python restatement of the
problem assembly of run_bundle_adjustment
(crates/sfm-reconstruction/src/lib.rs:1339-1474: BaScope::Global / Local,
the co-visibility window of LOCAL_BA_NUM_IMAGES = 6, held poses, gauge
fallback, cameras in order of first use, observations from tracks) and its
write-back (lib.rs:1697-1707), driving the oracle's solve policy.

The reconstruction is the Python ``sfmtory_amd.scope.Reconstruction``
(flat arrays); only its fields are read here.
"""
from __future__ import annotations

import numpy as np

import oracle
from oracle.policy import oracle_policy
from sfmtory_amd import _native as N
from sfmtory_amd import ba

LOCAL_BA_NUM_IMAGES = 6


def _tracks(rec):
    off = rec.track_offsets
    return [list(zip(rec.track_images[off[p]:off[p + 1]].tolist(), rec.track_keypoints[off[p]:off[p + 1]].tolist()))
            for p in range(len(rec.points))]


def build_window(rec, scope: str, center: int = 0):
    """Returns None when the reference returns early (fewer than 2 images or
    no observation), else a dict with the compact problem and the index
    maps."""
    tracks = _tracks(rec)
    registered = [bool(v) for v in rec.registered]
    n = len(registered)
    if scope == "global":
        point_ids = [p for p in range(len(tracks)) if len(tracks[p]) >= 2]
        free = None
    else:
        point_ids = [p for p in range(len(tracks))
                     if len(tracks[p]) >= 2 and any(img == center for img, _ in tracks[p])]
        covis = {}
        for p in point_ids:
            for img, _ in tracks[p]:
                if img != center and registered[img]:
                    covis[img] = covis.get(img, 0) + 1
        ranked = sorted(covis.items(), key=lambda kv: (-kv[1], kv[0]))
        free = {img for img, _ in ranked[:LOCAL_BA_NUM_IMAGES]}
        free.add(center)
    if free is None:
        image_ids = [i for i in range(n) if registered[i]]
    else:
        seen = [False] * n
        for p in point_ids:
            for img, _ in tracks[p]:
                if registered[img]:
                    seen[img] = True
        image_ids = [i for i in range(n) if seen[i]]
    image_pos = {orig: k for k, orig in enumerate(image_ids)}
    if len(image_ids) < 2:
        return None
    cam_entry_of_id = {int(cid): e for e, cid in enumerate(rec.camera_ids)}
    cam_compact, camera_entry, coi = {}, [], []
    for orig in image_ids:
        cid = int(rec.image_camera_id[orig])
        if cid not in cam_compact:
            cam_compact[cid] = len(camera_entry)
            camera_entry.append(cam_entry_of_id[cid])
        coi.append(cam_compact[cid])
    if free is None:
        fixed = [orig == rec.seed_image for orig in image_ids]
    else:
        fixed = [orig == rec.seed_image or orig not in free for orig in image_ids]
    if rec.pose_pinned is not None:
        fixed = [f or bool(rec.pose_pinned[orig]) for f, orig in zip(fixed, image_ids)]
    if not any(fixed):
        fixed[0] = True
    obs = []
    for cp, p in enumerate(point_ids):
        for img, kp in tracks[p]:
            if img not in image_pos:
                continue
            xy = rec.keypoints[int(rec.keypoint_offsets[img]) + kp]
            obs.append((image_pos[img], cp, float(xy[0]), float(xy[1])))
    if not obs:
        return None
    cams = rec.cameras[camera_entry].copy()
    caller = set(int(v) for v in rec.caller_fixed_camera_ids)
    return dict(
        image_ids=image_ids, point_ids=point_ids, camera_entry=camera_entry,
        caller_fixed=[int(rec.camera_ids[e]) in caller for e in camera_entry],
        input=ba.BaInput(np.array(coi, dtype=np.uint64), cams, rec.poses[image_ids].copy(),
                         rec.points[point_ids].copy().reshape(-1, 3), np.array(obs, dtype=N.OBSERVATION_DTYPE),
                         fixed_poses=np.array(fixed, dtype=bool), fixed_cameras=np.ones(len(camera_entry), dtype=bool),
                         fixed_camera_params=oracle.default_fixed_params_mask(cams)),
    )


def scatter(rec, win, out) -> None:
    """Write-back (lib.rs:1697-1707)."""
    for k, orig in enumerate(win["image_ids"]):
        rec.poses[orig] = out["poses"][k]
    for k, orig in enumerate(win["point_ids"]):
        rec.points[orig] = out["points"][k]
    for k, e in enumerate(win["camera_entry"]):
        rec.cameras[e] = out["cameras"][k]


def bundle_adjust_scope(rec, scope: str, center: int = 0, intrinsics: str = "fixed",
                        loss=ba.RobustLoss.HUBER, thr: float = 4.0):
    """run_bundle_adjustment (lib.rs:1323-1708) on the oracle, in place.
    Returns (window or None, solve output or None, policy report or None)."""
    win = build_window(rec, scope, center)
    if win is None:
        return None, None, None
    out, rep = oracle_policy(win["input"], intrinsics, thr=thr, loss=loss, caller_fixed=win["caller_fixed"])
    scatter(rec, win, out)
    return win, out, rep
