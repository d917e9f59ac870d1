"""COLMAP text models for the tests: the reference's own sample model
(crates/sfm-io/src/colmap.rs:331-361), synthetic scenes turned into models
with sparse ids, untracked keypoints and dangling track elements, and
hand-formatted input text exercising the reader's grammar."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from sfmtory_amd import ba

GOLDEN = Path(__file__).resolve().parent / "golden" / "colmap_sample"


def sample_model() -> dict:
    """sample_recon (colmap.rs:331-361): one SIMPLE_RADIAL camera, one image
    with two keypoints (one tracked), one point."""
    q = [0.98, 0.1, 0.05, 0.02]
    n = (((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]) + q[0] * q[0]) ** 0.5
    return dict(
        cameras={1: dict(model=2, width=1920, height=1080, params=[1600.0, 960.0, 540.0, -0.05])},
        images={1: dict(camera_id=1, q=[c / n for c in q], t=[0.1, -0.2, 1.5], name="img_0001.jpg",
                        kx=[np.float32(100.5), np.float32(300.0)], ky=[np.float32(200.25), np.float32(400.0)],
                        pid=[1, None])},
        points={1: dict(xyz=[1.0, 2.0, 3.0], rgb=[255, 128, 0], error=0.42, track=[(1, 0)])},
    )


# The sample model as a user might type it: raw quaternion, short numbers,
# comments and blank lines in odd places, CRLF line ends in one file.
SAMPLE_INPUT = {
    "cameras.txt": "# hand written\n\n1 SIMPLE_RADIAL 1920 1080 1600 960 540 -0.05\n",
    "images.txt": "\n# header\n1 0.98 0.1 0.05 0.02 0.1 -0.2 1.5 1 img_0001.jpg\n\n   # between\n"
                  "100.5 200.25 1 300 400.0 -1\n",
    "points3D.txt": "# points\r\n1 1 2 3 255 128 0 0.42 1 0\r\n",
}


def write_texts(d: Path, texts: dict) -> Path:
    d.mkdir(parents=True, exist_ok=True)
    for name, text in texts.items():
        (d / name).write_bytes(text.encode())
    return d


def scene_model(inp: ba.BaInput, seed: int = 0, extra_keypoints: int = 3) -> dict:
    """A COLMAP model of a synthetic BA problem: ids sparse and out of
    insertion order, keypoints in single precision, a few untracked keypoints
    per image, and per point one dangling track element (an unknown image or
    a keypoint index past the end) that the refine pass must skip."""
    rng = np.random.default_rng(seed)
    ni, nc, npt = len(inp.poses), len(inp.cameras), len(inp.points)
    cam_id = [10 + 3 * c for c in range(nc)][::-1]
    img_id = [2 + 5 * i for i in range(ni)]
    pt_id = [7 + 2 * k for k in range(npt)]
    cams = {}
    for c in range(nc):
        m = int(inp.cameras[c]["model_id"])
        cams[cam_id[c]] = dict(model=m, width=1024, height=768,
                               params=[float(v) for v in inp.cameras[c]["params"][: ba.MODEL_PARAM_COUNT[m]]])
    imgs = {}
    for i in range(ni):
        imgs[img_id[i]] = dict(camera_id=cam_id[int(inp.camera_of_image[i])], q=list(map(float, inp.poses[i][:4])),
                               t=list(map(float, inp.poses[i][4:])), name=f"frame_{i:05d}.png", kx=[], ky=[], pid=[])
    tracks = {k: [] for k in range(npt)}
    for o in inp.observations:
        i, k = int(o["image_idx"]), int(o["point_idx"])
        im = imgs[img_id[i]]
        tracks[k].append((img_id[i], len(im["kx"])))
        im["kx"].append(np.float32(o["x"]))
        im["ky"].append(np.float32(o["y"]))
        im["pid"].append(pt_id[k])
    for im in imgs.values():
        for _ in range(extra_keypoints):
            im["kx"].append(np.float32(rng.uniform(0, 1024)))
            im["ky"].append(np.float32(rng.uniform(0, 768)))
            im["pid"].append(None)
    pts = {}
    for k in range(npt):
        tr = list(tracks[k])
        if k % 3 == 0:
            tr.insert(int(rng.integers(0, len(tr) + 1)), (999_999, 0))
        elif k % 3 == 1:
            tr.append((img_id[int(rng.integers(0, ni))], 1_000_000))
        pts[pt_id[k]] = dict(xyz=list(map(float, inp.points[k])), rgb=[int(v) for v in rng.integers(0, 256, 3)],
                             error=float(rng.uniform(0.1, 3.0)), track=tr)
    return dict(cameras=cams, images=imgs, points=pts)
