"""COLMAP text model reader/writer (host code, no GPU): the native library's
sfmba_colmap_rewrite and the oracle restatement of
crates/sfm-io/src/colmap.rs against the golden text of the reference's own
sample model, against each other on synthetic models and hand-formatted
input, and on the parse errors the reference reports."""
import math
import shutil

import numpy as np
import pytest

import colmap_models as M
from oracle import colmap as OC
from sfmtory_amd import colmap, scenes

FILES = ("cameras.txt", "images.txt", "points3D.txt")


def texts(d):
    return {f: (d / f).read_bytes() for f in FILES}


def test_golden_sample_quaternion_digits():
    # The golden images.txt holds q / |q| of (0.98, 0.1, 0.05, 0.02) to 10
    # decimals; |q|^2 = 0.9733 exactly in decimal, so each component is
    # q_i / sqrt(0.9733): the digits are fixed by that closed form.
    line = (M.GOLDEN / "images.txt").read_text().splitlines()[4].split()
    got = [float(v) for v in line[1:5]]
    for g, q in zip(got, (0.98, 0.1, 0.05, 0.02)):
        assert abs(g - q / math.sqrt(0.9733)) <= 5e-11
    assert abs(sum(v * v for v in got) - 1.0) < 1e-9


def test_oracle_writer_reproduces_golden_sample(tmp_path):
    OC.write_model(M.sample_model(), tmp_path)
    assert texts(tmp_path) == texts(M.GOLDEN)


def test_native_rewrite_reproduces_golden_sample(tmp_path):
    src = M.write_texts(tmp_path / "in", M.SAMPLE_INPUT)
    counts = colmap.rewrite_model(src, tmp_path / "out")
    assert counts == (1, 1, 1)
    assert texts(tmp_path / "out") == texts(M.GOLDEN)


def test_round_trip_assertions_of_the_reference_test(tmp_path):
    # round_trips_through_text_files (colmap.rs:363-389), on the native writer
    src = M.write_texts(tmp_path / "in", M.SAMPLE_INPUT)
    colmap.rewrite_model(src, tmp_path / "out")
    m = OC.read_model(tmp_path / "out")
    assert len(m["cameras"]) == len(m["images"]) == len(m["points"]) == 1
    cam = m["cameras"][1]
    assert OC.NAMES[cam["model"]] == "SIMPLE_RADIAL" and cam["width"] == 1920
    img = m["images"][1]
    assert img["name"] == "img_0001.jpg" and len(img["kx"]) == 2 and img["pid"] == [1, None]
    p = m["points"][1]
    assert p["rgb"] == [255, 128, 0] and len(p["track"]) == 1 and abs(p["xyz"][0] - 1.0) < 1e-6


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_native_rewrite_matches_oracle_on_synthetic_models(tmp_path, seed):
    sc = scenes.generate("c0", num_images=12, num_points=300, seed=seed, shared_camera=seed % 2,
                         model_id=[2, 4, 5][seed - 1])
    OC.write_model(M.scene_model(sc.inp, seed), tmp_path / "in")
    colmap.rewrite_model(tmp_path / "in", tmp_path / "native")
    OC.write_model(OC.read_model(tmp_path / "in"), tmp_path / "oracle")
    assert texts(tmp_path / "native") == texts(tmp_path / "oracle")
    # a second pass is stable on both sides
    colmap.rewrite_model(tmp_path / "native", tmp_path / "native2")
    OC.write_model(OC.read_model(tmp_path / "native"), tmp_path / "oracle2")
    assert texts(tmp_path / "native2") == texts(tmp_path / "oracle2")


def test_reader_grammar_matches_oracle(tmp_path):
    t = {
        "cameras.txt": "7 PINHOLE 640 480 500 +501.5 320 2.4e2 99 98\n"
                       "3 OPENCV_FISHEYE 10 20 1 2 3 4 0.1 -0.2 1e-3 -0\n"
                       "4 RADIAL 1 1 1e400 1 1 nan -inf\n",
        "images.txt": "9 1 0 0 0 0 0 0 7 a.jpg extra tokens\n"
                      "1.0004 2.0005 5 3 4 -7 8 9\n"
                      "2 0 0 0 2 1 2 3 3 b.jpg\n"
                      "5 6\n",  # an incomplete triple: no keypoints
        "points3D.txt": "5 1 2 3 0 0 255 1e-7\n4 1.5 -2.5 3.5 1 2 3 0.5 9 0 2 1 9\n",
    }
    src = M.write_texts(tmp_path / "in", t)
    colmap.rewrite_model(src, tmp_path / "native")
    OC.write_model(OC.read_model(src), tmp_path / "oracle")
    assert texts(tmp_path / "native") == texts(tmp_path / "oracle")
    img = (tmp_path / "native" / "images.txt").read_text().splitlines()
    assert img[4].startswith("2 0.0000000000 0.0000000000 0.0000000000 1.0000000000 ")
    assert img[7] == "1.000 2.000 5 3.000 4.000 -1"  # negative ids print as -1
    assert img[5] == ""  # the empty POINTS2D line of image 2
    cams = (tmp_path / "native" / "cameras.txt").read_text().splitlines()
    assert cams[3] == "3 OPENCV_FISHEYE 10 20 1.0000000000 2.0000000000 3.0000000000 4.0000000000 " \
                      "0.1000000000 -0.2000000000 0.0010000000 -0.0000000000"
    assert cams[4] == "4 RADIAL 1 1 inf 1.0000000000 1.0000000000 NaN -inf"


BAD = [
    ("cameras.txt", "1 PINHOLE 640\n", "cameras.txt: malformed line"),
    ("cameras.txt", "x PINHOLE 640 480 1 1 1 1\n", "cameras.txt: bad CAMERA_ID"),
    ("cameras.txt", "-1 PINHOLE 640 480 1 1 1 1\n", "cameras.txt: bad CAMERA_ID"),
    ("cameras.txt", "1 PINHOLE 4294967296 480 1 1 1 1\n", "cameras.txt: bad WIDTH"),
    ("cameras.txt", "1 PINHOLE 640 480 1 1 0x10 1\n", "cameras.txt: bad PARAMS"),
    ("cameras.txt", "1 PINHOLE 640 480 1 1 1_0 1\n", "cameras.txt: bad PARAMS"),
    ("cameras.txt", "1 FOO 640 480 1 1 1 1\n", "unknown camera model FOO"),
    ("images.txt", "1 1 0 0 0 0 0 0 1\n\n", "images.txt: malformed header"),
    ("images.txt", "1 1 0 0 0 0 0 0 1 a.jpg\n", "images.txt: missing POINTS2D line"),
    ("images.txt", "1 1 0 0 q 0 0 0 1 a.jpg\n1 2 3\n", "images.txt: bad quaternion"),
    ("images.txt", "1 1 0 0 0 0 0 0 1 a.jpg\n1 2 3.5\n", "images.txt: bad POINT3D_ID"),
    # An image with no keypoints writes an empty POINTS2D line, which the
    # reader skips as blank: the next header is taken as the points line
    # (three triples, "b.jpg" left over) and its points line as a header.
    ("images.txt", "1 1 0 0 0 0 0 0 1 a.jpg\n\n2 1 0 0 0 0 0 0 1 b.jpg\n1 2 3\n",
     "images.txt: malformed header: 1 2 3"),
    ("points3D.txt", "1 1 2 3 0 0 0\n", "points3D.txt: malformed line"),
    ("points3D.txt", "1 1 2 3 0 0 256 0.5\n", "points3D.txt: bad B"),
    ("points3D.txt", "1 1 y 3 0 0 0 0.5\n", "points3D.txt: bad Y"),
    ("points3D.txt", "1 1 2 3 0 0 0 0.5 1 -2\n", "points3D.txt: bad TRACK POINT2D_IDX"),
]


@pytest.mark.parametrize("name,content,msg", BAD)
def test_parse_errors_match_oracle(tmp_path, name, content, msg):
    shutil.copytree(M.GOLDEN, tmp_path / "in")
    (tmp_path / "in" / name).write_text(content)
    with pytest.raises(ValueError) as native:
        colmap.rewrite_model(tmp_path / "in", tmp_path / "out")
    with pytest.raises(ValueError) as ref:
        OC.read_model(tmp_path / "in")
    assert msg in str(native.value) and msg in str(ref.value)


def test_missing_file_is_an_error(tmp_path):
    shutil.copytree(M.GOLDEN, tmp_path / "in")
    (tmp_path / "in" / "points3D.txt").unlink()
    with pytest.raises(ValueError, match="cannot read"):
        colmap.rewrite_model(tmp_path / "in", tmp_path / "out")


@pytest.mark.parametrize("case", ["one_image", "no_points", "no_observations"])
def test_refine_rejects_what_the_reference_rejects(tmp_path, case):
    # cmd_refine's checks (crates/sfm-cli/src/main.rs:957-1015) fail before
    # any device work.
    sc = scenes.generate("c0", num_images=4, num_points=40, seed=5)
    m = M.scene_model(sc.inp)
    if case == "one_image":
        first = min(m["images"])
        m["images"] = {first: m["images"][first]}
        msg = "at least 2 registered images"
    elif case == "no_points":
        m["points"] = {}
        msg = "no 3D points"
    else:
        for p in m["points"].values():
            p["track"] = [(999_999, 0)]
        msg = "no observations"
    OC.write_model(m, tmp_path / "in")
    with pytest.raises(ValueError, match=msg):
        colmap.refine_model(tmp_path / "in", tmp_path / "out")
    with pytest.raises(ValueError, match=msg):
        OC.refine_problem(OC.read_model(tmp_path / "in"), False)
    assert not (tmp_path / "out").exists()


def test_scene_model_keypoints_are_single_precision(tmp_path):
    sc = scenes.generate("c0", num_images=3, num_points=30, seed=6)
    OC.write_model(M.scene_model(sc.inp), tmp_path / "in")
    m = OC.read_model(tmp_path / "in")
    inp, *_ = OC.refine_problem(m, False)
    x = inp.observations["x"]
    assert np.array_equal(x, x.astype(np.float32).astype(np.float64))
