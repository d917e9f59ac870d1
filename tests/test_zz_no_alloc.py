"""sfmba_total_cost_noalloc (no device allocation: partial sums in the
library's static device array) against sfmba_total_cost on the same state,
through the C ABI: two calls in a row, the second with a smaller input, on
the null stream and on a caller's stream; the size limit and out-of-range
indices are rejected.  The test allocates the input buffers itself with the
HIP runtime; the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from sfmtory_amd import _native as N
from sfmtory_amd import ba, scenes

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs (device buffers for the
    inputs, a stream), from the runtime the engine links."""

    def __init__(self):
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        self.h = h

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value


HIP = None


class DeviceInput:
    """An sfmba_input whose arrays are device copies of a BaInput's (freed
    with this object)."""

    def __init__(self, inp):
        global HIP
        HIP = HIP or Hip()
        self.host = ba._CInput(inp)
        self.ptrs = []

        def dev(a):
            if a.size == 0:
                return None
            p = HIP.upload(a)
            self.ptrs.append(p)
            return p

        h = self.host
        s = N.Input()
        s.num_images = h.c.num_images
        s.num_cameras = h.c.num_cameras
        s.num_points = h.c.num_points
        s.num_observations = h.c.num_observations
        s.camera_of_image = dev(h.coi)
        s.num_camera_of_image = h.c.num_camera_of_image
        s.cameras = dev(h.cams)
        s.poses = dev(h.poses)
        s.points = dev(h.points)
        s.observations = dev(h.obs)
        self.c = s

    def __del__(self):
        if HIP is not None:
            HIP.h.hipDeviceSynchronize()
            for p in self.ptrs:
                HIP.h.hipFree(p)


def noalloc_cost(dinp, params, stream=None):
    lib = N.load_engine()
    out = C.c_double(-1.0)
    rc = lib.sfmba_total_cost_noalloc(C.byref(dinp.c), C.byref(ba._c_params(params)), stream, C.byref(out))
    return rc, out.value


def subset(inp, n_img, n_pt):
    """The observations of the first n_pt points seen by the first n_img images."""
    obs = inp.observations
    keep = (obs["image_idx"] < n_img) & (obs["point_idx"] < n_pt)
    return ba.BaInput(inp.camera_of_image[:n_img], inp.cameras[:n_img], inp.poses[:n_img],
                      inp.points[:n_pt], obs[keep].copy())


@pytest.mark.parametrize("loss", [ba.RobustLoss.HUBER, ba.RobustLoss.CAUCHY, ba.RobustLoss.NONE])
def test_noalloc_cost_matches_total_cost_two_calls(loss):
    sc = scenes.generate("c2", num_images=40, num_points=6000, seed=scenes.BASE_SEED + 700)
    p = ba.BaParams(robust_loss=loss)
    small = subset(sc.inp, 25, 2500)
    assert 0 < len(small.observations) < len(sc.inp.observations) // 2
    big = DeviceInput(sc.inp)  # the first call's inputs stay allocated during the second
    stream = C.c_void_p()
    assert HIP.h.hipStreamCreate(C.byref(stream)) == 0
    try:
        for d, inp, st in ((big, sc.inp, None), (DeviceInput(small), small, stream.value)):
            rc, got = noalloc_cost(d, p, st)
            assert rc == N.SFMBA_OK, N.load_engine().sfmba_last_error()
            assert got == pytest.approx(ba.total_cost(inp, p), rel=1e-12)
    finally:
        HIP.h.hipStreamDestroy(stream)


def test_noalloc_cost_edge_cases():
    sc = scenes.generate("c0", num_points=300, seed=31)
    p = ba.BaParams()
    # no observations: cost 0
    empty = ba.BaInput(sc.inp.camera_of_image, sc.inp.cameras, sc.inp.poses, sc.inp.points,
                       sc.inp.observations[:0])
    rc, got = noalloc_cost(DeviceInput(empty), p)
    assert rc == N.SFMBA_OK and got == 0.0
    # an observation behind its camera costs nothing (lib.rs:733-735): move
    # the first observation's point 5 units behind its camera
    inp = ba.BaInput(sc.inp.camera_of_image, sc.inp.cameras, sc.inp.poses.copy(), sc.inp.points.copy(),
                     sc.inp.observations)
    o0 = inp.observations[0]
    w, x, y, z, *t = np.asarray(inp.poses)[int(o0["image_idx"])]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    np.asarray(inp.points)[int(o0["point_idx"])] = -R.T @ np.asarray(t) - 5.0 * R.T[:, 2]
    rc, got = noalloc_cost(DeviceInput(inp), p)
    assert rc == N.SFMBA_OK and got == pytest.approx(ba.total_cost(inp, p), rel=1e-12)
    import oracle
    assert not oracle.residual_norms(inp)[1][0]  # the case is exercised: behind its camera
    # an out-of-range point index is rejected, as sfmba_total_cost does
    bad_obs = sc.inp.observations.copy()
    bad_obs[5]["point_idx"] = 10_000
    bad = ba.BaInput(sc.inp.camera_of_image, sc.inp.cameras, sc.inp.poses, sc.inp.points, bad_obs)
    rc, _ = noalloc_cost(DeviceInput(bad), p)
    assert rc == N.SFMBA_ERR_INVALID_ARGUMENT
    assert b"out of range" in N.load_engine().sfmba_last_error()
    # and the next good call is unaffected by the flag it left
    rc, got = noalloc_cost(DeviceInput(sc.inp), p)
    assert rc == N.SFMBA_OK and got == pytest.approx(ba.total_cost(sc.inp, p), rel=1e-12)


def test_noalloc_rejects_sizes_beyond_the_limit():
    # checked on the host before anything is launched
    d = N.Input()
    d.num_observations = N.NOALLOC_MAX_OBSERVATIONS + 1
    d.num_images = d.num_points = d.num_cameras = d.num_camera_of_image = 1
    d.observations = d.poses = d.points = d.cameras = d.camera_of_image = 8  # never dereferenced
    out = C.c_double()
    rc = N.load_engine().sfmba_total_cost_noalloc(C.byref(d), C.byref(ba._c_params(ba.BaParams())), None,
                                                  C.byref(out))
    assert rc == N.SFMBA_ERR_INVALID_ARGUMENT
    assert b"no-allocation limit" in N.load_engine().sfmba_last_error()
