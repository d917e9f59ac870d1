import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def _have_gpu() -> bool:
    try:
        from sfmtory_amd import ba

        return ba.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def c0_scene():
    from sfmtory_amd import scenes

    return scenes.generate("c0")


def pose_row_from_euler(roll, pitch, yaw, t):
    """nalgebra UnitQuaternion::from_euler_angles(roll, pitch, yaw):
    q = q_z(yaw) q_y(pitch) q_x(roll), as a (w x y z tx ty tz) row."""
    def qm(a, b):
        w1, x1, y1, z1 = a
        w2, x2, y2, z2 = b
        return (w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2)
    qx = (np.cos(roll / 2), np.sin(roll / 2), 0.0, 0.0)
    qy = (np.cos(pitch / 2), 0.0, np.sin(pitch / 2), 0.0)
    qz = (np.cos(yaw / 2), 0.0, 0.0, np.sin(yaw / 2))
    q = qm(qz, qm(qy, qx))
    return np.array([*q, *t], dtype=np.float64)
