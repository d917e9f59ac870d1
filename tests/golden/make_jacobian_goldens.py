"""Write tests/golden/jacobians.json: residual and Jacobians of one
observation for each of the six camera models, derived symbolically with
sympy from the projection formulas (crates/sfm-core/src/camera.rs:204-244)
and the world-to-camera transform (crates/sfm-core/src/pose.rs:46-48), and
evaluated with 40-digit arithmetic.

The pose perturbation is the reference's left-multiplicative one
(crates/sfm-ba/src/lib.rs:263-269): R(omega) = exp([omega]x) R0, t + dt.  The
derivative at omega = 0 only needs exp([omega]x) to first order.

Inputs are the reference's own Jacobian test case (lib.rs:1252-1363): pose
from_euler_angles(0.15, -0.3, 0.08) with t = (0.6, -0.2, 1.3), point
(0.35, -0.18, 2.7), observation (410, 260), and the six test cameras.
nalgebra's from_euler_angles(roll, pitch, yaw) applies roll, then pitch, then
yaw: q = q_z(yaw) q_y(pitch) q_x(roll).

Run: python tests/golden/make_jacobian_goldens.py
"""
import json
from pathlib import Path

import sympy as sp

DIG = 40

CAMERAS = [
    ("SIMPLE_PINHOLE", 0, [1420.0, 320.0, 240.0]),
    ("PINHOLE", 1, [1420.0, 1390.0, 320.0, 240.0]),
    ("SIMPLE_RADIAL", 2, [1420.0, 320.0, 240.0, -0.45]),
    ("RADIAL", 3, [1420.0, 320.0, 240.0, -0.35, 0.12]),
    ("OPENCV", 4, [1420.0, 1390.0, 320.0, 240.0, -0.28, 0.09, 0.015, -0.02]),
    ("OPENCV_FISHEYE", 5, [900.0, 880.0, 320.0, 240.0, -0.05, 0.02, -0.01, 0.005]),
]


def quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return (w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
            w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
            w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2)


def euler_quat(roll, pitch, yaw):
    h = sp.Rational(1, 2)
    qx = (sp.cos(roll * h), sp.sin(roll * h), 0, 0)
    qy = (sp.cos(pitch * h), 0, sp.sin(pitch * h), 0)
    qz = (sp.cos(yaw * h), 0, 0, sp.sin(yaw * h))
    return quat_mul(qz, quat_mul(qy, qx))


def rot(q):
    w, x, y, z = q
    return sp.Matrix([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def project(model, P, pc):
    x = pc[0] / pc[2]
    y = pc[1] / pc[2]
    if model == 0:
        return P[0] * x + P[1], P[0] * y + P[2]
    if model == 1:
        return P[0] * x + P[2], P[1] * y + P[3]
    if model == 2:
        d = 1 + P[3] * (x * x + y * y)
        return P[0] * x * d + P[1], P[0] * y * d + P[2]
    if model == 3:
        r2 = x * x + y * y
        d = 1 + P[3] * r2 + P[4] * r2 * r2
        return P[0] * x * d + P[1], P[0] * y * d + P[2]
    if model == 4:
        k1, k2, p1, p2 = P[4:8]
        r2 = x * x + y * y
        rad = 1 + k1 * r2 + k2 * r2 * r2
        xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
        yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
        return P[0] * xd + P[2], P[1] * yd + P[3]
    k1, k2, k3, k4 = P[4:8]
    r = sp.sqrt(x * x + y * y)
    th = sp.atan(r)
    t2 = th * th
    td = th * (1 + k1 * t2 + k2 * t2 ** 2 + k3 * t2 ** 3 + k4 * t2 ** 4)
    return P[0] * x * td / r + P[2], P[1] * y * td / r + P[3]


def main():
    q0 = euler_quat(sp.Float("0.15", DIG), sp.Float("-0.3", DIG), sp.Float("0.08", DIG))
    q0 = tuple(sp.N(v, DIG) for v in q0)
    R0 = rot(q0)
    t0 = [sp.Float(v, DIG) for v in ("0.6", "-0.2", "1.3")]
    X0 = [sp.Float(v, DIG) for v in ("0.35", "-0.18", "2.7")]
    obs = (sp.Float("410", DIG), sp.Float("260", DIG))
    w = sp.symbols("w0:3")
    dt = sp.symbols("dt0:3")
    X = sp.symbols("X0:3")
    W = sp.Matrix([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    Rw = (sp.eye(3) + W) * R0  # first order in omega suffices at omega = 0
    out = {"pose_wxyz": [float(v) for v in q0], "t": [float(v) for v in t0], "point": [float(v) for v in X0],
           "obs": [410.0, 260.0], "euler_roll_pitch_yaw": [0.15, -0.3, 0.08], "cameras": []}
    for name, model, params in CAMERAS:
        Ps = sp.symbols(f"p0:{len(params)}")
        pc = Rw * sp.Matrix(X) + sp.Matrix([t0[i] + dt[i] for i in range(3)])
        u, v = project(model, Ps, pc)
        res = [u - obs[0], v - obs[1]]
        subs = {**{w[i]: 0 for i in range(3)}, **{dt[i]: 0 for i in range(3)},
                **{X[i]: X0[i] for i in range(3)}, **{Ps[i]: sp.Float(repr(params[i]), DIG) for i in range(len(params))}}
        ev = lambda e: float(sp.N(e.subs(subs), DIG))
        r = [ev(e) for e in res]
        jp = [[ev(sp.diff(e, s)) for s in (*w, *dt)] for e in res]
        jx = [[ev(sp.diff(e, s)) for s in X] for e in res]
        jc = [[ev(sp.diff(e, s)) for s in Ps] for e in res]
        out["cameras"].append({"name": name, "model_id": model, "params": params, "r": r, "jp": jp, "jx": jx, "jc": jc})
        print(name, r)
    path = Path(__file__).with_name("jacobians.json")
    path.write_text(json.dumps(out, indent=1))
    print("wrote", path)


if __name__ == "__main__":
    main()
