"""The reference's own end-to-end test scenes, rebuilt as inputs.

crates/sfm-ba/src/lib.rs:1368-1476  recovers_ground_truth_from_perturbed_initialization
crates/sfm-ba/src/lib.rs:1487-1592  joint_optimization_recovers_shared_camera_focal_length
crates/sfm-ba/src/intrinsics.rs:252-383  refine_intrinsics tests

Projections use the oracle (test infrastructure) to build observations, as
the reference builds them with CameraModel::project.
"""
import numpy as np

import oracle
from conftest import pose_row_from_euler
from sfmtory_amd.ba import BaInput, CameraModel, cameras_to_array

IDENT = np.array([1.0, 0, 0, 0, 0, 0, 0])


def observe(cam: CameraModel, poses, points, zmin):
    obs = []
    for pi, X in enumerate(points):
        for ii, P in enumerate(poses):
            pc = oracle.transform_point(P, X)
            if pc[2] <= zmin:
                continue
            u, v = oracle.project(cam.model_id, cam.params_, pc)
            obs.append((ii, pi, u, v))
    arr = np.zeros(len(obs), dtype=[("image_idx", "<u8"), ("point_idx", "<u8"), ("x", "<f8"), ("y", "<f8")])
    for k, o in enumerate(obs):
        arr[k] = o
    return arr


def recover_scene():
    """lib.rs:1368-1436: two images, 25 points, SIMPLE_PINHOLE f=700."""
    cam = CameraModel.simple_pinhole(700.0, 320.0, 240.0)
    true_poses = [IDENT.copy(), pose_row_from_euler(0.02, 0.2, -0.01, (0.8, 0.05, 0.1))]
    pts = np.array([[0.5 * np.sin(t * 0.6), 0.4 * np.cos(t * 0.4), 3.0 + 0.1 * t] for t in map(float, range(25))])
    obs = observe(cam, true_poses, pts, 0.0)
    init_poses = [true_poses[0], oracle.perturb_pose(true_poses[1], [0.03, -0.02, 0.01, 0.05, -0.03, 0.04])]
    init_pts = pts.copy()
    for i in range(25):
        s = float(i)
        init_pts[i] += [0.05 * np.sin(s * 1.1), -0.04 * np.cos(s * 0.7), 0.06 * np.sin(s * 0.3)]
    inp = BaInput(camera_of_image=[0, 0], cameras=cameras_to_array([cam]), poses=np.array(init_poses),
                  points=init_pts, observations=obs, fixed_poses=[True, False], fixed_cameras=[True],
                  fixed_camera_params=[])
    return inp, np.array(true_poses), pts


def focal_scenes():
    """lib.rs:1487-1575: four images, 60 points, true f=1500, start f=1000.
    Returns (pre-converge input, builder of the joint input from its result)."""
    true_cam = CameraModel.simple_pinhole(1500.0, 500.0, 400.0)
    true_poses = [IDENT.copy(),
                  pose_row_from_euler(0.0, 0.25, 0.0, (1.2, 0.1, -0.3)),
                  pose_row_from_euler(0.1, -0.3, 0.05, (-0.9, 0.4, 0.5)),
                  pose_row_from_euler(-0.05, 0.4, -0.1, (0.5, -0.6, 1.1))]
    pts = np.array([[0.8 * np.sin(t * 0.31), 0.7 * np.cos(t * 0.47), 2.5 + 0.15 * t] for t in map(float, range(60))])
    obs = observe(true_cam, true_poses, pts, 0.1)
    wrong = cameras_to_array([CameraModel.simple_pinhole(1000.0, 500.0, 400.0)])
    pre = BaInput(camera_of_image=[0, 0, 0, 0], cameras=wrong, poses=np.array(true_poses), points=pts,
                  observations=obs, fixed_poses=[True, False, False, False], fixed_cameras=[True],
                  fixed_camera_params=[])

    def joint(poses, points):
        return BaInput(camera_of_image=[0, 0, 0, 0], cameras=wrong, poses=poses, points=points, observations=obs,
                       fixed_poses=[True, False, False, False], fixed_cameras=[False], fixed_camera_params=[])
    return pre, joint


def intrinsics_scene(n_points=40, wrong_f=960.0):
    """intrinsics.rs:252-286: identity pose, SIMPLE_PINHOLE true f=1500."""
    true_cam = CameraModel.simple_pinhole(1500.0, 500.0, 400.0)
    pts = np.array([[0.4 * np.sin(t * 0.6), 0.3 * np.cos(t * 0.4), 3.0 + 0.05 * t] for t in map(float, range(n_points))])
    obs = observe(true_cam, [IDENT], pts, -np.inf)
    cams = cameras_to_array([CameraModel.simple_pinhole(wrong_f, 500.0, 400.0)])
    return cams, [0], np.array([IDENT]), pts, obs
