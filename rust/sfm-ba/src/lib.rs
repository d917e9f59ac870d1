//! Schur-complement Levenberg-Marquardt bundle adjustment on an MI355X.
//!
//! Drop-in for the reference crate's public API (crates/sfm-ba/src/lib.rs:48-49
//! re-exports, :56-247 types, :708 mean_reprojection_error, :747
//! bundle_adjust): the same names, fields and signatures, with the work done
//! by the HIP engine behind include/sfmba.h (linked by build.rs).  Inputs
//! cross to the device once per call; the caller's data is not modified.
mod ffi;
mod intrinsics;

pub use intrinsics::{default_fixed_params_mask, refine_intrinsics, IntrinsicsRefineParams};

use nalgebra::{Quaternion, UnitQuaternion, Vector3};
use sfm_core::{CameraModel, Pose};

/// One 2D measurement of point `point_idx` in image `image_idx` (pixels).
#[derive(Debug, Clone, Copy)]
pub struct Observation {
    pub image_idx: usize,
    pub point_idx: usize,
    pub x: f64,
    pub y: f64,
}

/// The problem.  `fixed_poses` / `fixed_cameras` are index-aligned with
/// `poses` / `cameras`; a shorter vector leaves the rest free.
/// `fixed_camera_params[c][k]` holds parameter k of camera c fixed; a
/// shorter (or empty) inner vector leaves the rest free.
pub struct BaInput {
    pub camera_of_image: Vec<usize>,
    pub cameras: Vec<CameraModel>,
    pub poses: Vec<Pose>,
    pub points: Vec<Vector3<f64>>,
    pub observations: Vec<Observation>,
    pub fixed_poses: Vec<bool>,
    pub fixed_cameras: Vec<bool>,
    pub fixed_camera_params: Vec<Vec<bool>>,
}

#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum RobustLoss {
    None,
    Huber,
    Cauchy,
}

#[derive(Debug, Clone, Copy)]
pub struct BaParams {
    pub max_iterations: usize,
    pub initial_lambda: f64,
    pub robust_loss: RobustLoss,
    /// Loss scale in pixels.
    pub loss_scale_px: f64,
}

impl Default for BaParams {
    fn default() -> Self {
        BaParams { max_iterations: 50, initial_lambda: 1e-3, robust_loss: RobustLoss::Huber, loss_scale_px: 2.0 }
    }
}

pub struct BaOutput {
    pub poses: Vec<Pose>,
    pub points: Vec<Vector3<f64>>,
    pub cameras: Vec<CameraModel>,
    pub initial_cost: f64,
    pub final_cost: f64,
    pub iterations_run: usize,
}

pub(crate) fn loss_to_c(l: RobustLoss) -> i32 {
    match l {
        RobustLoss::None => 0,
        RobustLoss::Huber => 1,
        RobustLoss::Cauchy => 2,
    }
}

pub(crate) fn cam_to_c(c: &CameraModel) -> ffi::sfmba_camera {
    let p = c.params();
    let mut params = [0.0; 8];
    params[..p.len()].copy_from_slice(&p);
    ffi::sfmba_camera { model_id: c.model_id(), reserved: 0, params }
}

/// The engine writes parameters in CameraModel::params() order; the model
/// is the input camera's.
pub(crate) fn cam_from_c(c: &ffi::sfmba_camera, like: &CameraModel) -> CameraModel {
    let n = like.params().len();
    CameraModel::from_name_and_params(like.name(), &c.params[..n]).expect("camera model of the input")
}

pub(crate) fn pose_to_c(p: &Pose) -> ffi::sfmba_pose {
    let [qw, qx, qy, qz] = p.quaternion_wxyz();
    ffi::sfmba_pose { qw, qx, qy, qz, tx: p.translation.x, ty: p.translation.y, tz: p.translation.z }
}

/// The rotation comes back as the engine computed it, without
/// renormalization (the reference's update does not renormalize either).
pub(crate) fn pose_from_c(p: &ffi::sfmba_pose) -> Pose {
    Pose {
        rotation: UnitQuaternion::new_unchecked(Quaternion::new(p.qw, p.qx, p.qy, p.qz)),
        translation: Vector3::new(p.tx, p.ty, p.tz),
    }
}

pub(crate) fn obs_to_c(o: &Observation) -> ffi::sfmba_observation {
    ffi::sfmba_observation { image_idx: o.image_idx as u64, point_idx: o.point_idx as u64, x: o.x, y: o.y }
}

/// C views of a BaInput; the arrays live as long as this value.
struct CInput {
    camera_of_image: Vec<u64>,
    cameras: Vec<ffi::sfmba_camera>,
    poses: Vec<ffi::sfmba_pose>,
    points: Vec<f64>,
    observations: Vec<ffi::sfmba_observation>,
    fixed_poses: Vec<u8>,
    fixed_cameras: Vec<u8>,
    fixed_params: Vec<u8>,
    fixed_params_len: Vec<usize>,
}

impl CInput {
    fn new(i: &BaInput) -> Self {
        CInput {
            camera_of_image: i.camera_of_image.iter().map(|&c| c as u64).collect(),
            cameras: i.cameras.iter().map(cam_to_c).collect(),
            poses: i.poses.iter().map(pose_to_c).collect(),
            points: i.points.iter().flat_map(|p| [p.x, p.y, p.z]).collect(),
            observations: i.observations.iter().map(obs_to_c).collect(),
            fixed_poses: i.fixed_poses.iter().map(|&b| b as u8).collect(),
            fixed_cameras: i.fixed_cameras.iter().map(|&b| b as u8).collect(),
            fixed_params: i.fixed_camera_params.iter().flatten().map(|&b| b as u8).collect(),
            fixed_params_len: i.fixed_camera_params.iter().map(|m| m.len()).collect(),
        }
    }

    fn raw(&self) -> ffi::sfmba_input {
        ffi::sfmba_input {
            num_images: self.poses.len(),
            num_cameras: self.cameras.len(),
            num_points: self.points.len() / 3,
            num_observations: self.observations.len(),
            camera_of_image: self.camera_of_image.as_ptr(),
            num_camera_of_image: self.camera_of_image.len(),
            cameras: self.cameras.as_ptr(),
            poses: self.poses.as_ptr(),
            points: self.points.as_ptr(),
            observations: self.observations.as_ptr(),
            fixed_poses: self.fixed_poses.as_ptr(),
            num_fixed_poses: self.fixed_poses.len(),
            fixed_cameras: self.fixed_cameras.as_ptr(),
            num_fixed_cameras: self.fixed_cameras.len(),
            fixed_camera_params: self.fixed_params.as_ptr(),
            fixed_camera_params_len: self.fixed_params_len.as_ptr(),
            num_fixed_camera_params: self.fixed_params_len.len(),
        }
    }
}

/// Levenberg-Marquardt over poses, points and free intrinsics (the
/// reference's schedule, damping and stopping rule; see include/sfmba.h).
pub fn bundle_adjust(input: BaInput, params: &BaParams) -> BaOutput {
    let c = CInput::new(&input);
    let mut poses = c.poses.clone();
    let mut points = c.points.clone();
    let mut cameras = c.cameras.clone();
    let p = ffi::sfmba_params {
        max_iterations: params.max_iterations as u64,
        initial_lambda: params.initial_lambda,
        robust_loss: loss_to_c(params.robust_loss),
        reserved: 0,
        loss_scale_px: params.loss_scale_px,
    };
    let mut out = ffi::sfmba_output {
        poses: poses.as_mut_ptr(),
        points: points.as_mut_ptr(),
        cameras: cameras.as_mut_ptr(),
        initial_cost: 0.0,
        final_cost: 0.0,
        iterations_run: 0,
    };
    ffi::check(unsafe { ffi::sfmba_bundle_adjust(&c.raw(), &p, &mut out) });
    BaOutput {
        poses: poses.iter().map(pose_from_c).collect(),
        points: points.chunks_exact(3).map(|v| Vector3::new(v[0], v[1], v[2])).collect(),
        cameras: cameras.iter().zip(&input.cameras).map(|(c, like)| cam_from_c(c, like)).collect(),
        initial_cost: out.initial_cost,
        final_cost: out.final_cost,
        iterations_run: out.iterations_run as usize,
    }
}

/// Mean Euclidean reprojection error over the observations in front of
/// their camera (0 when there are none).
pub fn mean_reprojection_error(input: &BaInput) -> f64 {
    let c = CInput::new(input);
    let mut v = 0.0;
    ffi::check(unsafe { ffi::sfmba_mean_reprojection_error(&c.raw(), &mut v) });
    v
}
