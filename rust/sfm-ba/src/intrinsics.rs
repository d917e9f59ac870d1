//! Drop-in for crates/sfm-ba/src/intrinsics.rs:33-238 (the fixed-parameter
//! mask and the intrinsics-only refine), backed by the engine.
use crate::{cam_from_c, cam_to_c, ffi, loss_to_c, obs_to_c, pose_to_c, Observation, RobustLoss};
use nalgebra::Vector3;
use sfm_core::{CameraModel, Pose};

#[derive(Debug, Clone, Copy)]
pub struct IntrinsicsRefineParams {
    pub max_iterations: usize,
    pub initial_lambda: f64,
    pub robust_loss: RobustLoss,
    pub loss_scale_px: f64,
}

impl Default for IntrinsicsRefineParams {
    fn default() -> Self {
        IntrinsicsRefineParams {
            max_iterations: 30,
            initial_lambda: 1e-3,
            robust_loss: RobustLoss::Huber,
            loss_scale_px: 2.0,
        }
    }
}

/// Per camera, one flag per parameter (CameraModel::params() order): the
/// principal point held fixed, focal lengths and distortion free.
pub fn default_fixed_params_mask(cameras: &[CameraModel]) -> Vec<Vec<bool>> {
    let cams: Vec<_> = cameras.iter().map(cam_to_c).collect();
    let mut mask = vec![0u8; 8 * cams.len()];
    let mut lens = vec![0usize; cams.len()];
    ffi::check(unsafe {
        ffi::sfmba_default_fixed_params_mask(cams.as_ptr(), cams.len(), mask.as_mut_ptr(), lens.as_mut_ptr())
    });
    (0..cams.len()).map(|c| mask[8 * c..8 * c + lens[c]].iter().map(|&b| b != 0).collect()).collect()
}

/// Refines every camera's intrinsics in place with poses and points held
/// fixed (one small dense LM per camera with enough observations).
pub fn refine_intrinsics(
    cameras: &mut [CameraModel],
    camera_of_image: &[usize],
    poses: &[Pose],
    points: &[Vector3<f64>],
    observations: &[Observation],
    params: &IntrinsicsRefineParams,
) {
    let mut cams: Vec<_> = cameras.iter().map(cam_to_c).collect();
    let coi: Vec<u64> = camera_of_image.iter().map(|&c| c as u64).collect();
    let ps: Vec<_> = poses.iter().map(pose_to_c).collect();
    let pts: Vec<f64> = points.iter().flat_map(|p| [p.x, p.y, p.z]).collect();
    let obs: Vec<_> = observations.iter().map(obs_to_c).collect();
    let p = ffi::sfmba_intrinsics_params {
        max_iterations: params.max_iterations as u64,
        initial_lambda: params.initial_lambda,
        robust_loss: loss_to_c(params.robust_loss),
        reserved: 0,
        loss_scale_px: params.loss_scale_px,
    };
    ffi::check(unsafe {
        ffi::sfmba_refine_intrinsics(
            cams.as_mut_ptr(),
            cams.len(),
            coi.as_ptr(),
            coi.len(),
            ps.as_ptr(),
            pts.as_ptr(),
            points.len(),
            obs.as_ptr(),
            obs.len(),
            &p,
        )
    });
    for (dst, src) in cameras.iter_mut().zip(&cams) {
        *dst = cam_from_c(src, dst);
    }
}
