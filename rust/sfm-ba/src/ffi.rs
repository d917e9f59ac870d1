//! The subset of include/sfmba.h the facade binds: #[repr(C)] mirrors of its
//! structs (same field order and types; tests/test_rust_facade.py compares
//! them with the header) and the entry points.
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_int};

#[repr(C)]
#[derive(Clone, Copy)]
pub struct sfmba_camera {
    pub model_id: i32,
    pub reserved: i32,
    pub params: [f64; 8],
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct sfmba_pose {
    pub qw: f64,
    pub qx: f64,
    pub qy: f64,
    pub qz: f64,
    pub tx: f64,
    pub ty: f64,
    pub tz: f64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct sfmba_observation {
    pub image_idx: u64,
    pub point_idx: u64,
    pub x: f64,
    pub y: f64,
}

#[repr(C)]
pub struct sfmba_input {
    pub num_images: usize,
    pub num_cameras: usize,
    pub num_points: usize,
    pub num_observations: usize,
    pub camera_of_image: *const u64,
    pub num_camera_of_image: usize,
    pub cameras: *const sfmba_camera,
    pub poses: *const sfmba_pose,
    pub points: *const f64,
    pub observations: *const sfmba_observation,
    pub fixed_poses: *const u8,
    pub num_fixed_poses: usize,
    pub fixed_cameras: *const u8,
    pub num_fixed_cameras: usize,
    pub fixed_camera_params: *const u8,
    pub fixed_camera_params_len: *const usize,
    pub num_fixed_camera_params: usize,
}

#[repr(C)]
pub struct sfmba_params {
    pub max_iterations: u64,
    pub initial_lambda: f64,
    pub robust_loss: i32,
    pub reserved: i32,
    pub loss_scale_px: f64,
}

#[repr(C)]
pub struct sfmba_output {
    pub poses: *mut sfmba_pose,
    pub points: *mut f64,
    pub cameras: *mut sfmba_camera,
    pub initial_cost: f64,
    pub final_cost: f64,
    pub iterations_run: u64,
}

#[repr(C)]
pub struct sfmba_intrinsics_params {
    pub max_iterations: u64,
    pub initial_lambda: f64,
    pub robust_loss: i32,
    pub reserved: i32,
    pub loss_scale_px: f64,
}

extern "C" {
    pub fn sfmba_bundle_adjust(input: *const sfmba_input, params: *const sfmba_params, out: *mut sfmba_output) -> c_int;
    pub fn sfmba_mean_reprojection_error(input: *const sfmba_input, out: *mut f64) -> c_int;
    pub fn sfmba_default_fixed_params_mask(
        cameras: *const sfmba_camera,
        num_cameras: usize,
        mask: *mut u8,
        lens: *mut usize,
    ) -> c_int;
    pub fn sfmba_refine_intrinsics(
        cameras: *mut sfmba_camera,
        num_cameras: usize,
        camera_of_image: *const u64,
        num_images: usize,
        poses: *const sfmba_pose,
        points: *const f64,
        num_points: usize,
        observations: *const sfmba_observation,
        num_observations: usize,
        params: *const sfmba_intrinsics_params,
    ) -> c_int;
    pub fn sfmba_last_error() -> *const c_char;
}

/// Panics on a non-zero status with the engine's message.  The crate is
/// infallible for numerical trouble and the engine never reports numerical
/// trouble as an error: a non-zero status is a bad index (on which the
/// crate's own code would panic too) or a device failure, and there is no
/// CPU fallback.
pub fn check(code: c_int) {
    if code != 0 {
        let msg = unsafe { std::ffi::CStr::from_ptr(sfmba_last_error()) }.to_string_lossy().into_owned();
        panic!("sfm-ba engine error {code}: {msg}");
    }
}
