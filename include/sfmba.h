/*
 * sfmba.h - C ABI of the MI355X bundle-adjustment engine.
 *
 * This is the drop-in boundary for the `sfm-ba` crate of sfmtory.  Every
 * type and entry point below mirrors one public item of that crate's Rust
 * API (crates/sfm-ba/src/lib.rs, crates/sfm-ba/src/intrinsics.rs) or of the
 * sfm-core types it takes (crates/sfm-core/src/camera.rs, pose.rs).  A Rust
 * facade binds these with `extern "C"` and converts its Vec/nalgebra values
 * into the plain arrays used here; the library never sees Rust or torch
 * types.
 *
 * Conventions
 *  - All arrays are caller-owned host memory.  Inputs are read, never kept.
 *  - Entry points return 0 on success and a negative SFMBA_ERR_* code on
 *    failure; sfmba_last_error() then describes it (thread-local string).
 *    The Rust API is infallible for numerical trouble (a failed solve just
 *    stops LM and returns the last accepted state), and so is this ABI:
 *    a non-zero status only reports invalid arguments (an index out of
 *    range, which the Rust code would panic on) or a HIP/device error.
 *    There is no CPU fallback: without a usable gfx950 device the compute
 *    entry points fail with SFMBA_ERR_DEVICE.
 *  - Re-entrant: each call owns its stream and device workspace.
 */
#ifndef SFMBA_H
#define SFMBA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define SFMBA_OK 0
#define SFMBA_ERR_INVALID_ARGUMENT (-1)
#define SFMBA_ERR_DEVICE (-2)
#define SFMBA_ERR_OUT_OF_MEMORY (-3)
#define SFMBA_ERR_UNSUPPORTED (-4)

/* ------------------------------------------------------------ camera model
 * COLMAP numeric model ids and parameter order.
 * Replaces: enum CameraModel, CameraModel::model_id, CameraModel::params,
 *           CameraModel::from_name_and_params
 *           (crates/sfm-core/src/camera.rs:21-58, :74-83, :85-133, :136-163)
 */
#define SFMBA_SIMPLE_PINHOLE 0 /* f, cx, cy */
#define SFMBA_PINHOLE 1        /* fx, fy, cx, cy */
#define SFMBA_SIMPLE_RADIAL 2  /* f, cx, cy, k */
#define SFMBA_RADIAL 3         /* f, cx, cy, k1, k2 */
#define SFMBA_OPENCV 4         /* fx, fy, cx, cy, k1, k2, p1, p2 */
#define SFMBA_OPENCV_FISHEYE 5 /* fx, fy, cx, cy, k1, k2, k3, k4 */

typedef struct sfmba_camera {
  int32_t model_id;  /* one of SFMBA_* model ids above */
  int32_t reserved;  /* must be 0 */
  double params[8];  /* COLMAP order; entries past the model's count ignored */
} sfmba_camera;

/* World-to-camera pose, X_cam = R X_world + t, R a Hamilton quaternion.
 * Replaces: struct Pose (crates/sfm-core/src/pose.rs:15-20); the quaternion
 * is stored w,x,y,z as in Pose::quaternion_wxyz (pose.rs:69-72). */
typedef struct sfmba_pose {
  double qw, qx, qy, qz;
  double tx, ty, tz;
} sfmba_pose;

/* Replaces: struct Observation (crates/sfm-ba/src/lib.rs:56-62). */
typedef struct sfmba_observation {
  uint64_t image_idx;
  uint64_t point_idx;
  double x;
  double y;
} sfmba_observation;

/* Replaces: struct BaInput (crates/sfm-ba/src/lib.rs:64-96).
 * The three `fixed_*` vectors keep the Rust semantics: a vector shorter than
 * its index range means "free" for the missing entries (lib.rs:766, :772),
 * and a per-camera mask shorter than the model's parameter count fixes only
 * the entries it has (lib.rs:1052-1057).  fixed_camera_params is the
 * concatenation of the per-camera masks; fixed_camera_params_len[c] is the
 * length of mask c, for c < num_fixed_camera_params. */
typedef struct sfmba_input {
  size_t num_images;       /* == poses.len() */
  size_t num_cameras;      /* == cameras.len() */
  size_t num_points;       /* == points.len() */
  size_t num_observations; /* == observations.len() */
  const uint64_t* camera_of_image; /* num_camera_of_image entries */
  size_t num_camera_of_image;
  const sfmba_camera* cameras;
  const sfmba_pose* poses;
  const double* points; /* x,y,z per point */
  const sfmba_observation* observations;
  const uint8_t* fixed_poses;
  size_t num_fixed_poses;
  const uint8_t* fixed_cameras;
  size_t num_fixed_cameras;
  const uint8_t* fixed_camera_params;
  const size_t* fixed_camera_params_len;
  size_t num_fixed_camera_params;
} sfmba_input;

/* Replaces: enum RobustLoss (crates/sfm-ba/src/lib.rs:213-218). */
#define SFMBA_LOSS_NONE 0
#define SFMBA_LOSS_HUBER 1
#define SFMBA_LOSS_CAUCHY 2

/* Replaces: struct BaParams + Default (crates/sfm-ba/src/lib.rs:220-238).
 * Defaults: max_iterations 50, initial_lambda 1e-3, Huber, 2.0 px. */
typedef struct sfmba_params {
  uint64_t max_iterations;
  double initial_lambda;
  int32_t robust_loss;
  int32_t reserved;
  double loss_scale_px;
} sfmba_params;

/* Replaces: struct BaOutput (crates/sfm-ba/src/lib.rs:240-247).  The caller
 * provides the three arrays, sized like the input's. */
typedef struct sfmba_output {
  sfmba_pose* poses;
  double* points;
  sfmba_camera* cameras;
  double initial_cost;
  double final_cost;
  uint64_t iterations_run;
} sfmba_output;

/* Fills *p with BaParams::default() (lib.rs:229-238). */
void sfmba_params_default(sfmba_params* p);

/* --------------------------------------------------------- bundle adjust
 * Replaces: pub fn bundle_adjust(input: BaInput, params: &BaParams) -> BaOutput
 * (crates/sfm-ba/src/lib.rs:747-1232).  Schur-complement Levenberg-Marquardt
 * over free poses, free camera intrinsics and all points, computed on the
 * GPU in FP64 throughout (the reference's arithmetic).  The input is not
 * modified; the refined state is written to *out. */
int sfmba_bundle_adjust(const sfmba_input* input, const sfmba_params* params,
                        sfmba_output* out);

/* Engine options (no counterpart in the reference: the precision and solver
 * choices of this engine). */
#define SFMBA_PRECISION_FP64 0  /* everything in FP64 (parity mode) */
#define SFMBA_PRECISION_MIXED 1 /* FP32 Jacobian and Schur assembly while
                                   the Marquardt damping is >= 1e-5 (where
                                   an FP32-assembled system still factors),
                                   FP64 assembly from then on; FP64 point
                                   inverses, residuals, cost and Cholesky
                                   throughout; optional FP64 iterative
                                   refinement of the step */
#define SFMBA_MIXED_MIN_LAMBDA 1e-5 /* mixed: smallest damping assembled in FP32 */
#define SFMBA_SOLVER_CHOLESKY 0 /* dense FP64 MFMA Cholesky (LU fallback) */
#define SFMBA_SOLVER_PCG 1      /* block-Jacobi PCG on the reduced system */
#define SFMBA_SOLVER_PCG_MIXED 2 /* block-Jacobi PCG whose iterations read an FP32
                                    copy of the reduced system, with the true
                                    residual recomputed in FP64 every 16
                                    iterations and at convergence */

typedef struct sfmba_options {
  int32_t precision;  /* SFMBA_PRECISION_* */
  int32_t solver;     /* SFMBA_SOLVER_* */
  int32_t device;     /* HIP device ordinal */
  int32_t profile;    /* 1: time each phase with HIP events into stats */
  int32_t refine_steps; /* mixed precision, FP32-assembled tries: FP64 iterative-refinement
                           sweeps of the step against the assembled system (default 0) */
  int32_t pcg_max_iterations; /* PCG: iteration cap (default 500) */
  double pcg_tolerance;       /* PCG: relative residual target (default 1e-10) */
} sfmba_options;

void sfmba_options_default(sfmba_options* o);

/* Optional diagnostics filled by the _ex entry points.  Arrays are
 * caller-provided; *_len reports how many entries were written (capped by
 * *_capacity). */
typedef struct sfmba_stats {
  double* outer_costs;       /* accepted cost after each outer iteration */
  uint64_t outer_costs_capacity;
  uint64_t outer_costs_len;
  double* trial_costs;       /* cost of every trial step, in order */
  uint64_t trial_costs_capacity;
  uint64_t trial_costs_len;
  uint64_t dim;              /* size of the reduced camera system solved */
  uint64_t num_linearizations;
  uint64_t num_factorizations;
  uint64_t num_cholesky_failures;
  uint64_t num_lu_solves;
  uint64_t num_rejected_steps;
  uint64_t num_schur_pairs;  /* point-observation pairs reduced per try */
  double time_total_s;       /* wall-clock of the call, device-resident part */
  double time_upload_s;      /* host-side packing and host->device copies (at create) */
  double time_linearize_s;   /* profile=1: residual + Jacobian kernels */
  double time_schur_s;       /* profile=1: point inverses + Schur assembly */
  double time_factor_s;      /* profile=1: factorization */
  double time_solve_s;       /* profile=1: triangular solves + refinement */
  double time_update_s;      /* profile=1: back-substitution, update, cost */
  uint64_t count_linearize;
  uint64_t count_schur;
  uint64_t count_factor;
  /* profile=1: the factorization's trailing-update kernel, timed per launch
   * with HIP events on the solve stream, and the algorithmic FP64 FLOP of
   * those launches (lower-triangle update, rhs row, next diagonal block). */
  double time_trailing_update_s;
  uint64_t count_trailing_update;
  double flop_trailing_update;
  uint64_t num_pcg_iterations; /* solver = PCG: iterations over all tries */
  /* Every damping try in order (lib.rs:1076-1221): the Marquardt damping
   * its reduced system was built with, and its outcome (SFMBA_TRY_* flags).
   * Both arrays hold tries_capacity entries (either may be NULL). */
  double* try_lambdas;
  uint8_t* try_outcomes;
  uint64_t tries_capacity;
  uint64_t tries_len;
  uint64_t num_fp32_tries;   /* mixed: tries whose reduced system was assembled in FP32 */
  /* profile=1, mixed: the share of time_linearize_s / time_schur_s (and of
   * their counts) spent in FP32 stages */
  double time_linearize_fp32_s;
  double time_schur_fp32_s;
  uint64_t count_linearize_fp32;
  uint64_t count_schur_fp32;
} sfmba_stats;

/* Outcome of one damping try (sfmba_stats.try_outcomes): the low two bits
 * are the LM decision, the others how the step was formed. */
#define SFMBA_TRY_ACCEPTED 0   /* finite trial cost below the current cost */
#define SFMBA_TRY_REJECTED 1   /* trial cost not lower (or not finite) */
#define SFMBA_TRY_NO_STEP 2    /* neither Cholesky nor LU gave a step */
#define SFMBA_TRY_LU 4         /* the step came from the LU fallback */
#define SFMBA_TRY_FP32 8       /* mixed: the reduced system was assembled in FP32 */

/* bundle_adjust with engine options and diagnostics (stats may be NULL). */
int sfmba_bundle_adjust_ex(const sfmba_input* input, const sfmba_params* params,
                           const sfmba_options* options, sfmba_output* out,
                           sfmba_stats* stats);

/* A problem made resident on a device once and solved many times.  No
 * counterpart in the reference; for callers that re-solve one structure
 * (the fixed- and free-intrinsics solves of one bundle,
 * crates/sfm-reconstruction/src/lib.rs:1623, :1656) and for timing the
 * device-resident solve.  Every solve starts from the input's initial
 * state and has bundle_adjust's semantics; calls on one handle must not
 * overlap.  sfmba_bundle_adjust_ex == create + solve + destroy. */
typedef struct sfmba_problem sfmba_problem;

int sfmba_problem_create(const sfmba_input* input, int32_t device, sfmba_problem** problem);
int sfmba_problem_solve(sfmba_problem* problem, const sfmba_params* params,
                        const sfmba_options* options, sfmba_output* out, sfmba_stats* stats);
void sfmba_problem_destroy(sfmba_problem* problem);

/* Many independent problems (local-BA windows) solved together.
 * inputs/outputs/stats are arrays of `count` entries (stats may be NULL).  No counterpart in the reference, whose callers loop
 * over bundle_adjust (crates/sfm-reconstruction/src/lib.rs:708-719); each
 * problem gets exactly the bundle_adjust semantics. */
int sfmba_bundle_adjust_batch(const sfmba_input* inputs, size_t count,
                              const sfmba_params* params,
                              const sfmba_options* options,
                              sfmba_output* outputs, sfmba_stats* stats);

/* A batch plan: the same batched solve, prepared once and replayed.  No
 * counterpart in the reference (its callers loop over bundle_adjust,
 * crates/sfm-reconstruction/src/lib.rs:708-719); for callers that solve the
 * same batch of windows repeatedly (benchmarks, re-solves after outlier
 * edits made by rebuilding the plan).
 *
 * sfmba_batch_create packs and uploads the problems (the inputs need not
 * outlive the call), picks the window kernel's launch shape -- threads per
 * window and Schur point pairs in flight -- and captures one solve in a HIP
 * graph: restore every window's initial state, the persistent window
 * kernel, the gather of the final states and their copy to pinned host
 * memory.  autotune_iterations > 0 runs the autotuner first: every shape is
 * timed on this batch with the iteration count capped at that value and the
 * fastest is kept (and remembered for later batches of the same size class
 * in this process, sfmba_bundle_adjust_batch included); 0 uses the
 * remembered or default shape.  Problems too large for the window kernel
 * are solved by the single-problem path on each solve.
 * sfmba_batch_solve replays the graph from the initial state and writes
 * outputs/stats exactly as sfmba_bundle_adjust_batch does.  Solves of one
 * plan must not overlap. */
typedef struct sfmba_batch sfmba_batch;
#define SFMBA_MAX_WINDOW_SHAPES 8
typedef struct sfmba_batch_info {
  int32_t threads_per_window;  /* the launch shape in use: threads per window and */
  int32_t pairs_in_flight;     /* the Schur reduction's batch depth (load instructions
                                  per incidence-record side and batch) */
  int32_t tuned;               /* 1 when the autotuner chose it for this plan */
  int32_t num_shapes;          /* candidate shapes (below) */
  int32_t shape_threads[SFMBA_MAX_WINDOW_SHAPES];
  int32_t shape_pairs[SFMBA_MAX_WINDOW_SHAPES];  /* batch depth of each candidate */
  double shape_seconds[SFMBA_MAX_WINDOW_SHAPES]; /* tuning run time per shape (0: not timed) */
  double create_seconds;       /* packing, upload, buffers */
  double capture_seconds;      /* graph capture and instantiation */
  uint64_t num_windows;        /* solved by the captured window kernel */
  uint64_t num_large;          /* solved by the single-problem path */
} sfmba_batch_info;
int sfmba_batch_create(const sfmba_input* inputs, size_t count, const sfmba_params* params,
                       const sfmba_options* options, int32_t autotune_iterations,
                       sfmba_batch** batch);
int sfmba_batch_solve(sfmba_batch* batch, sfmba_output* outputs, sfmba_stats* stats);
int sfmba_batch_get_info(const sfmba_batch* batch, sfmba_batch_info* info);
void sfmba_batch_destroy(sfmba_batch* batch);

/* ------------------------------------------------ caller solve policy
 * The solve part of the reference's caller, from the assembled problem on.
 *
 * Replaces: fn filter_and_reoptimize
 * (crates/sfm-reconstruction/src/lib.rs:1710-1768): bundle_adjust, then up
 * to two rounds of dropping the observations behind their camera or with
 * plain residual norm > threshold_px and re-solving the survivors from the
 * current state; stops when nothing was dropped or fewer than 12
 * observations remain.  *num_observations_kept (may be NULL) receives the
 * survivor count. */
int sfmba_filter_and_reoptimize(const sfmba_input* input, const sfmba_params* params,
                                const sfmba_options* options, double threshold_px,
                                sfmba_output* out, uint64_t* num_observations_kept);

/* IntrinsicsMode (crates/sfm-reconstruction/src/lib.rs:1301-1315). */
#define SFMBA_INTRINSICS_FIXED 0        /* intrinsics held */
#define SFMBA_INTRINSICS_FREE 1         /* one joint solve, cameras with >= 5 images free */
#define SFMBA_INTRINSICS_FREE_GUARDED 2 /* fixed vs free solves, filtered, guarded */

typedef struct sfmba_policy {
  int32_t intrinsics;                 /* SFMBA_INTRINSICS_* (default FREE_GUARDED) */
  int32_t robust_loss;                /* loss of every solve (default Huber) */
  double max_reprojection_error_px;   /* outlier threshold of the guarded pass (default 4) */
  const uint8_t* camera_caller_fixed; /* per camera: intrinsics the caller pinned */
  size_t num_camera_caller_fixed;     /* shorter = the rest not pinned */
} sfmba_policy;

typedef struct sfmba_policy_report {
  int32_t chose_free;            /* the free-intrinsics result was kept */
  int32_t distortion_plausible;  /* every |distortion coefficient| < 0.5 (free result) */
  double fixed_mean_error_px;    /* plain mean reprojection error of each candidate, */
  double free_mean_error_px;     /* over all of the problem's observations */
  uint64_t observations_kept_fixed, observations_kept_free;
  uint64_t filter_rounds_fixed, filter_rounds_free;
} sfmba_policy_report;

void sfmba_policy_default(sfmba_policy* p);

/* Replaces: the solve of fn run_bundle_adjustment
 * (crates/sfm-reconstruction/src/lib.rs:1479-1695) on the problem the
 * caller assembled for its scope (images, points, observations and the
 * fixed poses of the gauge and of pinned extrinsics).  The input's
 * fixed_cameras and fixed_camera_params are replaced by the policy's:
 * principal points pinned (default_fixed_params_mask); FIXED holds every
 * camera; FREE frees the cameras with >= 5 images that the caller did not
 * pin, in one 50-iteration solve; FREE_GUARDED (200 iterations) solves
 * with all cameras fixed and, when a camera has >= 5 images, with those
 * cameras free -- concurrently, each through filter_and_reoptimize -- and
 * keeps the free result only if its distortion is plausible and its plain
 * mean reprojection error is lower.  A problem with fewer than 2 images or
 * no observations is returned unchanged.  report may be NULL. */
int sfmba_run_bundle_adjustment(const sfmba_input* input, const sfmba_policy* policy,
                                const sfmba_options* options, sfmba_output* out,
                                sfmba_policy_report* report);

/* ------------------------------------------------ mean reprojection error
 * Replaces: pub fn mean_reprojection_error(input: &BaInput) -> f64
 * (crates/sfm-ba/src/lib.rs:708-725): mean of the plain residual norm over
 * observations in front of their camera, 0 if there are none. */
int sfmba_mean_reprojection_error(const sfmba_input* input, double* out);

/* Robust cost sum_i w(|r_i|) |r_i|^2 over in-front observations, the
 * quantity bundle_adjust reports as initial/final cost.
 * Replaces: fn total_cost (crates/sfm-ba/src/lib.rs:727-745). */
int sfmba_total_cost(const sfmba_input* input, const sfmba_params* params,
                     double* out);

/* sfmba_total_cost of a state already in device memory, allocating no
 * device memory.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).
 *  - device_input is a host struct whose arrays (observations,
 *    camera_of_image, cameras, poses, points) are device pointers on the
 *    current HIP device, in this header's layouts; the fixed_* fields are
 *    ignored.  Observations are summed in the caller's order.
 *  - stream is a hipStream_t (NULL: the null stream); the call enqueues
 *    its kernels there and synchronizes it before returning *out (host).
 *  - Partial sums live in one static device array of the library, of
 *    SFMBA_NOALLOC_MAX_OBSERVATIONS / 2048 + 2 doubles (256 KB) per GPU;
 *    more observations than SFMBA_NOALLOC_MAX_OBSERVATIONS are rejected
 *    (SFMBA_ERR_INVALID_ARGUMENT).  Every call on one GPU shares that GPU's
 *    array: calls that overlap on one GPU race.
 *  - An out-of-range index fails with SFMBA_ERR_INVALID_ARGUMENT, as in
 *    sfmba_total_cost (checked on the device; nothing is read out of
 *    range). */
#define SFMBA_NOALLOC_MAX_OBSERVATIONS (1ull << 26)
int sfmba_total_cost_noalloc(const sfmba_input* device_input, const sfmba_params* params, void* stream,
                             double* out);

/* ------------------------------------------------ fixed-parameter mask
 * Replaces: pub fn default_fixed_params_mask(&[CameraModel]) -> Vec<Vec<bool>>
 * (crates/sfm-ba/src/intrinsics.rs:33-61).  Writes num_cameras masks of 8
 * entries each (mask[c*8 + k]) and the model's parameter count to
 * lens[c]; entries k >= lens[c] are 0. */
int sfmba_default_fixed_params_mask(const sfmba_camera* cameras,
                                    size_t num_cameras, uint8_t* mask,
                                    size_t* lens);

/* ----------------------------------------------- intrinsics-only refine
 * Replaces: struct IntrinsicsRefineParams + Default and pub fn
 * refine_intrinsics (crates/sfm-ba/src/intrinsics.rs:65-146, 148-238).
 * Poses and points fixed; each camera with >= 8 observations gets a small
 * dense LM over its own parameters with central-difference Jacobians.
 * `cameras` is updated in place. */
typedef struct sfmba_intrinsics_params {
  uint64_t max_iterations; /* default 30 */
  double initial_lambda;   /* default 1e-3 */
  int32_t robust_loss;     /* default SFMBA_LOSS_HUBER */
  int32_t reserved;
  double loss_scale_px;    /* default 2.0 */
} sfmba_intrinsics_params;

void sfmba_intrinsics_params_default(sfmba_intrinsics_params* p);

int sfmba_refine_intrinsics(sfmba_camera* cameras, size_t num_cameras,
                            const uint64_t* camera_of_image, size_t num_images,
                            const sfmba_pose* poses, const double* points,
                            size_t num_points,
                            const sfmba_observation* observations,
                            size_t num_observations,
                            const sfmba_intrinsics_params* params);

/* ------------------------------------------ per-observation linearization
 * The residual + analytic Jacobian kernel on its own, for checking it.
 * Replaces: fn analytic_jacobians (crates/sfm-ba/src/lib.rs:327-590) with
 * fn pose_point_jacobians (lib.rs:282-310).  For observation i writes
 * r[2i..], jp[12i..] (2x6 row-major: d(u,v)/d(omega, t)), jx[6i..] (2x3),
 * jc[16i..] (2x8 row-major, columns in CameraModel::params() order, unused
 * columns 0) and valid[i] (0 when the point is at or behind z = 1e-9).
 * precision selects the FP64 or the FP32 instantiation of the kernel. */
int sfmba_observation_jacobians(const sfmba_input* input, int32_t precision,
                                int32_t device, double* r, double* jp,
                                double* jx, double* jc, uint8_t* valid);

/* ------------------------------------------------ dense reduced-system solve
 * The factorization kernels on their own, for checking them.  Replaces the
 * reference's nalgebra Cholesky::new(S).solve(rhs) with its LU fallback
 * (crates/sfm-ba/src/lib.rs:1076-1094).  A is n x n column-major (only its
 * lower triangle is read: the matrix is symmetric), b has n entries.
 * method 0: Cholesky, *status = 1 when a pivot is not > 0;
 * method 1: LU with partial pivoting of the full symmetric matrix,
 * *status = 1 when U has a zero diagonal.  x receives the solution. */
int sfmba_dense_solve(const double* A, size_t n, const double* b, int32_t method,
                      int32_t device, double* x, int32_t* status);

/* ------------------------------------------- scoped BA over a reconstruction
 * The problem assembly of the reference's caller: which images, points and
 * observations of the growing reconstruction enter one bundle adjustment,
 * and which poses are held.
 *
 * Replaces: fn run_bundle_adjustment's scope handling
 * (crates/sfm-reconstruction/src/lib.rs:1339-1474) with BaScope
 * (:1287-1299), LOCAL_BA_NUM_IMAGES (:1320) and the write-back (:1697-1707).
 *
 * GLOBAL: every registered image; every point whose track has >= 2
 * elements; only the seed image's pose held.
 * LOCAL (center): the points with >= 2 track elements one of which is in
 * image `center`; the registered images ranked by how many of those
 * track elements they hold (descending, ties by lower index), the first
 * SFMBA_LOCAL_BA_NUM_IMAGES of them plus `center` free; every other
 * registered image observing one of the points included with its pose
 * held; the seed's pose held too.
 * Both: pinned poses held; if no pose is held, the lowest-indexed image's
 * is (gauge); observations are the track elements whose image is in the
 * problem, with the keypoint's single-precision coordinates; cameras in
 * order of first use.  Fewer than 2 images or no observation: nothing to
 * adjust (the window is empty and the state is left as it is). */
#define SFMBA_SCOPE_GLOBAL 0
#define SFMBA_SCOPE_LOCAL 1
#define SFMBA_LOCAL_BA_NUM_IMAGES 6

/* The incremental pipeline's working state (crates/sfm-reconstruction/
 * src/lib.rs:126-129 PointWork, the per-image poses, registered flags and
 * camera map).  Image-indexed arrays have num_images entries; keypoints are
 * the images' feature coordinates, image-major, keypoint_offsets[i] ..
 * keypoint_offsets[i+1] for image i; point p's track is
 * track_offsets[p] .. track_offsets[p+1] in track_images / track_keypoints. */
typedef struct sfmba_reconstruction {
  size_t num_images;
  const uint8_t* registered;
  sfmba_pose* poses;                 /* in/out; entries of unregistered images unused */
  const uint32_t* image_camera_id;   /* the id of each image's physical camera */
  const uint8_t* pose_pinned;        /* NULL or per image: caller-measured extrinsics, always held */
  const uint64_t* keypoint_offsets;  /* num_images + 1 */
  const float* keypoints;            /* x, y per keypoint */
  size_t num_points;
  double* points;                    /* in/out, x, y, z per point */
  const uint64_t* track_offsets;     /* num_points + 1 */
  const uint32_t* track_images;
  const uint32_t* track_keypoints;
  size_t num_cameras;
  const uint32_t* camera_ids;        /* id of each camera entry */
  sfmba_camera* cameras;             /* in/out */
  const uint32_t* caller_fixed_camera_ids; /* intrinsics the caller pinned, by id */
  size_t num_caller_fixed_camera_ids;
  uint64_t seed_image;               /* the gauge image */
} sfmba_reconstruction;

typedef struct sfmba_window_summary {
  uint64_t num_images, num_free_images, num_cameras, num_points, num_observations;
  double initial_cost, final_cost;
  uint64_t iterations_run;
} sfmba_window_summary;

/* One assembled problem, kept so that its result can be written back.  The
 * input it exposes holds every camera fixed with principal points masked,
 * the in-loop local bundle's semantics (lib.rs:708-719). */
typedef struct sfmba_window sfmba_window;

int sfmba_window_create(const sfmba_reconstruction* rec, int32_t scope,
                        uint64_t center, sfmba_window** out);
/* Pointers into the window; valid until sfmba_window_destroy.  An empty
 * window has num_images == 0. */
int sfmba_window_input(const sfmba_window* w, sfmba_input* input);
/* Original image / point index of each compact index (NULL pointers allowed). */
int sfmba_window_indices(const sfmba_window* w, const uint64_t** image_indices,
                         size_t* num_images, const uint64_t** point_indices,
                         size_t* num_points);
/* Writes poses, points and cameras of a solve of this window back into rec
 * (lib.rs:1697-1707); nothing for an empty window. */
int sfmba_window_scatter(const sfmba_window* w, const sfmba_output* out,
                         sfmba_reconstruction* rec);
void sfmba_window_destroy(sfmba_window* w);

/* Assemble (scope, center), solve it under `policy` (its camera_caller_fixed
 * is ignored: rec's caller_fixed_camera_ids are used) and write back.
 * summary and report may be NULL. */
int sfmba_reconstruction_bundle_adjust(sfmba_reconstruction* rec, int32_t scope,
                                       uint64_t center, const sfmba_policy* policy,
                                       const sfmba_options* options,
                                       sfmba_window_summary* summary,
                                       sfmba_policy_report* report);

/* Many independent reconstructions (seed trials, separate models), one
 * window each, with intrinsics held -- the in-loop local bundle -- solved in
 * one batched launch (sfmba_bundle_adjust_batch) and written back.
 * summaries may be NULL. */
int sfmba_reconstruction_bundle_adjust_batch(sfmba_reconstruction* recs, size_t count,
                                             const int32_t* scopes, const uint64_t* centers,
                                             const sfmba_params* params,
                                             const sfmba_options* options,
                                             sfmba_window_summary* summaries);

/* ------------------------------------------------ COLMAP text model refine
 * Replaces: fn cmd_refine (crates/sfm-cli/src/main.rs:946-1130) over the
 * plain-text model reader/writer (crates/sfm-io/src/colmap.rs:25-300).
 * Reads cameras.txt / images.txt / points3D.txt from model_dir, builds the
 * compact problem (one pose per image in id order, one camera block per
 * physical camera, one point per 3D point, observations from the tracks,
 * keypoints as single precision), fixes image 0's pose, solves with the
 * engine, recomputes each point's mean reprojection error against the
 * refined model and writes the three files to output_dir (which may be
 * model_dir, as the reference does).  Errors: fewer than 2 images, no
 * points, no observations, or a malformed file (SFMBA_ERR_INVALID_ARGUMENT
 * with the file and field named, as the reference's parse errors are). */
typedef struct sfmba_colmap_refine_options {
  int32_t refine_intrinsics; /* 0: cameras fixed; 1: free, principal point pinned */
  int32_t robust_loss;       /* default SFMBA_LOSS_HUBER */
  uint64_t max_iterations;   /* default 200 */
} sfmba_colmap_refine_options;

typedef struct sfmba_colmap_refine_report {
  uint64_t num_images;
  uint64_t num_points3d;
  uint64_t num_observations;
  double initial_cost;
  double final_cost;
  uint64_t iterations_run;
  double mean_reprojection_error_px_before; /* Reconstruction::mean_reprojection_error */
  double mean_reprojection_error_px_after;
} sfmba_colmap_refine_report;

void sfmba_colmap_refine_options_default(sfmba_colmap_refine_options* o);

int sfmba_colmap_refine(const char* model_dir, const char* output_dir,
                        const sfmba_colmap_refine_options* options,
                        const sfmba_options* engine_options,
                        sfmba_colmap_refine_report* report);

/* Read the model in model_dir and write it to output_dir unchanged, through
 * the same reader and writer (host only, no device call).  counts receives
 * the number of cameras, images and 3D points (3 entries; may be NULL).
 * Replaces: sfm_io::read_colmap_model + write_colmap_model
 * (crates/sfm-io/src/colmap.rs:25-43). */
int sfmba_colmap_rewrite(const char* model_dir, const char* output_dir,
                         uint64_t* counts);

/* Number of usable HIP devices (0 without a GPU). */
int sfmba_device_count(void);

/* Description of the last error on this thread ("" if none). */
const char* sfmba_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SFMBA_H */
