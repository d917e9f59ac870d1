// Internal interfaces of the MI355X bundle-adjustment engine (not installed).
//
// Device data layout (one problem):
//   observations are sorted point-major (CSR by point) and stored SoA:
//     obs_img u32, obs_pt u32, obs_xy double2;
//   an observation touches at most two "incidences": the block of the
//   reduced system its pose belongs to and the block its camera belongs to
//   (one merged block when the camera is used by this image alone);
//   every incidence has one record (rec_values below: the scaled Jacobian
//   rows of its block, the observation's scaled point Jacobian and
//   residual, and per damping try Y = X2 Cinv), in incidence order;
//   the reduced camera system S is dense, column-major, lower triangle,
//   (Dp + NB) x Dp with Dp = D rounded up to NB; padding is the identity and
//   row Dp carries the right-hand side, so the blocked Cholesky produces
//   L^{-1} rhs in that row.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sfmba.h"

namespace sfmba {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int NB = 128;           // dense factorization block
// Dense factor workspace per NB block: L^{-1} and its transpose (2 NB x NB),
// and, after all blocks, the panel-solve operands (36 16x16x4 MFMA operand
// sets of 4 x 64 doubles per block).
constexpr size_t LINV_BLOCK = 2 * (size_t)NB * NB;
constexpr size_t TOP_BLOCK = 36 * 4 * 64;
inline size_t dense_workspace_doubles(uint32_t Dp) {
  return (size_t)(Dp / NB) * (LINV_BLOCK + TOP_BLOCK);
}
// Mixed form: the smallest Marquardt damping at which a reduced system is
// assembled in FP32.  Each entry of an FP32-assembled system carries
// rounding of order 1e-7 of its diagonal, and over the D dimensions of the
// system that perturbs its small eigenvalues (the free scale of a problem
// gauged by one fixed pose has an eigenvalue of order lambda) by up to
// about sqrt(D) times that: on the 1,000-camera scene (D = 7,994) the first
// FP32 factorization fails at lambda = 1.4e-6.  Below this damping the
// single-problem solver continues in FP64; the window solver, whose
// precision is fixed per launch, damps FP32 systems at least this much (the
// LM schedule itself is unchanged).
constexpr double MIXED_MIN_LAMBDA = SFMBA_MIXED_MIN_LAMBDA;
constexpr int PAIR_CHUNK = 128;   // max point pairs reduced by one wavefront
constexpr uint32_t LINEARIZE_SPAN = 64;  // observations per linearization wavefront
constexpr int COST_BLOCK = 256;          // threads per block of the cost kernel
constexpr int COST_ITEMS = 8;            // observations per thread of the cost kernel
constexpr uint64_t SCHUR_TILE = 64; // destination tile edge (parameter blocks) of the chunk order

// Incidence record of the single-problem solve: everything the Schur
// assembly and the point back-substitution read of one incidence (one
// observation x one parameter block), in one place.  Values of T:
//   [0, 6)              X2 = sqrt(w) J_X, the observation's scaled point
//                       Jacobian (2 x 3, row-major)
//   [6, 8)              R2 = sqrt(w) r
//   [8, 8 + 2 DB)       J2 rows: row s = sqrt(w) (dr0/dx_s, dr1/dx_s) for
//                       the block's unknown s (rows past its dimension
//                       unused)
//   [8 + 2 DB, 14 + 2 DB)  Y = X2 Cinv (2 x 3, row-major), per damping try
//   [14 + 2 DB, 16 + 2 DB) z = Y g_point - R2, per damping try
// X2, R2 and J2 come from the linearization, Y and z from the damped point
// inverse.  A pair (a, b) of one point reads a's [8, 16 + 2 DB) and b's
// [0, 8 + 2 DB): two contiguous pieces of 8 + 2 DB values.  DB is the
// largest block dimension rounded up to 8, 12 or 16.  Records are stored in
// incidence (point-major) order, so the linearization, Y and the
// back-substitution stream them; measured against group-major records
// (which keep a destination block's records together for the pair
// gathers) on the 1,000-camera scene: the same assembly time and a 0.6%
// shorter solve.
constexpr uint32_t rec_values(uint32_t db) { return 16u + 2u * db; }
constexpr uint32_t rec_db_for(uint32_t max_grp_dim) {
  return max_grp_dim <= 8 ? 8u : max_grp_dim <= 12 ? 12u : 16u;
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw ::sfmba::Error(_e == hipErrorOutOfMemory ? SFMBA_ERR_OUT_OF_MEMORY           \
                                                     : SFMBA_ERR_DEVICE,                 \
                           std::string(#expr) + ": " + hipGetErrorString(_e));           \
  } while (0)

inline int model_param_count_host(int model) {
  switch (model) {
    case SFMBA_SIMPLE_PINHOLE: return 3;
    case SFMBA_PINHOLE: return 4;
    case SFMBA_SIMPLE_RADIAL: return 4;
    case SFMBA_RADIAL: return 5;
    case SFMBA_OPENCV: return 8;
    case SFMBA_OPENCV_FISHEYE: return 8;
    default: return -1;
  }
}

// Static structure of one problem on the device.
struct Layout {
  uint32_t n_img = 0, n_cam = 0, n_pt = 0, n_obs = 0, n_inc = 0, n_grp = 0;
  uint32_t D = 0, Dp = 0, ld = 0;
  uint64_t n_pairs = 0;
  uint32_t n_chunks = 0, n_dests = 0, n_multi = 0, n_parts = 0;
  uint32_t n_chunks_plain = 0;  // leading chunks without same-observation pairs
  uint32_t max_grp_dim = 0;
  uint32_t rec_db = 0;          // J2 rows per incidence record: 8, 12 or 16 (rec_values)
  // observations (point-major)
  uint32_t* obs_img = nullptr;
  uint32_t* obs_pt = nullptr;
  double2* obs_xy = nullptr;
  uint32_t* obs_inc = nullptr;  // 2 per obs, NONE if absent
  uint32_t* pt_ptr = nullptr;   // n_pt + 1
  uint32_t n_pt_fix = 0;
  uint32_t* pt_fix = nullptr;   // points with no observation or observations in two
                                // LINEARIZE_SPAN spans (summed by k_point_accumulate)
  // incidences
  uint32_t* inc_grp = nullptr;
  uint32_t* inc_pt = nullptr;
  uint32_t* inc_obs = nullptr;
  uint32_t* pt_inc_ptr = nullptr;  // incidences of each point (CSR, n_pt + 1)
  // cameras and images
  uint32_t* cam_of_img = nullptr;
  int32_t* cam_model = nullptr;
  uint32_t* img_slot = nullptr;   // first unknown of the pose, NONE if fixed
  uint32_t* cam_slot = nullptr;   // first unknown of the camera's free params
  uint8_t* cam_free = nullptr;    // 8 per camera: free param indices
  uint32_t* cam_nfree = nullptr;
  uint32_t* grp_off = nullptr;
  uint32_t* grp_dim = nullptr;
  // Schur pair lists (destination-sorted)
  uint32_t* pairs = nullptr;      // 2 per pair: incidences a and b (flag bits)
  uint4* chunks = nullptr;        // ((ga << 16) | gb, begin, end, part or NONE): the
                                  // n_chunks_plain chunks without same-observation pairs first
  uint32_t* multi_dest = nullptr; // destinations reduced from partials
  uint32_t* dest_part0 = nullptr; // first partial of each multi destination
  uint32_t* dest_nparts = nullptr;
  // host copies needed by the host-side LM loop
  std::vector<uint32_t> h_img_slot, h_cam_slot, h_cam_nfree, h_grp_off, h_grp_dim;
  std::vector<uint8_t> h_cam_free;
};

// Per-problem mutable state and scratch.
template <typename T>
struct Work {
  double* poses[2] = {nullptr, nullptr};   // 7 per image (w x y z tx ty tz)
  double* cams[2] = {nullptr, nullptr};    // 8 per camera
  double* pts[2] = {nullptr, nullptr};     // 3 per point
  T* rec = nullptr;            // rec_values(L.rec_db) per incidence, in incidence order
  T* obs_pc = nullptr;        // 9 per obs: w Jx^T Jx (6), w Jx^T r (3)
  uint8_t* obs_valid = nullptr;
  double* pt_C = nullptr;     // 6 per point (xx xy xz yy yz zz)
  double* pt_g = nullptr;     // 3 per point
  uint8_t* pt_active = nullptr;
  double* pt_cinv = nullptr;  // 9 per point
  uint8_t* pt_ok = nullptr;
  double* dpt = nullptr;      // 3 per point
  double* S = nullptr;        // ld x Dp
  double* S_copy = nullptr;   // refinement copy (mixed precision)
  double* Linv = nullptr;     // dense factor workspace (dense_workspace_doubles)
  double* delta = nullptr;    // Dp
  double* vec = nullptr;      // Dp scratch
  double* rhs = nullptr;      // Dp
  double* parts = nullptr;    // partial blocks of multi-chunk destinations
  double* inc_part = nullptr; // 3 per incidence: E^T delta (point back-substitution)
  double* cost_part = nullptr;
  double* cost = nullptr;     // [0] cost, [1] sum |r|, [2] count
  int* flags = nullptr;       // [0] cholesky failure, [1] lu failure, [2] back-solve hand-over timeout
  uint32_t* chain = nullptr;  // Dp / NB back-substitution flags
  uint32_t chain_epoch = 0;   // last epoch used with chain
  double* h_pinned = nullptr; // pinned host mirror of cost + flags
};

// ---- kernels (engine_kernels.hip) ----
template <typename T>
void launch_linearize(const Layout& L, Work<T>& W, int cur, int loss, double scale,
                      hipStream_t s);
template <typename T>
void launch_point_accumulate(const Layout& L, Work<T>& W, hipStream_t s);
template <typename T>
void launch_point_prepare(const Layout& L, Work<T>& W, double lambda, hipStream_t s);
template <typename T>
void launch_schur(const Layout& L, Work<T>& W, double lambda, hipStream_t s);
template <typename T>
void launch_backsub_update(const Layout& L, Work<T>& W, int cur, hipStream_t s);
void launch_cost(const Layout& L, const double* poses, const double* cams, const double* pts,
                 int loss, double scale, double* part, double* out, int mode, hipStream_t s);
// |residual| and in-front flag per observation, in the caller's order.
void launch_residual_norms(uint32_t n_obs, const sfmba_observation* obs, const uint64_t* cam_of_img,
                           const int32_t* cam_model, const double* poses, const double* cams,
                           const double* pts, double* norm, uint8_t* in_front, hipStream_t s);
// Host helpers shared by the entry points: the calling thread's error text,
// and |residual| / in-front flags of observations evaluated on `device`.
void set_error(const std::string& m);
void residual_norms(const std::vector<sfmba_observation>& obs, const std::vector<uint64_t>& coi,
                    const std::vector<sfmba_camera>& cams, const std::vector<sfmba_pose>& poses,
                    const std::vector<double>& pts, int device, std::vector<double>& norm,
                    std::vector<uint8_t>& front);
void launch_jacobians_debug(const Layout& L, const double* poses, const double* cams,
                            const double* pts, int fp32, double* r, double* jp, double* jx,
                            double* jc, uint8_t* valid, hipStream_t s);

// ---- intrinsics-only refinement (intrinsics_kernels.hip) ----
void launch_refine_intrinsics(uint32_t n_cams_refined, const uint32_t* cam_list,
                              const uint32_t* cam_obs_ptr, const uint32_t* obs_img,
                              const uint32_t* obs_pt, const double2* obs_xy, const double* poses,
                              const double* pts, const int32_t* cam_model, double* cams,
                              const sfmba_intrinsics_params* p, hipStream_t s);

// Per-launch timing of one kernel with HIP events on its stream (profiling).
struct KernelTimer {
  struct Rec {
    hipEvent_t a, b;
    double flop;
  };
  std::vector<Rec> recs;
  KernelTimer() = default;
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;
  ~KernelTimer() {
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
  }
  hipEvent_t before(hipStream_t s) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, s));
    return e;
  }
  void after(hipEvent_t a, hipStream_t s, double flop) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, s));
    recs.push_back(Rec{a, e, flop});
  }
  // Adds seconds, launches and FLOP of the recorded launches (stream synced by the caller).
  void collect(double& sec, uint64_t& n, double& flop) const {
    for (const auto& r : recs) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      sec += ms * 1e-3;
      ++n;
      flop += r.flop;
    }
  }
};

// ---- dense reduced-system solver (dense_kernels.hip) ----
// Cholesky of the lower triangle of S (ld x Dp) with the rhs row at Dp; on
// return row Dp holds y = L^{-1} rhs; flags[0] != 0 on a non-positive pivot.
// kt (optional) times every trailing-update launch.
void dense_cholesky(double* S, double* Linv, uint32_t Dp, uint32_t ld, uint32_t D, int* flags,
                    hipStream_t s, KernelTimer* kt = nullptr);
// Solve L^T x = y (Dp entries each) in one launch.  chain holds Dp / NB
// flags (zero at allocation); epoch must differ from every value passed
// before with the same chain (a per-buffer counter, never 0).  status
// receives a non-zero value if a block's hand-over timed out.
void dense_backsolve(const double* S, const double* Linv, uint32_t Dp, uint32_t ld, const double* y,
                     double* x, uint32_t* chain, uint32_t epoch, int* status, hipStream_t s);
// The next epoch of a chain buffer (skips 0, the value at allocation).
inline uint32_t next_epoch(uint32_t& e) {
  if (++e == 0) e = 1;
  return e;
}
// Solve L y = b; b (Dp) is consumed as scratch, y (Dp) receives the result.
void dense_forwardsolve(const double* S, const double* Linv, uint32_t Dp, uint32_t ld, double* b,
                        double* y, hipStream_t s);
// r = b - A x with A symmetric given by its lower triangle (ld x Dp).
void dense_symv_residual(const double* A, uint32_t Dp, uint32_t ld, const double* x,
                         const double* b, double* r, hipStream_t s);
void dense_copy(const double* src, double* dst, size_t n, hipStream_t s);
// dst (full symmetric, ld x Dp) <- lower triangle of src; src may equal dst.
void dense_mirror_copy(const double* src, double* dst, uint32_t Dp, uint32_t ld, hipStream_t s);
void dense_vec_add(double* x, const double* d, uint32_t n, hipStream_t s);
void dense_extract_row(const double* S, uint32_t Dp, uint32_t ld, uint32_t row, double* out,
                       hipStream_t s);
// LU with partial pivoting of the full symmetric matrix given by its lower
// triangle (mirrored first), then solve; rhs read from row Dp.
void dense_lu_solve(double* S, uint32_t Dp, uint32_t ld, uint32_t D, double* x, int* piv,
                    int* flags, hipStream_t s);

}  // namespace sfmba
