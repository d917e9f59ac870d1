// Total cost of a device-resident state with no device allocation
// (sfmba_total_cost_noalloc).  For callers that cannot allocate device
// memory: code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc.  The per-block partial sums and the result live in one
// file-scope device array of fixed size, `scratch`; every GPU holds its own
// copy until the process ends, and the host reaches it only through the
// address hipGetSymbolAddress returns for that device.
//
// The cost is the reference's total_cost (crates/sfm-ba/src/lib.rs:727-745)
// evaluated directly on the caller's arrays in the ABI's own layout
// (sfmba_observation, sfmba_pose, sfmba_camera, x/y/z points), in the
// caller's observation order: sum over observations in front of their
// camera (z > 1e-9) of w(|r|) |r|^2.  Index checks run in the kernel (an
// out-of-range observation is skipped, flagged, and the call fails with
// SFMBA_ERR_INVALID_ARGUMENT, as sfmba_total_cost does).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "camera_models.hpp"
#include "engine.hpp"

namespace sfmba {

namespace {

constexpr int NA_BLOCK = 256;                   // threads per block of the partial-sum kernel
constexpr int NA_ITEMS = 8;                     // observations per thread
constexpr uint32_t NA_PER_BLOCK = NA_BLOCK * NA_ITEMS;
constexpr uint64_t NA_MAX_OBS = SFMBA_NOALLOC_MAX_OBSERVATIONS;
constexpr uint32_t NA_MAX_BLOCKS = (uint32_t)(NA_MAX_OBS / NA_PER_BLOCK);
static_assert(NA_MAX_OBS % NA_PER_BLOCK == 0, "the limit is a whole number of blocks");
// scratch layout (doubles): [0, NA_MAX_BLOCKS) partial sums, then the
// result, then the index-check flag (an int in the next slot).
constexpr uint32_t NA_RESULT = NA_MAX_BLOCKS;
constexpr uint32_t NA_FLAG = NA_MAX_BLOCKS + 1;
constexpr uint32_t NA_SCRATCH = NA_MAX_BLOCKS + 2;

static __device__ __align__(256) double scratch[NA_SCRATCH];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[NA_MAX_DEVICES] = {};

// The device address of `scratch` on `device` (current), looked up once.
double* scratch_on(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES)
    throw Error(SFMBA_ERR_DEVICE, "device ordinal beyond the no-allocation entry point's table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(SFMBA_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<double*>(p);
  }
  return g_ptr[device];
}

__global__ void k_na_clear(double* __restrict__ sc) {
  if (threadIdx.x == 0) {
    sc[NA_RESULT] = 0.0;
    reinterpret_cast<int*>(sc + NA_FLAG)[0] = 0;
  }
}

// Partial sums of w(|r|) |r|^2, NA_ITEMS observations per thread, one
// fixed-order tree per block.
__global__ __launch_bounds__(NA_BLOCK) void k_na_cost(uint64_t n_obs, uint64_t n_img, uint64_t n_cam, uint64_t n_pt,
                                                      const sfmba_observation* __restrict__ obs,
                                                      const uint64_t* __restrict__ cam_of_img,
                                                      const sfmba_camera* __restrict__ cams,
                                                      const sfmba_pose* __restrict__ poses,
                                                      const double* __restrict__ pts, int loss, double scale,
                                                      double* __restrict__ sc) {
  __shared__ double red[NA_BLOCK];
  double acc = 0.0;
  bool bad = false;
  const uint64_t base = (uint64_t)blockIdx.x * NA_PER_BLOCK + threadIdx.x;
#pragma unroll
  for (int it = 0; it < NA_ITEMS; ++it) {
    const uint64_t o = base + (uint64_t)it * NA_BLOCK;
    if (o >= n_obs) continue;
    const sfmba_observation ob = obs[o];
    if (ob.image_idx >= n_img || ob.point_idx >= n_pt) {
      bad = true;
      continue;
    }
    const uint64_t c = cam_of_img[ob.image_idx];
    if (c >= n_cam) {
      bad = true;
      continue;
    }
    const int model = cams[c].model_id;
    if (model < SFMBA_SIMPLE_PINHOLE || model > SFMBA_OPENCV_FISHEYE) {
      bad = true;
      continue;
    }
    const double* X = pts + 3 * ob.point_idx;
    double c0, c1, c2;
    transform_point(&poses[ob.image_idx].qw, X[0], X[1], X[2], c0, c1, c2);
    if (c2 <= 1e-9) continue;  // behind the camera: no cost (lib.rs:733-735)
    double r0, r1;
    project_residual(model, cams[c].params, c0, c1, c2, ob.x, ob.y, r0, r1);
    const double n = sqrt(r0 * r0 + r1 * r1);
    acc += robust_weight(n, loss, scale) * n * n;
  }
  if (bad) atomicOr(reinterpret_cast<int*>(sc + NA_FLAG), 1);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = NA_BLOCK / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[blockIdx.x] = red[0];
}

// The partial sums in a fixed order: strided per thread, then a tree.
__global__ __launch_bounds__(1024) void k_na_final(uint32_t n_part, double* __restrict__ sc) {
  __shared__ double red[1024];
  double a = 0.0;
  for (uint32_t i = threadIdx.x; i < n_part; i += 1024) a += sc[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[NA_RESULT] = red[0];
}

void total_cost_noalloc(const sfmba_input* in, const sfmba_params* prm, hipStream_t s, double* out) {
  if (!in || !prm || !out) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "NULL argument");
  if (in->num_observations > NA_MAX_OBS)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT,
                "num_observations " + std::to_string(in->num_observations) +
                    " exceeds the no-allocation limit SFMBA_NOALLOC_MAX_OBSERVATIONS (" +
                    std::to_string(NA_MAX_OBS) + ")");
  if (in->num_observations &&
      (!in->observations || !in->poses || !in->points || !in->cameras || !in->camera_of_image))
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "input arrays are NULL");
  if (in->num_camera_of_image < in->num_images)
    throw Error(SFMBA_ERR_INVALID_ARGUMENT, "camera_of_image shorter than poses");
  int device = 0;
  HIP_CHECK(hipGetDevice(&device));
  double* sc = scratch_on(device);
  const uint32_t nb = (uint32_t)((in->num_observations + NA_PER_BLOCK - 1) / NA_PER_BLOCK);
  hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, s, sc);
  HIP_CHECK(hipGetLastError());
  if (nb) {
    hipLaunchKernelGGL(k_na_cost, dim3(nb), dim3(NA_BLOCK), 0, s, (uint64_t)in->num_observations,
                       (uint64_t)in->num_images, (uint64_t)in->num_cameras, (uint64_t)in->num_points,
                       in->observations, in->camera_of_image, in->cameras, in->poses, in->points,
                       prm->robust_loss, prm->loss_scale_px, sc);
    HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_na_final, dim3(1), dim3(1024), 0, s, nb, sc);
  HIP_CHECK(hipGetLastError());
  double h[2] = {0.0, 0.0};
  HIP_CHECK(hipMemcpyAsync(h, sc + NA_RESULT, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  int flag = 0;
  static_assert(sizeof(int) <= sizeof(double), "flag slot");
  __builtin_memcpy(&flag, &h[1], sizeof(int));
  if (flag) throw Error(SFMBA_ERR_INVALID_ARGUMENT, "an observation, image or camera index is out of range");
  *out = h[0];
}

}  // namespace

}  // namespace sfmba

int sfmba_total_cost_noalloc(const sfmba_input* device_input, const sfmba_params* params, void* stream,
                             double* out) {
  try {
    sfmba::set_error("");
    sfmba::total_cost_noalloc(device_input, params, static_cast<hipStream_t>(stream), out);
    return SFMBA_OK;
  } catch (const sfmba::Error& e) {
    sfmba::set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    sfmba::set_error(e.what());
    return SFMBA_ERR_DEVICE;
  }
}
