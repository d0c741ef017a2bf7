// Redwood depth noise model on the GPU (src/esp/sensor/RedwoodNoiseModel.cu,
// RedwoodNoiseModel.h; Python face redwood_depth_noise_model.py): pixel
// shuffle by N(0, 0.25 m) in x and y, 2x2 downsample, the 80x80x5
// undistortion table of the 640x480 Redwood sensor, disparity quantisation
// to 1/8 with N(0, 0.027778 m) noise, 0 beyond 10 m.
//
// One thread per output pixel of a batch of images.  The three standard
// normal deviates of a pixel come either from the caller or from a
// counter-based generator: one Philox4x32-10 block (Salmon et al., "Parallel
// random numbers: as easy as 1, 2, 3", SC'11) per (call, pixel) with the
// model's seed as key, turned into normals by the Box-Muller transform.  A
// counter generator needs no per-thread state in memory (the reference keeps
// a curand state per thread) and makes every call reproducible from (seed,
// call index).  The kernel is HBM-bound: 4 bytes read (the gathered depth
// mostly hits the cache lines of its neighbours) and 4 written per pixel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <mutex>
#include <new>
#include <string>

#include "gbr.h"
#include "gbr_host.h"

#pragma STDC FP_CONTRACT OFF

namespace gbr {
namespace {

constexpr int kRwCols = 80, kRwDims = 5;
constexpr int kNoiseBlock = 256;

__device__ __forceinline__ void philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t k0, uint32_t k1, uint32_t* out) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// uniform in (0, 1): the top 24 bits, centred in their interval
__device__ __forceinline__ float unit_open(uint32_t x) {
  return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// three N(0, 1) deviates of pixel `g` in call `call`
__device__ __forceinline__ void pixel_normals(uint64_t seed, uint64_t call, uint64_t g,
                                              float* n) {
  uint32_t r[4];
  philox4x32((uint32_t)g, (uint32_t)(g >> 32), (uint32_t)call, (uint32_t)(call >> 32),
             (uint32_t)seed, (uint32_t)(seed >> 32), r);
  // Box-Muller with the hardware transcendental instructions (~1 ulp;
  // deviates need a distribution, not correctly rounded values):
  // -2 ln u = -2 ln2 log2 u; v_sin/v_cos take the angle in revolutions
  const float m2ln2 = -1.38629436111989061883f;
  const float ra = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(unit_open(r[0])));
  const float rb = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(unit_open(r[2])));
  const float ta = unit_open(r[1]), tb = unit_open(r[3]);
  n[0] = ra * __builtin_amdgcn_cosf(ta);
  n[1] = ra * __builtin_amdgcn_sinf(ta);
  n[2] = rb * __builtin_amdgcn_cosf(tb);
}

__device__ __forceinline__ float undistort(int x_, int y_, float z, const float* __restrict__ model) {
  const int i2 = (int)((z + 1.0f) / 2.0f);
  const int i1 = i2 - 1;
  float a = (float)i1 * 2.0f;
  a = a + 1.0f;
  a = z - a;
  a = a / 2.0f;
  const int x = x_ / 8, y = y_ / 6;
  const float* row = model + (y * kRwCols + x) * kRwDims;
  const int k1 = min(max(i1, 0), 4), k2 = min(i2, 4);
  const float wa = 1.0f - a;
  const float p = wa * row[k1];
  const float q = a * row[k2];
  const float f = p + q;
  if (f < 1e-5f) return 0.0f;
  return z / f;
}

// images x rows x cols, consecutive; normals: 3 per pixel or null (generator)
__global__ __launch_bounds__(kNoiseBlock) void k_redwood(const float* __restrict__ depth,
                                                         long long total, int H, int W,
                                                         const float* __restrict__ model, float mult,
                                                         const float* __restrict__ normals,
                                                         uint64_t seed, uint64_t call,
                                                         float* __restrict__ out) {
  const long long g = (long long)blockIdx.x * kNoiseBlock + threadIdx.x;
  if (g >= total) return;
  const long long npx = (long long)H * W;
  const long long img = g / npx;
  const int idx = (int)(g - img * npx);
  const float* dimg = depth + img * npx;
  float n[3];
  if (normals) {
    n[0] = normals[3 * g];
    n[1] = normals[3 * g + 1];
    n[2] = normals[3 * g + 2];
  } else {
    pixel_normals(seed, call, (uint64_t)g, n);
  }
  const float ymax = (float)(H - 1), xmax = (float)(W - 1);
  const int j = idx / W, i = idx - j * W;
  float fy = n[0] * 0.25f;
  fy = fy * mult;
  fy = (float)j + fy;
  fy = fminf(fmaxf(fy, 0.0f), ymax);
  const int y = (int)(fy + 0.5f);
  float fx = n[1] * 0.25f;
  fx = fx * mult;
  fx = (float)i + fx;
  fx = fminf(fmaxf(fx, 0.0f), xmax);
  const int x = (int)(fx + 0.5f);
  const float d = dimg[(y - y % 2) * W + x - x % 2];
  float res = 0.0f;
  // beyond 10 m the sensor returns 0; negative and NaN depths give 0 too
  if (d < 10.0f && !(d < 0.0f)) {
    float ux = (float)x / xmax;
    ux = ux * 639.0f;
    ux = ux + 0.5f;
    float uy = (float)y / ymax;
    uy = uy * 479.0f;
    uy = uy + 0.5f;
    const float u = undistort((int)ux, (int)uy, d, model);
    if (u != 0.0f) {
      float nz = n[2] * 0.027778f;
      nz = nz * mult;
      double t = (double)35.130f / (double)u;
      t = t + (double)nz;
      t = t * 8.0;
      const double q = round(t);
      if (q > 1e-5) res = (float)((double)(35.130f * 8.0f) / q);
    }
  }
  out[g] = res;
}

__global__ __launch_bounds__(kNoiseBlock) void k_philox_bits(uint64_t seed, uint64_t call,
                                                             uint64_t first, uint64_t n,
                                                             uint32_t* __restrict__ out) {
  const uint64_t k = (uint64_t)blockIdx.x * kNoiseBlock + threadIdx.x;
  if (k >= n) return;
  const uint64_t g = first + k;
  uint32_t r[4];
  philox4x32((uint32_t)g, (uint32_t)(g >> 32), (uint32_t)call, (uint32_t)(call >> 32),
             (uint32_t)seed, (uint32_t)(seed >> 32), r);
  uint4 v;
  v.x = r[0];
  v.y = r[1];
  v.z = r[2];
  v.w = r[3];
  reinterpret_cast<uint4*>(out)[k] = v;
}

}  // namespace

// Staging of gbr_redwood_simulate_host_noalloc(): the input depth image in
// the first kRwScratchPixels floats, the output image in the second.  Every
// GPU the library runs on holds its own copy (8 MiB) until the process ends.
// Host code never names the array in a launch or copy: it uses the per-device
// address from hipGetSymbolAddress (scratch_address below).
constexpr long long kRwScratchPixels = GBR_REDWOOD_NOALLOC_MAX_PIXELS;
static __device__ __align__(256) float scratch[2 * kRwScratchPixels];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch[kScratchDevices] = {};

// the device address of `scratch` on device `dev`, which must be current;
// looked up once per device
gbr_status scratch_address(int dev, float** out) {
  if (dev < 0 || dev >= kScratchDevices)
    return fail(GBR_ERR_UNSUPPORTED, "redwood_create: device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GBR_ERR_DEVICE, std::string("redwood_create: hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
    g_scratch[dev] = static_cast<float*>(p);
  }
  *out = g_scratch[dev];
  return GBR_OK;
}

}  // namespace
}  // namespace gbr

struct gbr_redwood {
  int device = 0;
  float mult = 1.0f;
  uint64_t seed = 0;
  uint64_t calls = 0;  // generator calls so far (the counter's high word)
  gbr::DevBuf<float> d_model;
  gbr::DevBuf<float> d_in, d_out;  // staging of the host entry point
  float* scratch = nullptr;        // this device's `scratch` (no-allocation entry point)
};

using gbr::fail;

namespace {

gbr_status check_dims(int32_t images, int32_t rows, int32_t cols, const char* fn) {
  if (images < 0 || rows < 2 || cols < 2)
    return fail(GBR_ERR_ARGUMENT, std::string(fn) + ": need images >= 0 and images of at least 2x2 pixels");
  if ((long long)rows * cols >= (1ll << 31))
    return fail(GBR_ERR_SIZE, std::string(fn) + ": image of 2^31 pixels or more");
  return GBR_OK;
}

gbr_status launch(gbr_redwood* m, const float* depth, int32_t images, int32_t rows, int32_t cols,
                  const float* normals, float* out, hipStream_t s) {
  const long long total = (long long)images * rows * cols;
  if (total == 0) return GBR_OK;
  const long long blocks = (total + gbr::kNoiseBlock - 1) / gbr::kNoiseBlock;
  if (blocks > 0x7fffffffll) return fail(GBR_ERR_SIZE, "redwood: batch too large");
  hipLaunchKernelGGL(gbr::k_redwood, dim3((uint32_t)blocks), dim3(gbr::kNoiseBlock), 0, s, depth,
                     total, rows, cols, m->d_model.ptr, m->mult, normals, m->seed, m->calls, out);
  HIP_TRY(hipGetLastError());
  if (!normals) ++m->calls;
  return GBR_OK;
}

}  // namespace

extern "C" {

gbr_status gbr_redwood_create(const float* model, int32_t device, float noise_multiplier,
                              uint64_t seed, gbr_redwood** out) {
  if (!model || !out) return fail(GBR_ERR_ARGUMENT, "redwood_create: null argument");
  *out = nullptr;
  if (!std::isfinite(noise_multiplier))
    return fail(GBR_ERR_ARGUMENT, "redwood_create: noise multiplier is not finite");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  int dev = device;
  if (dev < 0) HIP_TRY(hipGetDevice(&dev));
  if (dev >= ndev) return fail(GBR_ERR_ARGUMENT, "redwood_create: device ordinal out of range");
  HIP_TRY(hipSetDevice(dev));
  gbr_redwood* m = new (std::nothrow) gbr_redwood();
  if (!m) return fail(GBR_ERR_OOM, "redwood_create: out of host memory");
  m->device = dev;
  m->mult = noise_multiplier;
  m->seed = seed;
  const size_t n = (size_t)GBR_REDWOOD_MODEL_FLOATS;
  gbr_status st = gbr::scratch_address(dev, &m->scratch);
  if (st == GBR_OK) st = m->d_model.reserve(n);
  if (st == GBR_OK) {
    const hipError_t e = hipMemcpy(m->d_model.ptr, model, n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) st = fail(GBR_ERR_DEVICE, std::string("redwood_create: ") + hipGetErrorString(e));
  }
  if (st != GBR_OK) {
    m->d_model.release();
    delete m;
    return st;
  }
  *out = m;
  return GBR_OK;
}

gbr_status gbr_redwood_destroy(gbr_redwood* m) {
  if (!m) return GBR_OK;
  (void)hipSetDevice(m->device);
  m->d_model.release();
  m->d_in.release();
  m->d_out.release();
  delete m;
  return GBR_OK;
}

gbr_status gbr_redwood_simulate(gbr_redwood* m, const float* depth, int32_t images, int32_t rows,
                                int32_t cols, float* out, void* stream) {
  if (!m || !depth || !out) return fail(GBR_ERR_ARGUMENT, "redwood_simulate: null argument");
  TRY(check_dims(images, rows, cols, "redwood_simulate"));
  HIP_TRY(hipSetDevice(m->device));
  return launch(m, depth, images, rows, cols, nullptr, out, static_cast<hipStream_t>(stream));
}

gbr_status gbr_redwood_simulate_normals(gbr_redwood* m, const float* depth, int32_t images,
                                        int32_t rows, int32_t cols, const float* normals,
                                        float* out, void* stream) {
  if (!m || !depth || !normals || !out)
    return fail(GBR_ERR_ARGUMENT, "redwood_simulate_normals: null argument");
  TRY(check_dims(images, rows, cols, "redwood_simulate_normals"));
  HIP_TRY(hipSetDevice(m->device));
  return launch(m, depth, images, rows, cols, normals, out, static_cast<hipStream_t>(stream));
}

gbr_status gbr_redwood_simulate_host(gbr_redwood* m, const float* depth, int32_t rows, int32_t cols,
                                     float* out) {
  if (!m || !depth || !out) return fail(GBR_ERR_ARGUMENT, "redwood_simulate_host: null argument");
  TRY(check_dims(1, rows, cols, "redwood_simulate_host"));
  HIP_TRY(hipSetDevice(m->device));
  const size_t n = (size_t)rows * cols;
  TRY(m->d_in.reserve(n));
  TRY(m->d_out.reserve(n));
  HIP_TRY(hipMemcpy(m->d_in.ptr, depth, n * sizeof(float), hipMemcpyHostToDevice));
  TRY(launch(m, m->d_in.ptr, 1, rows, cols, nullptr, m->d_out.ptr, nullptr));
  HIP_TRY(hipMemcpy(out, m->d_out.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
  return GBR_OK;
}

gbr_status gbr_redwood_simulate_host_noalloc(gbr_redwood* m, const float* depth, int32_t rows,
                                             int32_t cols, float* out) {
  if (!m || !depth || !out)
    return fail(GBR_ERR_ARGUMENT, "redwood_simulate_host_noalloc: null argument");
  TRY(check_dims(1, rows, cols, "redwood_simulate_host_noalloc"));
  const long long n = (long long)rows * cols;
  if (n > gbr::kRwScratchPixels)
    return fail(GBR_ERR_SIZE, "redwood_simulate_host_noalloc: rows x cols = " + std::to_string(n) +
                                  " exceeds GBR_REDWOOD_NOALLOC_MAX_PIXELS (" +
                                  std::to_string(gbr::kRwScratchPixels) + ")");
  if (!m->scratch) return fail(GBR_ERR_STATE, "redwood_simulate_host_noalloc: no scratch address");
  HIP_TRY(hipSetDevice(m->device));
  // the input region is written whole before the kernel reads it, and the
  // kernel writes every output pixel: nothing in the scratch needs clearing
  float* in = m->scratch;
  float* res = m->scratch + gbr::kRwScratchPixels;
  HIP_TRY(hipMemcpy(in, depth, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  TRY(launch(m, in, 1, rows, cols, nullptr, res, nullptr));
  HIP_TRY(hipMemcpy(out, res, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  return GBR_OK;
}

gbr_status gbr_redwood_calls(const gbr_redwood* m, uint64_t* out) {
  if (!m || !out) return fail(GBR_ERR_ARGUMENT, "redwood_calls: null argument");
  *out = m->calls;
  return GBR_OK;
}

gbr_status gbr_redwood_random_bits(gbr_redwood* m, uint64_t call, uint64_t first, uint64_t n,
                                   uint32_t* out, void* stream) {
  if (!m || (!out && n)) return fail(GBR_ERR_ARGUMENT, "redwood_random_bits: null argument");
  if (n == 0) return GBR_OK;
  const uint64_t blocks = (n + gbr::kNoiseBlock - 1) / gbr::kNoiseBlock;
  if (blocks > 0x7fffffffull) return fail(GBR_ERR_SIZE, "redwood_random_bits: too many blocks");
  HIP_TRY(hipSetDevice(m->device));
  hipLaunchKernelGGL(gbr::k_philox_bits, dim3((uint32_t)blocks), dim3(gbr::kNoiseBlock), 0,
                     static_cast<hipStream_t>(stream), m->seed, call, first, n, out);
  HIP_TRY(hipGetLastError());
  return GBR_OK;
}

}  // extern "C"
