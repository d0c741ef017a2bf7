/* CPU oracle of the Redwood depth noise model (test infrastructure only:
 * only tests/ and the benchmark's CPU baseline load it).
 *
 * Restates src/esp/sensor/RedwoodNoiseModel.cu:17-101 (the GPU model the
 * batch path's depth would feed; the numba CPU model of
 * src_python/habitat_sim/sensors/noise_models/redwood_depth_noise_model.py:31-98
 * computes the same in float64):
 *   per output pixel (j, i) of an H x W depth image, with standard normal
 *   deviates n0, n1, n2 and noise multiplier m,
 *   y = (int)(min(max(j + n0*0.25*m, 0), H-1) + 0.5)     (float)
 *   x = (int)(min(max(i + n1*0.25*m, 0), W-1) + 0.5)
 *   d = depth[y - y%2][x - x%2]                           (downsample)
 *   d >= 10 -> 0 (also d < 0 or NaN, where the reference reads out of bounds)
 *   u = undistort((int)(x/(W-1)*639 + 0.5), (int)(y/(H-1)*479 + 0.5), d)
 *   u == 0 -> 0
 *   q = round((35.130 / (double)u + n2*0.027778*m) * 8)    (double, half away)
 *   out = q > 1e-5 ? (float)(35.130f*8 / q) : 0
 * undistort (RedwoodNoiseModel.cu:17-36): i2 = (int)((z+1)/2), i1 = i2-1,
 *   a = (z - (2*i1 + 1))/2, f = (1-a)*M[y/6][x/8][clamp(i1,0,4)]
 *   + a*M[y/6][x/8][min(i2,4)], z/f, 0 when f < 1e-5.
 *
 * The normals come from the caller here; the HIP kernel draws them from a
 * Philox4x32-10 counter generator (Salmon et al., SC'11), restated below as
 * orc_philox4x32 so that its raw output can be compared bit for bit.
 */
#include <math.h>
#include <stdint.h>

#define RW_COLS 80
#define RW_DIMS 5

static float rw_undistort(int x_, int y_, float z, const float* model) {
  const int i2 = (int)((z + 1.0f) / 2.0f);
  const int i1 = i2 - 1;
  float a = (float)i1 * 2.0f;
  a = a + 1.0f;
  a = z - a;
  a = a / 2.0f;
  const int x = x_ / 8, y = y_ / 6;
  const float* row = model + (y * RW_COLS + x) * RW_DIMS;
  int k1 = i1 < 0 ? 0 : i1;
  if (k1 > 4) k1 = 4;
  const int k2 = i2 < 4 ? i2 : 4;
  const float wa = 1.0f - a;
  const float p = wa * row[k1];
  const float q = a * row[k2];
  const float f = p + q;
  if (f < 1e-5f) return 0.0f;
  return z / f;
}

/* one pixel; n = the pixel's three deviates */
static float rw_pixel(const float* depth, int H, int W, int idx, const float* model, float mult,
                      const float* n) {
  const float ymax = (float)(H - 1), xmax = (float)(W - 1);
  float fy = n[0] * 0.25f;
  fy = fy * mult;
  fy = (float)(idx / W) + fy;
  fy = fmaxf(fy, 0.0f);
  fy = fminf(fy, ymax);
  const int y = (int)(fy + 0.5f);
  float fx = n[1] * 0.25f;
  fx = fx * mult;
  fx = (float)(idx % W) + fx;
  fx = fmaxf(fx, 0.0f);
  fx = fminf(fx, xmax);
  const int x = (int)(fx + 0.5f);
  const float d = depth[(y - y % 2) * W + x - x % 2];
  /* beyond 10 m the sensor returns 0; so do negative and NaN depths here
     (the reference indexes its table out of bounds for them) */
  if (!(d < 10.0f) || d < 0.0f) return 0.0f;
  float ux = (float)x / xmax;
  ux = ux * 639.0f;
  ux = ux + 0.5f;
  float uy = (float)y / ymax;
  uy = uy * 479.0f;
  uy = uy + 0.5f;
  const float u = rw_undistort((int)ux, (int)uy, d, model);
  if (u == 0.0f) return 0.0f;
  float nz = n[2] * 0.027778f;
  nz = nz * mult;
  double t = (double)35.130f / (double)u;
  t = t + (double)nz;
  t = t * 8.0;
  const double q = round(t);
  if (!(q > 1e-5)) return 0.0f;
  const float num = 35.130f * 8.0f;
  return (float)((double)num / q);
}

/* images x H x W depth (float, consecutive images) -> out; normals
   [images*H*W][3] */
void orc_redwood(const float* depth, int images, int H, int W, const float* model, float mult,
                 const float* normals, float* out) {
  const long long npx = (long long)H * W;
#pragma omp parallel for schedule(static)
  for (long long g = 0; g < (long long)images * npx; ++g) {
    const long long img = g / npx;
    const int idx = (int)(g % npx);
    out[g] = rw_pixel(depth + img * npx, H, W, idx, model, mult, normals + 3 * g);
  }
}

/* Philox4x32-10: 10 rounds of (hi0, lo0) = M0*c0, (hi1, lo1) = M1*c2,
   c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0), key += (W0, W1) */
void orc_philox4x32(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
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
