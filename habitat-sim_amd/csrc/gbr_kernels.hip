// HIP kernels of the batch renderer for gfx950 (CDNA4, wave64).
//
// Pipeline per frame (all scenes in one launch per stage, on one stream):
//   k_nodes       per node: absolute transform = parent chain product
//                 (Renderer.cpp:1214-1220 moved to the GPU); resets the
//                 frame's counters
//   k_cull        culled modes, per work item (<= 1024 consecutive triangles
//                 of one draw): box-vs-frustum test with a float error bound,
//                 conservative pixel rectangle and nearest depth, phase-1 /
//                 candidate lists, learned mode: 32 feature rows for the
//                 visibility MLP (k_mlp_score in gbr_mlp.hip, then k_propose)
//   k_setup_bin   persistent, per work item from a device queue: vertex
//                 transform of the item's vertex range into LDS (coalesced
//                 SoA position streams), fixed-point triangle setup (8
//                 sub-pixel bits, top-left rule), trivial/back-face/empty
//                 rejection, per-triangle raster and resolve records, bin
//                 append with one global atomic per touched bin and item
//   k_setup_clip  triangles crossing the near/far plane or the guard band:
//                 Sutherland-Hodgman fan, each fan triangle binned on its own
//   k_raster_lane per (scene, bin): one binned triangle per lane, large
//                 rectangles from an LDS queue by the whole workgroup; depth
//                 test as a 64-bit LDS atomicMin of (window depth bits,
//                 primitive order); visibility keys to HBM, or, in the
//                 frame's last raster phase of an unlit renderer, the bin
//                 resolved in place from its LDS keys (k_raster_pairs:
//                 selectable pair-parallel form; k_raster_full: exact rescan
//                 of overflowed bins)
//   k_hiz, k_cull_hiz  two-phase culling: depth pyramid of the phase-1 keys,
//                 then phase-2 candidates not hidden by it go through setup
//                 and raster again
//   resolve_tile  per bin: each pixel once, key -> record -> perspective-
//                 correct attributes, mip-mapped texture sampling, RGBA8 /
//                 window depth / linear depth / semantic written once to HBM
//                 (inside the last raster phase, or k_resolve from the
//                 stored keys for lit renderers and k_raster_pairs)
//   k_gather      framebuffer-layout readback of a rectangle
//
// Semantics follow OpenGL as driven by the reference (Renderer.cpp:307-308,
// 1298-1351; RendererStandalone.cpp:155-159; DepthUnprojection.cpp:98-130);
// every float expression is evaluated in the same order as the CPU oracle
// so results agree bit for bit (the file is compiled with -ffp-contract=off).

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "gbr.h"
#include "gbr_device.h"

#pragma clang fp contract(off)

namespace gbr {

#define GL_NEAREST 9728
#define GL_LINEAR 9729
#define GL_NEAREST_MIPMAP_NEAREST 9984
#define GL_LINEAR_MIPMAP_NEAREST 9985
#define GL_NEAREST_MIPMAP_LINEAR 9986
#define GL_LINEAR_MIPMAP_LINEAR 9987
#define GL_CLAMP_TO_EDGE 33071
#define GL_MIRRORED_REPEAT 33648

constexpr unsigned long long kClearKey = 0x3F800000FFFFFFFFull;  // depth 1.0, no primitive
constexpr uint32_t kClearColor = 0xFF1F1F1Fu;                      // 0x1f1f1f, alpha 255

// --------------------------------------------------------------------------
// Arithmetic shared by setup, raster and resolve (op order = the oracle's)
// --------------------------------------------------------------------------

__device__ __forceinline__ void mat4_vec(const float* m, float x, float y, float z, float w,
                                         float* out) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float s = m[0 * 4 + r] * x;
    float p1 = m[1 * 4 + r] * y;
    s = s + p1;
    float p2 = m[2 * 4 + r] * z;
    s = s + p2;
    float p3 = m[3 * 4 + r] * w;
    s = s + p3;
    out[r] = s;
  }
}

// clip = PV * (M * (p, 1))
__device__ __forceinline__ float4 vertex_clip(const float* M, const float* PV, float x, float y,
                                              float z) {
  float q[4], c[4];
  mat4_vec(M, x, y, z, 1.0f, q);
  mat4_vec(PV, q[0], q[1], q[2], q[3], c);
  return make_float4(c[0], c[1], c[2], c[3]);
}

// a * b for operands that fit 24 signed bits, as one full-rate
// v_mul_i32_i24 (the compiler would otherwise emit the quarter-rate 32-bit
// multiply where it cannot prove the ranges)
__device__ __forceinline__ int32_t imul24(int32_t a, int32_t b) {
  int32_t r;
  asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// l / n for 0 <= l < 4096 and 1 <= n <= 64 from a float reciprocal
// inv = 1/n: (l + 0.5) / n lies at least 0.5 / n >= 2^-7 away from an
// integer and the float product errs by less than 2^-10 there, so the
// truncation is exact (an integer division costs ~25 VALU operations).
__device__ __forceinline__ int div_small(int l, float inv) {
  return (int)(((float)l + 0.5f) * inv);
}

struct Snap {
  int32_t X, Y;
  float z;
  uint32_t flags;  // bits 0-5 outcodes, bit 6 needs clipping
};

constexpr uint32_t kNeedClip = 1u << 6;

__device__ __forceinline__ uint32_t outcode(float4 c) {
  float nw = -c.w;
  uint32_t oc = 0;
  if (c.x > c.w) oc |= 1;
  if (c.x < nw) oc |= 2;
  if (c.y > c.w) oc |= 4;
  if (c.y < nw) oc |= 8;
  if (c.z < nw) oc |= 16;
  if (c.z > c.w) oc |= 32;
  return oc;
}

__device__ __forceinline__ float plane_dist(float4 v, int plane, float gx, float gy) {
  switch (plane) {
    case 0: return v.z + v.w;
    case 1: return v.w - v.z;
    case 2: { float g = gx * v.w; return g - v.x; }
    case 3: { float g = gx * v.w; return v.x + g; }
    case 4: { float g = gy * v.w; return g - v.y; }
    default: { float g = gy * v.w; return v.y + g; }
  }
}

__device__ __forceinline__ void project(float4 v, float hw, float hh, int32_t& X, int32_t& Y,
                                        float& Z) {
  float iw = 1.0f / v.w;
  float nx = v.x * iw, ny = v.y * iw, nz = v.z * iw;
  float xw = nx * hw;
  xw = xw + hw;
  float yw = ny * hh;
  yw = yw + hh;
  float zw = nz * 0.5f;
  zw = zw + 0.5f;
  X = (int32_t)rintf(xw * kSubpix);
  Y = (int32_t)rintf(yw * kSubpix);
  Z = zw;
}

__device__ __forceinline__ Snap snap_vertex(float4 c, float hw, float hh, float gx, float gy) {
  Snap s;
  s.flags = outcode(c);
  bool inside = true;
#pragma unroll
  for (int pl = 0; pl < 6; ++pl) inside = inside && (plane_dist(c, pl, gx, gy) >= 0.0f);
  inside = inside && (c.w > 0.0f);
  if (!inside) {
    s.flags |= kNeedClip;
    s.X = s.Y = 0;
    s.z = 0.0f;
  } else {
    project(c, hw, hh, s.X, s.Y, s.z);
  }
  return s;
}

// Sutherland-Hodgman against near, far and the guard-band planes;
// intersections are interpolated from the inside vertex.
__device__ __forceinline__ float4 lerp_from(float4 a, float4 b, float da, float db) {
  float den = da - db;
  float t = da / den;
  float4 r;
  float d;
  d = b.x - a.x; d = t * d; r.x = a.x + d;
  d = b.y - a.y; d = t * d; r.y = a.y + d;
  d = b.z - a.z; d = t * d; r.z = a.z + d;
  d = b.w - a.w; d = t * d; r.w = a.w + d;
  return r;
}

constexpr int kMaxPoly = 9;

__device__ int clip_poly(float4* poly, int n, float gx, float gy) {
  float4 tmp[kMaxPoly];
  for (int plane = 0; plane < 6; ++plane) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
      float4 a = poly[i];
      float4 b = poly[(i + 1) % n];
      float da = plane_dist(a, plane, gx, gy);
      float db = plane_dist(b, plane, gx, gy);
      bool ina = da >= 0.0f, inb = db >= 0.0f;
      if (ina && m < kMaxPoly) tmp[m++] = a;
      if (ina != inb && m < kMaxPoly) tmp[m++] = ina ? lerp_from(a, b, da, db) : lerp_from(b, a, db, da);
    }
    n = m;
    if (n < 3) return 0;
    for (int i = 0; i < n; ++i) poly[i] = tmp[i];
  }
  return n;
}

// The same clipper over polygons kept in LDS, vertex i of the calling thread
// at poly[i * S]: the dynamically indexed arrays stay out of scratch memory.
// Ping-pongs between poly and tmp; returns the vertex count and the buffer
// that holds the result in *out.
template <int S>
__device__ int clip_poly_lds(float4* poly, float4* tmp, int n, float gx, float gy, float4** out) {
  for (int plane = 0; plane < 6; ++plane) {
    int m = 0;
    const float4 first = poly[0];
    const float dfirst = plane_dist(first, plane, gx, gy);
    float4 a = first;
    float da = dfirst;
    for (int i = 0; i < n; ++i) {
      const bool wrap = i + 1 == n;
      const float4 b = wrap ? first : poly[(i + 1) * S];
      const float db = wrap ? dfirst : plane_dist(b, plane, gx, gy);
      const bool ina = da >= 0.0f, inb = db >= 0.0f;
      if (ina && m < kMaxPoly) tmp[(m++) * S] = a;
      if (ina != inb && m < kMaxPoly)
        tmp[(m++) * S] = ina ? lerp_from(a, b, da, db) : lerp_from(b, a, db, da);
      a = b;
      da = db;
    }
    n = m;
    if (n < 3) return 0;
    float4* t = poly;
    poly = tmp;
    tmp = t;
  }
  *out = poly;
  return n;
}

struct PixRect {
  int x0, y0, x1, y1;  // inclusive
};

// pixel centres px*256+128 inside the snapped bounding box, clipped to r
__device__ __forceinline__ PixRect tri_pixels(const int32_t* X, const int32_t* Y, PixRect r) {
  int32_t minx = min(X[0], min(X[1], X[2])), maxx = max(X[0], max(X[1], X[2]));
  int32_t miny = min(Y[0], min(Y[1], Y[2])), maxy = max(Y[0], max(Y[1], Y[2]));
  PixRect o;
  o.x0 = max((minx - 128 + 255) >> 8, r.x0);
  o.x1 = min((maxx - 128) >> 8, r.x1);
  o.y0 = max((miny - 128 + 255) >> 8, r.y0);
  o.y1 = min((maxy - 128) >> 8, r.y1);
  return o;
}

__device__ __forceinline__ long long tri_area(const int32_t* X, const int32_t* Y) {
  return (long long)(X[1] - X[0]) * (long long)(Y[2] - Y[0]) -
         (long long)(Y[1] - Y[0]) * (long long)(X[2] - X[0]);
}

// adds v to frame statistic `stat` (striped by workgroup)
__device__ __forceinline__ void add_stat(unsigned long long* stats, uint32_t stat,
                                         unsigned long long v) {
  atomicAdd(&stats[stat * kStatStripes + (blockIdx.x & (kStatStripes - 1))], v);
}

// block-wide sum of one value per thread, returned to thread 0
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v,
                                                        unsigned long long* s_red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += s_red[w];
  return t;
}

// --------------------------------------------------------------------------
// k_nodes: absolute node transforms
// --------------------------------------------------------------------------

__device__ void mat4_mul(const float* a, const float* b, float* out) {
  float t[16];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float p = a[k * 4 + r] * b[c * 4 + k];
        s = s + p;
      }
      t[c * 4 + r] = s;
    }
#pragma unroll
  for (int i = 0; i < 16; ++i) out[i] = t[i];
}

// absolute transform (and normal matrix) of node i
__device__ __forceinline__ void node_transform(const int32_t* __restrict__ parents,
                                               const float* __restrict__ local,
                                               float* __restrict__ abs_out,
                                               float* __restrict__ nrm_out, uint32_t i) {
  int depth = 0;
  for (int32_t q = parents[i]; q >= 0 && depth < 1024; q = parents[q]) ++depth;
  float acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = (k % 5 == 0) ? 1.0f : 0.0f;
  for (int lvl = depth; lvl >= 0; --lvl) {
    int32_t a = (int32_t)i;
    for (int k = 0; k < lvl; ++k) a = parents[a];
    float m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = local[(size_t)a * 16 + k];
    mat4_mul(acc, m, acc);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) abs_out[(size_t)i * 16 + k] = acc[k];
  if (nrm_out) {
    // normal matrix: inverse transpose of the upper 3x3 (cofactors over the
    // determinant), [col][row]
    const float a00 = acc[0], a01 = acc[4], a02 = acc[8];
    const float a10 = acc[1], a11 = acc[5], a12 = acc[9];
    const float a20 = acc[2], a21 = acc[6], a22 = acc[10];
    float c00 = a11 * a22; float t0 = a12 * a21; c00 = c00 - t0;
    float c01 = a12 * a20; float t1 = a10 * a22; c01 = c01 - t1;
    float c02 = a10 * a21; float t2 = a11 * a20; c02 = c02 - t2;
    float c10 = a02 * a21; float t3 = a01 * a22; c10 = c10 - t3;
    float c11 = a00 * a22; float t4 = a02 * a20; c11 = c11 - t4;
    float c12 = a01 * a20; float t5 = a00 * a21; c12 = c12 - t5;
    float c20 = a01 * a12; float t6 = a02 * a11; c20 = c20 - t6;
    float c21 = a02 * a10; float t7 = a00 * a12; c21 = c21 - t7;
    float c22 = a00 * a11; float t8 = a01 * a10; c22 = c22 - t8;
    float det = a00 * c00; float t9 = a01 * c01; det = det + t9; float t10 = a02 * c02;
    det = det + t10;
    float* nm = nrm_out + (size_t)i * 12;
    nm[0] = c00 / det; nm[3] = c01 / det; nm[6] = c02 / det;
    nm[1] = c10 / det; nm[4] = c11 / det; nm[7] = c12 / det;
    nm[2] = c20 / det; nm[5] = c21 / det; nm[8] = c22 / det;
  }
}

__global__ __launch_bounds__(kBlock) void k_nodes(const int32_t* __restrict__ parents,
                                                  const float* __restrict__ local,
                                                  float* __restrict__ abs_out, uint32_t n,
                                                  uint32_t* __restrict__ counters,
                                                  unsigned long long* __restrict__ stats,
                                                  uint32_t* __restrict__ frame,
                                                  float* __restrict__ nrm_out,
                                                  const uint32_t* __restrict__ dirty) {
  const uint32_t i0 = blockIdx.x * kBlock + threadIdx.x;
  // first kernel of every frame: reset the frame's counters, advance the
  // frame number (kept on the device, so a captured frame replays as is)
  if (i0 < kCtrTotal) counters[i0 * kCtrStride] = 0u;
  if (i0 == 0) *frame += 1u;
  for (uint32_t j = i0; j < kStatCount * kStatStripes; j += gridDim.x * kBlock) stats[j] = 0ull;
  // the host uploads *dirty with the cameras of every frame: 0 when no local
  // transform and no node list changed since the previous frame, whose
  // absolute transforms (and normal matrices) are still in place.  The
  // grid is capped (a frame without changes then dispatches few
  // workgroups); the nodes are taken grid-stride.
  if (*dirty == 0u) return;
  for (uint32_t i = i0; i < n; i += gridDim.x * kBlock)
    node_transform(parents, local, abs_out, nrm_out, i);
}

// --------------------------------------------------------------------------
// Edge functions and depth plane (shared by setup and raster)
// --------------------------------------------------------------------------

__device__ __forceinline__ void load_mat(const float* src, float* m) {
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = src[k];
}

struct TriSetup {
  long long dx[3], dy[3];
  long long bias[3];
  float z0, dzdx, dzdy;
};

__device__ __forceinline__ void setup_edges_only(const int32_t* X, const int32_t* Y, TriSetup& s) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3;
    s.dx[i] = (long long)X[j] - X[i];
    s.dy[i] = (long long)Y[j] - Y[i];
    s.bias[i] = (s.dy[i] < 0 || (s.dy[i] == 0 && s.dx[i] > 0)) ? 0 : -1;
  }
}

// depth plane relative to vertex 0 (same op order as the oracle)
__device__ __forceinline__ void depth_plane(const int32_t* X, const int32_t* Y, const float* Z,
                                            float& z0, float& dzdx_out, float& dzdy_out) {
  float fx1 = (float)(X[1] - X[0]), fy1 = (float)(Y[1] - Y[0]);
  float fx2 = (float)(X[2] - X[0]), fy2 = (float)(Y[2] - Y[0]);
  float fa = fx1 * fy2;
  float fa2 = fx2 * fy1;
  fa = fa - fa2;
  float dz1 = Z[1] - Z[0], dz2 = Z[2] - Z[0];
  float pp, qq;
  pp = dz1 * fy2; qq = dz2 * fy1;
  float dzdx = pp - qq;
  dzdx_out = dzdx / fa;
  pp = dz2 * fx1; qq = dz1 * fx2;
  float dzdy = pp - qq;
  dzdy_out = dzdy / fa;
  z0 = Z[0];
}

__device__ __forceinline__ void setup_edges(const int32_t* X, const int32_t* Y, const float* Z,
                                            TriSetup& s) {
  setup_edges_only(X, Y, s);
  depth_plane(X, Y, Z, s.z0, s.dzdx, s.dzdy);
}

__device__ __forceinline__ bool covers(const TriSetup& s, const int32_t* X, const int32_t* Y,
                                       long long Px, long long Py) {
  bool in = true;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    long long E = s.dx[i] * (Py - Y[i]) - s.dy[i] * (Px - X[i]);
    in = in && (E + s.bias[i] >= 0);
  }
  return in;
}

// smallest of three vertex depths (same selection order as the oracle)
__device__ __forceinline__ float min_depth(const float* Z) {
  float m = Z[0] < Z[1] ? Z[0] : Z[1];
  return Z[2] < m ? Z[2] : m;
}

// Px, Py: the pixel centre in sub-pixel units (px * 256 + 128); only the
// low 32 bits of the offsets from the plane's origin enter, so 32-bit
// wrapping arithmetic gives the same value the 64-bit difference did
__device__ __forceinline__ float depth_at(float z0, float dzdx, float dzdy, float zlo, int32_t X0,
                                          int32_t Y0, int32_t Px, int32_t Py) {
  float ox = (float)(int32_t)((uint32_t)Px - (uint32_t)X0);
  float oy = (float)(int32_t)((uint32_t)Py - (uint32_t)Y0);
  float z = dzdx * ox;
  z = z0 + z;
  float zy = dzdy * oy;
  z = z + zy;
  if (!(z >= zlo)) z = zlo;
  if (!(z > 0.0f)) z = 0.0f;
  if (z > 1.0f) z = 1.0f;
  return z;
}

__device__ __forceinline__ void vertex_uv(const FrameParams& p, uint32_t flags, uint32_t v,
                                          const float* tm, float& s, float& t) {
  float u, vv;
  if (flags & kDrawUvHalf) {
    uint32_t h2 = p.uv16[v];
    u = __half2float(__ushort_as_half((unsigned short)(h2 & 0xffff)));
    vv = __half2float(__ushort_as_half((unsigned short)(h2 >> 16)));
  } else {
    float2 t2 = p.uv32[v];
    u = t2.x;
    vv = t2.y;
  }
  // texture matrix in the vertex stage: ((c0*u + c1*v) + c2)
  float a = tm[0] * u; float a1 = tm[3] * vv; a = a + a1; a = a + tm[6];
  float b = tm[1] * u; float b1 = tm[4] * vv; b = b + b1; b = b + tm[7];
  s = a;
  t = b;
}

// a vertex's texture coordinates as stored (half pair, or two floats' bits)
__device__ __forceinline__ void load_uv_raw(const FrameParams& p, uint32_t flags, uint32_t v,
                                            uint32_t& a, uint32_t& b) {
  if (flags & kDrawUvHalf) {
    a = p.uv16[v];
    b = 0u;
  } else {
    const float2 t2 = p.uv32[v];
    a = __float_as_uint(t2.x);
    b = __float_as_uint(t2.y);
  }
}

// vertex_uv() on coordinates already loaded
__device__ __forceinline__ void uv_from_raw(uint32_t flags, uint32_t a, uint32_t b, const float* tm,
                                            float& s, float& t) {
  float u, vv;
  if (flags & kDrawUvHalf) {
    u = __half2float(__ushort_as_half((unsigned short)(a & 0xffff)));
    vv = __half2float(__ushort_as_half((unsigned short)(a >> 16)));
  } else {
    u = __uint_as_float(a);
    vv = __uint_as_float(b);
  }
  float x = tm[0] * u; float x1 = tm[3] * vv; x = x + x1; x = x + tm[6];
  float y = tm[1] * u; float y1 = tm[4] * vv; y = y + y1; y = y + tm[7];
  s = x;
  t = y;
}

constexpr int kVtxPerThread = (kVertexCache + kBlock - 1) / kBlock;  // vertex-cache slots per thread

// wave-aggregated append of one PrimRef per flagged lane (ballot + prefix
// popcount, one atomic per wave)
// Depth-pyramid test: true iff every fragment with depth >= zmin inside the
// pixel rectangle [x0, x1] x [y0, y1] of the scene fails LESS against the
// phase-1 depths (zmin above the largest of them over every covering tile).
// Level 0 (8x8 tiles) when the rectangle touches at most max_tiles of them,
// else level 1 (32x32).
__device__ __forceinline__ bool hiz_hidden(const FrameParams& p, uint32_t scene, int x0, int y0,
                                           int x1, int y1, float zmin, int max_tiles) {
  const float* h = p.hiz + (size_t)scene * p.hiz_stride;
  int w = p.hiz_w0;
  int bx0 = x0 >> kHiZ0, bx1 = x1 >> kHiZ0, by0 = y0 >> kHiZ0, by1 = y1 >> kHiZ0;
  if ((bx1 - bx0 + 1) * (by1 - by0 + 1) > max_tiles) {
    w = p.hiz_w1;
    h += p.hiz_w0 * p.hiz_h0;
    bx0 = x0 >> kHiZ1; bx1 = x1 >> kHiZ1; by0 = y0 >> kHiZ1; by1 = y1 >> kHiZ1;
  }
  for (int by = by0; by <= by1; ++by)
    for (int bx = bx0; bx <= bx1; ++bx)
      if (!(zmin > h[by * w + bx])) return false;
  return true;
}

// hiz_hidden() for a chunk's rectangle, which may cover many tiles: the
// tiles are read sixteen at a time, their loads issued together, so a
// hidden chunk costs one round trip per sixteen tiles rather than one per
// tile (same tiles, same answer).  Rows past the rectangle's last are read
// as its last row; their values are not used.
__device__ __forceinline__ bool hiz_hidden_wide(const FrameParams& p, uint32_t scene, int x0, int y0,
                                                int x1, int y1, float zmin, int max_tiles) {
  const float* h = p.hiz + (size_t)scene * p.hiz_stride;
  int w = p.hiz_w0;
  int bx0 = x0 >> kHiZ0, bx1 = x1 >> kHiZ0, by0 = y0 >> kHiZ0, by1 = y1 >> kHiZ0;
  if ((bx1 - bx0 + 1) * (by1 - by0 + 1) > max_tiles) {
    w = p.hiz_w1;
    h += p.hiz_w0 * p.hiz_h0;
    bx0 = x0 >> kHiZ1; bx1 = x1 >> kHiZ1; by0 = y0 >> kHiZ1; by1 = y1 >> kHiZ1;
  }
  const int nx = bx1 - bx0 + 1, n = nx * (by1 - by0 + 1);
  int c = 0, r = by0;
  for (int t0 = 0; t0 < n; t0 += 16) {
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      v[k] = h[min(r, by1) * w + bx0 + c];
      if (++c == nx) {
        c = 0;
        ++r;
      }
    }
    bool hidden = true;
#pragma unroll
    for (int k = 0; k < 16; ++k) hidden = hidden && (t0 + k >= n || zmin > v[k]);
    if (!hidden) return false;
  }
  return true;
}

// A load of data the kernel never writes, at an address uniform over the
// wave: through the constant address space it becomes a scalar load into
// SGPRs (no VGPRs, scalar-cache latency).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* ptr, uint32_t i) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
  const ConstWords w = (ConstWords)(ptr + __builtin_amdgcn_readfirstlane(i));
  T out;
  uint32_t* o = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) o[k] = w[k];
  return out;
}

__device__ __forceinline__ void append_ref(const FrameParams& p, uint32_t counter, bool flag,
                                           PrimRef ref, const ClipFast& fast, int lane,
                                           unsigned long long lt_mask) {
  const unsigned long long m = __ballot(flag);
  if (m == 0ull) return;
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&p.counters[counter], (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (flag) {
    const uint32_t slot = base + (uint32_t)__popcll(m & lt_mask);
    if (ref.fast && slot < p.clip_fast_cap)
      p.clip_fast[slot] = fast;
    else
      ref.fast = 0u;
    if (slot < p.clipped_cap) p.clipped[slot] = ref;
  }
}

// A bin's entry count went from `old` to `old + n` in this phase: the one
// reservation that crosses the heavy threshold lists the bin for the raster
// to take first, and in phase 2 the first reservation lists the bin as
// active (counts only grow within a phase, so a bin is listed at most once
// per list and no list exceeds the raster items).
__device__ __forceinline__ void note_heavy(const FrameParams& p, uint32_t gbin, uint32_t old,
                                           uint32_t n, int phase, uint32_t cap) {
  const bool active = phase == 2 && old == 0u && n != 0u;
  const uint32_t t = p.heavy_entries;
  const bool heavy = t != 0u && old < t && n >= t - old;
  // the reservation that takes the count past the bin's capacity lists the
  // bin for the exact rescan
  const bool over = old <= cap && n > cap - old;
  if (over)
    p.overflow[atomicAdd(&p.counters[(phase == 2 ? kCtrOverflow2 : kCtrOverflow1) * kCtrStride], 1u)] =
        gbin;
  if (!active && !heavy) return;
  // the bin's raster queue range and that range's section of the lists
  const uint32_t part = gbin / p.part_bins, first = part * p.part_bins;
  if (active)
    p.active[first + atomicAdd(&p.counters[raster_ctr(2, kPartActive, part) * kCtrStride], 1u)] = gbin;
  if (heavy) {
    p.heavy[first + atomicAdd(&p.counters[raster_ctr(phase, kPartHeavy, part) * kCtrStride], 1u)] =
        gbin;
    p.bin_heavy[gbin] = 1u;
  }
}

// --------------------------------------------------------------------------
// k_setup_bin: persistent workgroups draining a queue of work items (chunks
// of <= 1024 consecutive triangles of one draw)
//
//  1. vertex stage for the chunk's vertex range -> LDS: snapped window
//     coordinates, window depth, outcodes, clip-space x/y/w and the
//     texture coordinates after the draw's texture matrix
//  2. triangle setup, 4 triangles per lane; a survivor writes its record
//     (raster + resolve parts) straight from LDS and counts itself into
//     LDS per-bin counters (LDS atomics give its local slot)
//  3. one global atomicAdd per touched bin reserves the item's range in
//     the bin; 4. entries are scattered to base + local slot
// --------------------------------------------------------------------------

// at least 5 waves per SIMD (96 VGPRs): one register more costs a fifth
// of the resident waves and ~15% of the setup time
__global__ __launch_bounds__(kBlock, 5) void k_setup_bin(FrameParams p,
                                                      const uint32_t* __restrict__ list,
                                                      uint32_t list_ctr, uint32_t queue_ctr) {
  __shared__ int32_t s_x[kVertexCache];
  __shared__ int32_t s_y[kVertexCache];
  __shared__ float s_z[kVertexCache];
  __shared__ uint32_t s_f[kVertexCache];
  __shared__ float s_cx[kVertexCache];
  __shared__ float s_cy[kVertexCache];
  __shared__ float s_cw[kVertexCache];
  __shared__ float s_s[kVertexCache];
  __shared__ float s_t[kVertexCache];
  // per-bin counters and bases: 2 * p.lds_bins words of dynamic LDS
  extern __shared__ uint32_t s_bins[];
  uint32_t* const s_bcount = s_bins;
  uint32_t* const s_bbase = s_bins + p.lds_bins;
  __shared__ uint32_t s_pair[kPairCap];  // bin << 21 | local slot << 10 | triangle in chunk
  __shared__ uint32_t s_npairs;
  __shared__ uint32_t s_item;

  const int lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t written = 0;
  unsigned long long tri_acc = 0, vtx_acc = 0;  // triangles / vertices set up (uniform)
  uint32_t grab = 0, grab_end = 0;  // items claimed from the queue (uniform)
  const int phase = queue_ctr == kCtrQueue2 ? 2 : 1;
  const bool hiz_test = phase == 2 && p.hiz != nullptr;
  // the whole frame, or the work items listed for this phase
  const uint32_t n_items = list ? min(p.counters[list_ctr * kCtrStride], p.n_items) : p.n_items;

  bool have_pre = false;
  uint32_t pre_item = 0;
  WorkItem pre_wi = {};
  // items per claim: queue_grain, down to 1 when there are fewer than 8
  // claims' worth per workgroup (a short list balances better item by item)
  const uint32_t grain =
      max(1u, min((uint32_t)p.queue_grain, n_items / (gridDim.x * 8u)));
  // this workgroup's stripe of the list and its claim counter (every stripe
  // has workgroups: a grid smaller than kQueueStripes uses fewer stripes)
  const uint32_t nstripes = min(kQueueStripes, gridDim.x);
  const uint32_t stripe = blockIdx.x % nstripes;
  const uint32_t s_begin = (uint32_t)((unsigned long long)n_items * stripe / nstripes);
  const uint32_t s_end = (uint32_t)((unsigned long long)n_items * (stripe + 1) / nstripes);
  uint32_t* const claim_ctr =
      &p.counters[(kCtrStripes + (queue_ctr == kCtrQueue2 ? kQueueStripes : 0u) + stripe) * kCtrStride];
  // thread 0 claims the next chunk one chunk ahead, so the claim's round
  // trip overlaps the current chunk (the last, unused claim only moves the
  // counter past the end)
  uint32_t next_claim = 0;
  if (threadIdx.x == 0) next_claim = s_begin + atomicAdd(claim_ctr, grain);
  while (true) {
    if (grab == grab_end) {
      if (threadIdx.x == 0) {
        s_item = next_claim;
        if (next_claim < s_end) next_claim = s_begin + atomicAdd(claim_ctr, grain);
      }
      __syncthreads();
      grab = s_item;
      grab_end = min(grab + grain, s_end);
      __syncthreads();  // s_item may be rewritten only after everyone read it
      if (grab >= s_end) break;
    }
    uint32_t item;
    WorkItem wi;
    if (have_pre) {  // fetched during the previous item
      item = pre_item;
      wi = pre_wi;
    } else {
      item = list ? uniform_load(list, grab) : grab;
      wi = uniform_load(p.items, item);
    }
    ++grab;
    have_pre = false;
    // the next item of the chunk: its descriptor now (scalar loads, done by
    // the end of this item's vertex stage)
    const bool prefetch = grab < grab_end;
    if (prefetch) {
      pre_item = list ? uniform_load(list, grab) : grab;
      pre_wi = uniform_load(p.items, pre_item);
    }
    tri_acc += wi.tri_count;
    vtx_acc += wi.vcount;
    const SceneDesc sc = uniform_load(p.scenes, wi.scene);
    const DrawDesc d = uniform_load(p.draws, wi.draw);
    const MaterialDesc mat = uniform_load(p.materials, d.material);
    float M[16], PV[16], tm[9];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      M[q] = uniform_load(p.abs_nodes, wi.node * 16u + q);
      PV[q] = uniform_load(p.cameras, wi.scene * 16u + q);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) tm[q] = mat.texmat[q];
    const float hw = (float)sc.width * 0.5f, hh = (float)sc.height * 0.5f;
    const float gx = kGuardPx / hw, gy = kGuardPx / hh;
    const bool cached = wi.vcount <= (uint32_t)kVertexCache;
    // scenes with more bins than the LDS counters hold append every pair
    // straight to the global bins
    const int nbins = sc.bins_x * sc.bins_y;
    const int nlb = nbins <= (int)p.lds_bins ? nbins : 0;

    for (int b = threadIdx.x; b < nlb; b += kBlock) s_bcount[b] = 0u;
    if (threadIdx.x == 0) s_npairs = nlb ? 0u : (uint32_t)kPairCap;
    // the item's triangle indices, loaded while its vertex stage runs
    uint32_t ix[kChunkTris / kBlock][3];
#pragma unroll
    for (int k = 0; k < kChunkTris / kBlock; ++k) {
      const uint32_t j = (uint32_t)k * kBlock + threadIdx.x;
      if (cached && j < wi.tri_count) {
        const uint32_t* ip = p.indices + wi.first_index + 3 * j;
        ix[k][0] = ip[0];
        ix[k][1] = ip[1];
        ix[k][2] = ip[2];
      }
    }
    if (cached) {
#pragma unroll
      for (int q = 0; q < kVtxPerThread; ++q) {
        const uint32_t v = threadIdx.x + (uint32_t)q * kBlock;
        if (v >= wi.vcount) break;
        const uint32_t gv = wi.vmin + v;
        uint32_t ua, ub;
        load_uv_raw(p, d.flags, gv, ua, ub);
        const float4 c = vertex_clip(M, PV, p.pos_x[gv], p.pos_y[gv], p.pos_z[gv]);
        const Snap s = snap_vertex(c, hw, hh, gx, gy);
        s_x[v] = s.X;
        s_y[v] = s.Y;
        s_z[v] = s.z;
        s_f[v] = s.flags;
        s_cx[v] = c.x;
        s_cy[v] = c.y;
        s_cw[v] = c.w;
        float ss, tt;
        uv_from_raw(d.flags, ua, ub, tm, ss, tt);
        s_s[v] = ss;
        s_t[v] = tt;
      }
    }
    __syncthreads();

    const PixRect tile = {0, 0, sc.width - 1, sc.height - 1};
    const int sh = p.bin_shift;
#pragma unroll
    for (int k = 0; k < kChunkTris / kBlock; ++k) {
      const uint32_t j = (uint32_t)k * kBlock + threadIdx.x;
      bool clip_me = false;
      const uint32_t prim = wi.prim + j;
      if (j < wi.tri_count && !cached) {
        clip_me = true;  // vertex range beyond the LDS cache: k_setup_clip takes it whole
      } else if (j < wi.tri_count) {
        const uint32_t vi[3] = {ix[k][0], ix[k][1], ix[k][2]};
        int32_t X[3], Y[3];
        float Z[3];
        uint32_t f[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const uint32_t l = vi[q] - wi.vmin;
          X[q] = s_x[l];
          Y[q] = s_y[l];
          Z[q] = s_z[l];
          f[q] = s_f[l];
        }
        PixRect pr = {0, 0, -1, -1};
        bool active = false;
        if ((f[0] & f[1] & f[2] & 63u) == 0) {
          if (((f[0] | f[1] | f[2]) & kNeedClip) == 0) {
            if (tri_area(X, Y) > 0) {
              pr = tri_pixels(X, Y, tile);
              active = pr.x0 <= pr.x1 && pr.y0 <= pr.y1;
              // phase 2: a triangle whose nearest vertex depth (every
              // fragment's lower bound) lies behind the phase-1 pyramid over
              // its rectangle cannot change a pixel
              if (active && hiz_test &&
                  hiz_hidden(p, wi.scene, pr.x0, pr.y0, pr.x1, pr.y1, min_depth(Z), kHiZTiles))
                active = false;
            }
          } else {
            clip_me = true;  // rare: k_setup_clip handles the fan
          }
        }
        if (active) {
          RasterRec rr;
          pack_xy(X, Y, rr.xy);
          rr.z[0] = Z[0];
          rr.z[1] = Z[1];
          rr.z[2] = Z[2];
          p.rrecs[(size_t)sc.rec_base + prim] = rr;
        }
        // the resolve record, also of a primitive the clip pass takes (its
        // fan triangles carry their own raster records)
        if (active || clip_me) {
          ResolveRec rec;
          const uint32_t l0 = vi[0] - wi.vmin, l1 = vi[1] - wi.vmin, l2 = vi[2] - wi.vmin;
          {
            const float cx[3] = {s_cx[l0], s_cx[l1], s_cx[l2]};
            const float cy[3] = {s_cy[l0], s_cy[l1], s_cy[l2]};
            const float cw[3] = {s_cw[l0], s_cw[l1], s_cw[l2]};
            bary_rows_of(cx, cy, cw, rec.e);
          }
          rec.s0 = s_s[l0]; rec.t0 = s_t[l0];
          rec.s1 = s_s[l1]; rec.t1 = s_t[l1];
          rec.s2 = s_s[l2]; rec.t2 = s_t[l2];
          rec.item = item;
          p.srecs[(size_t)sc.rec_base + prim] = rec;
        }
        if (active) {
          // count into the LDS per-bin counters; pairs beyond the LDS list
          // go straight to the global bin
          for (int by = pr.y0 >> sh; by <= (pr.y1 >> sh); ++by)
            for (int bx = pr.x0 >> sh; bx <= (pr.x1 >> sh); ++bx) {
              const uint32_t b = (uint32_t)(by * sc.bins_x + bx);
              const uint32_t pi = s_npairs < (uint32_t)kPairCap ? atomicAdd(&s_npairs, 1u) : kPairCap;
              if (pi < (uint32_t)kPairCap) {
                const uint32_t slot = atomicAdd(&s_bcount[b], 1u);
                s_pair[pi] = (b << 21) | (slot << 10) | j;
              } else {
                const uint32_t slot = atomicAdd(&p.bin_count[sc.bin_base + b], 1u);
                note_heavy(p, sc.bin_base + b, slot, 1u, phase, sc.bin_cap);
                if (slot < sc.bin_cap) p.bin_entries[sc.entry_base + (size_t)b * sc.bin_cap + slot] = prim;
              }
              ++written;
            }
        }
      }
      // (a primitive of a cached item has its resolve record written above
      // and its vertex indices in ix)
      append_ref(p, kCtrClipped * kCtrStride, clip_me, PrimRef{wi.draw, prim, item, cached ? 1u : 0u},
                 ClipFast{wi.scene, wi.node, {ix[k][0], ix[k][1], ix[k][2]}, {0u, 0u, 0u}}, lane,
                 lt_mask);
    }
    have_pre = prefetch;
    __syncthreads();
    // one global reservation per touched bin
    for (int b = threadIdx.x; b < nlb; b += kBlock) {
      const uint32_t n = s_bcount[b];
      if (n) {
        const uint32_t old = atomicAdd(&p.bin_count[sc.bin_base + b], n);
        s_bbase[b] = old;
        note_heavy(p, sc.bin_base + b, old, n, phase, sc.bin_cap);
      }
    }
    __syncthreads();
    const uint32_t np = min(s_npairs, (uint32_t)kPairCap);
    for (uint32_t i = threadIdx.x; i < np; i += kBlock) {
      const uint32_t e = s_pair[i];
      const uint32_t b = e >> 21;
      const uint32_t slot = s_bbase[b] + ((e >> 10) & 0x7ffu);
      if (slot < sc.bin_cap)
        p.bin_entries[sc.entry_base + (size_t)b * sc.bin_cap + slot] = wi.prim + (e & 0x3ffu);
    }
    __syncthreads();  // LDS is reused by the next item
  }
  __shared__ unsigned long long s_red[kBlock / 64];
  const unsigned long long wsum = block_sum(written, s_red);
  if (threadIdx.x == 0) {
    const uint32_t st = queue_ctr == kCtrQueue2 ? 4u : 2u;
    if (wsum) add_stat(p.stats, 0, wsum);
    if (tri_acc) {
      add_stat(p.stats, st, tri_acc);
      add_stat(p.stats, st + 1, vtx_acc);
    }
  }
}

// --------------------------------------------------------------------------
// Rasterisation helpers: LDS visibility tile with 64-bit atomicMin keys
// --------------------------------------------------------------------------

// A triangle the whole workgroup rasterises, set up once by the lane that
// queued it: its three edge functions (fill-rule bias included) at the bin's
// first pixel centre, exact in 64 bits, and their steps per pixel right and
// down; narrow = every value over the bin (and a bin past it) fits in 32 bits.
struct BigTri {
  long long e[3];
  int32_t dx[3], dy[3];  // edge vectors (|d| < 2^23); per pixel right -dy * 256, down dx * 256
  float z0, dzdx, dzdy, zlo;
  int32_t X0, Y0;        // vertex 0, the depth plane's origin
  uint32_t prim;
  uint8_t x0, y0, x1, y1;  // covered pixel rectangle, relative to the bin
  uint16_t npx;            // its pixel count
  uint16_t narrow;
};

struct RasterCtx {
  unsigned long long* vis;
  BigTri* big;
  int* nbig;
  PixRect bin;
  int B;
  int lane_limit;       // rectangles above this many pixels go to the queue
  uint32_t* pairs;      // per-lane counters: pixels tested, fragments covered
  uint32_t* frags;
};

__device__ __forceinline__ bool small_box(const int32_t* X, const int32_t* Y) {
  const int32_t ex = max(X[0], max(X[1], X[2])) - min(X[0], min(X[1], X[2]));
  const int32_t ey = max(Y[0], max(Y[1], Y[2])) - min(Y[0], min(Y[1], Y[2]));
  // edge functions over the box, and one step past it, stay below 2^30
  return ex < 16384 && ey < 16384;
}

__device__ __forceinline__ void depth_key_min(const RasterCtx& rc, float z0, float dzdx, float dzdy,
                                              float zlo, const int32_t* X, const int32_t* Y,
                                              uint32_t prim, int px, int py) {
  const float z = depth_at(z0, dzdx, dzdy, zlo, X[0], Y[0], px * 256 + 128,
                           py * 256 + 128);
  if (z < 1.0f) {
    const unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | prim;
    atomicMin(&rc.vis[__mul24(py - rc.bin.y0, rc.B) + (px - rc.bin.x0)], key);
  }
}

// queue entry of a triangle with a known depth plane and pixel rectangle r
// (inside the bin)
__device__ __forceinline__ void make_big(const RasterCtx& rc, const int32_t* X, const int32_t* Y,
                                         float z0, float dzdx, float dzdy, float zlo,
                                         uint32_t prim, const PixRect& r, int npx, BigTri& b) {
  const long long Px0 = (long long)rc.bin.x0 * 256 + 128, Py0 = (long long)rc.bin.y0 * 256 + 128;
  bool narrow = true;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    const int32_t dx = X[j] - X[i], dy = Y[j] - Y[i];
    const long long bias = (dy < 0 || (dy == 0 && dx > 0)) ? 0 : -1;
    const long long e = (long long)dx * (Py0 - Y[i]) - (long long)dy * (Px0 - X[i]) + bias;
    // window coordinates stay within the guard band (|X|, |Y| < 2^22), so
    // |d| * 256 < 2^31
    b.e[i] = e;
    b.dx[i] = dx;
    b.dy[i] = dy;
    const long long reach = (llabs((long long)dx) + llabs((long long)dy)) * 256 * 64;
    narrow = narrow && llabs(e) + reach < (1ll << 31);
  }
  b.z0 = z0; b.dzdx = dzdx; b.dzdy = dzdy; b.zlo = zlo;
  b.X0 = X[0];
  b.Y0 = Y[0];
  b.prim = prim;
  b.x0 = (uint8_t)(r.x0 - rc.bin.x0);
  b.y0 = (uint8_t)(r.y0 - rc.bin.y0);
  b.x1 = (uint8_t)(r.x1 - rc.bin.x0);
  b.y1 = (uint8_t)(r.y1 - rc.bin.y0);
  b.npx = (uint16_t)npx;
  b.narrow = narrow ? 1 : 0;
}

// The queued triangles, a wave at a time each: a wave claims the next
// triangle (LDS counter `next`) and spreads its pixel rectangle over its
// lanes, pixel p = lane + 64 i of the rectangle in row-major order; each
// lane evaluates the edge functions at its pixel directly (exact: 32-bit
// when narrow) and merges a covered pixel with an LDS atomicMin, so the
// waves need no ownership of pixels and a triangle costs
// ceil(pixels / 64) wave steps.
__device__ void raster_big_waves(const RasterCtx& rc, const BigTri* big, int nb, int* next) {
  const int lane = (int)threadIdx.x & 63;
  if (threadIdx.x == 0)
    for (int q = 0; q < nb; ++q) *rc.pairs += big[q].npx;
  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(next, 1);
    q = __builtin_amdgcn_readfirstlane(q);
    if (q >= nb) break;
    const BigTri& t = big[q];
    const int x0 = t.x0, y0 = t.y0, w = t.x1 - t.x0 + 1, npx = t.npx;
    const float inv_w = 1.0f / (float)w;
    const int32_t X0 = t.X0, Y0 = t.Y0;
    const float z0 = t.z0, dzdx = t.dzdx, dzdy = t.dzdy, zlo = t.zlo;
    const uint32_t prim = t.prim;
    int32_t dx[3], dy[3];
    long long e[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      dx[i] = t.dx[i];
      dy[i] = t.dy[i];
      e[i] = t.e[i];
    }
    // integer products below are of a pixel offset (< 64) and an edge
    // vector component (|d| < 2^23) or a bin dimension: 24-bit multiplies
    const bool narrow = t.narrow;
    for (int pi = lane; pi < npx; pi += 64) {
      const int ry = div_small(pi, inv_w);
      const int lx = x0 + (pi - __mul24(ry, w)), ly = y0 + ry;
      bool in;
      if (narrow) {
        uint32_t v[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          v[i] = (uint32_t)e[i] + (uint32_t)(imul24(dx[i], ly) - imul24(dy[i], lx)) * 256u;
        in = (int32_t)(v[0] | v[1] | v[2]) >= 0;
      } else {
        long long v[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          v[i] = e[i] + (long long)(imul24(dx[i], ly) - imul24(dy[i], lx)) * 256;
        in = (v[0] | v[1] | v[2]) >= 0;
      }
      if (in) {
        ++*rc.frags;
        const float z = depth_at(z0, dzdx, dzdy, zlo, X0, Y0, ((rc.bin.x0 + lx) << 8) | 128,
                                 ((rc.bin.y0 + ly) << 8) | 128);
        if (z < 1.0f)
          atomicMin(&rc.vis[__mul24(ly, rc.B) + lx],
                    ((unsigned long long)__float_as_uint(z) << 32) | prim);
      }
    }
  }
}

// this lane walks the clipped rectangle r of one front-facing triangle with
// a known depth plane row by row, edge functions stepped by additions
__device__ void raster_walk(const RasterCtx& rc, const int32_t* X, const int32_t* Y, float z0,
                            float dzdx, float dzdy, float zlo, uint32_t prim, const PixRect& r) {
  if (small_box(X, Y)) {
    int32_t dx[3], dy[3], bias[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int j = (i + 1) % 3;
      dx[i] = X[j] - X[i];
      dy[i] = Y[j] - Y[i];
      bias[i] = (dy[i] < 0 || (dy[i] == 0 && dx[i] > 0)) ? 0 : -1;
    }
    const int32_t Px0 = r.x0 * 256 + 128;
    for (int py = r.y0; py <= r.y1; ++py) {
      const int32_t Py = py * 256 + 128;
      int32_t e[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) e[i] = dx[i] * (Py - Y[i]) - dy[i] * (Px0 - X[i]) + bias[i];
      for (int px = r.x0; px <= r.x1; ++px) {
        if ((e[0] | e[1] | e[2]) >= 0) {
          ++*rc.frags;
          depth_key_min(rc, z0, dzdx, dzdy, zlo, X, Y, prim, px, py);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) e[i] -= dy[i] * 256;
      }
    }
    return;
  }
  // large boxes: the same edge functions in 64 bits, stepped along the row
  int32_t dx[3], dy[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    dx[i] = X[j] - X[i];
    dy[i] = Y[j] - Y[i];
  }
  const long long Px0 = (long long)r.x0 * 256 + 128;
  for (int py = r.y0; py <= r.y1; ++py) {
    const long long Py = (long long)py * 256 + 128;
    long long e[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const long long bias = (dy[i] < 0 || (dy[i] == 0 && dx[i] > 0)) ? 0 : -1;
      e[i] = (long long)dx[i] * (Py - Y[i]) - (long long)dy[i] * (Px0 - X[i]) + bias;
    }
    for (int px = r.x0; px <= r.x1; ++px) {
      if ((e[0] | e[1] | e[2]) >= 0) {
        ++*rc.frags;
        depth_key_min(rc, z0, dzdx, dzdy, zlo, X, Y, prim, px, py);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) e[i] -= (long long)dy[i] * 256;
    }
  }
}

// The clipped rectangle of a triangle in this bin and its pixel count (0:
// none); rectangles above rc.lane_limit pixels go to the workgroup's queue
// (returns -1; a full queue leaves them to the lane)
__device__ __forceinline__ int raster_classify(const RasterCtx& rc, const int32_t* X,
                                               const int32_t* Y, float z0, float dzdx, float dzdy,
                                               float zlo, uint32_t prim, PixRect& r) {
  r = tri_pixels(X, Y, rc.bin);
  if (r.x0 > r.x1 || r.y0 > r.y1) return 0;
  const int npx = (r.x1 - r.x0 + 1) * (r.y1 - r.y0 + 1);
  if (npx > rc.lane_limit) {
    const int slot = atomicAdd(rc.nbig, 1);
    if (slot < kBigQueue) {
      make_big(rc, X, Y, z0, dzdx, dzdy, zlo, prim, r, npx, rc.big[slot]);
      return -1;
    }
  }
  *rc.pairs += (uint32_t)npx;
  return npx;
}

// rasterise one front-facing triangle with a known depth plane: queued, or
// walked by this lane
__device__ void raster_planed(const RasterCtx& rc, const int32_t* X, const int32_t* Y, float z0,
                              float dzdx, float dzdy, float zlo, uint32_t prim) {
  PixRect r;
  if (raster_classify(rc, X, Y, z0, dzdx, dzdy, zlo, prim, r) > 0)
    raster_walk(rc, X, Y, z0, dzdx, dzdy, zlo, prim, r);
}

__device__ __forceinline__ void raster_tri(const RasterCtx& rc, const int32_t* X, const int32_t* Y,
                                           const float* Z, uint32_t prim) {
  if (tri_area(X, Y) <= 0) return;
  PixRect r = tri_pixels(X, Y, rc.bin);
  if (r.x0 > r.x1 || r.y0 > r.y1) return;
  float z0, dzdx, dzdy;
  depth_plane(X, Y, Z, z0, dzdx, dzdy);
  raster_planed(rc, X, Y, z0, dzdx, dzdy, min_depth(Z), prim);
}

__device__ __forceinline__ uint32_t find_draw(const FrameParams& p, const SceneDesc& sc,
                                              const uint32_t* s_base, uint32_t prim) {
  // last draw with prim_base <= prim (upper bound - 1)
  uint32_t lo = 0, hi = sc.draw_count;
  const bool lds = sc.draw_count <= (uint32_t)kDrawCache;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t b = lds ? s_base[mid] : p.draws[sc.draw_begin + mid].prim_base;
    if (b <= prim) lo = mid + 1;
    else hi = mid;
  }
  return sc.draw_begin + lo - 1;
}

// full re-derivation of a primitive from the vertex streams (clipped
// primitives and the overflow fallback)
__device__ void raster_prim(const FrameParams& p, const SceneDesc& sc, const float* PV,
                            const RasterCtx& rc, const uint32_t* s_base, uint32_t prim) {
  const uint32_t di = find_draw(p, sc, s_base, prim);
  const DrawDesc d = p.draws[di];
  const uint32_t t = prim - d.prim_base;
  float M[16];
  load_mat(p.abs_nodes + (size_t)d.node * 16, M);
  const uint32_t* ip = p.indices + d.index_base + 3 * t;
  float4 c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t v = ip[k];
    c[k] = vertex_clip(M, PV, p.pos_x[v], p.pos_y[v], p.pos_z[v]);
  }
  if (outcode(c[0]) & outcode(c[1]) & outcode(c[2])) return;
  const float hw = (float)sc.width * 0.5f, hh = (float)sc.height * 0.5f;
  const float gx = kGuardPx / hw, gy = kGuardPx / hh;
  bool need = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    bool inside = true;
#pragma unroll
    for (int pl = 0; pl < 6; ++pl) inside = inside && (plane_dist(c[k], pl, gx, gy) >= 0.0f);
    inside = inside && (c[k].w > 0.0f);
    need = need || !inside;
  }
  if (!need) {
    int32_t X[3], Y[3];
    float Z[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) project(c[k], hw, hh, X[k], Y[k], Z[k]);
    raster_tri(rc, X, Y, Z, prim);
    return;
  }
  float4 poly[kMaxPoly];
  poly[0] = c[0]; poly[1] = c[1]; poly[2] = c[2];
  const int n = clip_poly(poly, 3, gx, gy);
  if (n < 3) return;
  int32_t PX[kMaxPoly], PY[kMaxPoly];
  float PZ[kMaxPoly];
  for (int k = 0; k < n; ++k) {
    if (!(poly[k].w > 0.0f)) return;
    project(poly[k], hw, hh, PX[k], PY[k], PZ[k]);
  }
  for (int k = 1; k + 1 < n; ++k) {
    int32_t tx[3] = {PX[0], PX[k], PX[k + 1]};
    int32_t ty[3] = {PY[0], PY[k], PY[k + 1]};
    float tz[3] = {PZ[0], PZ[k], PZ[k + 1]};
    raster_tri(rc, tx, ty, tz, prim);
  }
}


// ---- texture sampling (GL 4.6 section 8.14) -------------------------------

__device__ __forceinline__ float gbr_log2(float x) {
  uint32_t bits = __float_as_uint(x);
  int e = (int)((bits >> 23) & 0xffu);
  if (e == 0) return -126.0f;
  if (e == 0xff) return 128.0f;
  float m = __uint_as_float((bits & 0x7fffffu) | 0x3f800000u);
  float y = (m - 1.0f) / (m + 1.0f);
  float y2 = y * y;
  float s = y2 * (1.0f / 9.0f);
  s = s + (1.0f / 7.0f);
  s = s * y2;
  s = s + (1.0f / 5.0f);
  s = s * y2;
  s = s + (1.0f / 3.0f);
  s = s * y2;
  s = s + 1.0f;
  s = s * y;
  float l = s * 2.8853900817779268f;
  return (float)(e - 127) + l;
}

// mirrored repeat, and repeat of a size that is not a power of two
__device__ __noinline__ int wrap_coord_slow(int i, int size, int mode) {
  if (mode == GL_MIRRORED_REPEAT) {
    int pp = 2 * size;
    int m = i % pp;
    if (m < 0) m += pp;
    return m < size ? m : pp - 1 - m;
  }
  int m = i % size;
  if (m < 0) m += size;
  return m;
}

// GL wrap of texel indices along one axis of one level, decided once per
// level: clamp-to-edge and power-of-two repeat share one form,
// clamp(i & mask, 0, size - 1) (mask = size - 1 for repeat, all ones for
// clamp), the rest (mirrored repeat, repeat of other sizes) goes out of line.
struct AxisWrap {
  int size, mask, mode;
  bool fast;
};

__device__ __forceinline__ AxisWrap axis_wrap(int size, int mode) {
  const bool clamp = mode == GL_CLAMP_TO_EDGE;
  AxisWrap a;
  a.size = size;
  a.mode = mode;
  a.fast = clamp || (mode != GL_MIRRORED_REPEAT && (size & (size - 1)) == 0);
  a.mask = clamp ? -1 : size - 1;
  return a;
}

// min(max(x, 0), hi) in one instruction
__device__ __forceinline__ int clamp0(int x, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(x), "v"(hi));
  return r;
}

__device__ __forceinline__ int wrap_coord(const AxisWrap& a, int i) {
  if (__builtin_expect(a.fast, 1)) return clamp0(i & a.mask, a.size - 1);
  return wrap_coord_slow(i, a.size, a.mode);
}

// floor of u clamped to [-2^30, 2^30] (the clamp as one median; u is finite)
__device__ __forceinline__ int floor_to_int(float u) {
  u = __builtin_amdgcn_fmed3f(u, -1073741824.0f, 1073741824.0f);
  return (int)floorf(u);
}

// A texture's header held in registers (one 32-byte load), and the byte
// offset of the level being sampled: every field is read once per pixel
// instead of once per use.
struct TexHeader {
  int32_t width, height, layers, levels;
  int32_t min_filter, mag_filter, wrap_s, wrap_t;
};

__device__ __forceinline__ TexHeader load_tex_header(const TextureDesc* t) {
  const int4 a = *reinterpret_cast<const int4*>(t);
  const int4 b = *(reinterpret_cast<const int4*>(t) + 1);
  return TexHeader{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// texel `row + i` of a level whose first byte is `base`: row = (layer * h
// + j) * w in 32-bit arithmetic, as the texel index always was
__device__ __forceinline__ void fetch(const float* lut, const uint8_t* base, uint32_t row,
                                      int i, float* out) {
  const uint32_t idx = row + (uint32_t)i;
  const uint32_t texel = *reinterpret_cast<const uint32_t*>(base + (size_t)idx * 4);
  out[0] = lut[texel & 0xff];
  out[1] = lut[(texel >> 8) & 0xff];
  out[2] = lut[(texel >> 16) & 0xff];
  out[3] = lut[texel >> 24];
}

// (layer * h + j) * w mod 2^32 as layer_off + j * w, layer_off = layer * h * w
// mod 2^32 once per level; j < h and w below 2^24 (checked at import), so
// the product is a full-rate 24-bit multiply
__device__ __forceinline__ uint32_t texel_row(uint32_t layer_off, int j, int w) {
  return layer_off + __umul24((uint32_t)j, (uint32_t)w);
}

__device__ void sample_level(const float* lut, const TexHeader& t, const uint8_t* base, int level,
                             int layer, float s, float tc, bool linear, float* out) {
  int w = max(t.width >> level, 1), h = max(t.height >> level, 1);
  const AxisWrap ws = axis_wrap(w, t.wrap_s), wt = axis_wrap(h, t.wrap_t);
  float u = s * (float)w;
  float v = tc * (float)h;
  const uint32_t layer_off = (uint32_t)layer * (uint32_t)h * (uint32_t)w;
  if (!linear) {
    int i = wrap_coord(ws, floor_to_int(u));
    int j = wrap_coord(wt, floor_to_int(v));
    fetch(lut, base, texel_row(layer_off, j, w), i, out);
    return;
  }
  float uu = u - 0.5f;
  float vv = v - 0.5f;
  int i0 = floor_to_int(uu);
  int j0 = floor_to_int(vv);
  float a = uu - (float)i0;
  float b = vv - (float)j0;
  int i1 = wrap_coord(ws, i0 + 1);
  int j1 = wrap_coord(wt, j0 + 1);
  i0 = wrap_coord(ws, i0);
  j0 = wrap_coord(wt, j0);
  const uint32_t r0 = texel_row(layer_off, j0, w), r1 = texel_row(layer_off, j1, w);
  float t00[4], t10[4], t01[4], t11[4];
  fetch(lut, base, r0, i0, t00);
  fetch(lut, base, r0, i1, t10);
  fetch(lut, base, r1, i0, t01);
  fetch(lut, base, r1, i1, t11);
  float ia = 1.0f - a;
  float ib = 1.0f - b;
  float w00 = ia * ib, w10 = a * ib, w01 = ia * b, w11 = a * b;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float r = w00 * t00[c];
    float q1 = w10 * t10[c];
    r = r + q1;
    float q2 = w01 * t01[c];
    r = r + q2;
    float q3 = w11 * t11[c];
    r = r + q3;
    out[c] = r;
  }
}

__device__ __forceinline__ const uint8_t* level_base(const FrameParams& p, const TextureDesc* td,
                                                     int level) {
  return p.texels + td->level_offset[level];
}

__device__ void sample_tex(const FrameParams& p, const float* lut, const TextureDesc* td,
                           const TexHeader& t, int layer, float s, float tc, float lambda,
                           float* out) {
  if (layer < 0) layer = 0;
  if (layer >= t.layers) layer = t.layers - 1;
  float c = 0.0f;
  if (t.mag_filter == GL_LINEAR &&
      (t.min_filter == GL_NEAREST_MIPMAP_NEAREST || t.min_filter == GL_NEAREST_MIPMAP_LINEAR))
    c = 0.5f;
  if (lambda <= c) {
    sample_level(lut, t, level_base(p, td, 0), 0, layer, s, tc, t.mag_filter == GL_LINEAR, out);
    return;
  }
  const int q = t.levels - 1;
  switch (t.min_filter) {
    case GL_NEAREST:
    case GL_LINEAR:
      sample_level(lut, t, level_base(p, td, 0), 0, layer, s, tc, t.min_filter == GL_LINEAR, out);
      return;
    case GL_NEAREST_MIPMAP_NEAREST:
    case GL_LINEAR_MIPMAP_NEAREST: {
      int dlev;
      if (lambda <= 0.5f)
        dlev = 0;
      else {
        float f = ceilf(lambda + 0.5f) - 1.0f;
        dlev = f > (float)q ? q : (int)f;
      }
      sample_level(lut, t, level_base(p, td, dlev), dlev, layer, s, tc,
                   t.min_filter == GL_LINEAR_MIPMAP_NEAREST, out);
      return;
    }
    default: {
      bool lin = t.min_filter == GL_LINEAR_MIPMAP_LINEAR;
      if (lambda >= (float)q) {
        sample_level(lut, t, level_base(p, td, q), q, layer, s, tc, lin, out);
        return;
      }
      float fl = floorf(lambda);
      int d1 = (int)fl;
      float f = lambda - fl;
      const uint8_t* b1 = level_base(p, td, d1);
      const uint8_t* b2 = level_base(p, td, d1 + 1);
      float t1[4], t2[4];
      sample_level(lut, t, b1, d1, layer, s, tc, lin, t1);
      sample_level(lut, t, b2, d1 + 1, layer, s, tc, lin, t2);
      float g = 1.0f - f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float x1 = g * t1[k];
        float x2 = f * t2[k];
        out[k] = x1 + x2;
      }
      return;
    }
  }
}

// ---- resolve ----------------------------------------------------------------

struct TriAttr {
  float s[3], t[3];
  float e[3][3];  // homogeneous barycentric rows (bary_rows_of)
};

__device__ __forceinline__ void bary_at(const TriAttr& A, float hw, float hh, float ihw, float ihh,
                                        float fx, float fy, float* b) {
  float nx = fx - hw;
  nx = nx * ihw;
  float ny = fy - hh;
  ny = ny * ihh;
  float l[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float v = A.e[i][0] * nx;
    float v2 = A.e[i][1] * ny;
    v = v + v2;
    v = v + A.e[i][2];
    l[i] = v;
  }
  float S = l[0] + l[1];
  S = S + l[2];
  if (S == 0.0f) {
    b[0] = 1.0f; b[1] = 0.0f; b[2] = 0.0f;
    return;
  }
  float r = 1.0f / S;
  b[0] = l[0] * r;
  b[1] = l[1] * r;
  b[2] = l[2] * r;
}

__device__ __forceinline__ float interp3(const float* b, float a0, float a1, float a2) {
  float v = b[0] * a0;
  float v1 = b[1] * a1;
  v = v + v1;
  float v2 = b[2] * a2;
  return v + v2;
}

// clamp(c, 0, 1) (NaN to 0) as one median, then round to 8 bits
__device__ __forceinline__ uint32_t to_unorm8(float c) {
  c = __builtin_amdgcn_fmed3f(c, 0.0f, 1.0f);
  float v = c * 255.0f;
  v = v + 0.5f;
  return (uint32_t)(int)v;
}

// colour of the visible primitive at a pixel.  A holds the triangle's rows,
// b the pixel's barycentrics, (s, t) its texture coordinates; q00/q10/q01
// the texture coordinates at the centres of its 2x2 quad (derivatives for
// the level of detail) where a quad neighbour already computed them for the
// same primitive (has_*), else they are evaluated here.
struct QuadST {
  float s00, t00, s10, t10, s01, t01;
  bool h00, h10, h01;
};

// Phong diffuse term of the scene's lights (no specular) for a shaded
// material, in world space: per light, intensity = max(0, N.L) * attenuation
// with attenuation clamp(1 - (d/R)^4, 0, 1)^2 / (1 + d^2) for point lights
// (1 for directional ones), colour += diffuse * light colour * intensity; the
// diffuse alpha is added to the ambient alpha also with no lights.  diffuse =
// material diffuse * texture colour * vertex colour.  Vertex positions and
// normals come from the mesh through the draw's absolute transform and normal
// matrix, interpolated with the pixel's barycentrics b.
__device__ void add_lights(const FrameParams& p, const DrawDesc& d, uint32_t prim, uint32_t scene,
                           const float* b, const float* texc, const float* vc, float* c) {
  const MaterialDesc& mat = p.materials[d.material];
  float dd[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) dd[k] = mat.diffuse[k] * texc[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) dd[k] = dd[k] * vc[k];
  const uint32_t* ip = p.indices + d.index_base + 3 * (prim - d.prim_base);
  const float* M = p.abs_nodes + (size_t)d.node * 16;
  const float* NM = p.normal_mats + (size_t)d.node * 12;
  // world position and normal interpolated vertex by vertex: the sums
  // b0*a0 + b1*a1 + b2*a2 of interp3, in its order
  float n[3], pw[3];
#pragma unroll 1
  for (int k = 0; k < 3; ++k) {
    const uint32_t v = ip[k];
    const float bk = k == 0 ? b[0] : (k == 1 ? b[1] : b[2]);
    float q[4];
    mat4_vec(M, p.pos_x[v], p.pos_y[v], p.pos_z[v], 1.0f, q);
    const float n0 = p.normals[3 * (size_t)v], n1 = p.normals[3 * (size_t)v + 1],
                n2 = p.normals[3 * (size_t)v + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float x = NM[r] * n0; float y = NM[3 + r] * n1; x = x + y;
      float z = NM[6 + r] * n2; x = x + z;
      const float tn = bk * x, tp = bk * q[r];
      n[r] = k == 0 ? tn : n[r] + tn;
      pw[r] = k == 0 ? tp : pw[r] + tp;
    }
  }
  float nn = n[0] * n[0]; float n1 = n[1] * n[1]; nn = nn + n1; float n2 = n[2] * n[2];
  nn = nn + n2;
  const float nl = sqrtf(nn);
#pragma unroll
  for (int k = 0; k < 3; ++k) n[k] = n[k] / nl;
  const uint32_t first = p.scene_lights[2 * scene], count = p.scene_lights[2 * scene + 1];
  for (uint32_t l = first; l < first + count; ++l) {
    const LightDesc ld = p.lights[l];
    const float* T = p.abs_nodes + (size_t)ld.node * 16;
    float L[3], att = 1.0f;
    if (ld.type == GBR_LIGHT_POINT) {
#pragma unroll
      for (int k = 0; k < 3; ++k) L[k] = T[12 + k] - pw[k];
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) L[k] = -T[8 + k];
    }
    float ll = L[0] * L[0]; float l1 = L[1] * L[1]; ll = ll + l1; float l2 = L[2] * L[2];
    ll = ll + l2;
    const float dist = sqrtf(ll);
#pragma unroll
    for (int k = 0; k < 3; ++k) L[k] = L[k] / dist;
    if (ld.type == GBR_LIGHT_POINT) {
      float qd = dist / ld.range;
      qd = qd * qd;
      qd = qd * qd;
      float wdw = 1.0f - qd;
      if (!(wdw > 0.0f)) wdw = 0.0f;
      if (wdw > 1.0f) wdw = 1.0f;
      att = wdw * wdw;
      float den = dist * dist;
      den = 1.0f + den;
      att = att / den;
    }
    float ndl = n[0] * L[0]; float t1 = n[1] * L[1]; ndl = ndl + t1; float t2 = n[2] * L[2];
    ndl = ndl + t2;
    if (!(ndl > 0.0f)) ndl = 0.0f;
    const float inten = ndl * att;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = dd[k] * ld.color[k];
      v = v * inten;
      c[k] = c[k] + v;
    }
  }
  c[3] = c[3] + dd[3];
}

template <bool kLights>
__device__ uint32_t shade_pixel(const FrameParams& p, const float* lut, float hw, float hh,
                               float ihw, float ihh, const TriAttr& A, const float* b, float s,
                               float t, bool textured, const QuadST& q, const DrawDesc& d,
                               uint32_t prim, uint32_t scene, int px, int py) {
  float c[4] = {d.ambient[0], d.ambient[1], d.ambient[2], d.ambient[3]};
  float texc[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  float vc[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  if (textured) {
    const TextureDesc* tdp = p.textures + d.texture;
    const TexHeader td = load_tex_header(tdp);
    const int qx = px & ~1, qy = py & ~1;
    float s00 = q.s00, t00 = q.t00, s10 = q.s10, t10 = q.t10, s01 = q.s01, t01 = q.t01;
    float bq[3];
    if (!q.h00) {
      bary_at(A, hw, hh, ihw, ihh, (float)qx + 0.5f, (float)qy + 0.5f, bq);
      s00 = interp3(bq, A.s[0], A.s[1], A.s[2]);
      t00 = interp3(bq, A.t[0], A.t[1], A.t[2]);
    }
    if (!q.h10) {
      bary_at(A, hw, hh, ihw, ihh, (float)qx + 1.5f, (float)qy + 0.5f, bq);
      s10 = interp3(bq, A.s[0], A.s[1], A.s[2]);
      t10 = interp3(bq, A.t[0], A.t[1], A.t[2]);
    }
    if (!q.h01) {
      bary_at(A, hw, hh, ihw, ihh, (float)qx + 0.5f, (float)qy + 1.5f, bq);
      s01 = interp3(bq, A.s[0], A.s[1], A.s[2]);
      t01 = interp3(bq, A.t[0], A.t[1], A.t[2]);
    }
    const float fw = (float)td.width, fh = (float)td.height;
    float dux = s10 - s00; dux = dux * fw;
    float dvx = t10 - t00; dvx = dvx * fh;
    float duy = s01 - s00; duy = duy * fw;
    float dvy = t01 - t00; dvy = dvy * fh;
    float rx = dux * dux; float rx2 = dvx * dvx; rx = rx + rx2;
    float ry = duy * duy; float ry2 = dvy * dvy; ry = ry + ry2;
    const float r2 = rx > ry ? rx : ry;
    float lambda = gbr_log2(r2);
    lambda = lambda * 0.5f;
    float tc[4];
    sample_tex(p, lut, tdp, td, (int)d.layer, s, t, lambda, tc);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      texc[k] = tc[k];
      c[k] = c[k] * tc[k];
    }
  }
  const uint32_t cf = (d.flags >> kDrawColorShift) & 3u;
  if (cf != kColorNone) {
    const uint32_t* ip = p.indices + d.index_base + 3 * (prim - d.prim_base);
    float col[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t v = ip[k];
      if (cf == kColorU8) {
        const uint32_t u8 = p.col8[v];
        col[k][0] = lut[u8 & 0xff];
        col[k][1] = lut[(u8 >> 8) & 0xff];
        col[k][2] = lut[(u8 >> 16) & 0xff];
        col[k][3] = lut[u8 >> 24];
      } else {
        const float4 f = p.col32[v];
        col[k][0] = f.x; col[k][1] = f.y; col[k][2] = f.z; col[k][3] = f.w;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float v = interp3(b, col[0][k], col[1][k], col[2][k]);
      vc[k] = v;
      c[k] = c[k] * v;
    }
  }
  if (kLights && d.lit) add_lights(p, d, prim, scene, b, texc, vc, c);
  return to_unorm8(c[0]) | (to_unorm8(c[1]) << 8) | (to_unorm8(c[2]) << 16) |
         (to_unorm8(c[3]) << 24);
}


// --------------------------------------------------------------------------
// k_setup_clip: primitives crossing the near/far plane or the guard band.
// Sutherland-Hodgman fan, one raster record per front-facing fan triangle,
// each binned on its own (entry = kSubTri | record).
// --------------------------------------------------------------------------


// forces the exact rescan for every bin the rectangle touches
__device__ __forceinline__ void overflow_rect(const FrameParams& p, const SceneDesc& sc,
                                              const PixRect& r, int phase) {
  const int sh = p.bin_shift;
  for (int by = r.y0 >> sh; by <= (r.y1 >> sh); ++by)
    for (int bx = r.x0 >> sh; bx <= (r.x1 >> sh); ++bx) {
      const uint32_t gbin = sc.bin_base + (uint32_t)(by * sc.bins_x + bx);
      const uint32_t old = atomicAdd(&p.bin_count[gbin], sc.bin_cap + 1u);
      note_heavy(p, gbin, old, sc.bin_cap + 1u, phase, sc.bin_cap);
    }
}

// phase 1 takes the clipped list from its start and leaves its end in
// kCtrClipMark; phase 2 continues from there.  One wave per workgroup, one
// thread per primitive; its polygon and projected fan vertices live in LDS.
// The wave's fan triangles are then binned by all of its lanes, one fan
// each, so that a primitive with several fans does not bin them in turn.
constexpr int kClipBlock = 64;
constexpr int kClipFans = kClipBlock * (kMaxPoly - 2);  // fan triangles of one wave at most

// the scene fields the binning of a fan triangle needs
struct ClipBins {
  int32_t bins_x;
  uint32_t bin_base, bin_cap, pad;
  unsigned long long entry_base;
};

// the bins of a fan triangle's rectangle, in groups of 8 whose counter
// atomics are all in flight together (each returns the entry's slot)
__device__ __forceinline__ void bin_fan(const FrameParams& p, const ClipBins& sc, const PixRect& r,
                                        uint32_t entry, int phase) {
  const int sh = p.bin_shift;
  const int bx0 = r.x0 >> sh, by0 = r.y0 >> sh;
  const int nx = (r.x1 >> sh) - bx0 + 1, nb = nx * ((r.y1 >> sh) - by0 + 1);
  for (int j0 = 0; j0 < nb; j0 += 8) {
    uint32_t b[8], slot[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const int j = j0 + g;
      b[g] = (uint32_t)((by0 + j / nx) * sc.bins_x + bx0 + j % nx);
      slot[g] = j < nb ? atomicAdd(&p.bin_count[sc.bin_base + b[g]], 1u) : sc.bin_cap;
    }
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      if (j0 + g < nb) note_heavy(p, sc.bin_base + b[g], slot[g], 1u, phase, sc.bin_cap);
      if (slot[g] < sc.bin_cap) p.bin_entries[sc.entry_base + (size_t)b[g] * sc.bin_cap + slot[g]] = entry;
    }
  }
}

__global__ __launch_bounds__(kClipBlock) void k_setup_clip(FrameParams p, int phase) {
  __shared__ float4 s_poly[2][kMaxPoly * kClipBlock];
  __shared__ int32_t s_px[kMaxPoly * kClipBlock], s_py[kMaxPoly * kClipBlock];
  __shared__ float s_pz[kMaxPoly * kClipBlock];
  __shared__ PixRect s_frect[kClipFans];  // the wave's fan triangles: pixel rectangle,
  __shared__ uint32_t s_fentry[kClipFans];  // bin entry (~0u: no record slot left),
  __shared__ uint8_t s_fown[kClipFans];     // and the lane whose primitive it is
  __shared__ ClipBins s_bins[kClipBlock];
  const int tid = threadIdx.x;
  // phase 1 reads its first references beside the list length (the slots
  // exist whatever the length; the ones past it are not used)
  const uint32_t i_first = blockIdx.x * kClipBlock + tid;
  PrimRef ref_first = {};
  ClipFast cf_first = {};
  if (phase == 1) {
    if (i_first < p.clipped_cap) ref_first = p.clipped[i_first];
    if (i_first < p.clip_fast_cap) cf_first = p.clip_fast[i_first];
  }
  const uint32_t n = min(p.counters[kCtrClipped * kCtrStride], p.clipped_cap);
  const uint32_t start = phase == 1 ? 0u : p.counters[kCtrClipMark * kCtrStride];
  if (phase == 1 && blockIdx.x == 0 && threadIdx.x == 0) p.counters[kCtrClipMark * kCtrStride] = n;
  int32_t* PX = s_px + tid;
  int32_t* PY = s_py + tid;
  float* PZ = s_pz + tid;
  // wave-uniform trip count.  The fan triangles' records take fixed slots:
  // the (kMaxPoly - 2) of list slot i start at (kMaxPoly - 2) * i, so no
  // counter is waited for (slots past sub_cap fall back to the exact rescan)
  for (uint32_t base = start + blockIdx.x * kClipBlock; base < n; base += gridDim.x * kClipBlock) {
    const uint32_t i = base + tid;
    uint32_t fans = 0;  // bit q: fan triangle (0, q, q+1) is front-facing and covers a pixel
    PrimRef ref = {};
    SceneDesc sc = {};
    if (i < n) {
      // the reference and its fast entry (the entry is used only when the
      // reference says it was written)
      const bool first = phase == 1 && i == i_first;
      ref = first ? ref_first : p.clipped[i];
      ClipFast cf = first ? cf_first : ClipFast{};
      if (!first && i < p.clip_fast_cap) cf = p.clip_fast[i];
      const bool full = ref.fast == 0u;  // indices and resolve record left to this pass
      DrawDesc d = {};
      uint32_t scene = cf.scene, node = cf.node;
      uint32_t v[3] = {cf.v[0], cf.v[1], cf.v[2]};
      if (full) {
        d = p.draws[ref.draw];
        scene = d.scene;
        node = d.node;
        const uint32_t* ip = p.indices + d.index_base + 3 * (ref.prim - d.prim_base);
#pragma unroll
        for (int q = 0; q < 3; ++q) v[q] = ip[q];
      }
      sc = p.scenes[scene];
      float M[16], PV[16];
      load_mat(p.abs_nodes + (size_t)node * 16, M);
      load_mat(p.cameras + (size_t)scene * 16, PV);
      const float hw = (float)sc.width * 0.5f, hh = (float)sc.height * 0.5f;
      const float gx = kGuardPx / hw, gy = kGuardPx / hh;
      float4* poly = s_poly[0] + tid;
      float4 c[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        c[q] = vertex_clip(M, PV, p.pos_x[v[q]], p.pos_y[v[q]], p.pos_z[v[q]]);
        poly[q * kClipBlock] = c[q];
      }
      if (full) {
        // resolve record of the primitive (its fan triangles carry their own
        // raster records)
        const MaterialDesc& mat = p.materials[d.material];
        float tm[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) tm[q] = mat.texmat[q];
        ResolveRec rec;
        rec.item = ref.item;
        {
          const float cx[3] = {c[0].x, c[1].x, c[2].x};
          const float cy[3] = {c[0].y, c[1].y, c[2].y};
          const float cw[3] = {c[0].w, c[1].w, c[2].w};
          bary_rows_of(cx, cy, cw, rec.e);
        }
        vertex_uv(p, d.flags, v[0], tm, rec.s0, rec.t0);
        vertex_uv(p, d.flags, v[1], tm, rec.s1, rec.t1);
        vertex_uv(p, d.flags, v[2], tm, rec.s2, rec.t2);
        p.srecs[(size_t)sc.rec_base + ref.prim] = rec;
      }
      const int nv = clip_poly_lds<kClipBlock>(poly, s_poly[1] + tid, 3, gx, gy, &poly);
      bool ok = nv >= 3;
      for (int q = 0; q < nv && ok; ++q) {
        const float4 pv = poly[q * kClipBlock];
        if (!(pv.w > 0.0f)) {
          ok = false;
        } else {
          int32_t x, y;
          float z;
          project(pv, hw, hh, x, y, z);
          PX[q * kClipBlock] = x;
          PY[q * kClipBlock] = y;
          PZ[q * kClipBlock] = z;
        }
      }
      const PixRect tile = {0, 0, sc.width - 1, sc.height - 1};
      for (int q = 1; ok && q + 1 < nv; ++q) {
        const int a = q * kClipBlock, b = (q + 1) * kClipBlock;
        const int32_t tx[3] = {PX[0], PX[a], PX[b]};
        const int32_t ty[3] = {PY[0], PY[a], PY[b]};
        if (tri_area(tx, ty) <= 0) continue;
        const PixRect r = tri_pixels(tx, ty, tile);
        if (r.x0 > r.x1 || r.y0 > r.y1) continue;
        fans |= 1u << q;
      }
    }
    s_bins[tid] = ClipBins{sc.bins_x, sc.bin_base, sc.bin_cap, 0u, sc.entry_base};
    // exclusive wave prefix of the fan counts, one atomic for the wave
    const uint32_t cnt = (uint32_t)__popc(fans);
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o);
      if (tid >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, 63);
    if (tid == 63 && total) atomicAdd(&p.counters[kCtrSubTris * kCtrStride], total);  // (statistics)
    uint32_t k = incl - cnt;  // this lane's first fan within the wave
    unsigned long long slot = (unsigned long long)(kMaxPoly - 2) * base + k;
    if (fans) {
      const PixRect tile = {0, 0, sc.width - 1, sc.height - 1};
      for (uint32_t f = fans; f; f &= f - 1u, ++slot, ++k) {
        const int q = __ffs(f) - 1;
        const int a = q * kClipBlock, b = (q + 1) * kClipBlock;
        const int32_t tx[3] = {PX[0], PX[a], PX[b]};
        const int32_t ty[3] = {PY[0], PY[a], PY[b]};
        const PixRect r = tri_pixels(tx, ty, tile);
        s_frect[k] = r;
        s_fown[k] = (uint8_t)tid;
        if (slot < p.sub_cap) {
          RasterRec rr;
          pack_xy(tx, ty, rr.xy);
          rr.z[0] = PZ[0];
          rr.z[1] = PZ[a];
          rr.z[2] = PZ[b];
          p.sub_recs[slot] = rr;
          p.sub_prims[slot] = ref.prim;
          s_fentry[k] = kSubTri | (uint32_t)slot;
        } else {
          s_fentry[k] = ~0u;
          overflow_rect(p, sc, r, phase);
        }
      }
    }
    __syncthreads();  // the wave's fans are staged (the workgroup is one wave)
    // the wave's fan triangles into their bins, one per lane
    for (uint32_t f = (uint32_t)tid; f < total; f += kClipBlock) {
      const uint32_t e = s_fentry[f];
      if (e != ~0u) bin_fan(p, s_bins[s_fown[f]], s_frect[f], e, phase);
    }
    __syncthreads();  // LDS is rewritten by the next primitives
  }
}

// --------------------------------------------------------------------------
// Bin helpers shared by the raster and resolve kernels
// --------------------------------------------------------------------------

struct BinInfo {
  PixRect bin;
  uint32_t gbin, scene, bin_cap, plane_offset;
  int width, height, stride;  // viewport; row stride of the planes
  uint32_t planes;
  unsigned long long entries, rec_base;
};

__device__ __forceinline__ BinInfo bin_info(const RasterItem& it) {
  BinInfo b;
  b.bin.x0 = it.x0;
  b.bin.y0 = it.y0;
  b.bin.x1 = it.x1;
  b.bin.y1 = it.y1;
  b.gbin = it.gbin;
  b.scene = it.scene;
  b.bin_cap = it.bin_cap;
  b.plane_offset = it.plane_offset;
  b.width = it.width;
  b.height = it.height;
  b.entries = it.entries;
  b.rec_base = it.rec_base;
  b.stride = it.stride;
  b.planes = it.planes;
  return b;
}

// a bin entry's triangle: window coordinates, depth plane (same op order as
// everywhere) and nearest vertex depth; returns its primitive order
__device__ __forceinline__ uint32_t fetch_raster_rec(const FrameParams& p,
                                                     unsigned long long rec_base, uint32_t e,
                                                     RasterRec& rr) {
  const bool sub = (e & kSubTri) != 0;
  rr = sub ? p.sub_recs[e & ~kSubTri] : p.rrecs[(size_t)rec_base + e];
  return sub ? p.sub_prims[e & ~kSubTri] : e;
}

__device__ __forceinline__ void decode_raster_rec(const RasterRec& rr, int32_t* X, int32_t* Y,
                                                  float& z0, float& dzdx, float& dzdy, float& zlo) {
  unpack_xy(rr.xy, X, Y);
  depth_plane(X, Y, rr.z, z0, dzdx, dzdy);
  zlo = min_depth(rr.z);
}

__device__ __forceinline__ uint32_t load_raster_rec(const FrameParams& p, unsigned long long rec_base,
                                                    uint32_t e, int32_t* X, int32_t* Y, float& z0,
                                                    float& dzdx, float& dzdy, float& zlo) {
  RasterRec rr;
  const uint32_t prim = fetch_raster_rec(p, rec_base, e, rr);
  decode_raster_rec(rr, X, Y, z0, dzdx, dzdy, zlo);
  return prim;
}

__device__ __forceinline__ void store_vis(const FrameParams& p, const BinInfo& b,
                                          const unsigned long long* s_vis, int B) {
  const int bw = b.bin.x1 - b.bin.x0 + 1, bh = b.bin.y1 - b.bin.y0 + 1;
  const float inv_bw = 1.0f / (float)bw;
  for (int l = threadIdx.x; l < bw * bh; l += kBlock) {
    const int ly = div_small(l, inv_bw), lx = l - ly * bw;
    p.vis[(size_t)b.plane_offset + (size_t)(b.bin.y0 + ly) * b.stride + b.bin.x0 + lx] =
        s_vis[ly * B + lx];
  }
}

// One binned triangle staged in LDS for pixel-parallel rasterisation.
struct PairTri {
  int32_t X[3], Y[3];
  float z0, dzdx, dzdy, zlo;
  uint32_t prim;
  int16_t x0, y0;   // clipped pixel rectangle: origin
  uint16_t w;       //   and width
  uint16_t small;   // 1: bounding box under 128 px, edge functions fit in 32 bits
  uint32_t inv_w;   // ceil(65536 / w): row of a rectangle index without a divide
};

// coverage (top-left rule) and depth of one (triangle, pixel) pair
__device__ __forceinline__ bool raster_pair(const PairTri& t, unsigned long long* vis,
                                            const PixRect& bin, int B, int px, int py) {
  bool in = true;
  if (t.small) {
    const int32_t Px = px * 256 + 128, Py = py * 256 + 128;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int j = (i + 1) % 3;
      const int32_t dx = t.X[j] - t.X[i], dy = t.Y[j] - t.Y[i];
      const int32_t bias = (dy < 0 || (dy == 0 && dx > 0)) ? 0 : -1;
      in = in && (dx * (Py - t.Y[i]) - dy * (Px - t.X[i]) + bias >= 0);
    }
  } else {
    const long long Px = (long long)px * 256 + 128, Py = (long long)py * 256 + 128;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int j = (i + 1) % 3;
      const long long dx = (long long)t.X[j] - t.X[i], dy = (long long)t.Y[j] - t.Y[i];
      const long long bias = (dy < 0 || (dy == 0 && dx > 0)) ? 0 : -1;
      in = in && (dx * (Py - t.Y[i]) - dy * (Px - t.X[i]) + bias >= 0);
    }
  }
  if (!in) return false;
  const float z = depth_at(t.z0, t.dzdx, t.dzdy, t.zlo, t.X[0], t.Y[0], px * 256 + 128,
                           py * 256 + 128);
  if (z < 1.0f) {
    const unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | t.prim;
    atomicMin(&vis[(py - bin.y0) * B + (px - bin.x0)], key);
  }
  return true;
}

// LDS written by one lane and read by another lane of the same wave: LDS
// instructions of a wave execute in order, so only the compiler must not
// move accesses across this point
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// block-wide sums of the per-lane raster counters -> frame statistics
// (every thread of the block calls this)
__device__ __forceinline__ void add_raster_stats(const FrameParams& p, uint32_t pairs,
                                                 uint32_t frags) {
  __shared__ unsigned long long s_red[kBlock / 64];
  const unsigned long long a = block_sum(pairs, s_red);
  const unsigned long long b = block_sum(frags, s_red);
  if (threadIdx.x == 0 && a) {
    add_stat(p.stats, 6, a);
    add_stat(p.stats, 7, b);
  }
}

// Depth pyramid of one 32x32 bin (s_vis rows of kMaxBin keys) from its final
// phase-1 keys in LDS: max depth per level-0 block (2^kHiZ0 pixels square)
// and over the bin (level 1), the same values k_hiz computes from the
// stored keys.
__device__ void bin_hiz(const FrameParams& p, const BinInfo& b, const unsigned long long* s_vis) {
  constexpr int kN = kHiZPer * kHiZPer;
  __shared__ uint32_t s_max[kN];
  if (threadIdx.x < kN) s_max[threadIdx.x] = 0u;
  __syncthreads();
  const int bw = b.bin.x1 - b.bin.x0 + 1, bh = b.bin.y1 - b.bin.y0 + 1;
  for (int l = threadIdx.x; l < kMaxBin * kMaxBin; l += kBlock) {
    const int lx = l & 31, ly = l >> 5;
    if (lx < bw && ly < bh)
      atomicMax(&s_max[(ly >> kHiZ0) * kHiZPer + (lx >> kHiZ0)],
                (uint32_t)(s_vis[ly * kMaxBin + lx] >> 32));
  }
  __syncthreads();
  const int rx = b.bin.x0 >> kHiZ1, ry = b.bin.y0 >> kHiZ1;
  float* h = p.hiz + (size_t)b.scene * p.hiz_stride;
  if (threadIdx.x < kN) {
    const int bx = rx * kHiZPer + (int)(threadIdx.x % kHiZPer),
              by = ry * kHiZPer + (int)(threadIdx.x / kHiZPer);
    if (bx < p.hiz_w0 && by < p.hiz_h0) h[by * p.hiz_w0 + bx] = __uint_as_float(s_max[threadIdx.x]);
  }
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int k = 0; k < kN; ++k) m = max(m, s_max[k]);
    h[p.hiz_w0 * p.hiz_h0 + ry * p.hiz_w1 + rx] = __uint_as_float(m);
  }
}

// --------------------------------------------------------------------------
// k_raster_lane: persistent workgroups draining the phase's raster queue,
// split into one range of bins per XCD: within a range first the bins setup
// listed as heavy (most entries), then every other bin in order, so that
// the longest bins start first and the frame's tail is short.  Per bin, one
// binned triangle per lane walking its clipped rectangle; rectangles above
// the lane limit (gbr_tuning.lane_pixels) go to an LDS queue whose
// triangles the waves take one at a time, each spread over a wave's lanes.
// Phase 1 starts every bin from the clear key; phase 2 merges into the
// phase-1 keys and leaves bins without new entries untouched (or, when it
// resolves in place, resolves them from their phase-1 keys).
// --------------------------------------------------------------------------

struct RasterLds {
  unsigned long long vis[kMaxBin * kMaxBin];
  BigTri big[kBigQueue];
  float lut[256];  // unorm8 -> float, for resolving in place
  int nbig[2];     // queue fill of even / odd batches
  int next[2];     // next queue entry to claim, even / odd batches
};

template <bool kLights>
__device__ void resolve_tile(const FrameParams& p, const RasterItem& it,
                             const unsigned long long* tile, int B, const float* s_lut);

// the raster phase after which no pass changes the keys
__host__ __device__ __forceinline__ bool final_phase(const FrameParams& p, int phase) {
  return p.cull_mode >= kCullHiZ ? phase == 2 : phase == 1;
}

// whether the lane rasteriser resolves this phase's bins in place
__host__ __device__ __forceinline__ bool resolves_in_place(const FrameParams& p, int phase) {
  return p.fuse_resolve && final_phase(p, phase);
}

// kResolve: the phase resolves its bins in place (resolves_in_place)
template <bool kResolve>
__device__ void raster_bin(const FrameParams& p, const RasterItem& item, int phase, RasterLds& L,
                           uint32_t& pairs, uint32_t& frags) {
  const BinInfo b = bin_info(item);
  const uint32_t* entries = p.bin_entries + b.entries;
  // the first entries load beside the count (both addresses are known; an
  // entry at or past the count is never used)
  const uint32_t e_first = threadIdx.x < b.bin_cap ? entries[threadIdx.x] : 0u;
  const uint32_t n = p.bin_count[b.gbin];
  if (n > b.bin_cap) return;  // overflowed: k_raster_full rescans this bin
  // with kResolve every bin is resolved here, also those without new
  // entries; else the keys go to HBM for the next pass
  if (phase == 2 && n == 0 && !kResolve) return;
  const int B = p.bin_size;
  RasterCtx rc;
  rc.vis = L.vis;
  rc.big = L.big;
  rc.B = B;
  rc.lane_limit = p.lane_pixels;
  rc.bin = b.bin;
  rc.pairs = &pairs;
  rc.frags = &frags;
  // the first batch's records are in flight while the tile is set up
  RasterRec rr_first = {};
  uint32_t prim_first = 0;
  if (threadIdx.x < n) prim_first = fetch_raster_rec(p, b.rec_base, e_first, rr_first);
  const int bw = b.bin.x1 - b.bin.x0 + 1, bh = b.bin.y1 - b.bin.y0 + 1;
  if (phase == 2) {
    constexpr int kPer = kMaxBin * kMaxBin / kBlock;
    unsigned long long keys[kPer];
    const float inv_bw = 1.0f / (float)bw;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int l = threadIdx.x + q * kBlock;
      const int ly = div_small(l, inv_bw), lx = l - ly * bw;
      keys[q] = l < bw * bh
                    ? p.vis[(size_t)b.plane_offset + (size_t)(b.bin.y0 + ly) * b.stride + b.bin.x0 + lx]
                    : 0ull;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int l = threadIdx.x + q * kBlock;
      const int ly = div_small(l, inv_bw), lx = l - ly * bw;
      if (l < bw * bh) L.vis[ly * B + lx] = keys[q];
    }
  } else {
    for (int i = threadIdx.x; i < B * B; i += kBlock) L.vis[i] = kClearKey;
  }
  if (threadIdx.x == 0) L.nbig[0] = L.nbig[1] = L.next[0] = L.next[1] = 0;
  __syncthreads();
  // per batch of kBlock entries: lanes append large rectangles to the queue
  // (fill and claim counters of this batch's parity; the other ones, read
  // before the previous batch's second barrier, are cleared for the next
  // batch), then the waves take the queued triangles one at a time
  int parity = 0;
  for (uint32_t base = 0; base < n; base += kBlock, parity ^= 1) {
    const uint32_t k = base + threadIdx.x;
    rc.nbig = &L.nbig[parity];
    if (threadIdx.x == 0) L.nbig[parity ^ 1] = L.next[parity ^ 1] = 0;
    if (k < n) {
      int32_t X[3], Y[3];
      float z0, dzdx, dzdy, zlo;
      uint32_t prim;
      if (base == 0) {
        prim = prim_first;
        decode_raster_rec(rr_first, X, Y, z0, dzdx, dzdy, zlo);
      } else {
        prim = load_raster_rec(p, b.rec_base, entries[k], X, Y, z0, dzdx, dzdy, zlo);
      }
      raster_planed(rc, X, Y, z0, dzdx, dzdy, zlo, prim);
    }
    __syncthreads();
    raster_big_waves(rc, L.big, min(L.nbig[parity], kBigQueue), &L.next[parity]);
    __syncthreads();
  }
  if (kResolve) {
    resolve_tile<false>(p, item, L.vis, B, L.lut);
  } else {
    store_vis(p, b, L.vis, B);
    if (phase == 1 && p.hiz_fused) bin_hiz(p, b, L.vis);
  }
  if (threadIdx.x == 0) p.bin_count[b.gbin] = 0;  // ready for the next frame
}

// kWaves: the waves per SIMD the register allocation must allow.  Without
// resolve 6 (80 VGPRs, no spills): the bins' chains of dependent global
// loads are latency-bound, and 6 resident workgroups per CU rasterise the
// first phase ~7% faster than 5 (at 7 the walk spills).  With the resolve,
// 5 when the resolving phase is the only one (every bin rasterised: a few
// spilled registers cost less than the lost occupancy, plain frame -4%)
// and 4 (no spills) after the depth-pyramid phase, whose bins mostly only
// resolve.
template <bool kResolve, int kWaves>
__global__ __launch_bounds__(kBlock, kWaves) void k_raster_lane(FrameParams p,
                                                        const RasterItem* __restrict__ items,
                                                        uint32_t n_ritems, int phase) {
  __shared__ RasterLds L;
  __shared__ uint32_t s_claim;
  // phase 1 visits every bin (each gets its keys and pyramid); phase 2 only
  // the bins that received entries, unless it resolves (then every bin)
  const bool active_only = phase == 2 && !kResolve;
  static_assert(kBlock == 256, "one LUT entry per thread");
  if (kResolve) L.lut[threadIdx.x] = p.unorm_lut[threadIdx.x];
  uint32_t pairs = 0, frags = 0;
  // thread 0's position in the queue ranges: its own XCD's range first
  // (workgroups are dispatched to the XCDs round robin), then the next ones
  const uint32_t home = blockIdx.x % kRasterParts;
  uint32_t tried = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t claim = ~0u;  // every range is drained
      for (; tried < kRasterParts; ++tried) {
        const uint32_t part = (home + tried) % kRasterParts;
        const uint32_t first = part * p.part_bins;
        const uint32_t size = min(first + p.part_bins, n_ritems) - min(first, n_ritems);
        const uint32_t q = atomicAdd(&p.counters[raster_ctr(phase, kPartClaim, part) * kCtrStride], 1u);
        const uint32_t n_heavy = min(p.counters[raster_ctr(phase, kPartHeavy, part) * kCtrStride], size);
        const uint32_t n_rest =
            active_only ? min(p.counters[raster_ctr(2, kPartActive, part) * kCtrStride], size) : size;
        if (q < n_heavy) {
          claim = p.heavy[first + q];
          break;
        }
        if (q - n_heavy < n_rest) {
          claim = active_only ? p.active[first + q - n_heavy] : first + q - n_heavy;
          // a listed bin was (or is being) rasterised from the heavy list
          if (p.bin_heavy[claim]) {
            p.bin_heavy[claim] = 0u;
            claim |= 0x80000000u;
          }
          break;
        }
      }
      s_claim = claim;
    }
    __syncthreads();
    // uniform across the workgroup: the bin's fields then live in scalar registers
    const uint32_t claim = __builtin_amdgcn_readfirstlane(s_claim);
    __syncthreads();  // s_claim is rewritten by the next claim; the LDS tile is free
    if (claim == ~0u) break;
    if (claim & 0x80000000u) continue;
    raster_bin<kResolve>(p, items[claim], phase, L, pairs, frags);
  }
  add_raster_stats(p, pairs, frags);
}

// --------------------------------------------------------------------------
// k_raster_pairs: per (scene, bin), binned triangles -> LDS visibility tile ->
// visibility keys.  Waves claim 64 entries at a time, prefix-sum the pixel
// counts of their clipped rectangles and spread the (triangle, pixel) pairs
// over the lanes, so small and large triangles keep every lane busy.
// Phase 1 starts every bin from the clear key; phase 2 merges into the
// phase-1 keys and leaves bins without new entries untouched.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_raster_pairs(FrameParams p,
                                                         const RasterItem* __restrict__ items,
                                                         int phase) {
  __shared__ unsigned long long s_vis[kMaxBin * kMaxBin];
  __shared__ PairTri s_tri[kBlock];
  __shared__ uint32_t s_pref[kBlock];
  __shared__ uint32_t s_next;

  const BinInfo b = bin_info(items[blockIdx.x]);
  const uint32_t n = p.bin_count[b.gbin];
  if (n > b.bin_cap) return;  // overflowed: k_raster_full rescans this bin
  if (phase == 2 && n == 0) return;
  const int B = p.bin_size;
  if (phase == 2) {
    const int bw = b.bin.x1 - b.bin.x0 + 1, bh = b.bin.y1 - b.bin.y0 + 1;
    const float inv_bw = 1.0f / (float)bw;
    for (int l = threadIdx.x; l < bw * bh; l += kBlock) {
      const int ly = div_small(l, inv_bw), lx = l - ly * bw;
      s_vis[ly * B + lx] =
          p.vis[(size_t)b.plane_offset + (size_t)(b.bin.y0 + ly) * b.stride + b.bin.x0 + lx];
    }
  } else {
    for (int i = threadIdx.x; i < B * B; i += kBlock) s_vis[i] = kClearKey;
  }
  if (threadIdx.x == 0) s_next = 0u;
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wbase = threadIdx.x & ~63;
  uint32_t pairs = 0, frags = 0;
  const uint32_t* entries = p.bin_entries + b.entries;
  while (true) {
    uint32_t chunk = 0;
    if (lane == 0) chunk = atomicAdd(&s_next, 64u);
    chunk = __shfl(chunk, 0);
    if (chunk >= n) break;
    const uint32_t k = chunk + (uint32_t)lane;
    uint32_t npx = 0;
    if (k < n) {
      const uint32_t e = entries[k];
      int32_t X[3], Y[3];
      float z0, dzdx, dzdy, zlo;
      const uint32_t prim = load_raster_rec(p, b.rec_base, e, X, Y, z0, dzdx, dzdy, zlo);
      const PixRect r = tri_pixels(X, Y, b.bin);
      if (r.x0 <= r.x1 && r.y0 <= r.y1) {
        const int w = r.x1 - r.x0 + 1;
        npx = (uint32_t)(w * (r.y1 - r.y0 + 1));
        PairTri& t = s_tri[threadIdx.x];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          t.X[q] = X[q];
          t.Y[q] = Y[q];
        }
        t.z0 = z0;
        t.dzdx = dzdx;
        t.dzdy = dzdy;
        t.zlo = zlo;
        t.prim = prim;
        t.x0 = (int16_t)r.x0;
        t.y0 = (int16_t)r.y0;
        t.w = (uint16_t)w;
        const int32_t ex = max(X[0], max(X[1], X[2])) - min(X[0], min(X[1], X[2]));
        const int32_t ey = max(Y[0], max(Y[1], Y[2])) - min(Y[0], min(Y[1], Y[2]));
        t.small = (ex < 32768 && ey < 32768) ? 1 : 0;
        t.inv_w = (65536u + (uint32_t)w - 1u) / (uint32_t)w;
      }
    }
    uint32_t incl = npx;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    s_pref[threadIdx.x] = incl;
    const uint32_t total = __shfl(incl, 63);
    wave_lds_sync();
    for (uint32_t base = 0; base < total; base += 64) {
      const uint32_t g = base + (uint32_t)lane;
      if (g < total) {
        int lo = 0, hi = 63;  // first triangle whose inclusive prefix exceeds g
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_pref[wbase + mid] > g) hi = mid;
          else lo = mid + 1;
        }
        const PairTri& t = s_tri[wbase + lo];
        const uint32_t local = g - (lo ? s_pref[wbase + lo - 1] : 0u);
        const uint32_t ly = (local * t.inv_w) >> 16;
        const uint32_t lx = local - ly * t.w;
        frags += raster_pair(t, s_vis, b.bin, B, t.x0 + (int)lx, t.y0 + (int)ly) ? 1u : 0u;
      }
    }
    pairs += lane == 0 ? total : 0u;
    wave_lds_sync();  // this wave's staging is rewritten by its next chunk
  }
  __syncthreads();
  store_vis(p, b, s_vis, B);
  if (threadIdx.x == 0) p.bin_count[b.gbin] = 0;  // ready for the next frame
  add_raster_stats(p, pairs, frags);
}

// exact fallback for bins whose list overflowed: every primitive of the
// scene.  Setup lists the overflowed bins of the phase (the reservation that
// crosses a bin's capacity appends it); kFullBlocks workgroups take them in
// turn (overflow is rare: the workgroups exit after reading the count).
constexpr int kFullBlocks = 64;

__global__ __launch_bounds__(kBlock) void k_raster_full(FrameParams p,
                                                        const RasterItem* __restrict__ items,
                                                        uint32_t n, int phase) {
  __shared__ unsigned long long s_vis[kMaxBin * kMaxBin];
  __shared__ BigTri s_big[kBigQueue];
  __shared__ uint32_t s_base[kDrawCache];
  __shared__ int s_nbig, s_next;
  __shared__ float s_lut[256];

  const uint32_t n_over =
      min(p.counters[(phase == 2 ? kCtrOverflow2 : kCtrOverflow1) * kCtrStride], n);
  if (blockIdx.x >= n_over) return;  // (whole workgroup)
  if (p.fuse_resolve) s_lut[threadIdx.x] = p.unorm_lut[threadIdx.x];  // (kBlock == 256)
  __syncthreads();
  const int B = p.bin_size;
  uint32_t pairs = 0, frags = 0;
  for (uint32_t o = blockIdx.x; o < n_over; o += gridDim.x) {
    const RasterItem it = items[p.overflow[o]];
    const BinInfo b = bin_info(it);
    const SceneDesc sc = p.scenes[it.scene];
    RasterCtx rc;
    rc.vis = s_vis;
    rc.big = s_big;
    rc.nbig = &s_nbig;
    rc.B = B;
    rc.lane_limit = p.lane_pixels;
    rc.bin = b.bin;
    rc.pairs = &pairs;
    rc.frags = &frags;
    for (int i = threadIdx.x; i < B * B; i += kBlock) s_vis[i] = kClearKey;
    if (sc.draw_count <= (uint32_t)kDrawCache)
      for (uint32_t i = threadIdx.x; i < sc.draw_count; i += kBlock)
        s_base[i] = p.draws[sc.draw_begin + i].prim_base;
    if (threadIdx.x == 0) s_nbig = s_next = 0;
    float PV[16];
    load_mat(p.cameras + (size_t)it.scene * 16, PV);
    __syncthreads();
    for (uint32_t base = 0; base < sc.prim_count; base += kBlock) {
      const uint32_t k = base + threadIdx.x;
      if (k < sc.prim_count) raster_prim(p, sc, PV, rc, s_base, k);
      __syncthreads();
      raster_big_waves(rc, s_big, min(s_nbig, kBigQueue), &s_next);
      __syncthreads();
      if (threadIdx.x == 0) s_nbig = s_next = 0;
      __syncthreads();
    }
    if (resolves_in_place(p, phase)) {
      resolve_tile<false>(p, it, s_vis, B, s_lut);
    } else {
      store_vis(p, b, s_vis, B);
      if (phase == 1 && p.hiz_fused) bin_hiz(p, b, s_vis);
    }
    if (threadIdx.x == 0) {
      add_stat(p.stats, 1, 1ull);
      p.bin_count[b.gbin] = 0;
    }
    __syncthreads();  // the next bin reuses the LDS tiles
  }
  add_raster_stats(p, pairs, frags);
}

// --------------------------------------------------------------------------
// k_resolve: per pixel, visibility key -> record -> colour, window depth,
// linear depth, semantic id; every output written once
// --------------------------------------------------------------------------

// The outputs of one bin from its final visibility keys: `tile` holds them
// in rows of B keys (an LDS tile), s_lut the unorm8 -> float table.  Called
// by every thread of the workgroup.  kLights: the renderer has lights
// enabled (a separate instance, so that renderers without lights keep the
// smaller register footprint).
template <bool kLights>
__device__ void resolve_tile(const FrameParams& p, const RasterItem& it,
                             const unsigned long long* tile, int B, const float* s_lut) {
  const BinInfo b = bin_info(it);
  const uint32_t frame = *p.frame;
  const float un0 = p.unprojections[2 * it.scene], un1 = p.unprojections[2 * it.scene + 1];
  const float far_depth = un1 / (1.0f + un0);
  const int bw = b.bin.x1 - b.bin.x0 + 1, bh = b.bin.y1 - b.bin.y0 + 1;
  const int npx = bw * bh;
  // in a full bin the pixels of a 2x2 quad sit in lanes {l, l^1, l^bw, l^1^bw}
  // of one wave (rows of bw pixels, bw in {8, 16, 32}, even bin origin)
  const bool quads = bw == p.bin_size && bh == p.bin_size;
  const int lane = threadIdx.x & 63;
  const float hw = (float)b.width * 0.5f, hh = (float)b.height * 0.5f;
  const float ihw = 1.0f / hw, ihh = 1.0f / hh;
  const float inv_bw = 1.0f / (float)bw;
  // the scene's sensor set
  const bool want_color = p.out_color && (b.planes & GBR_PLANE_COLOR);
  const bool want_depth = p.out_depth && (b.planes & GBR_PLANE_DEPTH);
  const bool want_lin = p.out_lindepth && (b.planes & GBR_PLANE_LINEAR_DEPTH);
  const bool want_sem = p.out_semantic && (b.planes & GBR_PLANE_SEMANTIC);
  for (int base = 0; base < npx; base += kBlock) {
    const int l = base + (int)threadIdx.x;
    const bool valid = l < npx;
    const int ly = div_small(l, inv_bw), lx = l - ly * bw;
    const int px = b.bin.x0 + lx, py = b.bin.y0 + ly;
    const unsigned long long key = valid ? tile[ly * B + lx] : kClearKey;
    const bool drawn = key != kClearKey;
    const uint32_t prim = (uint32_t)key;
    ResolveRec rs;
    DrawDesc d;
    TriAttr A;
    float bc[3] = {1.0f, 0.0f, 0.0f};
    float s = 0.0f, t = 0.0f;
    bool textured = false;
    if (drawn) {
      rs = p.srecs[(size_t)b.rec_base + prim];
      // every drawn primitive's record was written this frame; the clamp
      // only keeps the lookups in bounds
      const uint32_t item = min(rs.item, p.n_items - 1u);
      d = p.item_draws[item];
      if (p.item_vis) p.item_vis[item] = frame;  // visibility feedback for the next frame
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) A.e[i][k] = rs.e[3 * i + k];
      A.s[0] = rs.s0; A.t[0] = rs.t0;
      A.s[1] = rs.s1; A.t[1] = rs.t1;
      A.s[2] = rs.s2; A.t[2] = rs.t2;
      bary_at(A, hw, hh, ihw, ihh, (float)px + 0.5f, (float)py + 0.5f, bc);
      textured = !p.no_textures && d.texture > 0;
      if (textured) {
        s = interp3(bc, A.s[0], A.s[1], A.s[2]);
        t = interp3(bc, A.t[0], A.t[1], A.t[2]);
      }
    }
    QuadST q = {};
    if (quads) {
      const uint32_t tag = textured ? prim : 0xffffffffu;
      const int l00 = lane & ~1 & ~bw, l10 = (lane | 1) & ~bw, l01 = (lane & ~1) | bw;
      const uint32_t g00 = __shfl(tag, l00), g10 = __shfl(tag, l10), g01 = __shfl(tag, l01);
      q.s00 = __shfl(s, l00); q.t00 = __shfl(t, l00);
      q.s10 = __shfl(s, l10); q.t10 = __shfl(t, l10);
      q.s01 = __shfl(s, l01); q.t01 = __shfl(t, l01);
      q.h00 = textured && g00 == prim;
      q.h10 = textured && g10 == prim;
      q.h01 = textured && g01 == prim;
    }
    if (!valid) continue;
    const int row = p.top_down ? (b.height - 1 - py) : py;
    const size_t o = (size_t)b.plane_offset + (size_t)row * b.stride + px;
    if (!drawn) {
      if (want_color) p.out_color[o] = kClearColor;
      if (want_depth) p.out_depth[o] = 1.0f;
      if (want_lin) p.out_lindepth[o] = 0.0f;
      if (want_sem) p.out_semantic[o] = 0u;
      continue;
    }
    const float z = __uint_as_float((uint32_t)(key >> 32));
    if (want_color)
      p.out_color[o] = shade_pixel<kLights>(p, s_lut, hw, hh, ihw, ihh, A, bc, s, t, textured, q, d, prim,
                                            it.scene, px, py);
    if (want_depth) p.out_depth[o] = z;
    if (want_lin) {
      const float lin = un1 / (z + un0);
      p.out_lindepth[o] = lin == far_depth ? 0.0f : lin;
    }
    if (want_sem) p.out_semantic[o] = d.semantic;
  }
}

// a bin's stored keys into an LDS tile of rows of B keys
__device__ __forceinline__ void load_vis(const FrameParams& p, const BinInfo& b,
                                         unsigned long long* tile, int B) {
  constexpr int kPer = kMaxBin * kMaxBin / kBlock;
  const int bw = b.bin.x1 - b.bin.x0 + 1, bh = b.bin.y1 - b.bin.y0 + 1;
  unsigned long long keys[kPer];
  const float inv_bw = 1.0f / (float)bw;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int l = threadIdx.x + q * kBlock;
    const int ly = div_small(l, inv_bw), lx = l - ly * bw;
    keys[q] = l < bw * bh
                  ? p.vis[(size_t)b.plane_offset + (size_t)(b.bin.y0 + ly) * b.stride + b.bin.x0 + lx]
                  : 0ull;
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int l = threadIdx.x + q * kBlock;
    const int ly = div_small(l, inv_bw), lx = l - ly * bw;
    if (l < bw * bh) tile[ly * B + lx] = keys[q];
  }
}

// k_resolve: per bin, the stored visibility keys -> colour, window depth,
// linear depth, semantic id (the renderers whose final raster pass does not
// resolve in place: lights, the pair rasteriser)
template <bool kLights>
__global__ __launch_bounds__(kBlock) void k_resolve(FrameParams p,
                                                    const RasterItem* __restrict__ items) {
  __shared__ float s_lut[256];
  __shared__ unsigned long long s_tile[kMaxBin * kMaxBin];
  static_assert(kBlock >= 256, "one LUT entry per thread");
  if (threadIdx.x < 256) s_lut[threadIdx.x] = p.unorm_lut[threadIdx.x];
  const RasterItem it = items[blockIdx.x];
  load_vis(p, bin_info(it), s_tile, p.bin_size);
  __syncthreads();
  resolve_tile<kLights>(p, it, s_tile, p.bin_size, s_lut);
}

// --------------------------------------------------------------------------
// Chunk culling
//
// k_cull tests each work item's mesh-local bounding box against the view
// volume.  The box corners go through the same transform as the vertices;
// eps bounds the float error of one transformed coordinate (8 ulp of the
// |PV|.|M|.|p| magnitude, with an 8x margin), so an item rejected here has
// every vertex outside one clip plane as setup computes it: its triangles
// would all be trivially rejected, and skipping them changes nothing.
//
// Items that pass and lie strictly inside the near, far and guard-band
// planes (so none of their triangles is clipped) get a conservative pixel
// rectangle and the smallest window depth any fragment can have (fragments
// are never in front of their triangle's nearest vertex).  If that depth is
// greater than the largest phase-1 depth over the rectangle, every fragment
// of the item fails the LESS test against what phase 1 already stored, so
// the item is skipped in phase 2 without changing a pixel.
// --------------------------------------------------------------------------

constexpr float kEps = 4e-6f;  // relative error bound per clip coordinate (with margin)

__device__ __forceinline__ void abs_bound(const float* M, const float* PV, float ax, float ay,
                                          float az, float* e) {
  float m[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    m[r] = fabsf(M[r]) * ax + fabsf(M[4 + r]) * ay + fabsf(M[8 + r]) * az + fabsf(M[12 + r]);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float a = fabsf(PV[r]) * m[0] + fabsf(PV[4 + r]) * m[1] + fabsf(PV[8 + r]) * m[2] +
                    fabsf(PV[12 + r]) * m[3];
    e[r] = kEps * a * 1.001f + 1e-30f;
  }
}

// Block-aggregated appends to L lists: one device atomic per list and block.
// (Single-address atomics from every wave serialise at the memory side and
// cost more than the work they count.)  Every thread of the block calls it
// once, with its flags; slot[l] is valid where flag[l] is set.
constexpr int kListBlock = 1024;

template <int L>
__device__ __forceinline__ void block_slots(uint32_t* const (&counter)[L], const bool (&flag)[L],
                                            uint32_t (&slot)[L]) {
  __shared__ uint32_t s_ofs[L][kListBlock / 64];
  __shared__ uint32_t s_base[L];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int waves = (int)(blockDim.x >> 6);
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned long long m[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    m[l] = __ballot(flag[l]);
    if (lane == 0) s_ofs[l][wave] = (uint32_t)__popcll(m[l]);
  }
  __syncthreads();
  if (threadIdx.x < L) {
    const int l = threadIdx.x;
    uint32_t total = 0;
    for (int w = 0; w < waves; ++w) {
      const uint32_t c = s_ofs[l][w];
      s_ofs[l][w] = total;
      total += c;
    }
    s_base[l] = total ? atomicAdd(counter[l], total) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int l = 0; l < L; ++l) slot[l] = s_base[l] + s_ofs[l][wave] + (uint32_t)__popcll(m[l] & lt_mask);
}

__device__ __forceinline__ uint32_t f2bf_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0x7fffffu) ? 0x40u : 0u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// features only steer the proposal (any proposal renders exactly), so they
// use the hardware log2 and reciprocal (~1 ulp)
__device__ __forceinline__ float log2c(float x) { return __builtin_amdgcn_logf(fmaxf(x, 1e-3f)); }

// The 32 features of a (camera, work item) pair for the visibility model,
// history feature set (an ablation of the pose set below):
//  0-2   visibility history: visible last frame, within 8 frames, age/64
//  3-7   conservative pixel rectangle (normalised) and its area fraction
//  8     log2 of the nearest linear depth / 8
//  9-13  box centre in NDC x, y; log2 of centre, nearest, farthest clip w / 8
//  14-15 log2 triangles / 10, log2 world-space box diagonal / 4
//  16-18 world box centre / 8, 19-21 camera position / 8, 22-24 camera forward
//  25-28 centre minus camera / 8 and its length / 8
//  29    cosine between forward and the direction to the centre
//  30    rectangle touches the tile border, 31 constant 1
__device__ void item_features_history(const FrameParams& p, const WorkItem& wi, const SceneDesc& sc,
                              const ItemBounds& bb, const float* M, const float* PV,
                              const ItemCull& ic, float wmin, float wmax, uint32_t item,
                              uint4* row) {
  float f[32];
  const uint32_t seen = p.item_vis[item], frame = *p.frame;
  f[0] = (seen + 1u == frame) ? 1.0f : 0.0f;
  f[1] = (seen != 0u && frame - seen <= 8u) ? 1.0f : 0.0f;
  f[2] = seen == 0u ? 1.0f : fminf((float)(frame - seen), 64.0f) / 64.0f;
  const float iW = 1.0f / (float)sc.width, iH = 1.0f / (float)sc.height;
  f[3] = ic.x0 * iW;
  f[4] = ic.y0 * iH;
  f[5] = (ic.x1 + 1) * iW;
  f[6] = (ic.y1 + 1) * iH;
  f[7] = (f[5] - f[3]) * (f[6] - f[4]);
  const float un0 = p.unprojections[2 * wi.scene], un1 = p.unprojections[2 * wi.scene + 1];
  f[8] = log2c(__fdividef(un1, fmaxf(ic.zmin, 0.0f) + un0)) / 8.0f;
  const float cx = 0.5f * (bb.lo[0] + bb.hi[0]), cy = 0.5f * (bb.lo[1] + bb.hi[1]),
              cz = 0.5f * (bb.lo[2] + bb.hi[2]);
  const float4 cc = vertex_clip(M, PV, cx, cy, cz);
  const float icw = __builtin_amdgcn_rcpf(cc.w);
  f[9] = fmaxf(fminf(cc.x * icw, 4.0f), -4.0f);
  f[10] = fmaxf(fminf(cc.y * icw, 4.0f), -4.0f);
  f[11] = log2c(cc.w) / 8.0f;
  f[12] = log2c(wmin) / 8.0f;
  f[13] = log2c(wmax) / 8.0f;
  f[14] = log2c((float)wi.tri_count) / 10.0f;
  float wc[4], wd[4];
  mat4_vec(M, cx, cy, cz, 1.0f, wc);
  mat4_vec(M, bb.hi[0] - bb.lo[0], bb.hi[1] - bb.lo[1], bb.hi[2] - bb.lo[2], 0.0f, wd);
  f[15] = log2c(__builtin_amdgcn_sqrtf(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2])) / 4.0f;
  const float* cam = p.cam_info + (size_t)wi.scene * 8;
  float d[3], dd = 0.0f, cosf_ = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f[16 + k] = wc[k] / 8.0f;
    f[19 + k] = cam[k] / 8.0f;
    f[22 + k] = cam[3 + k];
    d[k] = wc[k] - cam[k];
    f[25 + k] = d[k] / 8.0f;
    dd += d[k] * d[k];
    cosf_ += d[k] * cam[3 + k];
  }
  dd = __builtin_amdgcn_sqrtf(dd);
  f[28] = dd / 8.0f;
  f[29] = dd > 0.0f ? cosf_ * __builtin_amdgcn_rcpf(dd) : 1.0f;
  f[30] = (ic.x0 == 0 || ic.y0 == 0 || ic.x1 == sc.width - 1 || ic.y1 == sc.height - 1) ? 1.0f : 0.0f;
  f[31] = 1.0f;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint4 v;
    v.x = f2bf_bits(f[8 * q + 0]) | (f2bf_bits(f[8 * q + 1]) << 16);
    v.y = f2bf_bits(f[8 * q + 2]) | (f2bf_bits(f[8 * q + 3]) << 16);
    v.z = f2bf_bits(f[8 * q + 4]) | (f2bf_bits(f[8 * q + 5]) << 16);
    v.w = f2bf_bits(f[8 * q + 6]) | (f2bf_bits(f[8 * q + 7]) << 16);
    row[q] = v;
  }
}

// The 32 features of a (camera, work item) pair, pose feature set: the
// camera pose and the chunk's world box relative to the camera, nothing
// from earlier frames.
//  0-2   camera position / 8
//  3-4   sin, cos of the camera yaw; 5-6 sin, cos of its pitch
//  7-9   box centre - camera position / 8
//  10-12 box extents (world axis-aligned) / 4
//  13-15 centre - camera in the camera frame (right, up, forward) / 8
//  16    distance to the centre / 8, 17 its log2 / 4
//  18    cosine between the view direction and the direction to the centre
//  19-20 centre direction in the camera frame: right/forward, up/forward
//        (clamped to +-4)
//  21    box half-diagonal / distance (angular size, clamped to 4)
//  22-24 extents along the camera's right, up and forward axes / 4
//  25-27 box corner nearest the camera - camera position / 8 (per axis)
//  28-30 farthest corner - camera position / 8, 31 constant 1
__device__ void item_features_pose(const FrameParams& p, const WorkItem& wi,
                                   const ItemBounds& bb, const float* M, uint4* row) {
  float f[32];
  const float* cam = p.cam_info + (size_t)wi.scene * 8;
  const float fx = cam[3], fy = cam[4], fz = cam[5];
  // yaw about +y with the camera looking down -z at yaw 0; pitch up positive
  const float hz = __builtin_amdgcn_sqrtf(fx * fx + fz * fz);
  const float ih = hz > 1e-6f ? __builtin_amdgcn_rcpf(hz) : 0.0f;
  const float syaw = hz > 1e-6f ? -fx * ih : 0.0f, cyaw = hz > 1e-6f ? -fz * ih : 1.0f;
  // camera frame: right = forward x world up (normalised), up = right x forward
  const float rx = -fz * ih, rz = fx * ih;  // (ry = 0)
  const float ux = -rz * fy, uy = rz * fx - rx * fz, uz = rx * fy;
  const float cx = 0.5f * (bb.lo[0] + bb.hi[0]), cy = 0.5f * (bb.lo[1] + bb.hi[1]),
              cz = 0.5f * (bb.lo[2] + bb.hi[2]);
  const float ex = bb.hi[0] - bb.lo[0], ey = bb.hi[1] - bb.lo[1], ez = bb.hi[2] - bb.lo[2];
  float wc[4];
  mat4_vec(M, cx, cy, cz, 1.0f, wc);
  float we[3];  // world axis-aligned extents of the transformed box
#pragma unroll
  for (int k = 0; k < 3; ++k)
    we[k] = fabsf(M[k]) * ex + fabsf(M[4 + k]) * ey + fabsf(M[8 + k]) * ez;
  float d[3], lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    d[k] = wc[k] - cam[k];
    const float a = d[k] - 0.5f * we[k], b = d[k] + 0.5f * we[k];
    // per axis, the box coordinate nearest to / farthest from the camera
    lo[k] = (a > 0.0f) ? a : (b < 0.0f ? b : 0.0f);
    hi[k] = fabsf(a) > fabsf(b) ? a : b;
  }
  const float dr = d[0] * rx + d[2] * rz;
  const float du = d[0] * ux + d[1] * uy + d[2] * uz;
  const float df = d[0] * fx + d[1] * fy + d[2] * fz;
  const float dist = __builtin_amdgcn_sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const float idist = dist > 1e-6f ? __builtin_amdgcn_rcpf(dist) : 0.0f;
  const float half_diag = 0.5f * __builtin_amdgcn_sqrtf(we[0] * we[0] + we[1] * we[1] + we[2] * we[2]);
  const float idf = fabsf(df) > 1e-6f ? __builtin_amdgcn_rcpf(df) : 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f[k] = cam[k] * 0.125f;
    f[7 + k] = d[k] * 0.125f;
    f[10 + k] = we[k] * 0.25f;
    f[25 + k] = lo[k] * 0.125f;
    f[28 + k] = hi[k] * 0.125f;
  }
  f[3] = syaw;
  f[4] = cyaw;
  f[5] = fy;  // sin(pitch) of a unit forward vector
  f[6] = hz;  // cos(pitch)
  f[13] = dr * 0.125f;
  f[14] = du * 0.125f;
  f[15] = df * 0.125f;
  f[16] = dist * 0.125f;
  f[17] = log2c(dist) * 0.25f;
  f[18] = dist > 1e-6f ? df * idist : 1.0f;
  f[19] = fmaxf(fminf(dr * idf, 4.0f), -4.0f);
  f[20] = fmaxf(fminf(du * idf, 4.0f), -4.0f);
  f[21] = dist > 1e-6f ? fminf(half_diag * idist, 4.0f) : 4.0f;
  f[22] = (fabsf(rx) * we[0] + fabsf(rz) * we[2]) * 0.25f;
  f[23] = (fabsf(ux) * we[0] + fabsf(uy) * we[1] + fabsf(uz) * we[2]) * 0.25f;
  f[24] = (fabsf(fx) * we[0] + fabsf(fy) * we[1] + fabsf(fz) * we[2]) * 0.25f;
  f[31] = 1.0f;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint4 v;
    v.x = f2bf_bits(f[8 * q + 0]) | (f2bf_bits(f[8 * q + 1]) << 16);
    v.y = f2bf_bits(f[8 * q + 2]) | (f2bf_bits(f[8 * q + 3]) << 16);
    v.z = f2bf_bits(f[8 * q + 4]) | (f2bf_bits(f[8 * q + 5]) << 16);
    v.w = f2bf_bits(f[8 * q + 6]) | (f2bf_bits(f[8 * q + 7]) << 16);
    row[q] = v;
  }
}

__global__ __launch_bounds__(kListBlock) void k_cull(FrameParams p) {
  const uint32_t item = blockIdx.x * kListBlock + threadIdx.x;
  bool to1 = false, toc = false, feat = false;
  // the feature row, packed BF16, computed while the item's data is in
  // registers and stored once its row index is known
  uint4 frow[4] = {};
  if (item < p.n_items) {
    const WorkItem wi = p.items[item];
    const SceneDesc sc = p.scenes[wi.scene];
    const ItemBounds bb = p.bounds[item];
    float M[16], PV[16];
    load_mat(p.abs_nodes + (size_t)wi.node * 16, M);
    load_mat(p.cameras + (size_t)wi.scene * 16, PV);
    float e[4];
    abs_bound(M, PV, fmaxf(fabsf(bb.lo[0]), fabsf(bb.hi[0])), fmaxf(fabsf(bb.lo[1]), fabsf(bb.hi[1])),
              fmaxf(fabsf(bb.lo[2]), fabsf(bb.hi[2])), e);
    const float hw = (float)sc.width * 0.5f, hh = (float)sc.height * 0.5f;
    const float gx = kGuardPx / hw, gy = kGuardPx / hh;
    const float mx = 2.0f * (e[0] + e[3]), my = 2.0f * (e[1] + e[3]), mz = 2.0f * (e[2] + e[3]);
    uint32_t out_all = 63u;
    bool inner = true;  // strictly inside near, far and the guard band
    float xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    float zmin = INFINITY, wmin = INFINITY, wmax = 0.0f;
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
      const float4 c = vertex_clip(M, PV, (k & 1) ? bb.hi[0] : bb.lo[0], (k & 2) ? bb.hi[1] : bb.lo[1],
                                   (k & 4) ? bb.hi[2] : bb.lo[2]);
      uint32_t o = 0;
      if (c.x - c.w > mx) o |= 1;
      if (-c.x - c.w > mx) o |= 2;
      if (c.y - c.w > my) o |= 4;
      if (-c.y - c.w > my) o |= 8;
      if (-c.z - c.w > mz) o |= 16;
      if (c.z - c.w > mz) o |= 32;
      out_all &= o;
      inner = inner && c.w > 2.0f * e[3] && c.z + c.w > mz && c.w - c.z > mz &&
              gx * c.w - fabsf(c.x) > 2.0f * (gx * e[3] + e[0]) &&
              gy * c.w - fabsf(c.y) > 2.0f * (gy * e[3] + e[1]);
      if (inner) {
        const float iw = 1.0f / c.w;
        const float nx = c.x * iw, ny = c.y * iw, nz = c.z * iw;
        xmin = fminf(xmin, nx); xmax = fmaxf(xmax, nx);
        ymin = fminf(ymin, ny); ymax = fmaxf(ymax, ny);
        zmin = fminf(zmin, nz); wmin = fminf(wmin, c.w); wmax = fmaxf(wmax, c.w);
      }
    }
    // learned mode with history: a chunk that passed the frustum test in the
    // previous frame without being visible was occluded there; it is not
    // scored and waits for the depth-pyramid test, as in the Hi-Z mode (the
    // model scores the chunks that come into view)
    bool occluded_before = false;
    if (out_all == 0u && p.cull_mode == kCullLearned) {
      const uint32_t fr = *p.frame;
      occluded_before = p.proposal_history && !p.collecting && p.item_view[item] + 1u == fr;
      p.item_view[item] = fr;
    }
    if (out_all == 0u) {
      if (!inner) {
        to1 = true;  // crosses a clip plane: always drawn in phase 1
      } else {
        // error of a projected coordinate: (dx + |x/w| dw) / (w - dw), plus rounding
        const float den = wmin - 2.0f * e[3];
        const float ax = fmaxf(fabsf(xmin), fabsf(xmax)), ay = fmaxf(fabsf(ymin), fabsf(ymax));
        const float ex = (2.0f * e[0] + ax * 2.0f * e[3]) / den + 1e-6f * (1.0f + ax);
        const float ey = (2.0f * e[1] + ay * 2.0f * e[3]) / den + 1e-6f * (1.0f + ay);
        const float ez = (2.0f * e[2] + 2.0f * e[3]) / den + 4e-6f;
        const int x0 = max((int)floorf((xmin - ex) * hw + hw - 0.5f) - 1, 0);
        const int x1 = min((int)ceilf((xmax + ex) * hw + hw - 0.5f) + 1, sc.width - 1);
        const int y0 = max((int)floorf((ymin - ey) * hh + hh - 0.5f) - 1, 0);
        const int y1 = min((int)ceilf((ymax + ey) * hh + hh - 0.5f) + 1, sc.height - 1);
        if (x0 <= x1 && y0 <= y1) {  // else: no pixel centre of the tile can be covered
          ItemCull ic;
          ic.x0 = (int16_t)x0; ic.y0 = (int16_t)y0; ic.x1 = (int16_t)x1; ic.y1 = (int16_t)y1;
          ic.zmin = (zmin - ez) * 0.5f + 0.5f;
          ic.pad[0] = wi.scene;
          ic.pad[1] = 0u;
          // learned mode with history: a chunk visible in the previous frame
          // goes to phase 1 without a score (unless rows are collected, which
          // needs every chunk's row)
          const bool seen = p.cull_mode == kCullLearned && p.proposal_history && !p.collecting &&
                            p.item_vis[item] + 1u == *p.frame;
          if (seen) occluded_before = false;
          if (p.features && !seen && !occluded_before) {
            feat = true;
            if (p.feature_set == kFeaturesHistory)
              item_features_history(p, wi, sc, bb, M, PV, ic, wmin, wmax, item, frow);
            else
              item_features_pose(p, wi, bb, M, frow);
          }
          if (seen) {
            to1 = true;
          } else if (occluded_before) {
            p.cull[item] = ic;
            toc = true;
          } else if (p.cull_mode == kCullLearned) {
            p.cull[item] = ic;  // k_propose decides from the model's score
          } else {
            const bool proposed =
                p.cull_mode == kCullFrustum || p.item_vis[item] + 1u == *p.frame;
            if (proposed) {
              to1 = true;
            } else {
              p.cull[item] = ic;
              toc = true;
            }
          }
        }
      }
    }
  }
  uint32_t* const ctr[3] = {&p.counters[kCtrList1 * kCtrStride], &p.counters[kCtrCand * kCtrStride],
                            &p.counters[kCtrProp * kCtrStride]};
  const bool flag[3] = {to1, toc, feat};
  uint32_t slot[3];
  block_slots<3>(ctr, flag, slot);
  if (to1) p.list1[slot[0]] = item;
  if (toc) p.cand[slot[1]] = item;
  if (feat) {
    const uint32_t row = slot[2];
    p.prop[row] = item;
    uint4* dst = reinterpret_cast<uint4*>(p.features + (size_t)row * 32);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = frow[q];
  }
}

// learned mode: the model's score (with proposal_history, also visibility
// in the previous frame) splits the feature rows into phase 1 and the
// depth-pyramid candidates
__global__ __launch_bounds__(kListBlock) void k_propose(FrameParams p) {
  const uint32_t row = blockIdx.x * kListBlock + threadIdx.x;
  const uint32_t n = min(p.counters[kCtrProp * kCtrStride], p.n_items);
  if (blockIdx.x * kListBlock >= n) return;  // (whole block: no barrier is skipped)
  bool to1 = false, toc = false;
  uint32_t item = 0;
  if (row < n) {
    item = p.prop[row];
    to1 = p.logits[row] > p.proposal_threshold ||
          (p.proposal_history && p.item_vis[item] + 1u == *p.frame);
    toc = !to1;
  }
  uint32_t* const ctr[2] = {&p.counters[kCtrList1 * kCtrStride], &p.counters[kCtrCand * kCtrStride]};
  const bool flag[2] = {to1, toc};
  uint32_t slot[2];
  block_slots<2>(ctr, flag, slot);
  if (to1) p.list1[slot[0]] = item;
  if (toc) p.cand[slot[1]] = item;
}

// training rows of this frame: features and whether the item ended up with
// a visible pixel, appended at dataset row base (rows past cap dropped)
__global__ __launch_bounds__(kBlock) void k_collect(FrameParams p, uint16_t* __restrict__ X,
                                                    uint8_t* __restrict__ Y, uint64_t base,
                                                    uint64_t cap, uint32_t* __restrict__ positives) {
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n = min(p.counters[kCtrProp * kCtrStride], p.n_items);
  bool pos = false;
  if (row < n && base + row < cap) {
    const uint32_t item = p.prop[row];
    pos = p.item_vis[item] == *p.frame;
    const uint4* src = reinterpret_cast<const uint4*>(p.features + (size_t)row * 32);
    uint4* dst = reinterpret_cast<uint4*>(X + (size_t)(base + row) * 32);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = src[q];
    Y[base + row] = pos ? 1 : 0;
  }
  const unsigned long long m = __ballot(pos);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(positives, (uint32_t)__popcll(m));
}

// depth pyramid from the phase-1 visibility keys: one workgroup per 32x32
// region of a scene
__global__ __launch_bounds__(kBlock) void k_hiz(FrameParams p, uint32_t regions_per_scene) {
  constexpr int kN = kHiZPer * kHiZPer;
  __shared__ uint32_t s_max[kN];
  const uint32_t scene = blockIdx.x / regions_per_scene;
  const uint32_t reg = blockIdx.x % regions_per_scene;
  const SceneDesc sc = p.scenes[scene];
  const int rx = (int)(reg % (uint32_t)p.hiz_w1), ry = (int)(reg / (uint32_t)p.hiz_w1);
  if (threadIdx.x < kN) s_max[threadIdx.x] = 0u;
  __syncthreads();
  const int lx = threadIdx.x & 31;
  const int px = rx * 32 + lx;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ly = (threadIdx.x >> 5) + 8 * q;
    const int py = ry * 32 + ly;
    if (px < sc.width && py < sc.height) {
      const unsigned long long key = p.vis[(size_t)sc.plane_offset + (size_t)py * sc.stride + px];
      atomicMax(&s_max[(ly >> kHiZ0) * kHiZPer + (lx >> kHiZ0)], (uint32_t)(key >> 32));  // depth >= 0
    }
  }
  __syncthreads();
  float* h = p.hiz + (size_t)scene * p.hiz_stride;
  if (threadIdx.x < kN) {
    const int bx = rx * kHiZPer + (int)(threadIdx.x % kHiZPer),
              by = ry * kHiZPer + (int)(threadIdx.x / kHiZPer);
    if (bx < p.hiz_w0 && by < p.hiz_h0) h[by * p.hiz_w0 + bx] = __uint_as_float(s_max[threadIdx.x]);
  }
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int k = 0; k < kN; ++k) m = max(m, s_max[k]);
    h[p.hiz_w0 * p.hiz_h0 + ry * p.hiz_w1 + rx] = __uint_as_float(m);
  }
}

// candidates against the depth pyramid -> phase-2 list
__global__ __launch_bounds__(kListBlock) void k_cull_hiz(FrameParams p) {
  const uint32_t i = blockIdx.x * kListBlock + threadIdx.x;
  const uint32_t n = min(p.counters[kCtrCand * kCtrStride], p.n_items);
  if (blockIdx.x * kListBlock >= n) return;  // (whole block: no barrier is skipped)
  bool keep = false;
  uint32_t item = 0;
  if (i < n) {
    item = p.cand[i];
    const ItemCull ic = p.cull[item];
    keep = !hiz_hidden_wide(p, ic.pad[0], ic.x0, ic.y0, ic.x1, ic.y1, ic.zmin, kHiZTiles);
  }
  uint32_t* const ctr[1] = {&p.counters[kCtrList2 * kCtrStride]};
  const bool flag[1] = {keep};
  uint32_t slot[1];
  block_slots<1>(ctr, flag, slot);
  if (keep) p.list2[slot[0]] = item;
}

// --------------------------------------------------------------------------
// Readback helpers: gather a framebuffer rectangle (tile grid, bottom-up)
// from the per-scene planes.
// --------------------------------------------------------------------------



__global__ __launch_bounds__(kBlock) void k_gather(GatherArgs g) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (long long)g.w * g.h) return;
  const int lx = (int)(i % g.w), ly = (int)(i / g.w);
  const int fx = g.x0 + lx, fy = g.y0 + ly;
  const int tx = fx / g.tile_w, ty = fy / g.tile_h;
  const int scene = ty * g.tiles_x + tx;
  const int px = fx - tx * g.tile_w, py = fy - ty * g.tile_h;
  uint32_t v = (uint32_t)g.fill;
  if (scene < g.scene_count) {
    // the scene's viewport is the bottom-left of its tile; the rest is clear
    const int vw = g.scenes[scene].width, vh = g.scenes[scene].height;
    if (px < vw && py < vh) {
      const int row = g.top_down ? (vh - 1 - py) : py;
      v = *reinterpret_cast<const uint32_t*>(
          g.plane + (((long long)scene * g.tile_h + row) * g.tile_w + px) * 4);
    }
  }
  uint8_t* d = g.dst + (long long)ly * g.dst_stride + (long long)lx * g.dst_bytes;
  if (g.dst_bytes == 4) {
    *reinterpret_cast<uint32_t*>(d) = v;
  } else {
    d[0] = v & 0xff;
    d[1] = (v >> 8) & 0xff;
    d[2] = (v >> 16) & 0xff;
  }
}

// --------------------------------------------------------------------------
// Host-side launchers
// --------------------------------------------------------------------------

hipError_t launch_nodes(const int32_t* parents, const float* local, float* abs_out, uint32_t n,
                        uint32_t* counters, unsigned long long* stats, uint32_t* frame,
                        float* nrm_out, const uint32_t* dirty, uint32_t max_blocks, hipStream_t s) {
  // always runs (it resets the counters): at least one workgroup
  uint32_t blocks = n == 0 ? 1 : (n + kBlock - 1) / kBlock;
  if (max_blocks > 0 && blocks > max_blocks) blocks = max_blocks;
  hipLaunchKernelGGL(k_nodes, dim3(blocks), dim3(kBlock), 0, s, parents, local, abs_out, n,
                     counters, stats, frame, nrm_out, dirty);
  return hipGetLastError();
}

int setup_blocks_per_cu(uint32_t lds_bins) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_setup_bin, kBlock,
                                                   2 * lds_bins * sizeof(uint32_t)) != hipSuccess ||
      b < 1)
    b = 4;
  return b;
}

// list == nullptr: every work item of the frame; else the items listed in
// list[0 .. counters[list_ctr])
hipError_t launch_setup(const FrameParams& p, const uint32_t* list, uint32_t list_ctr,
                        uint32_t queue_ctr, uint32_t grid, hipStream_t s) {
  if (p.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_bin, dim3(grid < p.n_items ? grid : p.n_items), dim3(kBlock),
                     2 * p.lds_bins * sizeof(uint32_t), s, p, list, list_ctr, queue_ctr);
  return hipGetLastError();
}

hipError_t launch_clip(const FrameParams& p, uint32_t grid, int phase, hipStream_t s) {
  hipLaunchKernelGGL(k_setup_clip, dim3(grid), dim3(kClipBlock), 0, s, p, phase);
  return hipGetLastError();
}

// the lane rasteriser's workgroups per CU: kind 0 stores the keys, 1
// resolves a frame's only raster phase, 2 resolves after the depth pyramid
int raster_blocks_per_cu(int kind) {
  int b = 0;
  const hipError_t e =
      kind == 1   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_raster_lane<true, 5>, kBlock, 0)
      : kind == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_raster_lane<true, 4>, kBlock, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_raster_lane<false, 6>, kBlock, 0);
  if (e != hipSuccess || b < 1) b = 4;
  return b;
}

// raster_mode 0: `grid` persistent workgroups (at most one per bin) drain
// the phase's queue; mode 1: one workgroup per bin
hipError_t launch_raster(const FrameParams& p, const RasterItem* items, uint32_t n, int phase,
                         uint32_t grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (p.raster_mode == 1)
    hipLaunchKernelGGL(k_raster_pairs, dim3(n), dim3(kBlock), 0, s, p, items, phase);
  else if (resolves_in_place(p, phase) && phase == 1)
    hipLaunchKernelGGL((k_raster_lane<true, 5>), dim3(grid < n ? grid : n), dim3(kBlock), 0, s, p,
                       items, n, phase);
  else if (resolves_in_place(p, phase))
    hipLaunchKernelGGL((k_raster_lane<true, 4>), dim3(grid < n ? grid : n), dim3(kBlock), 0, s, p,
                       items, n, phase);
  else
    hipLaunchKernelGGL((k_raster_lane<false, 6>), dim3(grid < n ? grid : n), dim3(kBlock), 0, s, p,
                       items, n, phase);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_raster_full, dim3(n < (uint32_t)kFullBlocks ? n : (uint32_t)kFullBlocks),
                     dim3(kBlock), 0, s, p,
                     items, n, phase);
  return hipGetLastError();
}

hipError_t launch_resolve(const FrameParams& p, const RasterItem* items, uint32_t n,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (p.lights)
    hipLaunchKernelGGL(k_resolve<true>, dim3(n), dim3(kBlock), 0, s, p, items);
  else
    hipLaunchKernelGGL(k_resolve<false>, dim3(n), dim3(kBlock), 0, s, p, items);
  return hipGetLastError();
}

hipError_t launch_cull(const FrameParams& p, hipStream_t s) {
  if (p.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cull, dim3((p.n_items + kListBlock - 1) / kListBlock), dim3(kListBlock), 0, s,
                     p);
  return hipGetLastError();
}

hipError_t launch_propose(const FrameParams& p, hipStream_t s) {
  if (p.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_propose, dim3((p.n_items + kListBlock - 1) / kListBlock), dim3(kListBlock), 0, s,
                     p);
  return hipGetLastError();
}

hipError_t launch_collect(const FrameParams& p, uint16_t* X, uint8_t* Y, uint64_t base,
                          uint64_t cap, uint32_t* positives, hipStream_t s) {
  if (p.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_collect, dim3((p.n_items + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p, X,
                     Y, base, cap, positives);
  return hipGetLastError();
}

hipError_t launch_hiz(const FrameParams& p, uint32_t scenes, hipStream_t s) {
  const uint32_t regions = (uint32_t)(p.hiz_w1 * p.hiz_h1);
  if (scenes == 0 || regions == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hiz, dim3(scenes * regions), dim3(kBlock), 0, s, p, regions);
  return hipGetLastError();
}

hipError_t launch_cull_hiz(const FrameParams& p, hipStream_t s) {
  if (p.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cull_hiz, dim3((p.n_items + kListBlock - 1) / kListBlock), dim3(kListBlock), 0, s,
                     p);
  return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs& g, hipStream_t s) {
  long long n = (long long)g.w * g.h;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, g);
  return hipGetLastError();
}

}  // namespace gbr
