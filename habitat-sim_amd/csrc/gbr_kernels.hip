// HIP kernels of the batch renderer for gfx950 (CDNA4, wave64).
//
// Pipeline per frame (all scenes in one launch per stage):
//   k_nodes      per node: absolute transform = parent chain product
//                (Renderer.cpp:1214-1220 moved to the GPU)
//   k_setup_bin  per work item (scene, draw, chunk of <= 256 triangles):
//                vertex transform of the chunk's vertex range into LDS
//                (coalesced SoA position streams), fixed-point triangle setup
//                (8 sub-pixel bits, top-left rule), trivial/back-face/empty
//                rejection, screen-bin append with wave-aggregated atomics
//                (ballot + prefix popcount)
//   k_raster     per (scene, bin): re-setup of the bin's triangles, depth test
//                as a 64-bit atomicMin of (window depth bits, primitive order)
//                in an LDS visibility tile, then the resolve in the same
//                workgroup: perspective-correct attributes, mip-mapped
//                texture sampling, RGBA8 / window depth / linear depth /
//                semantic written once to HBM
//
// Semantics follow OpenGL as driven by the reference (Renderer.cpp:307-308,
// 1298-1351; RendererStandalone.cpp:155-159; DepthUnprojection.cpp:98-130);
// every float expression is evaluated in the same order as the CPU oracle
// so results agree bit for bit (the file is compiled with -ffp-contract=off).

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

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
constexpr int kSmallRaster = 48;  // pixels a single lane rasterises alone

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

__global__ __launch_bounds__(kBlock) void k_nodes(const int32_t* __restrict__ parents,
                                                  const float* __restrict__ local,
                                                  float* __restrict__ abs_out, uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
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
}

// --------------------------------------------------------------------------
// k_setup_bin
// --------------------------------------------------------------------------

__device__ __forceinline__ void load_mat(const float* src, float* m) {
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = src[k];
}

__global__ __launch_bounds__(kBlock) void k_setup_bin(FrameParams p,
                                                      const WorkItem* __restrict__ items) {
  __shared__ int32_t s_x[kVertexCache];
  __shared__ int32_t s_y[kVertexCache];
  __shared__ uint32_t s_f[kVertexCache];

  const WorkItem wi = items[blockIdx.x];
  const DrawDesc d = p.draws[wi.draw];
  const SceneDesc sc = p.scenes[wi.scene];
  float M[16], PV[16];
  load_mat(p.abs_nodes + (size_t)d.node * 16, M);
  load_mat(p.cameras + (size_t)wi.scene * 16, PV);
  const float hw = (float)sc.width * 0.5f, hh = (float)sc.height * 0.5f;
  const float gx = kGuardPx / hw, gy = kGuardPx / hh;
  const bool cached = wi.vcount <= (uint32_t)kVertexCache;

  if (cached) {
    for (uint32_t v = threadIdx.x; v < wi.vcount; v += kBlock) {
      uint32_t gv = wi.vmin + v;
      float4 c = vertex_clip(M, PV, p.pos_x[gv], p.pos_y[gv], p.pos_z[gv]);
      Snap s = snap_vertex(c, hw, hh, gx, gy);
      s_x[v] = s.X;
      s_y[v] = s.Y;
      s_f[v] = s.flags;
    }
  }
  __syncthreads();

  const PixRect tile = {0, 0, sc.width - 1, sc.height - 1};
  bool active = false;
  PixRect pr = {0, 0, -1, -1};
  uint32_t prim = 0;
  if (threadIdx.x < wi.tri_count) {
    const uint32_t t = wi.tri_start + threadIdx.x;
    prim = d.prim_base + t;
    const uint32_t* ip = p.indices + d.index_base + 3 * t;
    uint32_t vi[3] = {ip[0], ip[1], ip[2]};
    int32_t X[3], Y[3];
    uint32_t f[3];
    if (cached) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        uint32_t l = vi[k] - wi.vmin;
        X[k] = s_x[l];
        Y[k] = s_y[l];
        f[k] = s_f[l];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float4 c = vertex_clip(M, PV, p.pos_x[vi[k]], p.pos_y[vi[k]], p.pos_z[vi[k]]);
        Snap s = snap_vertex(c, hw, hh, gx, gy);
        X[k] = s.X;
        Y[k] = s.Y;
        f[k] = s.flags;
      }
    }
    if ((f[0] & f[1] & f[2] & 63u) == 0) {
      if (((f[0] | f[1] | f[2]) & kNeedClip) == 0) {
        if (tri_area(X, Y) > 0) {
          pr = tri_pixels(X, Y, tile);
          active = pr.x0 <= pr.x1 && pr.y0 <= pr.y1;
        }
      } else {
        // rare path: clip against near/far/guard planes, bin the union of
        // the front-facing fan triangles
        float4 poly[kMaxPoly];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          poly[k] = vertex_clip(M, PV, p.pos_x[vi[k]], p.pos_y[vi[k]], p.pos_z[vi[k]]);
        int n = clip_poly(poly, 3, gx, gy);
        bool ok = n >= 3;
        int32_t PX[kMaxPoly], PY[kMaxPoly];
        float PZ[kMaxPoly];
        for (int k = 0; k < n && ok; ++k) {
          if (!(poly[k].w > 0.0f)) ok = false;
          else project(poly[k], hw, hh, PX[k], PY[k], PZ[k]);
        }
        if (ok) {
          PixRect u = {1 << 30, 1 << 30, -(1 << 30), -(1 << 30)};
          for (int k = 1; k + 1 < n; ++k) {
            int32_t tx[3] = {PX[0], PX[k], PX[k + 1]};
            int32_t ty[3] = {PY[0], PY[k], PY[k + 1]};
            if (tri_area(tx, ty) <= 0) continue;
            PixRect q = tri_pixels(tx, ty, tile);
            if (q.x0 > q.x1 || q.y0 > q.y1) continue;
            u.x0 = min(u.x0, q.x0); u.y0 = min(u.y0, q.y0);
            u.x1 = max(u.x1, q.x1); u.y1 = max(u.y1, q.y1);
          }
          pr = u;
          active = u.x0 <= u.x1 && u.y0 <= u.y1;
        }
      }
    }
  }

  // bin append, aggregated per wave: lanes targeting the leader's bin take
  // consecutive slots from one atomicAdd (slot = base + prefix popcount)
  const int sh = p.bin_shift;
  int bx0 = pr.x0 >> sh, bx1 = pr.x1 >> sh, by0 = pr.y0 >> sh, by1 = pr.y1 >> sh;
  int cbx = bx0, cby = by0;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t written = 0;
  while (true) {
    unsigned long long act = __ballot(active);
    if (act == 0ull) break;
    int leader = __ffsll((long long)act) - 1;
    int mybin = cby * sc.bins_x + cbx;
    int lbin = __shfl(mybin, leader);
    bool match = active && (mybin == lbin);
    unsigned long long mm = __ballot(match);
    uint32_t cnt = (uint32_t)__popcll(mm);
    uint32_t base = 0;
    const uint32_t gbin = sc.bin_base + (uint32_t)lbin;
    if (lane == leader) base = atomicAdd(&p.bin_count[gbin], cnt);
    base = __shfl(base, leader);
    if (match) {
      uint32_t slot = base + (uint32_t)__popcll(mm & lt_mask);
      if (slot < sc.bin_cap) p.bin_entries[sc.entry_base + (size_t)lbin * sc.bin_cap + slot] = prim;
      ++written;
      if (++cbx > bx1) {
        cbx = bx0;
        if (++cby > by1) active = false;
      }
    }
  }
  // count bin entries per wave (frame statistics)
  for (int o = 32; o > 0; o >>= 1) written += __shfl_xor(written, o);
  if (lane == 0 && written) atomicAdd(&p.stats[0], (unsigned long long)written);
}

// --------------------------------------------------------------------------
// k_raster: per (scene, bin) raster into LDS + resolve
// --------------------------------------------------------------------------

struct BigTri {
  int32_t X[3], Y[3];
  float z0, dzdx, dzdy;
  uint32_t prim;
  PixRect r;
};

struct TriSetup {
  long long dx[3], dy[3];
  long long bias[3];
  float z0, dzdx, dzdy;
};

__device__ __forceinline__ void setup_edges(const int32_t* X, const int32_t* Y, const float* Z,
                                            TriSetup& s) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3;
    s.dx[i] = (long long)X[j] - X[i];
    s.dy[i] = (long long)Y[j] - Y[i];
    s.bias[i] = (s.dy[i] < 0 || (s.dy[i] == 0 && s.dx[i] > 0)) ? 0 : -1;
  }
  float fx1 = (float)(X[1] - X[0]), fy1 = (float)(Y[1] - Y[0]);
  float fx2 = (float)(X[2] - X[0]), fy2 = (float)(Y[2] - Y[0]);
  float fa = fx1 * fy2;
  float fa2 = fx2 * fy1;
  fa = fa - fa2;
  float dz1 = Z[1] - Z[0], dz2 = Z[2] - Z[0];
  float pp, qq;
  pp = dz1 * fy2; qq = dz2 * fy1;
  float dzdx = pp - qq;
  s.dzdx = dzdx / fa;
  pp = dz2 * fx1; qq = dz1 * fx2;
  float dzdy = pp - qq;
  s.dzdy = dzdy / fa;
  s.z0 = Z[0];
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

__device__ __forceinline__ float depth_at(float z0, float dzdx, float dzdy, int32_t X0, int32_t Y0,
                                          long long Px, long long Py) {
  float ox = (float)(int32_t)(Px - X0);
  float oy = (float)(int32_t)(Py - Y0);
  float z = dzdx * ox;
  z = z0 + z;
  float zy = dzdy * oy;
  z = z + zy;
  if (!(z > 0.0f)) z = 0.0f;
  if (z > 1.0f) z = 1.0f;
  return z;
}

struct RasterCtx {
  unsigned long long* vis;
  BigTri* big;
  int* nbig;
  PixRect bin;
  int B;
};

__device__ void raster_tri(const RasterCtx& rc, const int32_t* X, const int32_t* Y, const float* Z,
                           uint32_t prim) {
  if (tri_area(X, Y) <= 0) return;
  PixRect r = tri_pixels(X, Y, rc.bin);
  if (r.x0 > r.x1 || r.y0 > r.y1) return;
  TriSetup s;
  setup_edges(X, Y, Z, s);
  const int npx = (r.x1 - r.x0 + 1) * (r.y1 - r.y0 + 1);
  if (npx > kSmallRaster) {
    int slot = atomicAdd(rc.nbig, 1);
    if (slot < kBigQueue) {
      BigTri& b = rc.big[slot];
#pragma unroll
      for (int k = 0; k < 3; ++k) { b.X[k] = X[k]; b.Y[k] = Y[k]; }
      b.z0 = s.z0; b.dzdx = s.dzdx; b.dzdy = s.dzdy;
      b.prim = prim;
      b.r = r;
      return;
    }
  }
  for (int py = r.y0; py <= r.y1; ++py) {
    long long Py = (long long)py * 256 + 128;
    for (int px = r.x0; px <= r.x1; ++px) {
      long long Px = (long long)px * 256 + 128;
      if (!covers(s, X, Y, Px, Py)) continue;
      float z = depth_at(s.z0, s.dzdx, s.dzdy, X[0], Y[0], Px, Py);
      if (z < 1.0f) {
        unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | prim;
        atomicMin(&rc.vis[(py - rc.bin.y0) * rc.B + (px - rc.bin.x0)], key);
      }
    }
  }
}

__device__ void raster_big(const RasterCtx& rc, const BigTri& b) {
  TriSetup s;
  // edges from the stored vertices; depth plane stored
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3;
    s.dx[i] = (long long)b.X[j] - b.X[i];
    s.dy[i] = (long long)b.Y[j] - b.Y[i];
    s.bias[i] = (s.dy[i] < 0 || (s.dy[i] == 0 && s.dx[i] > 0)) ? 0 : -1;
  }
  const int w = b.r.x1 - b.r.x0 + 1;
  const int n = w * (b.r.y1 - b.r.y0 + 1);
  for (int k = threadIdx.x; k < n; k += kBlock) {
    int px = b.r.x0 + k % w, py = b.r.y0 + k / w;
    long long Px = (long long)px * 256 + 128, Py = (long long)py * 256 + 128;
    if (!covers(s, b.X, b.Y, Px, Py)) continue;
    float z = depth_at(b.z0, b.dzdx, b.dzdy, b.X[0], b.Y[0], Px, Py);
    if (z < 1.0f) {
      unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | b.prim;
      atomicMin(&rc.vis[(py - rc.bin.y0) * rc.B + (px - rc.bin.x0)], key);
    }
  }
}

__device__ __forceinline__ uint32_t find_draw(const FrameParams& p, const SceneDesc& sc,
                                              const uint32_t* s_base, uint32_t prim) {
  // last draw with prim_base <= prim (upper bound - 1)
  uint32_t lo = 0, hi = sc.draw_count;
  const bool lds = sc.draw_count <= (uint32_t)kDrawCache;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t b = lds ? s_base[mid] : p.draws[sc.draw_begin + mid].prim_base;
    if (b <= prim) lo = mid + 1;
    else hi = mid;
  }
  return sc.draw_begin + lo - 1;
}

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
    uint32_t v = ip[k];
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
  int n = clip_poly(poly, 3, gx, gy);
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

__device__ __forceinline__ int wrap_coord(int i, int size, int mode) {
  if (mode == GL_CLAMP_TO_EDGE) return i < 0 ? 0 : (i >= size ? size - 1 : i);
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

__device__ __forceinline__ int floor_to_int(float u) {
  if (u > 1073741824.0f) u = 1073741824.0f;
  if (u < -1073741824.0f) u = -1073741824.0f;
  return (int)floorf(u);
}

__device__ __forceinline__ void fetch(const FrameParams& p, const TextureDesc& t, int level,
                                      int layer, int i, int j, float* out) {
  int w = max(t.width >> level, 1), h = max(t.height >> level, 1);
  const uint32_t texel = *reinterpret_cast<const uint32_t*>(
      p.texels + t.level_offset[level] + (((long long)layer * h + j) * w + i) * 4);
  out[0] = p.unorm_lut[texel & 0xff];
  out[1] = p.unorm_lut[(texel >> 8) & 0xff];
  out[2] = p.unorm_lut[(texel >> 16) & 0xff];
  out[3] = p.unorm_lut[texel >> 24];
}

__device__ void sample_level(const FrameParams& p, const TextureDesc& t, int level, int layer,
                             float s, float tc, bool linear, float* out) {
  int w = max(t.width >> level, 1), h = max(t.height >> level, 1);
  float u = s * (float)w;
  float v = tc * (float)h;
  if (!linear) {
    int i = wrap_coord(floor_to_int(u), w, t.wrap_s);
    int j = wrap_coord(floor_to_int(v), h, t.wrap_t);
    fetch(p, t, level, layer, i, j, out);
    return;
  }
  float uu = u - 0.5f;
  float vv = v - 0.5f;
  int i0 = floor_to_int(uu);
  int j0 = floor_to_int(vv);
  float a = uu - (float)i0;
  float b = vv - (float)j0;
  int i1 = wrap_coord(i0 + 1, w, t.wrap_s);
  int j1 = wrap_coord(j0 + 1, h, t.wrap_t);
  i0 = wrap_coord(i0, w, t.wrap_s);
  j0 = wrap_coord(j0, h, t.wrap_t);
  float t00[4], t10[4], t01[4], t11[4];
  fetch(p, t, level, layer, i0, j0, t00);
  fetch(p, t, level, layer, i1, j0, t10);
  fetch(p, t, level, layer, i0, j1, t01);
  fetch(p, t, level, layer, i1, j1, t11);
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

__device__ void sample_tex(const FrameParams& p, const TextureDesc& t, int layer, float s,
                           float tc, float lambda, float* out) {
  if (layer < 0) layer = 0;
  if (layer >= t.layers) layer = t.layers - 1;
  float c = 0.0f;
  if (t.mag_filter == GL_LINEAR &&
      (t.min_filter == GL_NEAREST_MIPMAP_NEAREST || t.min_filter == GL_NEAREST_MIPMAP_LINEAR))
    c = 0.5f;
  if (lambda <= c) {
    sample_level(p, t, 0, layer, s, tc, t.mag_filter == GL_LINEAR, out);
    return;
  }
  const int q = t.levels - 1;
  switch (t.min_filter) {
    case GL_NEAREST:
    case GL_LINEAR:
      sample_level(p, t, 0, layer, s, tc, t.min_filter == GL_LINEAR, out);
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
      sample_level(p, t, dlev, layer, s, tc, t.min_filter == GL_LINEAR_MIPMAP_NEAREST, out);
      return;
    }
    default: {
      bool lin = t.min_filter == GL_LINEAR_MIPMAP_LINEAR;
      if (lambda >= (float)q) {
        sample_level(p, t, q, layer, s, tc, lin, out);
        return;
      }
      float fl = floorf(lambda);
      int d1 = (int)fl;
      float f = lambda - fl;
      float t1[4], t2[4];
      sample_level(p, t, d1, layer, s, tc, lin, t1);
      sample_level(p, t, d1 + 1, layer, s, tc, lin, t2);
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
  float cx[3], cy[3], cw[3];
  float s[3], t[3];
  float e[3][3];
};

__device__ __forceinline__ void bary_rows(TriAttr& A) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    float pp, qq;
    pp = A.cy[j] * A.cw[k]; qq = A.cw[j] * A.cy[k]; A.e[i][0] = pp - qq;
    pp = A.cw[j] * A.cx[k]; qq = A.cx[j] * A.cw[k]; A.e[i][1] = pp - qq;
    pp = A.cx[j] * A.cy[k]; qq = A.cy[j] * A.cx[k]; A.e[i][2] = pp - qq;
  }
}

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

__device__ __forceinline__ uint32_t to_unorm8(float c) {
  if (!(c > 0.0f)) c = 0.0f;
  if (c > 1.0f) c = 1.0f;
  float v = c * 255.0f;
  v = v + 0.5f;
  return (uint32_t)(int)v;
}

__device__ uint32_t shade_pixel(const FrameParams& p, const SceneDesc& sc, const float* PV,
                                const DrawDesc& d, uint32_t tri, int px, int py) {
  float M[16];
  load_mat(p.abs_nodes + (size_t)d.node * 16, M);
  const MaterialDesc mat = p.materials[d.material];
  const uint32_t* ip = p.indices + d.index_base + 3 * tri;
  TriAttr A;
  uint32_t vi[3] = {ip[0], ip[1], ip[2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t v = vi[k];
    float4 c = vertex_clip(M, PV, p.pos_x[v], p.pos_y[v], p.pos_z[v]);
    A.cx[k] = c.x; A.cy[k] = c.y; A.cw[k] = c.w;
    float u, vv;
    if (d.flags & kDrawUvHalf) {
      uint32_t h2 = p.uv16[v];
      u = __half2float(__ushort_as_half((unsigned short)(h2 & 0xffff)));
      vv = __half2float(__ushort_as_half((unsigned short)(h2 >> 16)));
    } else {
      float2 t2 = p.uv32[v];
      u = t2.x; vv = t2.y;
    }
    float s = mat.texmat[0] * u; float s1 = mat.texmat[3] * vv; s = s + s1; s = s + mat.texmat[6];
    float tt = mat.texmat[1] * u; float t1 = mat.texmat[4] * vv; tt = tt + t1; tt = tt + mat.texmat[7];
    A.s[k] = s; A.t[k] = tt;
  }
  bary_rows(A);
  const float hw = (float)sc.width * 0.5f, hh = (float)sc.height * 0.5f;
  const float ihw = 1.0f / hw, ihh = 1.0f / hh;
  float b[3];
  bary_at(A, hw, hh, ihw, ihh, (float)px + 0.5f, (float)py + 0.5f, b);
  float c[4] = {mat.ambient[0], mat.ambient[1], mat.ambient[2], mat.ambient[3]};
  if (!p.no_textures && mat.texture > 0) {
    const TextureDesc& td = p.textures[mat.texture];
    float s = interp3(b, A.s[0], A.s[1], A.s[2]);
    float t = interp3(b, A.t[0], A.t[1], A.t[2]);
    int qx = px & ~1, qy = py & ~1;
    float b00[3], b10[3], b01[3];
    bary_at(A, hw, hh, ihw, ihh, (float)qx + 0.5f, (float)qy + 0.5f, b00);
    bary_at(A, hw, hh, ihw, ihh, (float)qx + 1.5f, (float)qy + 0.5f, b10);
    bary_at(A, hw, hh, ihw, ihh, (float)qx + 0.5f, (float)qy + 1.5f, b01);
    float s00 = interp3(b00, A.s[0], A.s[1], A.s[2]);
    float t00 = interp3(b00, A.t[0], A.t[1], A.t[2]);
    float s10 = interp3(b10, A.s[0], A.s[1], A.s[2]);
    float t10 = interp3(b10, A.t[0], A.t[1], A.t[2]);
    float s01 = interp3(b01, A.s[0], A.s[1], A.s[2]);
    float t01 = interp3(b01, A.t[0], A.t[1], A.t[2]);
    float fw = (float)td.width, fh = (float)td.height;
    float dux = s10 - s00; dux = dux * fw;
    float dvx = t10 - t00; dvx = dvx * fh;
    float duy = s01 - s00; duy = duy * fw;
    float dvy = t01 - t00; dvy = dvy * fh;
    float rx = dux * dux; float rx2 = dvx * dvx; rx = rx + rx2;
    float ry = duy * duy; float ry2 = dvy * dvy; ry = ry + ry2;
    float r2 = rx > ry ? rx : ry;
    float lambda = gbr_log2(r2);
    lambda = lambda * 0.5f;
    float tc[4];
    sample_tex(p, td, (int)mat.layer, s, t, lambda, tc);
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = c[k] * tc[k];
  }
  const uint32_t cf = (d.flags >> kDrawColorShift) & 3u;
  if (cf != kColorNone) {
    float col[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (cf == kColorU8) {
        uint32_t u8 = p.col8[vi[k]];
        col[k][0] = p.unorm_lut[u8 & 0xff];
        col[k][1] = p.unorm_lut[(u8 >> 8) & 0xff];
        col[k][2] = p.unorm_lut[(u8 >> 16) & 0xff];
        col[k][3] = p.unorm_lut[u8 >> 24];
      } else {
        float4 f = p.col32[vi[k]];
        col[k][0] = f.x; col[k][1] = f.y; col[k][2] = f.z; col[k][3] = f.w;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float v = interp3(b, col[0][k], col[1][k], col[2][k]);
      c[k] = c[k] * v;
    }
  }
  return to_unorm8(c[0]) | (to_unorm8(c[1]) << 8) | (to_unorm8(c[2]) << 16) |
         (to_unorm8(c[3]) << 24);
}

__global__ __launch_bounds__(kBlock) void k_raster(FrameParams p,
                                                   const RasterItem* __restrict__ items) {
  __shared__ unsigned long long s_vis[kMaxBin * kMaxBin];
  __shared__ BigTri s_big[kBigQueue];
  __shared__ uint32_t s_base[kDrawCache];
  __shared__ int s_nbig;

  const RasterItem it = items[blockIdx.x];
  const SceneDesc sc = p.scenes[it.scene];
  const int B = p.bin_size;
  const int bx = (int)(it.bin % (uint32_t)sc.bins_x), by = (int)(it.bin / (uint32_t)sc.bins_x);
  RasterCtx rc;
  rc.vis = s_vis;
  rc.big = s_big;
  rc.nbig = &s_nbig;
  rc.B = B;
  rc.bin.x0 = bx * B;
  rc.bin.y0 = by * B;
  rc.bin.x1 = min(bx * B + B, sc.width) - 1;
  rc.bin.y1 = min(by * B + B, sc.height) - 1;

  for (int i = threadIdx.x; i < B * B; i += kBlock) s_vis[i] = kClearKey;
  if (sc.draw_count <= (uint32_t)kDrawCache)
    for (uint32_t i = threadIdx.x; i < sc.draw_count; i += kBlock)
      s_base[i] = p.draws[sc.draw_begin + i].prim_base;
  if (threadIdx.x == 0) s_nbig = 0;
  float PV[16];
  load_mat(p.cameras + (size_t)it.scene * 16, PV);
  __syncthreads();

  const uint32_t gbin = sc.bin_base + it.bin;
  const uint32_t n = p.bin_count[gbin];
  const bool overflow = n > sc.bin_cap;
  const uint32_t total = overflow ? sc.prim_count : n;
  for (uint32_t base = 0; base < total; base += kBlock) {
    const uint32_t k = base + threadIdx.x;
    if (k < total) {
      const uint32_t prim = overflow ? k : p.bin_entries[sc.entry_base + (size_t)it.bin * sc.bin_cap + k];
      raster_prim(p, sc, PV, rc, s_base, prim);
    }
    __syncthreads();
    const int nb = min(s_nbig, kBigQueue);
    for (int q = 0; q < nb; ++q) raster_big(rc, s_big[q]);
    __syncthreads();
    if (threadIdx.x == 0) s_nbig = 0;
    __syncthreads();
  }

  // resolve: every pixel of the bin, written once
  const float un0 = p.unprojections[2 * it.scene], un1 = p.unprojections[2 * it.scene + 1];
  const float far_depth = un1 / (1.0f + un0);
  const int bw = rc.bin.x1 - rc.bin.x0 + 1, bh = rc.bin.y1 - rc.bin.y0 + 1;
  for (int l = threadIdx.x; l < bw * bh; l += kBlock) {
    const int lx = l % bw, ly = l / bw;
    const int px = rc.bin.x0 + lx, py = rc.bin.y0 + ly;
    const unsigned long long key = s_vis[ly * B + lx];
    const int row = p.top_down ? (sc.height - 1 - py) : py;
    const size_t o = (size_t)sc.plane_offset + (size_t)row * sc.width + px;
    if (key == kClearKey) {
      if (p.out_color) p.out_color[o] = kClearColor;
      if (p.out_depth) p.out_depth[o] = 1.0f;
      if (p.out_lindepth) p.out_lindepth[o] = 0.0f;
      if (p.out_semantic) p.out_semantic[o] = 0u;
      continue;
    }
    const uint32_t prim = (uint32_t)key;
    const float z = __uint_as_float((uint32_t)(key >> 32));
    const uint32_t di = find_draw(p, sc, s_base, prim);
    const DrawDesc d = p.draws[di];
    if (p.out_color) p.out_color[o] = shade_pixel(p, sc, PV, d, prim - d.prim_base, px, py);
    if (p.out_depth) p.out_depth[o] = z;
    if (p.out_lindepth) {
      float lin = un1 / (z + un0);
      p.out_lindepth[o] = lin == far_depth ? 0.0f : lin;
    }
    if (p.out_semantic) p.out_semantic[o] = d.semantic;
  }
  if (threadIdx.x == 0) {
    if (overflow) atomicAdd(&p.stats[1], 1ull);
    p.bin_count[gbin] = 0;  // ready for the next frame
  }
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
  const int row = g.top_down ? (g.tile_h - 1 - py) : py;
  uint32_t v = (uint32_t)g.fill;
  if (scene < g.scene_count)
    v = *reinterpret_cast<const uint32_t*>(
        g.plane + (((long long)scene * g.tile_h + row) * g.tile_w + px) * 4);
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
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodes, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, parents,
                     local, abs_out, n);
  return hipGetLastError();
}

hipError_t launch_setup(const FrameParams& p, const WorkItem* items, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_bin, dim3(n), dim3(kBlock), 0, s, p, items);
  return hipGetLastError();
}

hipError_t launch_raster(const FrameParams& p, const RasterItem* items, uint32_t n,
                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raster, dim3(n), dim3(kBlock), 0, s, p, items);
  return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs& g, hipStream_t s) {
  long long n = (long long)g.w * g.h;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, g);
  return hipGetLastError();
}

}  // namespace gbr
