/*
 * CPU oracle for the batch sensor renderer -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain, sequential C restatement of what the reference batch renderer asks
 * OpenGL to do per tile, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the product links
 * or calls this file.
 *
 * Reference behaviour restated here (file:line relative to the habitat-sim
 * source tree):
 *   - hierarchy: abs[i+1] = abs[parents[i]+1] * T[i], abs[0] = identity
 *       src/esp/gfx_batch/Renderer.cpp:1214-1220 (Magnum Matrix4 product,
 *       accumulated from +0 over k in order, no FMA contraction)
 *   - per-draw vertex stage of Magnum PhongGL (projection * (transformation *
 *       position)), invoked at src/esp/gfx_batch/Renderer.cpp:1345-1348 with
 *       the camera matrix stored at Renderer.cpp:1085
 *   - GL state: back-face culling (CCW front), depth test LESS against a
 *       cleared 1.0, src/esp/gfx_batch/Renderer.cpp:307-308 and
 *       RendererStandalone.cpp:155-159; colour clear 0x1f1f1f / alpha 255
 *       (pinned by GfxBatchRendererTestDefaults.png)
 *   - flat PhongGL fragment: ambient * texture(layer, texMatrix * uv) *
 *       vertexColor, materials from Renderer.cpp:579-620, per-draw texture
 *       matrix + layer from Renderer.cpp:941-950
 *   - depth unprojection: a' = (P[2][2]-1)/2, b' = P[3][2]/2,
 *       d_lin = b'/(d + a'), d_lin == b'/(1 + a') -> 0
 *       src/esp/gfx_batch/DepthUnprojection.cpp:98-130
 *
 * GL-spec rules restated (OpenGL 4.6 core, sections 13.7 and 14.6, 8.14):
 *   - homogeneous clipping against near (z >= -w) and far (z <= w) planes,
 *     plus a guard band of 8192 px around the viewport centre in x and y;
 *     intersections interpolate from the inside vertex (watertight)
 *   - window coordinates tile-local, snapped to 8 sub-pixel bits
 *     (round-to-nearest-even), pixel centres at +0.5
 *   - point-sampling coverage with a top-left tie rule in framebuffer-row
 *     order (rows increase with GL window y), i.e. a centre exactly on an
 *     edge belongs to the triangle iff the edge is a left edge or a
 *     horizontal edge with the interior at larger y
 *   - window depth interpolated linearly in screen space from the snapped
 *     vertices, raised to the triangle's smallest vertex depth (a NaN or a
 *     rounding overshoot in a sliver never lands in front of its triangle)
 *     and clamped to [0,1]
 *   - attributes perspective-correct, evaluated from the unclipped
 *     triangle's homogeneous barycentrics at the pixel centre
 *   - texture LOD from coarse 2x2-quad derivatives of the same triangle's
 *     interpolated texture coordinates, lambda = log2(rho); REPEAT /
 *     CLAMP_TO_EDGE / MIRRORED_REPEAT wrapping, NEAREST / LINEAR filters and
 *     the four mipmap modes
 *
 * Every float operation is written out in a fixed order; the file must be
 * compiled with -ffp-contract=off so that no FMA is formed.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_SUBPIX 256.0f
#define ORC_GUARD_PX 8192.0f

/* GL enums used by samplers */
#define GL_NEAREST 9728
#define GL_LINEAR 9729
#define GL_NEAREST_MIPMAP_NEAREST 9984
#define GL_LINEAR_MIPMAP_NEAREST 9985
#define GL_NEAREST_MIPMAP_LINEAR 9986
#define GL_LINEAR_MIPMAP_LINEAR 9987
#define GL_REPEAT 10497
#define GL_CLAMP_TO_EDGE 33071
#define GL_MIRRORED_REPEAT 33648

#define ORC_FLAG_NO_TEXTURES 1

typedef struct {
  /* geometry, shared by all envs */
  const float* pos;     /* [nv][3] */
  const float* uv;      /* [nv][2] */
  const float* col;     /* [nv][4] */
  const uint32_t* tris; /* [nt][3] absolute vertex indices */
  /* textures */
  int32_t ntex;
  const int32_t* tex_desc;      /* [ntex][8] w h layers levels min mag wrapS wrapT */
  const int64_t* tex_level_off; /* [ntex][16] byte offset of each level */
  const uint8_t* texels;        /* RGBA8, level = [layers][h][w][4] */
  /* draws, in submission order per env */
  const int32_t* draw_i; /* [nd][8] triStart triCount node tex layer semantic vcolor pad */
  const float* draw_f;   /* [nd][16] ambient[4] texMatrix[9] pad[3] */
  /* nodes */
  const int32_t* node_parent; /* [nn] parent within env, -1 = root */
  const float* node_local;    /* [nn][16] column-major */
  /* envs */
  int32_t nenv;
  const int32_t* env_i; /* [ne][8] w h nodeStart nodeCount drawStart drawCount flags pixelOffset */
  const float* env_f;   /* [ne][32] camera (P*V)[16], projection P[16] */
  /* outputs, env-major, rows bottom-up; any may be NULL */
  uint8_t* out_color;      /* [px][4] */
  float* out_depth;        /* window depth */
  float* out_lindepth;     /* unprojected depth, far -> 0 */
  uint32_t* out_semantic;  /* draw semantic id, 0 = background */
  uint32_t* out_count;     /* optional: covered-fragment count (no depth test) */
  int32_t* out_draw;       /* optional: winning draw index (global), -1 = none */
  int32_t* out_tri;        /* optional: winning triangle index within draw */
  /* lights (Renderer.cpp:990-1015, 1263-1289); may be NULL when no scene
     has lights */
  const float* nrm;          /* [nv][3] vertex normals */
  const float* draw_diffuse; /* [nd][4] diffuse colour; all zero: flat material */
  const int32_t* env_light;  /* [ne][2] first light, light count (capped at the maximum) */
  const int32_t* light_i;    /* [nl][2] node (env-local), type (0 directional, 1 point) */
  const float* light_f;      /* [nl][4] colour rgb, range */
} orc_job;

/* ------------------------------------------------------------------ */
/* Matrix helpers                                                      */
/* ------------------------------------------------------------------ */

/* Magnum Matrix4 product: out[c][r] = sum_k a[k][r]*b[c][k], accumulated
   from +0 over k = 0..3. */
void orc_mat4_mul(const float* a, const float* b, float* out) {
  float t[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      float s = 0.0f;
      for (int k = 0; k < 4; ++k) {
        float p = a[k * 4 + r] * b[c * 4 + k];
        s = s + p;
      }
      t[c * 4 + r] = s;
    }
  memcpy(out, t, sizeof t);
}

/* GLSL mat4 * vec4 evaluated as ((c0*x + c1*y) + c2*z) + c3*w. */
static void mat4_vec(const float* m, const float* v, float* out) {
  for (int r = 0; r < 4; ++r) {
    float s = m[0 * 4 + r] * v[0];
    float p1 = m[1 * 4 + r] * v[1];
    s = s + p1;
    float p2 = m[2 * 4 + r] * v[2];
    s = s + p2;
    float p3 = m[3 * 4 + r] * v[3];
    s = s + p3;
    out[r] = s;
  }
}

/* calculateDepthUnprojection(), DepthUnprojection.cpp:98-101 */
void orc_depth_unprojection(const float* P, float* out2) {
  float a = P[2 * 4 + 2] - 1.0f;
  float b = P[3 * 4 + 2];
  out2[0] = a * 0.5f;
  out2[1] = b * 0.5f;
}

/* unprojectDepth(), DepthUnprojection.cpp:107-130, on a contiguous array */
void orc_unproject_depth(const float* unproj, float* depth, int64_t n) {
  const float farDepth = unproj[1] / (1.0f + unproj[0]);
  for (int64_t i = 0; i < n; ++i) {
    float d = unproj[1] / (depth[i] + unproj[0]);
    depth[i] = d == farDepth ? 0.0f : d;
  }
}

/* ------------------------------------------------------------------ */
/* log2 for texture LOD: exponent + 2*atanh series of the mantissa     */
/* (ln m = 2(y + y^3/3 + y^5/5 + y^7/7 + y^9/9), y = (m-1)/(m+1))     */
/* ------------------------------------------------------------------ */
static float orc_log2(float x) {
  uint32_t bits;
  memcpy(&bits, &x, 4);
  int e = (int)((bits >> 23) & 0xffu);
  if (e == 0) return -126.0f; /* zero / denormal: deep magnification */
  if (e == 0xff) return 128.0f;
  uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
  float m;
  memcpy(&m, &mb, 4);
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
  /* 2/ln 2 */
  float l = s * 2.8853900817779268f;
  return (float)(e - 127) + l;
}

/* ------------------------------------------------------------------ */
/* Texture sampling (GL 4.6 section 8.14)                              */
/* ------------------------------------------------------------------ */

static float g_unorm[256];

static void init_unorm(void) {
  for (int i = 0; i < 256; ++i) g_unorm[i] = (float)i / 255.0f;
}

static int wrap_coord(int i, int size, int mode) {
  if (mode == GL_CLAMP_TO_EDGE) return i < 0 ? 0 : (i >= size ? size - 1 : i);
  if (mode == GL_MIRRORED_REPEAT) {
    int p = 2 * size;
    int m = i % p;
    if (m < 0) m += p;
    return m < size ? m : p - 1 - m;
  }
  int m = i % size;
  if (m < 0) m += size;
  return m;
}

static int floor_to_int(float u) {
  if (u > 1073741824.0f) u = 1073741824.0f;
  if (u < -1073741824.0f) u = -1073741824.0f;
  return (int)floorf(u);
}

typedef struct {
  int w, h, layers, levels, minf, magf, wraps, wrapt;
  const int64_t* off;
  const uint8_t* texels;
} tex_view;

static void fetch(const tex_view* t, int level, int layer, int i, int j,
                  float* out) {
  int w = t->w >> level;
  if (w < 1) w = 1;
  int h = t->h >> level;
  if (h < 1) h = 1;
  const uint8_t* p =
      t->texels + t->off[level] + (((int64_t)layer * h + j) * w + i) * 4;
  out[0] = g_unorm[p[0]];
  out[1] = g_unorm[p[1]];
  out[2] = g_unorm[p[2]];
  out[3] = g_unorm[p[3]];
}

static void sample_level(const tex_view* t, int level, int layer, float s,
                         float tc, int linear, float* out) {
  int w = t->w >> level;
  if (w < 1) w = 1;
  int h = t->h >> level;
  if (h < 1) h = 1;
  float u = s * (float)w;
  float v = tc * (float)h;
  if (!linear) {
    int i = wrap_coord(floor_to_int(u), w, t->wraps);
    int j = wrap_coord(floor_to_int(v), h, t->wrapt);
    fetch(t, level, layer, i, j, out);
    return;
  }
  float uu = u - 0.5f;
  float vv = v - 0.5f;
  int i0 = floor_to_int(uu);
  int j0 = floor_to_int(vv);
  float a = uu - (float)i0;
  float b = vv - (float)j0;
  int i1 = wrap_coord(i0 + 1, w, t->wraps);
  int j1 = wrap_coord(j0 + 1, h, t->wrapt);
  i0 = wrap_coord(i0, w, t->wraps);
  j0 = wrap_coord(j0, h, t->wrapt);
  float t00[4], t10[4], t01[4], t11[4];
  fetch(t, level, layer, i0, j0, t00);
  fetch(t, level, layer, i1, j0, t10);
  fetch(t, level, layer, i0, j1, t01);
  fetch(t, level, layer, i1, j1, t11);
  float ia = 1.0f - a;
  float ib = 1.0f - b;
  float w00 = ia * ib, w10 = a * ib, w01 = ia * b, w11 = a * b;
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

/* lambda given as the LOD; selects magnification / minification and mip
   levels per GL 4.6 section 8.14.2-8.14.3. */
static void sample_tex(const tex_view* t, int layer, float s, float tc,
                       float lambda, float* out) {
  if (layer < 0) layer = 0;
  if (layer >= t->layers) layer = t->layers - 1;
  float c = 0.0f;
  if (t->magf == GL_LINEAR &&
      (t->minf == GL_NEAREST_MIPMAP_NEAREST || t->minf == GL_NEAREST_MIPMAP_LINEAR))
    c = 0.5f;
  if (lambda <= c) {
    sample_level(t, 0, layer, s, tc, t->magf == GL_LINEAR, out);
    return;
  }
  int q = t->levels - 1;
  switch (t->minf) {
    case GL_NEAREST:
    case GL_LINEAR:
      sample_level(t, 0, layer, s, tc, t->minf == GL_LINEAR, out);
      return;
    case GL_NEAREST_MIPMAP_NEAREST:
    case GL_LINEAR_MIPMAP_NEAREST: {
      int d;
      if (lambda <= 0.5f)
        d = 0;
      else {
        float f = ceilf(lambda + 0.5f) - 1.0f;
        d = f > (float)q ? q : (int)f;
      }
      sample_level(t, d, layer, s, tc, t->minf == GL_LINEAR_MIPMAP_NEAREST,
                   out);
      return;
    }
    default: { /* *_MIPMAP_LINEAR */
      int lin = t->minf == GL_LINEAR_MIPMAP_LINEAR;
      if (lambda >= (float)q) {
        sample_level(t, q, layer, s, tc, lin, out);
        return;
      }
      float fl = floorf(lambda);
      int d1 = (int)fl;
      float f = lambda - fl;
      float t1[4], t2[4];
      sample_level(t, d1, layer, s, tc, lin, t1);
      sample_level(t, d1 + 1, layer, s, tc, lin, t2);
      float g = 1.0f - f;
      for (int k = 0; k < 4; ++k) {
        float x1 = g * t1[k];
        float x2 = f * t2[k];
        out[k] = x1 + x2;
      }
      return;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Clipping                                                            */
/* ------------------------------------------------------------------ */

typedef struct {
  float x, y, z, w;
} cvert;

static float plane_dist(const cvert* v, int plane, float gx, float gy) {
  switch (plane) {
    case 0: return v->z + v->w;            /* near */
    case 1: return v->w - v->z;            /* far */
    case 2: { float g = gx * v->w; return g - v->x; }
    case 3: { float g = gx * v->w; return v->x + g; }
    case 4: { float g = gy * v->w; return g - v->y; }
    default: { float g = gy * v->w; return v->y + g; }
  }
}

static cvert lerp_from(const cvert* a, const cvert* b, float da, float db) {
  /* a inside (da >= 0), b outside */
  float den = da - db;
  float t = da / den;
  cvert r;
  float d;
  d = b->x - a->x; d = t * d; r.x = a->x + d;
  d = b->y - a->y; d = t * d; r.y = a->y + d;
  d = b->z - a->z; d = t * d; r.z = a->z + d;
  d = b->w - a->w; d = t * d; r.w = a->w + d;
  return r;
}

static int clip_poly(cvert* poly, int n, float gx, float gy) {
  cvert tmp[16];
  for (int plane = 0; plane < 6; ++plane) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
      const cvert* a = &poly[i];
      const cvert* b = &poly[(i + 1) % n];
      float da = plane_dist(a, plane, gx, gy);
      float db = plane_dist(b, plane, gx, gy);
      int ina = da >= 0.0f, inb = db >= 0.0f;
      if (ina) tmp[m++] = *a;
      if (ina != inb) tmp[m++] = ina ? lerp_from(a, b, da, db) : lerp_from(b, a, db, da);
    }
    n = m;
    if (n < 3) return 0;
    memcpy(poly, tmp, sizeof(cvert) * (size_t)n);
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* Per-env rendering                                                   */
/* ------------------------------------------------------------------ */

/* largest light count per env (the renderer's maximum) */
#define ORC_MAX_LIGHTS 64

typedef struct {
  /* unclipped triangle in clip space + per-vertex attributes */
  float cx[3], cy[3], cw[3];
  float s[3], t[3];
  float col[3][4];
  int has_col;
  /* shaded materials: world-space positions and normals of the vertices */
  float wp[3][3], wn[3][3];
  /* homogeneous barycentric rows */
  float e[3][3];
} tri_attr;

typedef struct {
  const orc_job* J;
  int W, H;
  float hw, hh, inv_hw, inv_hh;
  float* zbuf;
  uint8_t* color;
  uint32_t* sem;
  uint32_t* count;
  int32_t* odraw;
  int32_t* otri;
  int no_textures;
  /* the env's lights in world space: position (point) or direction towards
     the light (directional), colour, range, type */
  int nlights;
  float lpos[ORC_MAX_LIGHTS][3], lcol[ORC_MAX_LIGHTS][3], lrange[ORC_MAX_LIGHTS];
  int ltype[ORC_MAX_LIGHTS];
} env_ctx;

static void bary_rows(tri_attr* A) {
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    float p, q;
    p = A->cy[j] * A->cw[k]; q = A->cw[j] * A->cy[k]; A->e[i][0] = p - q;
    p = A->cw[j] * A->cx[k]; q = A->cx[j] * A->cw[k]; A->e[i][1] = p - q;
    p = A->cx[j] * A->cy[k]; q = A->cy[j] * A->cx[k]; A->e[i][2] = p - q;
  }
}

static void bary_at(const env_ctx* C, const tri_attr* A, float fx, float fy,
                    float* b) {
  float nx = fx - C->hw;
  nx = nx * C->inv_hw;
  float ny = fy - C->hh;
  ny = ny * C->inv_hh;
  float l[3];
  for (int i = 0; i < 3; ++i) {
    float v = A->e[i][0] * nx;
    float v2 = A->e[i][1] * ny;
    v = v + v2;
    v = v + A->e[i][2];
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

static float interp3(const float* b, float a0, float a1, float a2) {
  float v = b[0] * a0;
  float v1 = b[1] * a1;
  v = v + v1;
  float v2 = b[2] * a2;
  return v + v2;
}

static uint8_t to_unorm8(float c) {
  if (!(c > 0.0f)) c = 0.0f;
  if (c > 1.0f) c = 1.0f;
  float v = c * 255.0f;
  v = v + 0.5f;
  return (uint8_t)(int)v;
}

static void shade(const env_ctx* C, const tri_attr* A, int draw, int px,
                  int py, uint8_t* out) {
  const orc_job* J = C->J;
  const int32_t* di = J->draw_i + (int64_t)draw * 8;
  const float* df = J->draw_f + (int64_t)draw * 16;
  float b[3];
  float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  bary_at(C, A, fx, fy, b);
  float c[4] = {df[0], df[1], df[2], df[3]};
  float texc[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  int tex = di[3];
  if (!C->no_textures && tex >= 0) {
    const int32_t* td = J->tex_desc + (int64_t)tex * 8;
    tex_view tv = {td[0], td[1], td[2], td[3], td[4], td[5], td[6], td[7],
                   J->tex_level_off + (int64_t)tex * 16, J->texels};
    float s = interp3(b, A->s[0], A->s[1], A->s[2]);
    float t = interp3(b, A->t[0], A->t[1], A->t[2]);
    /* coarse quad derivatives from the same triangle */
    int qx = px & ~1, qy = py & ~1;
    float b00[3], b10[3], b01[3];
    bary_at(C, A, (float)qx + 0.5f, (float)qy + 0.5f, b00);
    bary_at(C, A, (float)qx + 1.5f, (float)qy + 0.5f, b10);
    bary_at(C, A, (float)qx + 0.5f, (float)qy + 1.5f, b01);
    float s00 = interp3(b00, A->s[0], A->s[1], A->s[2]);
    float t00 = interp3(b00, A->t[0], A->t[1], A->t[2]);
    float s10 = interp3(b10, A->s[0], A->s[1], A->s[2]);
    float t10 = interp3(b10, A->t[0], A->t[1], A->t[2]);
    float s01 = interp3(b01, A->s[0], A->s[1], A->s[2]);
    float t01 = interp3(b01, A->t[0], A->t[1], A->t[2]);
    float fw = (float)tv.w, fh = (float)tv.h;
    float dux = s10 - s00; dux = dux * fw;
    float dvx = t10 - t00; dvx = dvx * fh;
    float duy = s01 - s00; duy = duy * fw;
    float dvy = t01 - t00; dvy = dvy * fh;
    float rx = dux * dux; float rx2 = dvx * dvx; rx = rx + rx2;
    float ry = duy * duy; float ry2 = dvy * dvy; ry = ry + ry2;
    float r2 = rx > ry ? rx : ry;
    float lambda = orc_log2(r2);
    lambda = lambda * 0.5f;
    float tc[4];
    sample_tex(&tv, di[4], s, t, lambda, tc);
    for (int k = 0; k < 4; ++k) {
      texc[k] = tc[k];
      c[k] = c[k] * tc[k];
    }
  }
  float vc[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  if (A->has_col) {
    for (int k = 0; k < 4; ++k) {
      float v = interp3(b, A->col[0][k], A->col[1][k], A->col[2][k]);
      vc[k] = v;
      c[k] = c[k] * v;
    }
  }
  /* shaded material in a renderer with lights enabled: Phong diffuse term
     per light (no specular), in world space, attenuation
     clamp(1 - (d/R)^4, 0, 1)^2 / (1 + d^2) for point lights; the diffuse
     alpha is added to the ambient alpha also when the scene has no lights
     (the shader is compiled for the maximum light count) */
  const float* dd = J->draw_diffuse ? J->draw_diffuse + (int64_t)draw * 4 : NULL;
  if (dd && (dd[0] != 0.0f || dd[1] != 0.0f || dd[2] != 0.0f || dd[3] != 0.0f)) {
    float d[4] = {dd[0], dd[1], dd[2], dd[3]};
    for (int k = 0; k < 4; ++k) d[k] = d[k] * texc[k];
    for (int k = 0; k < 4; ++k) d[k] = d[k] * vc[k];
    float n[3], pw[3];
    for (int k = 0; k < 3; ++k) {
      n[k] = interp3(b, A->wn[0][k], A->wn[1][k], A->wn[2][k]);
      pw[k] = interp3(b, A->wp[0][k], A->wp[1][k], A->wp[2][k]);
    }
    float nn = n[0] * n[0]; float n1 = n[1] * n[1]; nn = nn + n1; float n2 = n[2] * n[2]; nn = nn + n2;
    float nl = sqrtf(nn);
    for (int k = 0; k < 3; ++k) n[k] = n[k] / nl;
    for (int l = 0; l < C->nlights; ++l) {
      float L[3], att = 1.0f;
      if (C->ltype[l] == 1) {
        for (int k = 0; k < 3; ++k) L[k] = C->lpos[l][k] - pw[k];
      } else {
        for (int k = 0; k < 3; ++k) L[k] = C->lpos[l][k];
      }
      float ll = L[0] * L[0]; float l1 = L[1] * L[1]; ll = ll + l1; float l2 = L[2] * L[2]; ll = ll + l2;
      float dist = sqrtf(ll);
      for (int k = 0; k < 3; ++k) L[k] = L[k] / dist;
      if (C->ltype[l] == 1) {
        float q = dist / C->lrange[l];
        q = q * q;
        q = q * q;
        float wdw = 1.0f - q;
        if (!(wdw > 0.0f)) wdw = 0.0f;
        if (wdw > 1.0f) wdw = 1.0f;
        att = wdw * wdw;
        float den = dist * dist;
        den = 1.0f + den;
        att = att / den;
      }
      float ndl = n[0] * L[0]; float t1 = n[1] * L[1]; ndl = ndl + t1; float t2 = n[2] * L[2]; ndl = ndl + t2;
      if (!(ndl > 0.0f)) ndl = 0.0f;
      float inten = ndl * att;
      for (int k = 0; k < 3; ++k) {
        float v = d[k] * C->lcol[l][k];
        v = v * inten;
        c[k] = c[k] + v;
      }
    }
    c[3] = c[3] + d[3];
  }
  for (int k = 0; k < 4; ++k) out[k] = to_unorm8(c[k]);
}

static int is_top_left(int64_t dx, int64_t dy) {
  return dy < 0 || (dy == 0 && dx > 0);
}

static void raster_tri(env_ctx* C, const tri_attr* A, int draw_global,
                       int tri_local, const int32_t* X, const int32_t* Y,
                       const float* Z) {
  int64_t area = (int64_t)(X[1] - X[0]) * (int64_t)(Y[2] - Y[0]) -
                 (int64_t)(Y[1] - Y[0]) * (int64_t)(X[2] - X[0]);
  if (area <= 0) return; /* back-facing or degenerate: culled */
  int32_t minx = X[0], maxx = X[0], miny = Y[0], maxy = Y[0];
  for (int i = 1; i < 3; ++i) {
    if (X[i] < minx) minx = X[i];
    if (X[i] > maxx) maxx = X[i];
    if (Y[i] < miny) miny = Y[i];
    if (Y[i] > maxy) maxy = Y[i];
  }
  /* pixel centres px*256+128 inside [min, max] */
  int x0 = (minx - 128 + 255) >> 8, x1 = (maxx - 128) >> 8;
  int y0 = (miny - 128 + 255) >> 8, y1 = (maxy - 128) >> 8;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 > C->W - 1) x1 = C->W - 1;
  if (y1 > C->H - 1) y1 = C->H - 1;
  if (x0 > x1 || y0 > y1) return;

  int64_t dx[3], dy[3], bias[3];
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3;
    dx[i] = (int64_t)X[j] - X[i];
    dy[i] = (int64_t)Y[j] - Y[i];
    bias[i] = is_top_left(dx[i], dy[i]) ? 0 : -1;
  }
  /* depth plane relative to vertex 0 */
  float fx1 = (float)(X[1] - X[0]), fy1 = (float)(Y[1] - Y[0]);
  float fx2 = (float)(X[2] - X[0]), fy2 = (float)(Y[2] - Y[0]);
  float fa = fx1 * fy2;
  float fa2 = fx2 * fy1;
  fa = fa - fa2;
  float dz1 = Z[1] - Z[0], dz2 = Z[2] - Z[0];
  float p, q;
  p = dz1 * fy2; q = dz2 * fy1;
  float dzdx = p - q;
  dzdx = dzdx / fa;
  p = dz2 * fx1; q = dz1 * fx2;
  float dzdy = p - q;
  dzdy = dzdy / fa;
  float zlo = Z[0] < Z[1] ? Z[0] : Z[1];
  zlo = Z[2] < zlo ? Z[2] : zlo;

  for (int py = y0; py <= y1; ++py) {
    int64_t Py = (int64_t)py * 256 + 128;
    for (int px = x0; px <= x1; ++px) {
      int64_t Px = (int64_t)px * 256 + 128;
      int in = 1;
      for (int i = 0; i < 3; ++i) {
        int64_t E = dx[i] * (Py - Y[i]) - dy[i] * (Px - X[i]);
        if (E + bias[i] < 0) { in = 0; break; }
      }
      if (!in) continue;
      int64_t pix = (int64_t)py * C->W + px;
      if (C->count) C->count[pix] += 1;
      float ox = (float)(int32_t)(Px - X[0]);
      float oy = (float)(int32_t)(Py - Y[0]);
      float z = dzdx * ox;
      z = Z[0] + z;
      float zy = dzdy * oy;
      z = z + zy;
      if (!(z >= zlo)) z = zlo;
      if (!(z > 0.0f)) z = 0.0f;
      if (z > 1.0f) z = 1.0f;
      if (z < C->zbuf[pix]) {
        C->zbuf[pix] = z;
        if (C->color) shade(C, A, draw_global, px, py, C->color + pix * 4);
        C->sem[pix] = (uint32_t)C->J->draw_i[(int64_t)draw_global * 8 + 5];
        if (C->odraw) C->odraw[pix] = draw_global;
        if (C->otri) C->otri[pix] = tri_local;
      }
    }
  }
}

static void process_triangle(env_ctx* C, const tri_attr* A, const float* cz,
                             int draw_global, int tri_local) {
  /* trivial reject against the view volume */
  unsigned oc_and = 0x3f;
  for (int i = 0; i < 3; ++i) {
    float x = A->cx[i], y = A->cy[i], z = cz[i], w = A->cw[i];
    float nw = -w;
    unsigned oc = 0;
    if (x > w) oc |= 1;
    if (x < nw) oc |= 2;
    if (y > w) oc |= 4;
    if (y < nw) oc |= 8;
    if (z < nw) oc |= 16;
    if (z > w) oc |= 32;
    oc_and &= oc;
  }
  if (oc_and) return;
  float gx = ORC_GUARD_PX / C->hw, gy = ORC_GUARD_PX / C->hh;
  cvert poly[16];
  for (int i = 0; i < 3; ++i) {
    poly[i].x = A->cx[i]; poly[i].y = A->cy[i]; poly[i].z = cz[i]; poly[i].w = A->cw[i];
  }
  int n = clip_poly(poly, 3, gx, gy);
  if (n < 3) return;
  int32_t X[16], Y[16];
  float Z[16];
  for (int i = 0; i < n; ++i) {
    if (!(poly[i].w > 0.0f)) return;
    float iw = 1.0f / poly[i].w;
    float nx = poly[i].x * iw, ny = poly[i].y * iw, nz = poly[i].z * iw;
    float xw = nx * C->hw; xw = xw + C->hw;
    float yw = ny * C->hh; yw = yw + C->hh;
    float zw = nz * 0.5f; zw = zw + 0.5f;
    X[i] = (int32_t)rintf(xw * ORC_SUBPIX);
    Y[i] = (int32_t)rintf(yw * ORC_SUBPIX);
    Z[i] = zw;
  }
  for (int i = 1; i + 1 < n; ++i) {
    int32_t tx[3] = {X[0], X[i], X[i + 1]};
    int32_t ty[3] = {Y[0], Y[i], Y[i + 1]};
    float tz[3] = {Z[0], Z[i], Z[i + 1]};
    raster_tri(C, A, draw_global, tri_local, tx, ty, tz);
  }
}

static void identity16(float* m) {
  memset(m, 0, 16 * sizeof(float));
  m[0] = m[5] = m[10] = m[15] = 1.0f;
}

static void render_env(const orc_job* J, int e) {
  const int32_t* ei = J->env_i + (int64_t)e * 8;
  const float* ef = J->env_f + (int64_t)e * 32;
  env_ctx C;
  memset(&C, 0, sizeof C);
  C.J = J;
  C.W = ei[0];
  C.H = ei[1];
  C.no_textures = (ei[6] & ORC_FLAG_NO_TEXTURES) != 0;
  C.hw = (float)C.W * 0.5f;
  C.hh = (float)C.H * 0.5f;
  C.inv_hw = 1.0f / C.hw;
  C.inv_hh = 1.0f / C.hh;
  int64_t npx = (int64_t)C.W * C.H;
  int64_t off = (int64_t)ei[7];
  C.zbuf = (float*)malloc(sizeof(float) * (size_t)npx);
  uint32_t* sem = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)npx);
  for (int64_t i = 0; i < npx; ++i) { C.zbuf[i] = 1.0f; sem[i] = 0; }
  C.sem = sem;
  if (J->out_color) {
    C.color = J->out_color + off * 4;
    for (int64_t i = 0; i < npx; ++i) {
      C.color[i * 4 + 0] = 0x1f; C.color[i * 4 + 1] = 0x1f;
      C.color[i * 4 + 2] = 0x1f; C.color[i * 4 + 3] = 0xff;
    }
  }
  if (J->out_count) {
    C.count = J->out_count + off;
    memset(C.count, 0, sizeof(uint32_t) * (size_t)npx);
  }
  if (J->out_draw) {
    C.odraw = J->out_draw + off;
    for (int64_t i = 0; i < npx; ++i) C.odraw[i] = -1;
  }
  if (J->out_tri) {
    C.otri = J->out_tri + off;
    for (int64_t i = 0; i < npx; ++i) C.otri[i] = -1;
  }

  /* absolute transforms, Renderer.cpp:1214-1220 */
  int nn = ei[3];
  float* abs_t = (float*)malloc(sizeof(float) * 16 * (size_t)(nn + 1));
  identity16(abs_t);
  for (int i = 0; i < nn; ++i) {
    int par = J->node_parent[ei[2] + i];
    orc_mat4_mul(abs_t + 16 * (par + 1), J->node_local + 16 * (int64_t)(ei[2] + i),
                 abs_t + 16 * (i + 1));
  }

  /* lights from their nodes' absolute transforms: point = translation,
     directional = -backward (the transformation's third column negated) */
  if (J->env_light) {
    const int32_t* el = J->env_light + (int64_t)e * 2;
    C.nlights = el[1] < ORC_MAX_LIGHTS ? el[1] : ORC_MAX_LIGHTS;
    for (int l = 0; l < C.nlights; ++l) {
      const int32_t* li = J->light_i + (int64_t)(el[0] + l) * 2;
      const float* lf = J->light_f + (int64_t)(el[0] + l) * 4;
      const float* T = abs_t + 16 * (li[0] + 1);
      C.ltype[l] = li[1];
      for (int k = 0; k < 3; ++k) {
        C.lpos[l][k] = li[1] == 1 ? T[12 + k] : -T[8 + k];
        C.lcol[l][k] = lf[k];
      }
      C.lrange[l] = lf[3];
    }
  }

  const float* PV = ef;
  for (int d = 0; d < ei[5]; ++d) {
    int dg = ei[4] + d;
    const int32_t* di = J->draw_i + (int64_t)dg * 8;
    const float* df = J->draw_f + (int64_t)dg * 16;
    const float* M = abs_t + 16 * (di[2] + 1);
    /* normal matrix: inverse transpose of M's upper 3x3 (cofactors over the
       determinant), [col][row] */
    float NM[9];
    {
      const float a00 = M[0], a01 = M[4], a02 = M[8];   /* row 0 */
      const float a10 = M[1], a11 = M[5], a12 = M[9];   /* row 1 */
      const float a20 = M[2], a21 = M[6], a22 = M[10];  /* row 2 */
      float c00 = a11 * a22; float t0 = a12 * a21; c00 = c00 - t0;
      float c01 = a12 * a20; float t1 = a10 * a22; c01 = c01 - t1;
      float c02 = a10 * a21; float t2 = a11 * a20; c02 = c02 - t2;
      float c10 = a02 * a21; float t3 = a01 * a22; c10 = c10 - t3;
      float c11 = a00 * a22; float t4 = a02 * a20; c11 = c11 - t4;
      float c12 = a01 * a20; float t5 = a00 * a21; c12 = c12 - t5;
      float c20 = a01 * a12; float t6 = a02 * a11; c20 = c20 - t6;
      float c21 = a02 * a10; float t7 = a00 * a12; c21 = c21 - t7;
      float c22 = a00 * a11; float t8 = a01 * a10; c22 = c22 - t8;
      float det = a00 * c00; float t9 = a01 * c01; det = det + t9; float t10 = a02 * c02; det = det + t10;
      /* inverse transpose = cofactor matrix / det; NM[col*3 + row] */
      NM[0] = c00 / det; NM[3] = c01 / det; NM[6] = c02 / det;
      NM[1] = c10 / det; NM[4] = c11 / det; NM[7] = c12 / det;
      NM[2] = c20 / det; NM[5] = c21 / det; NM[8] = c22 / det;
    }
    for (int t = 0; t < di[1]; ++t) {
      const uint32_t* tri = J->tris + 3 * ((int64_t)di[0] + t);
      tri_attr A;
      float cz[3];
      A.has_col = di[6] != 0;
      for (int k = 0; k < 3; ++k) {
        uint32_t v = tri[k];
        float p4[4] = {J->pos[3 * v], J->pos[3 * v + 1], J->pos[3 * v + 2], 1.0f};
        float q[4], c[4];
        mat4_vec(M, p4, q);
        mat4_vec(PV, q, c);
        A.cx[k] = c[0]; A.cy[k] = c[1]; cz[k] = c[2]; A.cw[k] = c[3];
        A.wp[k][0] = q[0]; A.wp[k][1] = q[1]; A.wp[k][2] = q[2];
        if (J->nrm) {
          const float* nv = J->nrm + 3 * (int64_t)v;
          for (int r = 0; r < 3; ++r) {
            float x = NM[r] * nv[0]; float y = NM[3 + r] * nv[1]; x = x + y;
            float z = NM[6 + r] * nv[2]; A.wn[k][r] = x + z;
          }
        } else {
          A.wn[k][0] = 0.0f; A.wn[k][1] = 0.0f; A.wn[k][2] = 1.0f;
        }
        float u = J->uv[2 * v], vv = J->uv[2 * v + 1];
        /* texture matrix applied in the vertex stage: ((c0*u + c1*v) + c2) */
        float s = df[4] * u; float s1 = df[7] * vv; s = s + s1; s = s + df[10];
        float tt = df[5] * u; float t1 = df[8] * vv; tt = tt + t1; tt = tt + df[11];
        A.s[k] = s; A.t[k] = tt;
        for (int ch = 0; ch < 4; ++ch) A.col[k][ch] = J->col[4 * v + ch];
      }
      bary_rows(&A);
      process_triangle(&C, &A, cz, dg, t);
    }
  }

  /* resolves */
  float un[2];
  orc_depth_unprojection(ef + 16, un);
  if (J->out_depth) memcpy(J->out_depth + off, C.zbuf, sizeof(float) * (size_t)npx);
  if (J->out_lindepth) {
    memcpy(J->out_lindepth + off, C.zbuf, sizeof(float) * (size_t)npx);
    orc_unproject_depth(un, J->out_lindepth + off, npx);
  }
  if (J->out_semantic) memcpy(J->out_semantic + off, sem, sizeof(uint32_t) * (size_t)npx);
  free(abs_t);
  free(sem);
  free(C.zbuf);
}

int orc_render(const orc_job* J, int nthreads) {
  init_unorm();
  if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int e = 0; e < J->nenv; ++e) render_env(J, e);
  return 0;
}

/* exported for tests of the LOD helper */
float orc_log2_export(float x) { return orc_log2(x); }

int orc_abi_version(void) { return 1; }
