// Device-side data layout of the batch renderer (shared by the host code and
// the HIP kernels).  All structs are POD, 16-byte aligned where they are read
// as vectors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbr {

constexpr int kBlock = 256;        // threads per workgroup, all kernels
constexpr int kMaxBin = 32;        // largest bin edge (LDS visibility tile)
constexpr int kChunkTris = 1024;   // triangles per setup work item (4 per lane)
constexpr int kVertexCache = 640;  // vertices per setup work item held in LDS
constexpr int kMaxSceneBins = 1024;  // bins per scene (LDS counters in setup, <= 2^11)
constexpr int kPairCap = 2048;     // (bin, triangle) pairs aggregated in LDS per item (<= 2^11)
constexpr int kBigQueue = 128;     // cooperative-raster queue entries per bin
constexpr int kDrawCache = 1024;   // draw prim bases cached in LDS per bin
constexpr float kSubpix = 256.0f;  // 8 sub-pixel bits
constexpr float kGuardPx = 8192.0f;

// per-vertex attribute formats
enum : uint32_t {
  kColorNone = 0,
  kColorU8 = 1,   // RGBA8 in col8[]
  kColorF32 = 2,  // float4 in col32[]
  kUvF32 = 0,     // float2 in uv32[]
  kUvF16 = 1      // half2 in uv16[]
};

// draw flags
enum : uint32_t {
  kDrawColorShift = 0,  // 2 bits: color format
  kDrawUvHalf = 1u << 2,
};

struct alignas(16) DrawDesc {
  uint32_t index_base;  // first index of the mesh view in indices[]
  uint32_t tri_count;
  uint32_t node;        // global node index (absolute transform)
  uint32_t material;
  uint32_t semantic;
  uint32_t flags;       // color format | uv format
  uint32_t prim_base;   // scene-local primitive order of the first triangle
  uint32_t scene;
  // the material's shading fields, copied so that the resolve pass reaches
  // them with one dependent load less
  float ambient[4];
  int32_t texture;      // 0 = untextured (white)
  uint32_t layer;
  uint32_t lit;         // MaterialDesc::lit
  uint32_t pad;
};

struct alignas(16) SceneDesc {
  int32_t width, height;
  int32_t bins_x, bins_y;
  uint32_t bin_base;    // first bin of this scene in the bin arrays
  uint32_t bin_cap;     // entries per bin
  uint32_t draw_begin;  // first draw in draws[]
  uint32_t draw_count;
  uint32_t plane_offset;  // pixel offset of this scene's planes
  uint32_t prim_count;
  unsigned long long entry_base;  // first bin entry of this scene (bins x bin_cap)
  unsigned long long rec_base;    // first primitive record of this scene
  int32_t stride;                 // row stride of the scene's planes (the tile width)
  uint32_t planes;                // GBR_PLANE_* written for this scene
};

// One chunk of <= kChunkTris consecutive triangles of a draw; everything a
// setup workgroup needs comes from this one 32-byte load.
struct alignas(16) WorkItem {
  uint32_t scene;
  uint32_t draw;         // global draw index
  uint32_t first_index;  // index of the chunk's first triangle in indices[]
  uint32_t tri_count;
  uint32_t prim;         // scene-local primitive order of the first triangle
  uint32_t vmin;         // global vertex index range referenced by the chunk
  uint32_t vcount;
  uint32_t node;         // global node index of the draw (absolute transform)
};

// Per-primitive records written by setup for triangles that reach a bin,
// indexed by (scene record base + primitive order), in two arrays so that
// the rasteriser and the resolve pass each read one aligned record.  Bin
// entries are the primitive order, or kSubTri | index of a clipped fan
// triangle in the sub-triangle buffers.
constexpr uint32_t kSubTri = 0x80000000u;

// Raster part (32 bytes): the snapped window coordinates X0 Y0 X1 Y1 X2 Y2
// (sub-pixel units) as 24-bit two's complement fields at bit offsets 0, 24,
// ..., 120 of xy[], and the vertices' window depths.  The depth plane is
// recomputed from these where it is used.
struct alignas(32) RasterRec {
  uint32_t xy[5];
  float z[3];
};

// Resolve part (64 bytes): the homogeneous barycentric rows of the
// unclipped triangle (e[i] = cross products of its clip-space (x, y, w),
// computed once per triangle by setup), its texture coordinates, and its
// work item (whose draw gives the material).
struct alignas(64) ResolveRec {
  float e[9];  // e[3*i + k]: row i, coefficient k (of nx, ny, 1)
  float s0, t0, s1, t1, s2, t2;
  uint32_t item;
};

// Rows of the homogeneous barycentric form from the clip-space x, y, w of
// the three vertices (op order fixed: products, then the difference).
__host__ __device__ inline void bary_rows_of(const float* cx, const float* cy, const float* cw,
                                             float* e) {
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    float pp, qq;
    pp = cy[j] * cw[k]; qq = cw[j] * cy[k]; e[3 * i + 0] = pp - qq;
    pp = cw[j] * cx[k]; qq = cx[j] * cw[k]; e[3 * i + 1] = pp - qq;
    pp = cx[j] * cy[k]; qq = cy[j] * cx[k]; e[3 * i + 2] = pp - qq;
  }
}

// 24-bit fields of RasterRec::xy; window coordinates stay within the guard
// band, |X|, |Y| < 2^23
__host__ __device__ inline void pack_xy(const int32_t* X, const int32_t* Y, uint32_t* xy) {
  const uint32_t v[6] = {(uint32_t)X[0] & 0xFFFFFFu, (uint32_t)Y[0] & 0xFFFFFFu,
                         (uint32_t)X[1] & 0xFFFFFFu, (uint32_t)Y[1] & 0xFFFFFFu,
                         (uint32_t)X[2] & 0xFFFFFFu, (uint32_t)Y[2] & 0xFFFFFFu};
  xy[0] = v[0] | (v[1] << 24);
  xy[1] = (v[1] >> 8) | (v[2] << 16);
  xy[2] = (v[2] >> 16) | (v[3] << 8);
  xy[3] = v[4] | (v[5] << 24);
  xy[4] = v[5] >> 8;
}

__host__ __device__ inline int32_t sext24(uint32_t v) { return (int32_t)(v << 8) >> 8; }

__host__ __device__ inline void unpack_xy(const uint32_t* xy, int32_t* X, int32_t* Y) {
  X[0] = sext24(xy[0]);
  Y[0] = sext24((xy[0] >> 24) | (xy[1] << 8));
  X[1] = sext24((xy[1] >> 16) | (xy[2] << 16));
  Y[1] = sext24(xy[2] >> 8);
  X[2] = sext24(xy[3]);
  Y[2] = sext24((xy[3] >> 24) | (xy[4] << 8));
}

// A primitive that needs clipping (crosses the near/far plane or the guard band).
struct alignas(16) PrimRef {
  uint32_t draw;   // global draw index
  uint32_t prim;   // scene-local primitive order
  uint32_t item;   // work item (chunk) the primitive belongs to
  uint32_t fast;   // 1: setup wrote the resolve record and the ClipFast entry of the same slot
};

// Beside the first clip_fast_cap list slots: what the clip pass needs of a
// primitive whose vertices were in setup's LDS cache, so that it reads the
// scene, the node transform and the three positions all at once instead of
// through the draw and the index buffer.
struct alignas(16) ClipFast {
  uint32_t scene, node;
  uint32_t v[3];   // vertex indices
  uint32_t pad[3];
};

// device-side counters, reset by k_nodes at the start of every frame; each
// on its own 128-byte line (index = counter * kCtrStride)
enum : uint32_t {
  kCtrClipped = 0,
  kCtrSubTris = 1,
  kCtrQueue1 = 2,    // phase-1 setup queue: first of kQueueStripes counters
  kCtrQueue2 = 3,    // phase 2 (the stripes live at kCtrStripes + ...)
  kCtrList1 = 4,     // work items set up in phase 1
  kCtrCand = 5,      // work items to test against the phase-1 depth pyramid
  kCtrList2 = 6,     // work items set up in phase 2
  kCtrClipMark = 7,  // clipped primitives handled by the phase-1 clip pass
  kCtrProp = 8,      // work items with a visibility-feature row this frame
  kCtrOverflow1 = 9,   // bins whose entries overflowed in phase 1 (exact rescan)
  kCtrOverflow2 = 10,  // and in phase 2
  kCtrCount = 11
};
constexpr uint32_t kCtrStride = 32;
// The setup work queue of each phase is split into kQueueStripes contiguous
// ranges of the item list, each with its own claim counter: single-address
// atomics serialise at the memory side (~20 ns each), so one counter for
// thousands of claims would cost more than the claims save.
constexpr uint32_t kQueueStripes = 16;
constexpr uint32_t kCtrStripes = kCtrCount;                       // [phase][stripe]
// The raster queue of each phase is split into kRasterParts contiguous
// ranges of the bins (raster items), one per XCD: a persistent workgroup
// drains the range of its own XCD (blockIdx mod 8) first, so that the bins
// of a range, whose scenes share triangle records, are rasterised out of
// one L2, and then moves on to the other ranges.  Per phase and range:
// bins listed as heavy by setup (taken first), bins with phase-2 entries,
// and the claim counter.
constexpr uint32_t kRasterParts = 8;
enum : uint32_t { kPartHeavy = 0, kPartActive = 1, kPartClaim = 2, kPartKinds = 3 };
constexpr uint32_t kCtrRasterParts = kCtrStripes + 2 * kQueueStripes;  // [phase][kind][range]
__host__ __device__ constexpr uint32_t raster_ctr(int phase, uint32_t kind, uint32_t part) {
  return kCtrRasterParts + ((uint32_t)(phase - 1) * kPartKinds + kind) * kRasterParts + part;
}
constexpr uint32_t kCtrTotal = kCtrRasterParts + 2 * kPartKinds * kRasterParts;  // reset per frame

// frame statistics (64-bit): bin entries, overflowed bins, then triangles and
// vertex-range entries set up in phase 1, the same for phase 2, then pixels
// tested by the rasteriser and fragments inside their triangles
constexpr uint32_t kStatCount = 8;
// each statistic is striped over kStatStripes words (stripe = workgroup index
// mod kStatStripes) so that workgroups rarely add to the same word; the host
// sums the stripes
constexpr uint32_t kStatStripes = 32;

// Mesh-local bounding box of the vertices a work item references.
struct alignas(16) ItemBounds {
  float lo[3];
  uint32_t pad0;
  float hi[3];
  uint32_t pad1;
};

// Per-frame screen footprint of a work item that passed the frustum test:
// conservative pixel rectangle (inclusive) and smallest window depth any of
// its fragments can have.
struct alignas(16) ItemCull {
  int16_t x0, y0, x1, y1;
  float zmin;
  uint32_t pad[2];
};

// culling modes (same values as GBR_CULL_*)
enum : int32_t { kCullNone = 0, kCullFrustum = 1, kCullHiZ = 2, kCullLearned = 3 };
// visibility-model feature sets (same values as GBR_VIS_FEATURES_*)
enum : int32_t { kFeaturesPose = 0, kFeaturesHistory = 1 };

// depth pyramid: level 0 = max depth per 8x8 pixels, level 1 per 32x32
constexpr int kHiZ0 = 3, kHiZ1 = 5;  // log2 block edges
constexpr int kHiZPer = 1 << (kHiZ1 - kHiZ0);  // level-0 blocks per level-1 block edge
// blocks a depth-pyramid test may read at level 0 before it uses level 1
constexpr int kHiZTiles = 16;


// One (scene, bin) of the raster and resolve grids, with the scene fields
// those kernels need, so that a workgroup's first load is its last
// dependent one before the bin's entries.
struct alignas(16) RasterItem {
  uint32_t scene;
  uint32_t gbin;            // global bin index (scene bin_base + bin)
  int16_t x0, y0, x1, y1;   // the bin's pixels (inclusive)
  uint32_t bin_cap;         // entries per bin of the scene
  uint32_t plane_offset;    // first pixel of the scene's planes
  int32_t width, height;    // the scene's viewport in pixels
  unsigned long long entries;   // the bin's first entry in bin_entries
  unsigned long long rec_base;  // the scene's first primitive record
  int32_t stride;           // row stride of the planes (the tile width)
  uint32_t planes;          // GBR_PLANE_* written for the scene
  uint32_t pad[2];
};

struct alignas(16) MaterialDesc {
  float ambient[4];
  int32_t texture;      // -1 = untextured
  uint32_t layer;
  uint32_t lit;         // diffuse term of the scene's lights applies (diffuse != 0)
  uint32_t pad;
  float texmat[9];      // column-major
  float pad2[3];
  float diffuse[4];     // base colour of a shaded material with lights enabled, else 0
};

// A light of a scene: the node whose absolute transform places it (global
// node index), its type (GBR_LIGHT_*), colour and range.
struct alignas(16) LightDesc {
  uint32_t node;
  int32_t type;
  float range;
  uint32_t pad;
  float color[3];
  float pad2;
};

struct alignas(16) TextureDesc {
  int32_t width, height, layers, levels;
  int32_t min_filter, mag_filter, wrap_s, wrap_t;
  int64_t level_offset[16];  // byte offsets into texels[]
};

struct FrameParams {
  // geometry
  const float* pos_x;
  const float* pos_y;
  const float* pos_z;
  const float2* uv32;
  const uint32_t* uv16;
  const uint32_t* col8;
  const float4* col32;
  const uint32_t* indices;
  // materials / textures
  const MaterialDesc* materials;
  const TextureDesc* textures;
  const uint8_t* texels;
  const float* unorm_lut;  // i/255.0f, 256 entries
  // scenes
  const DrawDesc* draws;
  const SceneDesc* scenes;
  const float* cameras;        // [scene][16] projection*view
  const float* unprojections;  // [scene][2]
  const float* abs_nodes;      // [node][16]
  // lights (null when the renderer has max_light_count 0)
  const float* normals;        // [vertex][3] mesh-local vertex normals
  const float* normal_mats;    // [node][12] inverse transpose of the upper 3x3, [col][row]
  const LightDesc* lights;
  const uint32_t* scene_lights;  // [scene][2] first light, light count
  // binning
  uint32_t* bin_count;
  uint32_t* bin_entries;
  RasterRec* rrecs;             // per primitive, indexed by rec_base + prim
  ResolveRec* srecs;            // per primitive, indexed by rec_base + prim
  RasterRec* sub_recs;          // clipped fan triangles
  uint32_t* sub_prims;          // their primitive order
  PrimRef* clipped;             // primitives that need clipping
  ClipFast* clip_fast;          // [clip_fast_cap] their scene, node and vertices
  uint32_t* counters;           // kCtr*
  unsigned long long* vis;      // per-pixel visibility keys, scene planes
  // heavy bins (entries >= heavy_entries in a phase): listed by setup when
  // a bin's count crosses the threshold, flagged per bin; the persistent
  // raster takes the list first (longest first), then every other bin
  uint32_t* heavy;              // [raster items] bins listed this phase, by range
  uint32_t* bin_heavy;          // [raster items] 1 = listed (cleared by raster)
  uint32_t* active;             // [raster items] bins with phase-2 entries, by range
  uint32_t* overflow;           // [raster items] bins whose entries overflowed this phase
  uint32_t n_ritems;            // raster items (bins of every scene)
  uint32_t part_bins;           // bins per raster queue range: ceil(n_ritems / kRasterParts)
  uint32_t heavy_entries;       // 0: no heavy list
  int32_t fuse_resolve;         // the final raster phase resolves its bins in place
  uint32_t clipped_cap, sub_cap, clip_fast_cap;
  uint32_t queue_grain;         // setup work items fetched per queue atomic
  uint32_t n_items;             // work items of the frame
  // culling
  const WorkItem* items;
  const DrawDesc* item_draws;   // per work item, its draw's descriptor
  const ItemBounds* bounds;     // per work item
  ItemCull* cull;               // per work item, this frame
  uint32_t* list1;              // work items of phase 1
  uint32_t* cand;               // work items awaiting the depth-pyramid test
  uint32_t* list2;              // work items of phase 2
  uint32_t* item_vis;           // frame number in which the item was last visible
  uint32_t* item_view;          // frame number in which the item last passed the frustum test
  float* hiz;                   // depth pyramid, hiz_stride floats per scene
  uint32_t* frame;              // device frame number (>= 2 while drawing), advanced by k_nodes
  int32_t cull_mode;
  int32_t hiz_w0, hiz_h0, hiz_w1, hiz_h1;  // pyramid level sizes in blocks
  uint32_t hiz_stride;
  float proposal_threshold;
  int32_t proposal_history;  // learned mode: phase 1 also takes last frame's visible items
  int32_t collecting;        // visibility rows are collected this frame (every chunk gets its row)
  int32_t raster_mode;  // 0: a lane per triangle, 1: a lane per (triangle, pixel)
  int32_t lane_pixels;  // mode 0: rectangles above this go to the workgroup queue
  uint32_t lds_bins;    // setup: scenes with at most this many bins count in LDS
  int32_t hiz_fused;    // phase-1 raster writes the depth pyramid (32-pixel bins)
  // learned proposal: feature rows [kCtrProp count][32] bf16 and their items
  const float* cam_info;   // [scene][8]: camera position xyz, forward xyz
  uint16_t* features;      // null: no feature rows this frame
  uint32_t* prop;
  const float* logits;     // model output per feature row (learned mode)
  int32_t feature_set;     // kFeatures*
  // outputs (env-major planes), any may be null
  uint32_t* out_color;
  float* out_depth;
  float* out_lindepth;
  uint32_t* out_semantic;
  // counters
  unsigned long long* stats;  // kStatCount x kStatStripes frame statistics
  int32_t bin_size;
  int32_t bin_shift;
  int32_t top_down;
  int32_t no_textures;
};

struct GatherArgs {
  const uint8_t* plane;
  const SceneDesc* scenes;  // per-scene viewports
  uint8_t* dst;
  long long dst_stride;
  int x0, y0, w, h;       // rectangle in framebuffer pixels
  int tile_w, tile_h, tiles_x, scene_count;
  int top_down;
  int src_bytes;          // 4
  int dst_bytes;          // 4 or 3 (RGB8 from RGBA8)
  int fill;               // value for pixels of tiles beyond scene_count
};

}  // namespace gbr
