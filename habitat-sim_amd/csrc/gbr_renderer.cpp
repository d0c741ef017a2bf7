// Host side of the MI355X batch renderer: the esp::gfx_batch::Renderer /
// RendererStandalone logic (src/esp/gfx_batch/Renderer.cpp,
// RendererStandalone.cpp) restated natively, device buffer management, and
// the C ABI declared in include/gbr.h.
//
// Scene state mirrors the reference: per scene a flat node list with
// parents[i] < i, local transforms, draws keyed into draw batches by
// (vertex-colour shader variant, mesh, texture) (Renderer.cpp:136-152,
// 880-961).  draw() sorts draws by batch (Renderer.cpp:1151-1188) to fix the
// primitive submission order that decides depth ties, uploads what changed,
// and launches the HIP kernels on the renderer's stream.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gbr.h"
#include "gbr_device.h"
#include "gbr_host.h"
#include "gbr_import.h"
#include "gbr_vis.h"

#pragma STDC FP_CONTRACT OFF

namespace gbr {
hipError_t launch_nodes(const int32_t* parents, const float* local, float* abs_out, uint32_t n,
                        uint32_t* counters, unsigned long long* stats, uint32_t* frame,
                        float* nrm_out, const uint32_t* dirty, uint32_t max_blocks, hipStream_t s);
hipError_t launch_clip(const FrameParams& p, uint32_t grid, int phase, hipStream_t s);
int setup_blocks_per_cu(uint32_t lds_bins);
hipError_t launch_setup(const FrameParams& p, const uint32_t* list, uint32_t list_ctr,
                        uint32_t queue_ctr, uint32_t grid, hipStream_t s);
int raster_blocks_per_cu(int kind);
hipError_t launch_raster(const FrameParams& p, const RasterItem* items, uint32_t n, int phase,
                         uint32_t grid,
                         hipStream_t s);
hipError_t launch_resolve(const FrameParams& p, const RasterItem* items, uint32_t n,
                          hipStream_t s);
hipError_t launch_cull(const FrameParams& p, hipStream_t s);
hipError_t launch_hiz(const FrameParams& p, uint32_t scenes, hipStream_t s);
hipError_t launch_cull_hiz(const FrameParams& p, hipStream_t s);
hipError_t launch_propose(const FrameParams& p, hipStream_t s);
hipError_t launch_collect(const FrameParams& p, uint16_t* X, uint8_t* Y, uint64_t base,
                          uint64_t cap, uint32_t* positives, hipStream_t s);
hipError_t launch_gather(const GatherArgs& g, hipStream_t s);
}  // namespace gbr

using namespace gbr;

namespace {
thread_local std::string g_error;
}  // namespace

gbr_status gbr::fail(gbr_status s, const std::string& msg) {
  g_error = msg;
  return s;
}

namespace {

struct Mat4 {
  float m[16];
};

Mat4 identity4() {
  Mat4 r;
  for (int i = 0; i < 16; ++i) r.m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  return r;
}

// Magnum Matrix4 product (accumulated from +0 over k), Renderer.cpp:1217-1220
Mat4 mul4(const Mat4& a, const Mat4& b) {
  Mat4 t;
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      float s = 0.0f;
      for (int k = 0; k < 4; ++k) {
        float p = a.m[k * 4 + r] * b.m[c * 4 + k];
        s = s + p;
      }
      t.m[c * 4 + r] = s;
    }
  return t;
}

Mat4 load4(const float* p) {
  Mat4 r;
  std::memcpy(r.m, p, sizeof r.m);
  return r;
}

struct MeshInfo {
  uint32_t vbase, vcount;
  uint32_t ibase, icount;
  uint32_t color_fmt;
  uint32_t uv_half;
};

// A run of consecutive triangles of a mesh view: the unit of setup work and
// of culling.
struct ChunkInfo {
  uint32_t t0, tn;        // first triangle (view-relative), triangle count
  uint32_t vmin, vcount;  // referenced global vertex range
  float lo[3], hi[3];     // mesh-local bounding box of the referenced vertices
};

struct MeshView {
  uint32_t mesh;
  uint32_t index_offset;  // in indices, relative to the mesh
  uint32_t index_count;
  uint32_t material;
  Mat4 xf;
  uint32_t ranges_chunk = 0;              // chunk size the chunks were computed for
  std::vector<ChunkInfo> chunks;
};


struct Draw {
  uint32_t batch, node, material, view, semantic;
};

struct Batch {
  uint32_t vcolor, mesh, texture;
};

struct Light {
  uint64_t node;
  int32_t type;
  float color[3];
  float range;
};

struct Scene {
  std::vector<int32_t> parents;
  std::vector<Mat4> transforms;
  std::vector<Batch> batches;
  std::vector<Draw> draws;
  std::vector<Light> lights;
  Mat4 camera = identity4();
  float unproj[2] = {0.0f, 0.0f};
  float cam_info[8] = {0, 0, 0, 0, 0, -1, 0, 0};  // camera position, forward (world)
  bool xform_dirty = true;
  int32_t vp_w = 0, vp_h = 0;  // viewport (0: the whole tile)
  uint32_t planes = 0;         // sensor set (0: every configured plane)
};

float unorm(int i) { return (float)i / 255.0f; }

}  // namespace

struct gbr_renderer {
  gbr_config cfg{};
  gbr_tuning tun{};
  int device = 0;
  int cu_count = 256;
  int setup_bpc = 4;      // setup workgroups per CU (occupancy with lds_bins)
  // persistent raster workgroups per CU (occupancy): storing, resolving the
  // only phase, resolving after the depth pyramid
  int raster_bpc[3] = {4, 4, 4};
  uint32_t lds_bins = 0;  // bins per scene counted in setup's LDS (0: none)
  hipStream_t stream = nullptr;

  // assets (host mirrors)
  std::vector<float> pos_x, pos_y, pos_z;
  std::vector<float2> uv32;
  std::vector<uint32_t> uv16;
  std::vector<uint32_t> col8;
  std::vector<float4> col32;
  std::vector<float> nrm;  // [vertex][3], kept only with lights enabled
  std::vector<uint32_t> indices;
  std::vector<uint8_t> texels;
  std::vector<TextureDesc> textures;
  std::vector<MaterialDesc> materials;
  std::vector<MeshInfo> meshes;
  std::vector<MeshView> views;
  std::unordered_map<std::string, std::pair<uint32_t, uint32_t>> names;
  bool files_added = false;
  bool assets_dirty = true;

  std::vector<Scene> scenes;
  bool structure_dirty = true;
  bool nodes_dirty = true;  // absolute transforms must be recomputed next frame
  DevBuf<uint32_t> d_nodes_dirty;  // the flag k_nodes reads, uploaded with the cameras

  // frame layout (rebuilt on structural change)
  std::vector<uint32_t> scene_node_base;
  std::vector<SceneDesc> scene_desc;
  uint32_t n_nodes = 0, n_items = 0, n_ritems = 0;
  uint64_t n_tris = 0;
  size_t n_bins = 0, entries_cap = 0;

  // device buffers
  DevBuf<float> d_pos_x, d_pos_y, d_pos_z;
  DevBuf<float2> d_uv32;
  DevBuf<uint32_t> d_uv16, d_col8;
  DevBuf<float4> d_col32;
  DevBuf<uint32_t> d_indices;
  DevBuf<uint8_t> d_texels;
  DevBuf<TextureDesc> d_textures;
  DevBuf<MaterialDesc> d_materials;
  DevBuf<float> d_lut;
  DevBuf<DrawDesc> d_draws;
  DevBuf<SceneDesc> d_scenes;
  DevBuf<WorkItem> d_items;
  DevBuf<DrawDesc> d_item_draws;  // each work item's draw, indexed by item
  DevBuf<ItemBounds> d_bounds;
  DevBuf<ItemCull> d_cull;
  DevBuf<uint32_t> d_list1, d_cand, d_list2, d_item_vis, d_item_view;
  DevBuf<float> d_hiz;
  DevBuf<float> d_nrm, d_nmat;       // vertex normals, per-node normal matrices
  DevBuf<LightDesc> d_lights;
  DevBuf<uint32_t> d_scene_lights;   // [scene][2] first light, count
  DevBuf<uint32_t> d_frame;  // device frame number (k_nodes advances it)

  // the last frame's kernel sequence as a HIP graph, replayed while its
  // launch parameters stay the same
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint8_t> graph_key;
  DevBuf<float> d_caminfo;
  VisModel vis;
  int32_t vis_features = GBR_VIS_FEATURES_POSE;  // feature set of the rows and the model
  bool collecting = false;
  DevBuf<RasterItem> d_ritems;
  DevBuf<int32_t> d_parents;
  DevBuf<float> d_local, d_abs;
  DevBuf<float> d_cameras, d_unproj;
  DevBuf<uint32_t> d_bin_count, d_bin_entries;
  DevBuf<uint32_t> d_heavy, d_bin_heavy;  // heavy-bin list and per-bin flags (raster items)
  DevBuf<uint32_t> d_active;              // bins with phase-2 entries (raster items)
  DevBuf<uint32_t> d_overflow;            // bins whose entries overflowed this phase
  DevBuf<RasterRec> d_rrecs;
  DevBuf<ResolveRec> d_srecs;
  DevBuf<RasterRec> d_sub_recs;
  DevBuf<uint32_t> d_sub_prims;
  DevBuf<PrimRef> d_clipped;
  DevBuf<ClipFast> d_clip_fast;
  DevBuf<uint32_t> d_counters;
  DevBuf<unsigned long long> d_vis;
  uint32_t clipped_cap = 0, sub_cap = 0, clip_fast_cap = 0;
  DevBuf<unsigned long long> d_stats;
  DevBuf<uint32_t> d_color, d_sem;
  DevBuf<float> d_depth, d_lin;
  DevBuf<uint8_t> d_fb_color, d_fb_depth, d_stage;
  // pinned per-frame staging for cameras
  float* h_cam = nullptr;
  size_t h_cam_cap = 0;
  uint8_t* h_light = nullptr;  // pinned per-frame staging for lights
  size_t h_light_cap = 0;
  hipEvent_t cam_event = nullptr;
  // optional per-stage timing of gbr_draw (tuning.timing): one event set per
  // frame since the last gbr_frame_timings(), up to kTimingFrames frames
  std::vector<std::array<hipEvent_t, 2 * GBR_TIMING_SLOTS>> t_pool;
  std::vector<std::array<bool, GBR_TIMING_SLOTS>> t_used;
  size_t t_frames = 0;
  int t_cur = -1;
  bool cam_event_pending = false;

  uint64_t tile_pixels() const { return (uint64_t)cfg.tile_size[0] * cfg.tile_size[1]; }
  uint32_t scene_count() const { return (uint32_t)scenes.size(); }
};

namespace {

gbr_status check_scene(const gbr_renderer* r, uint32_t scene, const char* fn) {
  if (!r) return fail(GBR_ERR_ARGUMENT, std::string(fn) + ": null renderer");
  if (scene >= r->scenes.size())
    return fail(GBR_ERR_RANGE, std::string(fn) + ": index " + std::to_string(scene) +
                                   " out of range for " + std::to_string(r->scenes.size()) +
                                   " scenes");
  return GBR_OK;
}

gbr_status set_device(const gbr_renderer* r) {
  HIP_TRY(hipSetDevice(r->device));
  return GBR_OK;
}

// glGenerateMipmap() stand-in: 2x2 rounded box filter down to min(w,h) == 1
std::vector<std::vector<uint8_t>> generate_mips(const uint8_t* level0, int w, int h, int layers,
                                                int& levels_out) {
  int n = 1;
  while ((1 << n) <= std::min(w, h)) ++n;  // log2(min)+1
  std::vector<std::vector<uint8_t>> out;
  out.emplace_back(level0, level0 + (size_t)w * h * layers * 4);
  int cw = w, ch = h;
  for (int l = 1; l < n; ++l) {
    int nw = std::max(cw / 2, 1), nh = std::max(ch / 2, 1);
    const std::vector<uint8_t>& src = out.back();
    std::vector<uint8_t> dst((size_t)nw * nh * layers * 4);
    for (int la = 0; la < layers; ++la)
      for (int y = 0; y < nh; ++y)
        for (int x = 0; x < nw; ++x)
          for (int c = 0; c < 4; ++c) {
            auto at = [&](int xx, int yy) {
              return (int)src[(((size_t)la * ch + yy) * cw + xx) * 4 + c];
            };
            int s = at(2 * x, 2 * y) + at(2 * x, 2 * y + 1) + at(2 * x + 1, 2 * y) +
                    at(2 * x + 1, 2 * y + 1);
            dst[(((size_t)la * nh + y) * nw + x) * 4 + c] = (uint8_t)((s + 2) >> 2);
          }
    out.push_back(std::move(dst));
    cw = nw;
    ch = nh;
  }
  levels_out = n;
  return out;
}

uint32_t batch_id(Scene& sc, uint32_t vcolor, uint32_t mesh, uint32_t tex) {
  for (size_t i = 0; i < sc.batches.size(); ++i)
    if (sc.batches[i].vcolor == vcolor && sc.batches[i].mesh == mesh && sc.batches[i].texture == tex)
      return (uint32_t)i;
  sc.batches.push_back({vcolor, mesh, tex});
  return (uint32_t)sc.batches.size() - 1;
}

// Chunks of a mesh view for the setup work items: runs of consecutive
// triangles, at most `chunk` long, cut where the referenced vertex range
// would outgrow the LDS vertex cache.
const std::vector<ChunkInfo>& view_chunks(gbr_renderer* r, MeshView& v, uint32_t chunk) {
  if (v.ranges_chunk == chunk) return v.chunks;
  v.chunks.clear();
  const MeshInfo& m = r->meshes[v.mesh];
  const uint32_t ntri = v.index_count / 3;
  const uint32_t* ix = r->indices.data() + m.ibase + v.index_offset;
  uint32_t t0 = 0;
  while (t0 < ntri) {
    uint32_t lo = 0xffffffffu, hi = 0, t = t0;
    while (t < ntri && t - t0 < chunk) {
      uint32_t nlo = lo, nhi = hi;
      for (int k = 0; k < 3; ++k) {
        nlo = std::min(nlo, ix[3 * t + k]);
        nhi = std::max(nhi, ix[3 * t + k]);
      }
      if (t > t0 && nhi - nlo + 1 > (uint32_t)kVertexCache) break;
      lo = nlo;
      hi = nhi;
      ++t;
    }
    ChunkInfo ci{};
    ci.t0 = t0;
    ci.tn = t - t0;
    ci.vmin = lo;
    ci.vcount = hi - lo + 1;
    for (int k = 0; k < 3; ++k) {
      ci.lo[k] = INFINITY;
      ci.hi[k] = -INFINITY;
    }
    for (uint32_t i = 3 * t0; i < 3 * t; ++i) {
      const uint32_t g = ix[i];
      const float q[3] = {r->pos_x[g], r->pos_y[g], r->pos_z[g]};
      for (int k = 0; k < 3; ++k) {
        ci.lo[k] = std::min(ci.lo[k], q[k]);
        ci.hi[k] = std::max(ci.hi[k], q[k]);
      }
    }
    v.chunks.push_back(ci);
    t0 = t;
  }
  v.ranges_chunk = chunk;
  return v.chunks;
}

gbr_status upload_assets(gbr_renderer* r) {
  hipStream_t s = r->stream;
  TRY(upload(r->d_pos_x, r->pos_x, s));
  TRY(upload(r->d_pos_y, r->pos_y, s));
  TRY(upload(r->d_pos_z, r->pos_z, s));
  TRY(upload(r->d_uv32, r->uv32, s));
  TRY(upload(r->d_uv16, r->uv16, s));
  TRY(upload(r->d_col8, r->col8, s));
  TRY(upload(r->d_col32, r->col32, s));
  if (r->cfg.max_light_count) TRY(upload(r->d_nrm, r->nrm, s));
  TRY(upload(r->d_indices, r->indices, s));
  TRY(upload(r->d_texels, r->texels, s));
  TRY(upload(r->d_textures, r->textures, s));
  TRY(upload(r->d_materials, r->materials, s));
  std::vector<float> lut(256);
  for (int i = 0; i < 256; ++i) lut[i] = unorm(i);
  TRY(upload(r->d_lut, lut, s));
  HIP_TRY(hipStreamSynchronize(s));  // host vectors may change after return
  r->assets_dirty = false;
  return GBR_OK;
}

gbr_status rebuild_structure(gbr_renderer* r) {
  hipStream_t s = r->stream;
  const uint32_t nsc = r->scene_count();
  const int W = r->cfg.tile_size[0], H = r->cfg.tile_size[1];
  const int B = r->tun.bin_size;
  const uint32_t chunk = (uint32_t)r->tun.chunk_triangles;
  // bin grids follow the scenes' viewports; setup counts a scene's pairs in
  // LDS when its grid fits the largest grid that does (kMaxSceneBins), and
  // appends them straight to the global bins otherwise
  r->lds_bins = 0;
  for (uint32_t si = 0; si < nsc; ++si) {
    const Scene& sc = r->scenes[si];
    const int vw = sc.vp_w ? sc.vp_w : W, vh = sc.vp_h ? sc.vp_h : H;
    const uint32_t nb = (uint32_t)(((vw + B - 1) / B) * ((vh + B - 1) / B));
    if (nb <= (uint32_t)kMaxSceneBins) r->lds_bins = std::max(r->lds_bins, nb);
  }
  r->setup_bpc = setup_blocks_per_cu(r->lds_bins);
  for (int k = 0; k < 3; ++k) r->raster_bpc[k] = raster_blocks_per_cu(k);

  std::vector<int32_t> parents;
  std::vector<float> local;
  std::vector<DrawDesc> draws;
  std::vector<WorkItem> items;
  std::vector<ItemBounds> bounds;
  std::vector<RasterItem> ritems;
  r->scene_node_base.assign(nsc, 0);
  r->scene_desc.assign(nsc, SceneDesc{});
  uint64_t bins = 0, entries = 0, tris = 0;
  for (uint32_t si = 0; si < nsc; ++si) {
    Scene& sc = r->scenes[si];
    const uint32_t nbase = (uint32_t)parents.size();
    r->scene_node_base[si] = nbase;
    for (size_t i = 0; i < sc.parents.size(); ++i) {
      parents.push_back(sc.parents[i] < 0 ? -1 : (int32_t)(nbase + sc.parents[i]));
      local.insert(local.end(), sc.transforms[i].m, sc.transforms[i].m + 16);
    }
    sc.xform_dirty = false;
    // counting sort of draws by batch id, stable (Renderer.cpp:1151-1188)
    std::vector<uint32_t> order;
    order.reserve(sc.draws.size());
    for (uint32_t b = 0; b < sc.batches.size(); ++b)
      for (uint32_t i = 0; i < sc.draws.size(); ++i)
        if (sc.draws[i].batch == b) order.push_back(i);
    SceneDesc& sd = r->scene_desc[si];
    const int vw = sc.vp_w ? sc.vp_w : W, vh = sc.vp_h ? sc.vp_h : H;
    const int bins_x = (vw + B - 1) / B, bins_y = (vh + B - 1) / B;
    sd.width = vw;
    sd.height = vh;
    sd.stride = W;
    sd.planes = sc.planes ? (sc.planes & r->cfg.planes) : r->cfg.planes;
    sd.bins_x = bins_x;
    sd.bins_y = bins_y;
    sd.draw_begin = (uint32_t)draws.size();
    sd.draw_count = (uint32_t)order.size();
    sd.plane_offset = (uint32_t)((uint64_t)si * W * H);
    uint32_t prim = 0;
    for (uint32_t oi : order) {
      const Draw& dr = sc.draws[oi];
      MeshView& v = r->views[dr.view];
      const MeshInfo& m = r->meshes[v.mesh];
      DrawDesc dd{};
      dd.index_base = m.ibase + v.index_offset;
      dd.tri_count = v.index_count / 3;
      dd.node = nbase + dr.node;
      dd.material = dr.material;
      dd.semantic = dr.semantic;
      dd.flags = (m.color_fmt << kDrawColorShift) | (m.uv_half ? kDrawUvHalf : 0u);
      dd.prim_base = prim;
      dd.scene = si;
      {
        const MaterialDesc& md = r->materials[dr.material];
        for (int c = 0; c < 4; ++c) dd.ambient[c] = md.ambient[c];
        dd.texture = md.texture;
        dd.layer = md.layer;
        dd.lit = md.lit;
      }
      const uint32_t gdraw = (uint32_t)draws.size();
      draws.push_back(dd);
      for (const ChunkInfo& ci : view_chunks(r, v, chunk)) {
        WorkItem wi{};
        wi.scene = si;
        wi.draw = gdraw;
        wi.first_index = dd.index_base + 3 * ci.t0;
        wi.tri_count = ci.tn;
        wi.prim = prim + ci.t0;
        wi.vmin = ci.vmin;  // global vertex indices
        wi.vcount = ci.vcount;
        wi.node = dd.node;
        items.push_back(wi);
        ItemBounds ib{};
        for (int k = 0; k < 3; ++k) {
          ib.lo[k] = ci.lo[k];
          ib.hi[k] = ci.hi[k];
        }
        bounds.push_back(ib);
      }
      prim += dd.tri_count;
    }
    sd.prim_count = prim;
    sd.rec_base = tris;
    tris += prim;
    const uint32_t nb = (uint32_t)(bins_x * bins_y);
    sd.bin_base = (uint32_t)bins;
    double per_bin = (double)prim / nb * r->tun.bin_capacity_factor;
    uint32_t cap = (uint32_t)std::min<double>(std::max<double>(per_bin, r->tun.min_bin_capacity), 1u << 24);
    cap = (cap + 63) & ~63u;
    sd.bin_cap = cap;
    sd.entry_base = entries;
    bins += nb;
    entries += (uint64_t)nb * cap;
    for (uint32_t b = 0; b < nb; ++b) {
      RasterItem ri{};
      ri.scene = si;
      ri.gbin = sd.bin_base + b;
      const int bx = (int)(b % (uint32_t)bins_x), by = (int)(b / (uint32_t)bins_x);
      ri.x0 = (int16_t)(bx * B);
      ri.y0 = (int16_t)(by * B);
      ri.x1 = (int16_t)(std::min(bx * B + B, vw) - 1);
      ri.y1 = (int16_t)(std::min(by * B + B, vh) - 1);
      ri.bin_cap = cap;
      ri.plane_offset = sd.plane_offset;
      ri.width = vw;
      ri.height = vh;
      ri.entries = sd.entry_base + (uint64_t)b * cap;
      ri.rec_base = sd.rec_base;
      ri.stride = W;
      ri.planes = sd.planes;
      ritems.push_back(ri);
    }
  }
  r->n_nodes = (uint32_t)parents.size();
  r->n_items = (uint32_t)items.size();
  r->n_ritems = (uint32_t)ritems.size();
  r->n_tris = tris;
  TRY(upload(r->d_parents, parents, s));
  TRY(upload(r->d_local, local, s));
  TRY(r->d_abs.reserve(std::max<size_t>(local.size(), 16)));
  if (r->cfg.max_light_count) TRY(r->d_nmat.reserve(std::max<size_t>(parents.size(), 1) * 12));
  TRY(upload(r->d_draws, draws, s));
  TRY(upload(r->d_scenes, r->scene_desc, s));
  TRY(upload(r->d_items, items, s));
  // (kept alive until the stream synchronisation below, like the other
  // host arrays of this upload)
  std::vector<DrawDesc> item_draws(items.size());
  for (size_t i = 0; i < items.size(); ++i) item_draws[i] = draws[items[i].draw];
  TRY(upload(r->d_item_draws, item_draws, s));
  TRY(upload(r->d_bounds, bounds, s));
  {
    const size_t n = std::max<size_t>(items.size(), 1);
    TRY(r->d_cull.reserve(n));
    TRY(r->d_list1.reserve(n));
    TRY(r->d_cand.reserve(n));
    TRY(r->d_list2.reserve(n));
    TRY(r->d_item_vis.reserve(n));
    TRY(r->d_item_view.reserve(n));
    // visibility and frustum history refer to the old item numbering
    HIP_TRY(hipMemsetAsync(r->d_item_vis.ptr, 0, n * sizeof(uint32_t), s));
    HIP_TRY(hipMemsetAsync(r->d_item_view.ptr, 0, n * sizeof(uint32_t), s));
    TRY(r->d_frame.reserve(1));
    HIP_TRY(hipMemsetD32Async(r->d_frame.ptr, 1u, 1, s));
  }
  TRY(upload(r->d_ritems, ritems, s));
  if (bins > r->n_bins || r->d_bin_count.ptr == nullptr) {
    TRY(r->d_bin_count.reserve(bins + 1));
    HIP_TRY(hipMemsetAsync(r->d_bin_count.ptr, 0, r->d_bin_count.cap * sizeof(uint32_t), s));
  }
  r->n_bins = bins;
  // heavy-bin list and flags: the raster clears every flag it meets, so they
  // are zero between phases; cleared here for a new layout
  TRY(r->d_heavy.reserve(std::max<size_t>(ritems.size(), 1)));
  TRY(r->d_bin_heavy.reserve(std::max<size_t>(ritems.size(), 1)));
  TRY(r->d_active.reserve(std::max<size_t>(ritems.size(), 1)));
  TRY(r->d_overflow.reserve(std::max<size_t>(ritems.size(), 1)));
  HIP_TRY(hipMemsetAsync(r->d_bin_heavy.ptr, 0, r->d_bin_heavy.cap * sizeof(uint32_t), s));
  TRY(r->d_bin_entries.reserve(std::max<uint64_t>(entries, 1)));
  r->entries_cap = entries;
  TRY(r->d_rrecs.reserve(std::max<uint64_t>(tris, 1)));
  TRY(r->d_srecs.reserve(std::max<uint64_t>(tris, 1)));
  // every primitive is listed at most once; the sub-triangle pool gives list
  // slot i the fan slots [7 i, 7 i + 7) and is bounded (bins of a fan past
  // it fall back to the exact rescan)
  r->clipped_cap = (uint32_t)std::max<uint64_t>(tris, 1);
  r->sub_cap = (uint32_t)std::min<uint64_t>(7 * std::max<uint64_t>(tris, 1024), 7u << 20);
  // the clip pass's fast entries cover the first slots of the list (beyond
  // them it reads a primitive through its draw)
  r->clip_fast_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(tris, 1), 1u << 20);
  TRY(r->d_clipped.reserve(r->clipped_cap));
  TRY(r->d_clip_fast.reserve(r->clip_fast_cap));
  TRY(r->d_sub_recs.reserve(r->sub_cap));
  TRY(r->d_sub_prims.reserve(r->sub_cap));
  HIP_TRY(hipStreamSynchronize(s));
  r->structure_dirty = false;
  r->nodes_dirty = true;  // new node arrays: every absolute transform is stale
  return GBR_OK;
}

gbr_status upload_dirty_transforms(gbr_renderer* r) {
  for (uint32_t si = 0; si < r->scene_count(); ++si) {
    Scene& sc = r->scenes[si];
    if (!sc.xform_dirty) continue;
    sc.xform_dirty = false;
    if (sc.transforms.empty()) continue;
    r->nodes_dirty = true;
    HIP_TRY(hipMemcpyAsync(r->d_local.ptr + (size_t)r->scene_node_base[si] * 16, sc.transforms.data(),
                           sc.transforms.size() * sizeof(Mat4), hipMemcpyHostToDevice, r->stream));
  }
  return GBR_OK;
}

gbr_status upload_cameras(gbr_renderer* r) {
  const size_t n = r->scene_count();
  // projection*view, unprojection, camera position + forward, then the
  // frame's node-recompute flag
  const size_t need = n * 26 + 1;
  if (r->cam_event_pending) {
    HIP_TRY(hipEventSynchronize(r->cam_event));
    r->cam_event_pending = false;
  }
  if (r->h_cam_cap < need) {
    if (r->h_cam) (void)hipHostFree(r->h_cam);
    r->h_cam = nullptr;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r->h_cam), need * sizeof(float)));
    r->h_cam_cap = need;
  }
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(r->h_cam + i * 16, r->scenes[i].camera.m, 64);
    r->h_cam[n * 16 + 2 * i] = r->scenes[i].unproj[0];
    r->h_cam[n * 16 + 2 * i + 1] = r->scenes[i].unproj[1];
    std::memcpy(r->h_cam + n * 18 + 8 * i, r->scenes[i].cam_info, 32);
  }
  TRY(r->d_cameras.reserve(n * 16 + 16));
  TRY(r->d_unproj.reserve(n * 2 + 2));
  TRY(r->d_caminfo.reserve(n * 8 + 8));
  HIP_TRY(hipMemcpyAsync(r->d_cameras.ptr, r->h_cam, n * 64, hipMemcpyHostToDevice, r->stream));
  HIP_TRY(hipMemcpyAsync(r->d_unproj.ptr, r->h_cam + n * 16, n * 8, hipMemcpyHostToDevice, r->stream));
  HIP_TRY(hipMemcpyAsync(r->d_caminfo.ptr, r->h_cam + n * 18, n * 32, hipMemcpyHostToDevice,
                         r->stream));
  {
    const uint32_t flag = r->nodes_dirty ? 1u : 0u;
    std::memcpy(r->h_cam + n * 26, &flag, sizeof flag);
    TRY(r->d_nodes_dirty.reserve(1));
    HIP_TRY(hipMemcpyAsync(r->d_nodes_dirty.ptr, r->h_cam + n * 26, sizeof flag,
                           hipMemcpyHostToDevice, r->stream));
    r->nodes_dirty = false;
  }
  if (r->cfg.max_light_count) {
    // per scene its first light and light count (at most max_light_count,
    // the count the reference's shaders are compiled for), then the lights
    // with their nodes as global indices
    size_t nl = 0;
    for (size_t i = 0; i < n; ++i)
      nl += std::min<size_t>(r->scenes[i].lights.size(), r->cfg.max_light_count);
    const size_t head = (n * 2 * sizeof(uint32_t) + 15) & ~size_t(15);
    const size_t bytes = head + std::max<size_t>(nl, 1) * sizeof(LightDesc);
    if (r->h_light_cap < bytes) {
      if (r->h_light) (void)hipHostFree(r->h_light);
      r->h_light = nullptr;
      HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r->h_light), bytes));
      r->h_light_cap = bytes;
    }
    uint32_t* ranges = reinterpret_cast<uint32_t*>(r->h_light);
    LightDesc* lights = reinterpret_cast<LightDesc*>(r->h_light + head);
    uint32_t k = 0;
    for (size_t i = 0; i < n; ++i) {
      const Scene& sc = r->scenes[i];
      const uint32_t cnt = (uint32_t)std::min<size_t>(sc.lights.size(), r->cfg.max_light_count);
      ranges[2 * i] = k;
      ranges[2 * i + 1] = cnt;
      for (uint32_t l = 0; l < cnt; ++l, ++k) {
        const Light& L = sc.lights[l];
        LightDesc& d = lights[k];
        d = LightDesc{};
        d.node = r->scene_node_base[i] + (uint32_t)L.node;
        d.type = L.type;
        d.range = L.range;
        std::memcpy(d.color, L.color, sizeof d.color);
      }
    }
    TRY(r->d_scene_lights.reserve(n * 2 + 2));
    TRY(r->d_lights.reserve(std::max<size_t>(nl, 1)));
    HIP_TRY(hipMemcpyAsync(r->d_scene_lights.ptr, ranges, n * 2 * sizeof(uint32_t),
                           hipMemcpyHostToDevice, r->stream));
    if (nl)
      HIP_TRY(hipMemcpyAsync(r->d_lights.ptr, lights, nl * sizeof(LightDesc),
                             hipMemcpyHostToDevice, r->stream));
  }
  HIP_TRY(hipEventRecord(r->cam_event, r->stream));
  r->cam_event_pending = true;
  return GBR_OK;
}

gbr_status ensure_planes(gbr_renderer* r) {
  const size_t px = (size_t)r->scene_count() * r->tile_pixels();
  const uint32_t pl = r->cfg.planes;
  if (pl & GBR_PLANE_COLOR) TRY(r->d_color.reserve(px));
  if (pl & GBR_PLANE_DEPTH) TRY(r->d_depth.reserve(px));
  if (pl & GBR_PLANE_LINEAR_DEPTH) TRY(r->d_lin.reserve(px));
  if (pl & GBR_PLANE_SEMANTIC) TRY(r->d_sem.reserve(px));
  return GBR_OK;
}

constexpr size_t kTimingFrames = 4096;

gbr_status ensure_timing_events(gbr_renderer* r) {
  r->t_cur = -1;
  if (!r->tun.timing || r->t_frames >= kTimingFrames) return GBR_OK;
  if (r->t_frames == r->t_pool.size()) {
    std::array<hipEvent_t, 2 * GBR_TIMING_SLOTS> set{};
    for (hipEvent_t& e : set) HIP_TRY(hipEventCreate(&e));
    r->t_pool.push_back(set);
    r->t_used.emplace_back();
  }
  r->t_cur = (int)r->t_frames++;
  r->t_used[r->t_cur].fill(false);
  return GBR_OK;
}

// Brackets one stage of the frame with events on the renderer's stream.
struct StageTimer {
  gbr_renderer* r;
  int slot;
  StageTimer(gbr_renderer* rr, int s) : r(rr), slot(s) {
    if (r->t_cur >= 0) (void)hipEventRecord(r->t_pool[r->t_cur][2 * slot], r->stream);
  }
  ~StageTimer() {
    if (r->t_cur >= 0) {
      (void)hipEventRecord(r->t_pool[r->t_cur][2 * slot + 1], r->stream);
      r->t_used[r->t_cur][slot] = true;
    }
  }
};

gbr_status enqueue_frame(gbr_renderer* r, const FrameParams& p);

// The executable graph owns the kernel-argument memory of its launches; the
// last launch may still be running on the stream, so it finishes first.
void release_graph(gbr_renderer* r) {
  if (r->graph_exec) {
    (void)hipStreamSynchronize(r->stream);
    (void)hipGraphExecDestroy(r->graph_exec);
  }
  if (r->graph) (void)hipGraphDestroy(r->graph);
  r->graph_exec = nullptr;
  r->graph = nullptr;
  r->graph_key.clear();
}

gbr_status do_draw(gbr_renderer* r) {
  TRY(set_device(r));
  TRY(ensure_planes(r));
  if (r->assets_dirty) TRY(upload_assets(r));
  if (r->structure_dirty) TRY(rebuild_structure(r));
  else TRY(upload_dirty_transforms(r));
  TRY(upload_cameras(r));
  TRY(r->d_stats.reserve(kStatCount * kStatStripes));
  TRY(r->d_counters.reserve(kCtrTotal * kCtrStride));
  TRY(r->d_vis.reserve((size_t)r->scene_count() * r->tile_pixels()));
  TRY(ensure_timing_events(r));

  FrameParams p{};
  p.pos_x = r->d_pos_x.ptr;
  p.pos_y = r->d_pos_y.ptr;
  p.pos_z = r->d_pos_z.ptr;
  p.uv32 = r->d_uv32.ptr;
  p.uv16 = r->d_uv16.ptr;
  p.col8 = r->d_col8.ptr;
  p.col32 = r->d_col32.ptr;
  p.indices = r->d_indices.ptr;
  p.materials = r->d_materials.ptr;
  p.textures = r->d_textures.ptr;
  p.texels = r->d_texels.ptr;
  p.unorm_lut = r->d_lut.ptr;
  p.draws = r->d_draws.ptr;
  p.scenes = r->d_scenes.ptr;
  p.cameras = r->d_cameras.ptr;
  p.unprojections = r->d_unproj.ptr;
  p.abs_nodes = r->d_abs.ptr;
  if (r->cfg.max_light_count) {
    p.normals = r->d_nrm.ptr;
    p.normal_mats = r->d_nmat.ptr;
    p.lights = r->d_lights.ptr;
    p.scene_lights = r->d_scene_lights.ptr;
  }
  p.bin_count = r->d_bin_count.ptr;
  p.heavy = r->d_heavy.ptr;
  p.bin_heavy = r->d_bin_heavy.ptr;
  p.active = r->d_active.ptr;
  p.overflow = r->d_overflow.ptr;
  p.n_ritems = r->n_ritems;
  p.part_bins = std::max<uint32_t>((r->n_ritems + kRasterParts - 1) / kRasterParts, 1u);
  p.heavy_entries = r->tun.raster_mode == 0 ? (uint32_t)r->tun.heavy_entries : 0u;
  // the lane rasteriser resolves its final phase in place (lit renderers
  // keep the separate resolve and its larger register footprint)
  p.fuse_resolve = (r->tun.raster_mode == 0 && r->cfg.max_light_count == 0) ? 1 : 0;
  p.bin_entries = r->d_bin_entries.ptr;
  p.rrecs = r->d_rrecs.ptr;
  p.srecs = r->d_srecs.ptr;
  p.sub_recs = r->d_sub_recs.ptr;
  p.sub_prims = r->d_sub_prims.ptr;
  p.clipped = r->d_clipped.ptr;
  p.clip_fast = r->d_clip_fast.ptr;
  p.counters = r->d_counters.ptr;
  p.vis = r->d_vis.ptr;
  p.clipped_cap = r->clipped_cap;
  p.clip_fast_cap = r->clip_fast_cap;
  p.queue_grain = (uint32_t)r->tun.queue_grain;
  p.n_items = r->n_items;
  p.items = r->d_items.ptr;
  p.item_draws = r->d_item_draws.ptr;
  p.bounds = r->d_bounds.ptr;
  p.cull = r->d_cull.ptr;
  p.list1 = r->d_list1.ptr;
  p.cand = r->d_cand.ptr;
  p.list2 = r->d_list2.ptr;
  p.cull_mode = r->tun.cull_mode;
  // visibility feedback is kept in every mode, so switching to the
  // previous-frame proposal starts from the last frame drawn
  p.item_vis = r->d_item_vis.ptr;
  p.item_view = r->d_item_view.ptr;
  p.frame = r->d_frame.ptr;
  p.proposal_threshold = r->tun.proposal_threshold;
  p.proposal_history = r->tun.proposal_history;
  p.collecting = r->collecting ? 1 : 0;
  p.cam_info = r->d_caminfo.ptr;
  const bool learned = p.cull_mode == kCullLearned;
  if (learned || r->collecting) {
    TRY(r->vis.reserve_infer(r->n_items));
    p.features = r->vis.Xi.ptr;
    p.prop = r->vis.prop.ptr;
    p.logits = r->vis.logits_i.ptr;
  }
  p.feature_set = r->vis_features;
  p.raster_mode = r->tun.raster_mode;
  p.lane_pixels = r->tun.lane_pixels;
  {
    const int W = r->cfg.tile_size[0], H = r->cfg.tile_size[1];
    p.hiz_w0 = (W + (1 << kHiZ0) - 1) >> kHiZ0;
    p.hiz_h0 = (H + (1 << kHiZ0) - 1) >> kHiZ0;
    p.hiz_w1 = (W + 31) >> 5;
    p.hiz_h1 = (H + 31) >> 5;
    p.hiz_stride = (uint32_t)(p.hiz_w0 * p.hiz_h0 + p.hiz_w1 * p.hiz_h1);
    if (p.cull_mode >= kCullHiZ) {
      TRY(r->d_hiz.reserve((size_t)p.hiz_stride * std::max<uint32_t>(r->scene_count(), 1)));
      p.hiz = r->d_hiz.ptr;
      // the lane rasteriser builds the pyramid from its 32x32 LDS tiles
      p.hiz_fused = (r->tun.bin_size == 32 && r->tun.raster_mode != 1) ? 1 : 0;
    }
  }
  p.sub_cap = r->sub_cap;
  const uint32_t pl = r->cfg.planes;
  p.out_color = (pl & GBR_PLANE_COLOR) ? r->d_color.ptr : nullptr;
  p.out_depth = (pl & GBR_PLANE_DEPTH) ? r->d_depth.ptr : nullptr;
  p.out_lindepth = (pl & GBR_PLANE_LINEAR_DEPTH) ? r->d_lin.ptr : nullptr;
  p.out_semantic = (pl & GBR_PLANE_SEMANTIC) ? r->d_sem.ptr : nullptr;
  p.stats = r->d_stats.ptr;
  p.bin_size = r->tun.bin_size;
  int sh = 0;
  while ((1 << sh) < r->tun.bin_size) ++sh;
  p.bin_shift = sh;
  p.lds_bins = r->lds_bins;
  p.top_down = r->cfg.top_down ? 1 : 0;
  p.no_textures = (r->cfg.flags & GBR_FLAG_NO_TEXTURES) ? 1 : 0;

  // every launch below reads its per-frame inputs from device memory, so
  // with timing and row collection off the sequence is captured once as a
  // HIP graph and replayed while the parameters stay the same
  const bool graphable = r->tun.graph && !r->tun.timing && !r->collecting;
  if (!graphable) return enqueue_frame(r, p);
  // the key holds every value a captured launch takes as an argument: the
  // frame parameters, the node buffers, and the scorer's weight buffers
  constexpr int kExtra = 13;
  std::vector<uint8_t> key(sizeof p + kExtra * sizeof(uint64_t));
  {
    const uint64_t extra[kExtra] = {
        (uint64_t)(uintptr_t)r->d_parents.ptr, (uint64_t)(uintptr_t)r->d_local.ptr,
        (uint64_t)(uintptr_t)r->d_abs.ptr, (uint64_t)(uintptr_t)r->d_ritems.ptr,
        ((uint64_t)r->n_nodes << 32) | r->n_ritems,
        ((uint64_t)r->cu_count << 32) | (uint32_t)r->setup_bpc,
        (uint64_t)(uintptr_t)r->vis.Xi.ptr, (uint64_t)(uintptr_t)r->vis.logits_i.ptr,
        (uint64_t)(uintptr_t)r->vis.w.ptr, (uint64_t)(uintptr_t)r->vis.W1b.ptr,
        (uint64_t)(uintptr_t)r->vis.W2b.ptr, (uint64_t)(uintptr_t)r->vis.w3b.ptr,
        (uint64_t)(uintptr_t)r->d_nodes_dirty.ptr};
    std::memcpy(key.data(), &p, sizeof p);
    std::memcpy(key.data() + sizeof p, extra, sizeof extra);
  }
  if (!r->graph_exec || key != r->graph_key) {
    release_graph(r);
    HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeRelaxed));
    const gbr_status st = enqueue_frame(r, p);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(r->stream, &g);
    if (st != GBR_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HIP_TRY(e);
    r->graph = g;
    HIP_TRY(hipGraphInstantiate(&r->graph_exec, r->graph, nullptr, nullptr, 0));
    r->graph_key = std::move(key);
  }
  HIP_TRY(hipGraphLaunch(r->graph_exec, r->stream));
  return GBR_OK;
}

// The frame's kernel sequence on the renderer's stream (no host
// synchronisation: list lengths live in device counters).
gbr_status enqueue_frame(gbr_renderer* r, const FrameParams& p) {
  const bool learned = p.cull_mode == kCullLearned;
  {
    StageTimer t(r, GBR_TIMING_NODES);
    HIP_TRY(launch_nodes(r->d_parents.ptr, r->d_local.ptr, r->d_abs.ptr, r->n_nodes, p.counters,
                         p.stats, p.frame, const_cast<float*>(p.normal_mats),
                         r->d_nodes_dirty.ptr, (uint32_t)r->cu_count * 2u, r->stream));
  }
  const uint32_t setup_grid = (uint32_t)(r->cu_count * r->setup_bpc);
  const uint32_t clip_grid = (uint32_t)r->cu_count * 4u;
  // the lane rasteriser's phase that resolves in place runs at lower occupancy
  const bool fuse2 = p.fuse_resolve && p.cull_mode >= kCullHiZ;
  const uint32_t raster_grid1 = (uint32_t)(r->cu_count * r->raster_bpc[p.fuse_resolve && !fuse2 ? 1 : 0]);
  const uint32_t raster_grid2 = (uint32_t)(r->cu_count * r->raster_bpc[fuse2 ? 2 : 0]);
  const bool culled = p.cull_mode != kCullNone;
  if (culled || r->collecting) {  // (unculled frames run it for the feature rows only)
    // phase 1: items that pass the frustum test and are proposed visible
    // (all of them in frustum mode); the rest become phase-2 candidates
    StageTimer t(r, GBR_TIMING_CULL);
    HIP_TRY(launch_cull(p, r->stream));
    if (learned) {
      const int rows = (int)((r->n_items + 31) & ~31u);
      TRY(r->vis.score(r->vis.Xi.ptr, rows, p.counters + kCtrProp * kCtrStride,
                       r->vis.logits_i.ptr, r->cu_count * 2, r->stream));
      HIP_TRY(launch_propose(p, r->stream));
    }
  }
  {
    StageTimer t(r, GBR_TIMING_SETUP1);
    HIP_TRY(launch_setup(p, culled ? p.list1 : nullptr, kCtrList1, kCtrQueue1, setup_grid,
                         r->stream));
  }
  {
    StageTimer t(r, GBR_TIMING_CLIP1);
    HIP_TRY(launch_clip(p, clip_grid, 1, r->stream));
  }
  {
    StageTimer t(r, GBR_TIMING_RASTER1);
    HIP_TRY(launch_raster(p, r->d_ritems.ptr, r->n_ritems, 1, raster_grid1, r->stream));
  }
  if (p.cull_mode >= kCullHiZ) {
    // phase 2: candidates not hidden behind the phase-1 depth pyramid
    {
      StageTimer t(r, GBR_TIMING_HIZ);
      if (!p.hiz_fused) HIP_TRY(launch_hiz(p, r->scene_count(), r->stream));
      HIP_TRY(launch_cull_hiz(p, r->stream));
    }
    {
      StageTimer t(r, GBR_TIMING_SETUP2);
      HIP_TRY(launch_setup(p, p.list2, kCtrList2, kCtrQueue2, setup_grid, r->stream));
    }
    {
      StageTimer t(r, GBR_TIMING_CLIP2);
      HIP_TRY(launch_clip(p, clip_grid, 2, r->stream));
    }
    {
      StageTimer t(r, GBR_TIMING_RASTER2);
      HIP_TRY(launch_raster(p, r->d_ritems.ptr, r->n_ritems, 2, raster_grid2, r->stream));
    }
  }
  if (!p.fuse_resolve) {
    StageTimer t(r, GBR_TIMING_RESOLVE);
    HIP_TRY(launch_resolve(p, r->d_ritems.ptr, r->n_ritems, r->stream));
  }
  if (r->collecting)
    HIP_TRY(launch_collect(p, r->vis.Xd.ptr, r->vis.Yd.ptr, r->vis.rows, r->vis.cap,
                           r->vis.positives.ptr, r->stream));
  return GBR_OK;
}

gbr_status gather_plane(gbr_renderer* r, uint32_t plane, const int32_t* rect, void* dst,
                        int64_t row_stride, int dst_bytes, int32_t memory) {
  if (!r || !rect || !dst) return fail(GBR_ERR_ARGUMENT, "image_into: null argument");
  TRY(set_device(r));
  const int fw = r->cfg.tile_size[0] * r->cfg.tile_count[0];
  const int fh = r->cfg.tile_size[1] * r->cfg.tile_count[1];
  if (rect[0] < 0 || rect[1] < 0 || rect[2] > fw || rect[3] > fh || rect[2] < rect[0] ||
      rect[3] < rect[1])
    return fail(GBR_ERR_SIZE, "image_into: rectangle {" + std::to_string(rect[0]) + ", " +
                                  std::to_string(rect[1]) + ", " + std::to_string(rect[2]) + ", " +
                                  std::to_string(rect[3]) + "} doesn't fit in a size of {" +
                                  std::to_string(fw) + ", " + std::to_string(fh) + "}");
  if (!(r->cfg.planes & plane))
    return fail(GBR_ERR_STATE, "image_into: the plane is not produced by this renderer (gbr_config.planes)");
  const int w = rect[2] - rect[0], h = rect[3] - rect[1];
  if (w == 0 || h == 0) return GBR_OK;
  const uint8_t* src = nullptr;
  uint32_t fill = 0;
  switch (plane) {
    case GBR_PLANE_COLOR: src = reinterpret_cast<const uint8_t*>(r->d_color.ptr); fill = 0xFF1F1F1Fu; break;
    case GBR_PLANE_DEPTH: src = reinterpret_cast<const uint8_t*>(r->d_depth.ptr); fill = 0x3F800000u; break;
    case GBR_PLANE_LINEAR_DEPTH: src = reinterpret_cast<const uint8_t*>(r->d_lin.ptr); break;
    default: src = reinterpret_cast<const uint8_t*>(r->d_sem.ptr); break;
  }
  if (!src) return fail(GBR_ERR_STATE, "image_into: nothing drawn yet");
  GatherArgs g{};
  g.plane = src;
  g.scenes = r->d_scenes.ptr;
  g.x0 = rect[0];
  g.y0 = rect[1];
  g.w = w;
  g.h = h;
  g.tile_w = r->cfg.tile_size[0];
  g.tile_h = r->cfg.tile_size[1];
  g.tiles_x = r->cfg.tile_count[0];
  g.scene_count = (int)r->scene_count();
  g.top_down = r->cfg.top_down ? 1 : 0;
  g.src_bytes = 4;
  g.dst_bytes = dst_bytes;
  g.fill = (int)fill;
  const int64_t tight = (int64_t)w * dst_bytes;
  if (row_stride == 0) row_stride = tight;
  if (row_stride < tight) return fail(GBR_ERR_SIZE, "image_into: row stride smaller than a row");
  if (memory == GBR_MEMORY_DEVICE) {
    g.dst = reinterpret_cast<uint8_t*>(dst);
    g.dst_stride = row_stride;
    HIP_TRY(launch_gather(g, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return GBR_OK;
  }
  TRY(r->d_stage.reserve((size_t)tight * h));
  g.dst = r->d_stage.ptr;
  g.dst_stride = tight;
  HIP_TRY(launch_gather(g, r->stream));
  HIP_TRY(hipMemcpy2DAsync(dst, row_stride, r->d_stage.ptr, tight, tight, h, hipMemcpyDeviceToHost,
                           r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return GBR_OK;
}

gbr_status framebuffer_pointer(gbr_renderer* r, uint32_t plane, DevBuf<uint8_t>& buf,
                               const void** out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "device_pointer: null argument");
  const int fw = r->cfg.tile_size[0] * r->cfg.tile_count[0];
  const int fh = r->cfg.tile_size[1] * r->cfg.tile_count[1];
  TRY(buf.reserve((size_t)fw * fh * 4));
  int32_t rect[4] = {0, 0, fw, fh};
  TRY(gather_plane(r, plane, rect, buf.ptr, (int64_t)fw * 4, 4, GBR_MEMORY_DEVICE));
  *out = buf.ptr;
  return GBR_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int32_t gbr_abi_version(void) { return GBR_ABI_VERSION; }

const char* gbr_last_error(void) { return g_error.c_str(); }

gbr_status gbr_config_init(gbr_config* c) {
  if (!c) return fail(GBR_ERR_ARGUMENT, "gbr_config_init: null");
  std::memset(c, 0, sizeof *c);
  c->tile_size[0] = c->tile_size[1] = 128;
  c->tile_count[0] = c->tile_count[1] = 1;
  c->ambient_factor = 0.1f;
  c->device = -1;
  c->planes = GBR_PLANE_ALL;
  return GBR_OK;
}

gbr_status gbr_tuning_init(gbr_tuning* t) {
  if (!t) return fail(GBR_ERR_ARGUMENT, "gbr_tuning_init: null");
  std::memset(t, 0, sizeof *t);
  t->bin_size = 32;
  t->chunk_triangles = kChunkTris;
  t->bin_capacity_factor = 2.0f;
  t->min_bin_capacity = 256;
  t->cull_mode = GBR_CULL_NONE;
  t->queue_grain = 4;
  t->lane_pixels = 64;
  t->graph = 1;
  t->heavy_entries = 256;
  return GBR_OK;
}

static gbr_status validate_tuning(const gbr_tuning* t) {
  if (t->bin_size != 8 && t->bin_size != 16 && t->bin_size != 32)
    return fail(GBR_ERR_ARGUMENT, "tuning: bin_size must be 8, 16 or 32");
  if (t->chunk_triangles < 1 || t->chunk_triangles > kChunkTris)
    return fail(GBR_ERR_ARGUMENT, "tuning: chunk_triangles must be in [1, 1024]");
  if (!(t->bin_capacity_factor > 0.0f) || t->min_bin_capacity < 1)
    return fail(GBR_ERR_ARGUMENT, "tuning: capacity must be positive");
  if (t->raster_mode < 0 || t->raster_mode > 1)
    return fail(GBR_ERR_ARGUMENT, "tuning: raster_mode must be 0 or 1");
  if (t->graph != 0 && t->graph != 1)
    return fail(GBR_ERR_ARGUMENT, "tuning: graph must be 0 or 1");
  if (t->lane_pixels < 1 || t->lane_pixels > kMaxBin * kMaxBin)
    return fail(GBR_ERR_ARGUMENT, "tuning: lane_pixels must be in [1, 1024]");
  if (t->timing != 0 && t->timing != 1)
    return fail(GBR_ERR_ARGUMENT, "tuning: timing must be 0 or 1");
  if (t->queue_grain < 1 || t->queue_grain > 64)
    return fail(GBR_ERR_ARGUMENT, "tuning: queue_grain must be in [1, 64]");
  if (t->cull_mode < GBR_CULL_NONE || t->cull_mode > GBR_CULL_LEARNED)
    return fail(GBR_ERR_ARGUMENT, "tuning: unknown cull_mode");
  if (!std::isfinite(t->proposal_threshold))
    return fail(GBR_ERR_ARGUMENT, "tuning: proposal_threshold must be finite");
  if (t->proposal_history != 0 && t->proposal_history != 1)
    return fail(GBR_ERR_ARGUMENT, "tuning: proposal_history must be 0 or 1");
  if (t->heavy_entries < 0)
    return fail(GBR_ERR_ARGUMENT, "tuning: heavy_entries must be >= 0");
  return GBR_OK;
}

gbr_status gbr_create(const gbr_config* config, gbr_renderer** out) {
  if (!config || !out) return fail(GBR_ERR_ARGUMENT, "gbr_create: null argument");
  *out = nullptr;
  if (config->tile_size[0] < 1 || config->tile_size[1] < 1 || config->tile_count[0] < 1 ||
      config->tile_count[1] < 1)
    return fail(GBR_ERR_ARGUMENT, "gbr_create: tile size and count must be positive");
  if (config->tile_size[0] > 8192 || config->tile_size[1] > 8192)
    return fail(GBR_ERR_ARGUMENT, "gbr_create: tile size above 8192 pixels (guard band)");
  if (config->max_light_count > GBR_MAX_LIGHTS)
    return fail(GBR_ERR_UNSUPPORTED, "gbr_create: max_light_count above " +
                                         std::to_string(GBR_MAX_LIGHTS));
  if (config->planes & ~GBR_PLANE_ALL) return fail(GBR_ERR_ARGUMENT, "gbr_create: unknown plane bits");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  int dev = config->device;
  if (dev < 0) HIP_TRY(hipGetDevice(&dev));
  if (dev >= ndev) return fail(GBR_ERR_ARGUMENT, "gbr_create: device ordinal out of range");
  HIP_TRY(hipSetDevice(dev));
  gbr_renderer* r = new (std::nothrow) gbr_renderer();
  if (!r) return fail(GBR_ERR_OOM, "gbr_create: out of host memory");
  r->cfg = *config;
  r->device = dev;
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      r->cu_count = cus;
  }
  gbr_tuning_init(&r->tun);
  const uint64_t n = (uint64_t)config->tile_count[0] * config->tile_count[1];
  r->scenes.resize(n);
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r->cam_event, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete r;
    return fail(GBR_ERR_DEVICE, std::string("gbr_create: ") + hipGetErrorString(e));
  }
  // texture 0 = white pixel, material 0 = white ambient (Renderer.cpp:287-304)
  TextureDesc white{};
  white.width = white.height = white.layers = white.levels = 1;
  white.min_filter = 9728;
  white.mag_filter = 9728;
  white.wrap_s = white.wrap_t = 10497;
  white.level_offset[0] = 0;
  r->textures.push_back(white);
  r->texels.insert(r->texels.end(), {255, 255, 255, 255});
  MaterialDesc m0{};
  for (int i = 0; i < 4; ++i) m0.ambient[i] = 1.0f;
  m0.texture = 0;
  m0.texmat[0] = m0.texmat[4] = m0.texmat[8] = 1.0f;
  r->materials.push_back(m0);
  *out = r;
  return GBR_OK;
}

gbr_status gbr_destroy(gbr_renderer* r) {
  if (!r) return GBR_OK;
  (void)hipSetDevice(r->device);
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  release_graph(r);
  r->d_frame.release();
  r->d_nodes_dirty.release();
  DevBuf<float>* fb[] = {&r->d_pos_x, &r->d_pos_y, &r->d_pos_z, &r->d_lut, &r->d_local, &r->d_abs,
                         &r->d_cameras, &r->d_unproj, &r->d_depth, &r->d_lin};
  for (auto* b : fb) b->release();
  DevBuf<uint32_t>* ub[] = {&r->d_uv16, &r->d_col8, &r->d_indices, &r->d_bin_count,
                            &r->d_bin_entries, &r->d_color, &r->d_sem, &r->d_heavy,
                            &r->d_bin_heavy, &r->d_active, &r->d_overflow};
  for (auto* b : ub) b->release();
  r->d_uv32.release();
  r->d_col32.release();
  r->d_texels.release();
  r->d_textures.release();
  r->d_materials.release();
  r->d_draws.release();
  r->d_rrecs.release();
  r->d_srecs.release();
  r->d_sub_prims.release();
  r->d_sub_recs.release();
  r->d_clipped.release();
  r->d_clip_fast.release();
  r->d_counters.release();
  r->d_vis.release();
  r->d_scenes.release();
  r->d_items.release();
  r->d_item_draws.release();
  r->d_ritems.release();
  r->d_parents.release();
  r->d_stats.release();
  r->d_caminfo.release();
  r->vis.release();
  r->d_fb_color.release();
  r->d_fb_depth.release();
  r->d_stage.release();
  r->d_hiz.release();
  r->d_nrm.release();
  r->d_nmat.release();
  r->d_lights.release();
  r->d_scene_lights.release();
  if (r->h_cam) (void)hipHostFree(r->h_cam);
  if (r->h_light) (void)hipHostFree(r->h_light);
  if (r->cam_event) (void)hipEventDestroy(r->cam_event);
  for (auto& set : r->t_pool)
    for (hipEvent_t e : set)
      if (e) (void)hipEventDestroy(e);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
  return GBR_OK;
}

gbr_status gbr_get_config(const gbr_renderer* r, gbr_config* out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "gbr_get_config: null argument");
  *out = r->cfg;
  out->device = r->device;  // the ordinal in use (resolved when -1 was given)
  return GBR_OK;
}

gbr_status gbr_scene_count(const gbr_renderer* r, uint64_t* out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "gbr_scene_count: null argument");
  *out = r->scenes.size();
  return GBR_OK;
}

// Renderer::addFile(), Renderer.cpp:354-867: the native importer parses the
// glTF composite, then the parsed file goes through gbr_add_file_data()
gbr_status gbr_add_file(gbr_renderer* r, const char* filename, uint32_t flags, const char* name) {
  if (!r || !filename) return fail(GBR_ERR_ARGUMENT, "gbr_add_file: null argument");
  imp::File f;
  std::string err;
  if (!imp::load_gltf(filename, f, err))
    return fail(GBR_ERR_IO, std::string("Renderer::addFile(): can't open the file ") + filename +
                                ": " + err);
  const gbr_file_desc d = f.desc();
  return gbr_add_file_data(r, &d, flags, name, filename);
}

// Renderer::addFile(), Renderer.cpp:354-867, on a parsed file
gbr_status gbr_add_file_data(gbr_renderer* r, const gbr_file_desc* f, uint32_t flags,
                             const char* name, const char* filename) {
  if (!r || !f) return fail(GBR_ERR_ARGUMENT, "gbr_add_file_data: null argument");
  if (flags & ~(GBR_FILE_WHOLE | GBR_FILE_GENERATE_MIPMAP))
    return fail(GBR_ERR_ARGUMENT, "gbr_add_file_data: unknown file flags");
  const std::string fname = filename ? filename : "";
  const uint32_t tex_off = (uint32_t)r->textures.size();
  const uint32_t mesh_off = (uint32_t)r->meshes.size();
  const uint32_t view_off = (uint32_t)r->views.size();
  const uint32_t mat_off = (uint32_t)r->materials.size();

  // validate references up front
  for (uint32_t i = 0; i < f->material_count; ++i)
    if (f->materials[i].texture >= (int32_t)f->texture_count)
      return fail(GBR_ERR_IO, "addFile(): material references a missing texture in " + fname);
  for (uint32_t i = 0; i < f->mesh_count; ++i) {
    const gbr_mesh_desc& m = f->meshes[i];
    if (!m.positions || (m.index_count && !m.indices) || m.index_count % 3)
      return fail(GBR_ERR_IO, "addFile(): can't import mesh " + std::to_string(i) + " of " + fname);
    for (uint32_t k = 0; k < m.index_count; ++k)
      if (m.indices[k] >= m.vertex_count)
        return fail(GBR_ERR_IO, "addFile(): mesh " + std::to_string(i) + " index out of range in " + fname);
  }

  // textures (Renderer.cpp:417-534)
  if (!(r->cfg.flags & GBR_FLAG_NO_TEXTURES)) {
    for (uint32_t i = 0; i < f->texture_count; ++i) {
      const gbr_texture_desc& t = f->textures[i];
      if (!t.texels || t.width < 1 || t.height < 1 || t.layers < 1 || t.levels < 1 || t.levels > 16)
        return fail(GBR_ERR_IO, "addFile(): can't import texture " + std::to_string(i) + " of " + fname);
      // the kernels index a level's texels with 32 bits (rows with 24-bit products)
      if (t.width >= (1u << 24) || t.height >= (1u << 24))
        return fail(GBR_ERR_SIZE, "addFile(): texture " + std::to_string(i) + " of " + fname +
                                      " is 2^24 or more texels wide or high");
      if ((uint64_t)t.width * t.height * t.layers >= (1ull << 32))
        return fail(GBR_ERR_SIZE, "addFile(): texture " + std::to_string(i) + " of " + fname +
                                      " has 2^32 or more texels per level");
      TextureDesc td{};
      td.width = (int)t.width;
      td.height = (int)t.height;
      td.layers = (int)t.layers;
      td.min_filter = (int)t.min_filter;
      td.mag_filter = (int)t.mag_filter;
      td.wrap_s = (int)t.wrap_s;
      td.wrap_t = (int)t.wrap_t;
      const bool gen = t.levels == 1 && (flags & GBR_FILE_GENERATE_MIPMAP) && !t.compressed;
      if (gen) {
        int nl = 1;
        auto lv = generate_mips(t.texels, td.width, td.height, td.layers, nl);
        td.levels = nl;
        for (int l = 0; l < nl; ++l) {
          td.level_offset[l] = (int64_t)r->texels.size();
          r->texels.insert(r->texels.end(), lv[l].begin(), lv[l].end());
        }
      } else {
        td.levels = (int)t.levels;
        size_t off = 0;
        for (int l = 0; l < td.levels; ++l) {
          const size_t sz = (size_t)std::max(td.width >> l, 1) * std::max(td.height >> l, 1) * td.layers * 4;
          td.level_offset[l] = (int64_t)r->texels.size();
          r->texels.insert(r->texels.end(), t.texels + off, t.texels + off + sz);
          off += sz;
        }
      }
      r->textures.push_back(td);
    }
  }

  // meshes (Renderer.cpp:536-557)
  for (uint32_t i = 0; i < f->mesh_count; ++i) {
    const gbr_mesh_desc& m = f->meshes[i];
    MeshInfo mi{};
    mi.vbase = (uint32_t)r->pos_x.size();
    mi.vcount = m.vertex_count;
    mi.ibase = (uint32_t)r->indices.size();
    mi.icount = m.index_count;
    mi.uv_half = m.texcoords_half ? 1u : 0u;
    mi.color_fmt = m.colors ? (m.colors_u8 ? kColorU8 : kColorF32) : kColorNone;
    for (uint32_t v = 0; v < m.vertex_count; ++v) {
      r->pos_x.push_back(m.positions[3 * v]);
      r->pos_y.push_back(m.positions[3 * v + 1]);
      r->pos_z.push_back(m.positions[3 * v + 2]);
      if (r->cfg.max_light_count) {  // meshes without normals shade with +Z
        for (int k = 0; k < 3; ++k)
          r->nrm.push_back(m.normals ? m.normals[3 * v + k] : (k == 2 ? 1.0f : 0.0f));
      }
      float u = m.texcoords ? m.texcoords[2 * v] : 0.0f;
      float vv = m.texcoords ? m.texcoords[2 * v + 1] : 0.0f;
      r->uv32.push_back(make_float2(u, vv));
      __half hu = __float2half(u), hv = __float2half(vv);
      r->uv16.push_back((uint32_t)__half_as_ushort(hu) | ((uint32_t)__half_as_ushort(hv) << 16));
      if (mi.color_fmt == kColorU8) {
        const uint8_t* c = static_cast<const uint8_t*>(m.colors) + 4 * v;
        r->col8.push_back((uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) |
                          ((uint32_t)c[3] << 24));
        r->col32.push_back(make_float4(0, 0, 0, 0));
      } else if (mi.color_fmt == kColorF32) {
        const float* c = static_cast<const float*>(m.colors) + 4 * v;
        r->col8.push_back(0);
        r->col32.push_back(make_float4(c[0], c[1], c[2], c[3]));
      } else {
        r->col8.push_back(0);
        r->col32.push_back(make_float4(1, 1, 1, 1));
      }
    }
    for (uint32_t k = 0; k < m.index_count; ++k) r->indices.push_back(mi.vbase + m.indices[k]);
    r->meshes.push_back(mi);
  }

  // materials (Renderer.cpp:559-626): flat if marked flat or no lights
  for (uint32_t i = 0; i < f->material_count; ++i) {
    const gbr_material_desc& md = f->materials[i];
    MaterialDesc m{};
    for (int c = 0; c < 4; ++c) {
      float b = md.base_color[c];
      const bool shaded = !md.flat && r->cfg.max_light_count;
      m.ambient[c] = shaded ? b * r->cfg.ambient_factor : b;
      m.diffuse[c] = shaded ? b : 0.0f;
    }
    m.lit = (m.diffuse[0] != 0.0f || m.diffuse[1] != 0.0f || m.diffuse[2] != 0.0f ||
             m.diffuse[3] != 0.0f) ? 1u : 0u;
    if (md.texture < 0) {
      m.texture = 0;
      m.layer = 0;
      m.texmat[0] = m.texmat[4] = m.texmat[8] = 1.0f;
    } else {
      m.texture = (int32_t)(md.texture + tex_off);
      m.layer = md.layer;
      std::memcpy(m.texmat, md.texture_matrix, sizeof m.texmat);
    }
    // with NoTextures nothing is imported: the id only keys draw batches and
    // the kernels never sample (Renderer.cpp:417-418, 842-849)
    r->materials.push_back(m);
  }
  r->files_added = true;
  r->assets_dirty = true;
  // cameras are re-created by addFile(), Renderer.cpp:630-631
  for (Scene& sc : r->scenes) sc.camera = identity4();

  if (!f->has_scene) {
    if (f->mesh_count != 1 || f->material_count)
      return fail(GBR_ERR_IO, "addFile(): expected exactly one mesh and no material for a scene-less file " + fname);
    if (!(flags & GBR_FILE_WHOLE))
      return fail(GBR_ERR_ARGUMENT, "addFile(): scene-less file " + fname + " has to be added with GBR_FILE_WHOLE");
    MeshView v;
    v.mesh = mesh_off;
    v.index_offset = 0;
    v.index_count = f->meshes[0].index_count;
    v.material = 0;
    v.xf = identity4();
    r->views.push_back(std::move(v));
    const std::string used = (name && *name) ? name : fname;
    if (!r->names.emplace(used, std::make_pair(view_off, view_off + 1)).second)
      return fail(GBR_ERR_STATE, "addFile(): node name " + used + " in " + fname + " already exists");
    return GBR_OK;
  }

  bool has_views = false;
  for (uint32_t i = 0; i < f->mesh_entry_count; ++i) has_views |= f->mesh_entries[i].has_view != 0;
  std::vector<MeshView> nv;
  for (uint32_t i = 0; i < f->mesh_entry_count; ++i) {
    const gbr_mesh_entry_desc& e = f->mesh_entries[i];
    if (e.mesh >= f->mesh_count || e.node >= f->node_count)
      return fail(GBR_ERR_IO, "addFile(): mesh field out of range in " + fname);
    const gbr_mesh_desc& m = f->meshes[e.mesh];
    MeshView v;
    v.mesh = e.mesh + mesh_off;
    int32_t mat;
    if (has_views) {
      const uint32_t isz = m.source_index_size ? m.source_index_size : 4;
      v.index_offset = e.view_index_offset_bytes / isz;
      v.index_count = e.view_index_count;
      mat = e.view_material;
    } else {
      v.index_offset = 0;
      v.index_count = m.index_count;
      mat = e.material;
    }
    if (v.index_offset + v.index_count > m.index_count || v.index_offset % 3 || v.index_count % 3)
      return fail(GBR_ERR_IO, "addFile(): mesh view out of range in " + fname);
    if (mat >= (int32_t)f->material_count)
      return fail(GBR_ERR_IO, "addFile(): mesh view material out of range in " + fname);
    v.material = mat == -1 ? 0 : (uint32_t)mat + mat_off;
    nv.push_back(std::move(v));
  }
  if (!(flags & GBR_FILE_WHOLE)) {
    for (uint32_t i = 0; i < f->mesh_entry_count; ++i)
      nv[i].xf = load4(f->nodes[f->mesh_entries[i].node].transform);
    for (auto& v : nv) r->views.push_back(std::move(v));
    uint32_t offset = view_off;
    for (uint32_t root = 0; root < f->node_count; ++root) {
      if (f->nodes[root].parent != -1) continue;
      uint32_t children = 0;
      for (uint32_t c = 0; c < f->node_count; ++c) children += f->nodes[c].parent == (int32_t)root;
      const char* nm = f->nodes[root].name;
      if (!nm || !*nm)
        return fail(GBR_ERR_IO, "addFile(): node " + std::to_string(root) + " in " + fname + " has no name");
      if (!r->names.emplace(nm, std::make_pair(offset, offset + children)).second)
        return fail(GBR_ERR_STATE, "addFile(): node name " + std::string(nm) + " in " + fname + " already exists");
      offset += children;
    }
    if (offset != r->views.size())
      return fail(GBR_ERR_IO, "addFile(): mesh views don't match the template children in " + fname);
  } else {
    std::vector<int> state(f->node_count, 0);
    std::vector<Mat4> absm(f->node_count);
    for (uint32_t i = 0; i < f->node_count; ++i) {
      // parents precede children in a glTF-ordered file only by convention:
      // resolve recursively with an explicit stack
      std::vector<uint32_t> stack{i};
      while (!stack.empty()) {
        uint32_t n = stack.back();
        if (state[n] == 2) { stack.pop_back(); continue; }
        int32_t p = f->nodes[n].parent;
        if (p >= (int32_t)f->node_count) return fail(GBR_ERR_IO, "addFile(): bad parent in " + fname);
        if (p >= 0 && state[p] != 2) {
          if (state[p] == 1) return fail(GBR_ERR_IO, "addFile(): cyclic hierarchy in " + fname);
          state[n] = 1;
          stack.push_back((uint32_t)p);
          continue;
        }
        absm[n] = mul4(p >= 0 ? absm[p] : identity4(), load4(f->nodes[n].transform));
        state[n] = 2;
        stack.pop_back();
      }
    }
    for (uint32_t i = 0; i < f->mesh_entry_count; ++i) nv[i].xf = absm[f->mesh_entries[i].node];
    const uint32_t cnt = (uint32_t)nv.size();
    for (auto& v : nv) r->views.push_back(std::move(v));
    const std::string used = (name && *name) ? name : fname;
    if (!r->names.emplace(used, std::make_pair(view_off, view_off + cnt)).second)
      return fail(GBR_ERR_STATE, "addFile(): node name " + used + " already exists");
  }
  return GBR_OK;
}

gbr_status gbr_has_node_hierarchy(const gbr_renderer* r, const char* name, int32_t* out) {
  if (!r || !name || !out) return fail(GBR_ERR_ARGUMENT, "gbr_has_node_hierarchy: null argument");
  *out = r->names.count(name) ? 1 : 0;
  return GBR_OK;
}

// Renderer::addNodeHierarchy(), Renderer.cpp:880-961
gbr_status gbr_add_node_hierarchy_semantic(gbr_renderer* r, uint32_t scene, const char* name,
                                           const float* bake, uint32_t semantic,
                                           uint64_t* out_node) {
  TRY(check_scene(r, scene, "addNodeHierarchy()"));
  if (!name) return fail(GBR_ERR_ARGUMENT, "addNodeHierarchy(): null name");
  auto found = r->names.find(name);
  if (found == r->names.end())
    return fail(GBR_ERR_NOT_FOUND, std::string("Renderer::add(): name ") + name + " not found");
  Scene& sc = r->scenes[scene];
  const Mat4 bk = bake ? load4(bake) : identity4();
  const uint32_t top = (uint32_t)sc.transforms.size();
  sc.parents.push_back(-1);
  sc.transforms.push_back(identity4());
  for (uint32_t i = found->second.first; i != found->second.second; ++i) {
    const MeshView& v = r->views[i];
    const uint32_t id = (uint32_t)sc.transforms.size();
    sc.parents.push_back((int32_t)top);
    sc.transforms.push_back(mul4(bk, v.xf));
    const MaterialDesc& m = r->materials[v.material];
    const uint32_t b = batch_id(sc, r->meshes[v.mesh].color_fmt != kColorNone ? 1u : 0u, v.mesh,
                                (uint32_t)m.texture);
    sc.draws.push_back(Draw{b, id, v.material, i, semantic});
  }
  sc.xform_dirty = true;
  r->structure_dirty = true;
  if (out_node) *out_node = top;
  return GBR_OK;
}

gbr_status gbr_add_node_hierarchy(gbr_renderer* r, uint32_t scene, const char* name,
                                  const float* bake, uint64_t* out_node) {
  return gbr_add_node_hierarchy_semantic(r, scene, name, bake, 0u, out_node);
}

gbr_status gbr_add_node_hierarchies(gbr_renderer* r, uint32_t scene, uint32_t count,
                                    const char* const* names, const float* bakes,
                                    const uint32_t* semantics, uint64_t* out_nodes) {
  TRY(check_scene(r, scene, "addNodeHierarchy()"));
  if (count && !names) return fail(GBR_ERR_ARGUMENT, "addNodeHierarchy(): null names");
  for (uint32_t i = 0; i < count; ++i) {
    uint64_t node = 0;
    TRY(gbr_add_node_hierarchy_semantic(r, scene, names[i], bakes ? bakes + 16 * (size_t)i : nullptr,
                                        semantics ? semantics[i] : 0u, &node));
    if (out_nodes) out_nodes[i] = node;
  }
  return GBR_OK;
}

gbr_status gbr_set_node_hierarchy_semantic(gbr_renderer* r, uint32_t scene, uint64_t node,
                                           uint32_t semantic) {
  TRY(check_scene(r, scene, "setNodeHierarchySemantic()"));
  Scene& sc = r->scenes[scene];
  if (node >= sc.parents.size())
    return fail(GBR_ERR_RANGE, "setNodeHierarchySemantic(): node " + std::to_string(node) +
                                   " out of range for " + std::to_string(sc.parents.size()) +
                                   " nodes");
  // the hierarchy's draws sit on the top-level node's children
  for (Draw& d : sc.draws)
    if (d.node == node || sc.parents[d.node] == (int32_t)node) d.semantic = semantic;
  r->structure_dirty = true;
  return GBR_OK;
}

gbr_status gbr_add_empty_node(gbr_renderer* r, uint32_t scene, uint64_t* out_node) {
  TRY(check_scene(r, scene, "addEmptyNode()"));
  Scene& sc = r->scenes[scene];
  sc.parents.push_back(-1);
  sc.transforms.push_back(identity4());
  sc.xform_dirty = true;
  r->structure_dirty = true;  // node arrays grew
  if (out_node) *out_node = sc.transforms.size() - 1;
  return GBR_OK;
}

gbr_status gbr_add_light(gbr_renderer* r, uint32_t scene, uint64_t node, int32_t type,
                         uint64_t* out) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "addLight(): null renderer");
  if (!r->cfg.max_light_count) return fail(GBR_ERR_STATE, "Renderer::addLight(): max light count is zero");
  TRY(check_scene(r, scene, "addLight()"));
  Scene& sc = r->scenes[scene];
  if (node >= sc.parents.size())
    return fail(GBR_ERR_RANGE, "Renderer::addLight(): index " + std::to_string(node) +
                                   " out of range for " + std::to_string(sc.parents.size()) +
                                   " nodes in scene " + std::to_string(scene));
  if (type != GBR_LIGHT_DIRECTIONAL && type != GBR_LIGHT_POINT)
    return fail(GBR_ERR_ARGUMENT, "Renderer::addLight(): unknown light type " + std::to_string(type));
  Light L{};
  L.node = node;
  L.type = type;
  L.color[0] = L.color[1] = L.color[2] = 1.0f;  // 0xffffff
  L.range = std::numeric_limits<float>::infinity();
  sc.lights.push_back(L);
  if (out) *out = sc.lights.size() - 1;
  return GBR_OK;
}

// lightColors() / lightRanges(), Renderer.cpp:1101-1121: strided views into
// the scene's light records
gbr_status gbr_light_colors(gbr_renderer* r, uint32_t scene, float** out_data, uint64_t* out_count,
                            uint64_t* out_stride) {
  TRY(check_scene(r, scene, "Renderer::lightColors()"));
  if (!out_data || !out_count || !out_stride) return fail(GBR_ERR_ARGUMENT, "lightColors(): null output");
  Scene& sc = r->scenes[scene];
  *out_data = sc.lights.empty() ? nullptr : sc.lights.data()->color;
  *out_count = sc.lights.size();
  *out_stride = sizeof(Light);
  return GBR_OK;
}

gbr_status gbr_light_ranges(gbr_renderer* r, uint32_t scene, float** out_data, uint64_t* out_count,
                            uint64_t* out_stride) {
  TRY(check_scene(r, scene, "Renderer::lightRanges()"));
  if (!out_data || !out_count || !out_stride) return fail(GBR_ERR_ARGUMENT, "lightRanges(): null output");
  Scene& sc = r->scenes[scene];
  *out_data = sc.lights.empty() ? nullptr : &sc.lights.data()->range;
  *out_count = sc.lights.size();
  *out_stride = sizeof(Light);
  return GBR_OK;
}

// Renderer::clear(), Renderer.cpp:1017-1040 (camera kept)
gbr_status gbr_clear(gbr_renderer* r, uint32_t scene) {
  TRY(check_scene(r, scene, "Renderer::clear()"));
  Scene& sc = r->scenes[scene];
  sc.parents.clear();
  sc.transforms.clear();
  sc.lights.clear();
  sc.draws.clear();
  sc.batches.clear();
  r->structure_dirty = true;
  return GBR_OK;
}

gbr_status gbr_set_scene_viewport(gbr_renderer* r, uint32_t scene, int32_t width, int32_t height) {
  TRY(check_scene(r, scene, "set_scene_viewport"));
  if (width < 1 || height < 1 || width > r->cfg.tile_size[0] || height > r->cfg.tile_size[1])
    return fail(GBR_ERR_SIZE, "set_scene_viewport: " + std::to_string(width) + "x" +
                                  std::to_string(height) + " doesn't fit a tile of " +
                                  std::to_string(r->cfg.tile_size[0]) + "x" +
                                  std::to_string(r->cfg.tile_size[1]));
  Scene& sc = r->scenes[scene];
  const int32_t w = width == r->cfg.tile_size[0] && height == r->cfg.tile_size[1] ? 0 : width;
  const int32_t h = w ? height : 0;
  if (sc.vp_w != w || sc.vp_h != h) r->structure_dirty = true;
  sc.vp_w = w;
  sc.vp_h = h;
  return GBR_OK;
}

gbr_status gbr_scene_viewport(const gbr_renderer* r, uint32_t scene, int32_t* out2) {
  TRY(check_scene(r, scene, "scene_viewport"));
  if (!out2) return fail(GBR_ERR_ARGUMENT, "scene_viewport: null output");
  const Scene& sc = r->scenes[scene];
  out2[0] = sc.vp_w ? sc.vp_w : r->cfg.tile_size[0];
  out2[1] = sc.vp_h ? sc.vp_h : r->cfg.tile_size[1];
  return GBR_OK;
}

gbr_status gbr_set_scene_planes(gbr_renderer* r, uint32_t scene, uint32_t planes) {
  TRY(check_scene(r, scene, "set_scene_planes"));
  if (planes & ~r->cfg.planes)
    return fail(GBR_ERR_ARGUMENT, "set_scene_planes: planes outside gbr_config.planes");
  if (r->scenes[scene].planes != planes) r->structure_dirty = true;
  r->scenes[scene].planes = planes;
  return GBR_OK;
}

gbr_status gbr_clear_lights(gbr_renderer* r, uint32_t scene) {
  TRY(check_scene(r, scene, "Renderer::clearLights()"));
  r->scenes[scene].lights.clear();
  return GBR_OK;
}

gbr_status gbr_camera(const gbr_renderer* r, uint32_t scene, float* out16) {
  TRY(check_scene(r, scene, "Renderer::camera()"));
  if (!out16) return fail(GBR_ERR_ARGUMENT, "camera(): null output");
  std::memcpy(out16, r->scenes[scene].camera.m, 64);
  return GBR_OK;
}

gbr_status gbr_camera_depth_unprojection(const gbr_renderer* r, uint32_t scene, float* out2) {
  TRY(check_scene(r, scene, "Renderer::cameraDepthUnprojection()"));
  if (!out2) return fail(GBR_ERR_ARGUMENT, "cameraDepthUnprojection(): null output");
  out2[0] = r->scenes[scene].unproj[0];
  out2[1] = r->scenes[scene].unproj[1];
  return GBR_OK;
}

gbr_status gbr_calculate_depth_unprojection(const float* P, float* out2) {
  if (!P || !out2) return fail(GBR_ERR_ARGUMENT, "calculateDepthUnprojection(): null argument");
  float a = P[2 * 4 + 2] - 1.0f;
  float b = P[3 * 4 + 2];
  out2[0] = a * 0.5f;
  out2[1] = b * 0.5f;
  return GBR_OK;
}

// Renderer::updateCamera(), Renderer.cpp:1077-1088
gbr_status gbr_update_camera(gbr_renderer* r, uint32_t scene, const float* projection16,
                             const float* view16) {
  TRY(check_scene(r, scene, "Renderer::updateCamera()"));
  if (!projection16 || !view16) return fail(GBR_ERR_ARGUMENT, "updateCamera(): null matrix");
  Scene& sc = r->scenes[scene];
  sc.camera = mul4(load4(projection16), load4(view16));
  gbr_calculate_depth_unprojection(projection16, sc.unproj);
  // camera position -R^T t and forward -R^T e_z of a rigid view (features of
  // the visibility model)
  const float* v = view16;
  for (int i = 0; i < 3; ++i) {
    sc.cam_info[i] = -(v[i * 4 + 0] * v[12] + v[i * 4 + 1] * v[13] + v[i * 4 + 2] * v[14]);
    sc.cam_info[3 + i] = -v[i * 4 + 2];
  }
  return GBR_OK;
}

gbr_status gbr_update_cameras(gbr_renderer* r, uint32_t first, uint32_t count,
                              const float* projections16, const float* views16) {
  if (!r || !projections16 || !views16) return fail(GBR_ERR_ARGUMENT, "updateCameras(): null argument");
  if ((uint64_t)first + count > r->scenes.size())
    return fail(GBR_ERR_RANGE, "updateCameras(): scene range out of range");
  for (uint32_t i = 0; i < count; ++i)
    TRY(gbr_update_camera(r, first + i, projections16 + 16 * (size_t)i, views16 + 16 * (size_t)i));
  return GBR_OK;
}

gbr_status gbr_transformations(gbr_renderer* r, uint32_t scene, float** out_data,
                               uint64_t* out_count) {
  TRY(check_scene(r, scene, "Renderer::transformations()"));
  if (!out_data || !out_count) return fail(GBR_ERR_ARGUMENT, "transformations(): null output");
  Scene& sc = r->scenes[scene];
  *out_data = sc.transforms.empty() ? nullptr : sc.transforms.data()->m;
  *out_count = sc.transforms.size();
  sc.xform_dirty = true;
  return GBR_OK;
}

gbr_status gbr_mark_transformations_dirty(gbr_renderer* r, uint32_t scene) {
  TRY(check_scene(r, scene, "markTransformationsDirty()"));
  r->scenes[scene].xform_dirty = true;
  return GBR_OK;
}

gbr_status gbr_scene_stats(const gbr_renderer* r, uint32_t scene, gbr_scene_stats_info* out) {
  TRY(check_scene(r, scene, "Renderer::sceneStats()"));
  if (!out) return fail(GBR_ERR_ARGUMENT, "sceneStats(): null output");
  const Scene& sc = r->scenes[scene];
  out->node_count = sc.transforms.size();
  out->draw_count = sc.draws.size();
  out->draw_batch_count = sc.batches.size();
  return GBR_OK;
}

gbr_status gbr_draw(gbr_renderer* r) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "draw(): null renderer");
  return do_draw(r);
}

gbr_status gbr_synchronize(gbr_renderer* r) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "synchronize(): null renderer");
  TRY(set_device(r));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return GBR_OK;
}

gbr_status gbr_stream(gbr_renderer* r, void** out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "stream(): null argument");
  *out = r->stream;
  return GBR_OK;
}

gbr_status gbr_color_image_into(gbr_renderer* r, const int32_t* rect, void* dst, int64_t stride,
                                int32_t format, int32_t memory) {
  if (format != GBR_FORMAT_RGBA8 && format != GBR_FORMAT_RGB8)
    return fail(GBR_ERR_ARGUMENT, "colorImageInto(): format must be RGBA8 or RGB8");
  return gather_plane(r, GBR_PLANE_COLOR, rect, dst, stride, format == GBR_FORMAT_RGB8 ? 3 : 4, memory);
}

gbr_status gbr_depth_image_into(gbr_renderer* r, const int32_t* rect, float* dst, int64_t stride,
                                int32_t memory) {
  return gather_plane(r, GBR_PLANE_DEPTH, rect, dst, stride, 4, memory);
}

gbr_status gbr_linear_depth_image_into(gbr_renderer* r, const int32_t* rect, float* dst,
                                       int64_t stride, int32_t memory) {
  return gather_plane(r, GBR_PLANE_LINEAR_DEPTH, rect, dst, stride, 4, memory);
}

gbr_status gbr_semantic_image_into(gbr_renderer* r, const int32_t* rect, uint32_t* dst,
                                   int64_t stride, int32_t memory) {
  return gather_plane(r, GBR_PLANE_SEMANTIC, rect, dst, stride, 4, memory);
}

gbr_status gbr_color_device_pointer(gbr_renderer* r, const void** out) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "color_device_pointer: null renderer");
  return framebuffer_pointer(r, GBR_PLANE_COLOR, r->d_fb_color, out);
}

gbr_status gbr_depth_device_pointer(gbr_renderer* r, const void** out) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "depth_device_pointer: null renderer");
  return framebuffer_pointer(r, GBR_PLANE_DEPTH, r->d_fb_depth, out);
}

gbr_status gbr_plane_device_pointer(gbr_renderer* r, uint32_t plane, const void** out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "plane_device_pointer: null argument");
  TRY(ensure_planes(r));
  if (!(r->cfg.planes & plane)) return fail(GBR_ERR_STATE, "plane_device_pointer: plane not produced");
  switch (plane) {
    case GBR_PLANE_COLOR: *out = r->d_color.ptr; break;
    case GBR_PLANE_DEPTH: *out = r->d_depth.ptr; break;
    case GBR_PLANE_LINEAR_DEPTH: *out = r->d_lin.ptr; break;
    case GBR_PLANE_SEMANTIC: *out = r->d_sem.ptr; break;
    default: return fail(GBR_ERR_ARGUMENT, "plane_device_pointer: unknown plane");
  }
  return GBR_OK;
}

// DepthUnprojection.cpp:107-130
gbr_status gbr_unproject_depth(const float* un, float* depth, int64_t rows, int64_t cols,
                               int64_t row_stride) {
  if (!un || (!depth && rows * cols)) return fail(GBR_ERR_ARGUMENT, "unprojectDepth(): null argument");
  if (row_stride == 0) row_stride = cols * 4;
  const float far_depth = un[1] / (1.0f + un[0]);
  for (int64_t y = 0; y < rows; ++y) {
    float* row = reinterpret_cast<float*>(reinterpret_cast<char*>(depth) + y * row_stride);
    for (int64_t x = 0; x < cols; ++x) {
      float d = un[1] / (row[x] + un[0]);
      row[x] = d == far_depth ? 0.0f : d;
    }
  }
  return GBR_OK;
}

gbr_status gbr_get_tuning(const gbr_renderer* r, gbr_tuning* out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "get_tuning: null argument");
  *out = r->tun;
  return GBR_OK;
}

gbr_status gbr_set_tuning(gbr_renderer* r, const gbr_tuning* t) {
  if (!r || !t) return fail(GBR_ERR_ARGUMENT, "set_tuning: null argument");
  TRY(validate_tuning(t));
  if (t->cull_mode == GBR_CULL_LEARNED && !r->vis.ready)
    return fail(GBR_ERR_STATE, "tuning: GBR_CULL_LEARNED needs visibility-model weights "
                               "(gbr_vis_init / gbr_vis_set_weights)");
  // the frame layout depends on these; culling and launch shapes do not
  if (t->bin_size != r->tun.bin_size || t->chunk_triangles != r->tun.chunk_triangles ||
      t->bin_capacity_factor != r->tun.bin_capacity_factor ||
      t->min_bin_capacity != r->tun.min_bin_capacity)
    r->structure_dirty = true;
  r->tun = *t;
  return GBR_OK;
}

gbr_status gbr_autotune_params_init(gbr_autotune_params* params) {
  if (!params) return fail(GBR_ERR_ARGUMENT, "autotune_params_init: null");
  params->warmup = 2;
  params->frames = 4;
  params->passes = 1;
  return GBR_OK;
}

namespace {

// GPU milliseconds per draw() under tuning t (warmup draws first)
gbr_status time_tuning(gbr_renderer* r, const gbr_tuning& t, const gbr_autotune_params& ap,
                       float* ms) {
  TRY(gbr_set_tuning(r, &t));
  for (int i = 0; i < ap.warmup; ++i) TRY(do_draw(r));
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  hipError_t e = hipEventCreate(&b);
  if (e == hipSuccess) e = hipEventRecord(a, r->stream);
  gbr_status st = GBR_OK;
  for (int i = 0; e == hipSuccess && st == GBR_OK && i < ap.frames; ++i) st = do_draw(r);
  if (e == hipSuccess && st == GBR_OK) e = hipEventRecord(b, r->stream);
  if (e == hipSuccess && st == GBR_OK) e = hipEventSynchronize(b);
  float total = 0.0f;
  if (e == hipSuccess && st == GBR_OK) e = hipEventElapsedTime(&total, a, b);
  (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (st != GBR_OK) return st;
  HIP_TRY(e);
  *ms = total / (float)ap.frames;
  return GBR_OK;
}

}  // namespace

gbr_status gbr_autotune(gbr_renderer* r, const gbr_autotune_params* params,
                        gbr_autotune_report* report) {
  if (!r || !params || !report) return fail(GBR_ERR_ARGUMENT, "autotune: null argument");
  if (params->warmup < 1 || params->frames < 1 || params->passes < 1)
    return fail(GBR_ERR_ARGUMENT, "autotune: warmup, frames and passes must be >= 1");
  TRY(set_device(r));
  std::memset(report, 0, sizeof *report);
  gbr_tuning best = r->tun;
  float best_ms = 0.0f;
  TRY(time_tuning(r, best, *params, &best_ms));
  report->initial_ms = best_ms;
  report->candidates = 1;
  // candidate values per parameter
  const int bins[] = {8, 16, 32};
  const float caps[] = {1.0f, 1.5f, 2.0f, 3.0f};
  const int lanes[] = {32, 64, 96, 128, 192};
  const int grains[] = {1, 2, 4, 8, 16};
  const int modes[] = {0, 1};
  const int heavies[] = {0, 64, 128, 256, 512};
  const float thresholds[] = {-2.0f, -1.0f, 0.0f, 1.0f};
  // the proposal threshold only matters for the learned proposal
  const int nparams = best.cull_mode == GBR_CULL_LEARNED ? 7 : 6;
  for (int pass = 0; pass < params->passes; ++pass) {
    for (int param = 0; param < nparams; ++param) {
      const int n = param == 0 ? 3 : param == 1 ? 4 : param == 2 ? 5 : param == 3 ? 5 : param == 4 ? 2
                    : param == 5 ? 5 : 4;
      for (int k = 0; k < n; ++k) {
        gbr_tuning t = best;
        switch (param) {
          case 0: t.bin_size = bins[k]; break;
          case 1: t.bin_capacity_factor = caps[k]; break;
          case 2: t.lane_pixels = lanes[k]; break;
          case 3: t.queue_grain = grains[k]; break;
          case 4: t.raster_mode = modes[k]; break;
          case 5: t.heavy_entries = heavies[k]; break;
          default: t.proposal_threshold = thresholds[k]; break;
        }
        if (std::memcmp(&t, &best, sizeof t) == 0) continue;
        float ms = 0.0f;
        TRY(time_tuning(r, t, *params, &ms));
        ++report->candidates;
        // a candidate has to beat the best by more than the frame-to-frame
        // noise (~1%) to replace it
        if (ms < best_ms * 0.99f) {
          best_ms = ms;
          best = t;
        }
      }
    }
  }
  TRY(gbr_set_tuning(r, &best));
  report->best_ms = best_ms;
  report->best = best;
  return GBR_OK;
}

// Visible work items of the last frame (those owning a pixel) and how many
// of them phase 1 set up: the proposal's recall is visible_phase1 / visible.
gbr_status gbr_frame_visibility(gbr_renderer* r, uint64_t* visible, uint64_t* visible_phase1) {
  if (!r || !visible || !visible_phase1) return fail(GBR_ERR_ARGUMENT, "frame_visibility: null argument");
  *visible = *visible_phase1 = 0;
  if (!r->d_frame.ptr || !r->d_item_vis.ptr || !r->n_items) return GBR_OK;
  TRY(set_device(r));
  uint32_t frame = 0;
  std::vector<uint32_t> vis(r->n_items);
  HIP_TRY(hipMemcpyAsync(&frame, r->d_frame.ptr, 4, hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipMemcpyAsync(vis.data(), r->d_item_vis.ptr, vis.size() * 4, hipMemcpyDeviceToHost,
                         r->stream));
  std::vector<uint32_t> list;
  if (r->tun.cull_mode != GBR_CULL_NONE && r->d_counters.ptr && r->d_list1.ptr) {
    uint32_t n1 = 0;
    HIP_TRY(hipMemcpyAsync(&n1, r->d_counters.ptr + kCtrList1 * kCtrStride, 4, hipMemcpyDeviceToHost,
                           r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    list.resize(std::min(n1, r->n_items));
    if (!list.empty())
      HIP_TRY(hipMemcpyAsync(list.data(), r->d_list1.ptr, list.size() * 4, hipMemcpyDeviceToHost,
                             r->stream));
  }
  HIP_TRY(hipStreamSynchronize(r->stream));
  uint64_t nv = 0;
  for (uint32_t v : vis) nv += v == frame;
  *visible = nv;
  if (r->tun.cull_mode == GBR_CULL_NONE) {
    *visible_phase1 = nv;
  } else {
    uint64_t n = 0;
    for (uint32_t it : list) n += it < r->n_items && vis[it] == frame;
    *visible_phase1 = n;
  }
  return GBR_OK;
}

gbr_status gbr_frame_stats(gbr_renderer* r, gbr_frame_stats_info* out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "frame_stats: null argument");
  std::memset(out, 0, sizeof *out);
  out->work_items = r->n_items;
  out->triangles = r->n_tris;
  out->raster_items = r->n_ritems;
  if (r->d_stats.ptr) {
    unsigned long long striped[kStatCount * kStatStripes] = {};
    unsigned long long s[kStatCount] = {};
    TRY(set_device(r));
    HIP_TRY(hipMemcpyAsync(striped, r->d_stats.ptr, sizeof striped, hipMemcpyDeviceToHost,
                           r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    for (uint32_t k = 0; k < kStatCount; ++k)
      for (uint32_t j = 0; j < kStatStripes; ++j) s[k] += striped[k * kStatStripes + j];
    out->bin_entries = s[0];
    out->overflowed_bins = s[1];
    out->triangles_setup = s[2] + s[4];
    out->vertices_setup = s[3] + s[5];
    out->triangles_phase2 = s[4];
    out->vertices_phase2 = s[5];
    out->raster_pixels = s[6];
    out->fragments = s[7];
  }
  if (r->d_counters.ptr && r->tun.cull_mode != GBR_CULL_NONE) {
    std::vector<uint32_t> c(kCtrCount * kCtrStride);
    HIP_TRY(hipMemcpyAsync(c.data(), r->d_counters.ptr, c.size() * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    out->items_phase1 = c[kCtrList1 * kCtrStride];
    out->items_candidates = c[kCtrCand * kCtrStride];
    out->items_phase2 = r->tun.cull_mode >= GBR_CULL_HIZ ? c[kCtrList2 * kCtrStride] : 0;
    out->triangles_clipped = c[kCtrClipped * kCtrStride];
    out->clip_fan_triangles = c[kCtrSubTris * kCtrStride];
    out->items_scored = r->tun.cull_mode == GBR_CULL_LEARNED ? c[kCtrProp * kCtrStride] : 0;
  } else {
    out->items_phase1 = r->n_items;
  }
  if (r->d_counters.ptr && r->tun.cull_mode == GBR_CULL_NONE) {
    uint32_t c[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&c[0], r->d_counters.ptr + kCtrClipped * kCtrStride, 4,
                           hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipMemcpyAsync(&c[1], r->d_counters.ptr + kCtrSubTris * kCtrStride, 4,
                           hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    out->triangles_clipped = c[0];
    out->clip_fan_triangles = c[1];
  }
  return GBR_OK;
}

gbr_status gbr_vis_fit_params_init(gbr_vis_fit_params* p) {
  if (!p) return fail(GBR_ERR_ARGUMENT, "vis_fit_params_init: null");
  std::memset(p, 0, sizeof *p);
  p->steps = 1000;
  p->batch = 8192;
  p->learning_rate = 1e-3f;
  p->beta1 = 0.9f;
  p->beta2 = 0.999f;
  p->eps = 1e-8f;
  p->pos_weight = 1.0f;
  p->seed = 1;
  return GBR_OK;
}

gbr_status gbr_vis_init(gbr_renderer* r, uint64_t seed) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "vis_init: null renderer");
  TRY(set_device(r));
  return r->vis.init(seed, r->stream);
}

gbr_status gbr_vis_collect(gbr_renderer* r, uint64_t* rows_added) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "vis_collect: null renderer");
  TRY(set_device(r));
  if (r->structure_dirty) {
    TRY(ensure_planes(r));
    if (r->assets_dirty) TRY(upload_assets(r));
    TRY(rebuild_structure(r));
  }
  TRY(r->vis.reserve_dataset(r->vis.rows + r->n_items + 1));
  r->collecting = true;
  const gbr_status st = do_draw(r);
  r->collecting = false;
  TRY(st);
  uint32_t n = 0;
  HIP_TRY(hipMemcpyAsync(&n, r->d_counters.ptr + kCtrProp * kCtrStride, sizeof n,
                         hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  const uint64_t added = std::min<uint64_t>(n, r->vis.cap - r->vis.rows);
  r->vis.rows += added;
  if (rows_added) *rows_added = added;
  return GBR_OK;
}

gbr_status gbr_vis_set_feature_set(gbr_renderer* r, int32_t feature_set) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "vis_set_feature_set: null renderer");
  if (feature_set != GBR_VIS_FEATURES_POSE && feature_set != GBR_VIS_FEATURES_HISTORY)
    return fail(GBR_ERR_ARGUMENT, "vis_set_feature_set: unknown feature set");
  if (feature_set != r->vis_features) {
    // rows of the other set would mix two meanings of every column
    r->vis.rows = 0;
    if (r->vis.positives.ptr) {
      TRY(set_device(r));
      HIP_TRY(hipMemsetAsync(r->vis.positives.ptr, 0, sizeof(uint32_t), r->stream));
    }
  }
  r->vis_features = feature_set;
  return GBR_OK;
}

gbr_status gbr_vis_feature_set(const gbr_renderer* r, int32_t* out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "vis_feature_set: null argument");
  *out = r->vis_features;
  return GBR_OK;
}

gbr_status gbr_vis_clear_rows(gbr_renderer* r) {
  if (!r) return fail(GBR_ERR_ARGUMENT, "vis_clear_rows: null renderer");
  TRY(set_device(r));
  r->vis.rows = 0;
  if (r->vis.positives.ptr) HIP_TRY(hipMemset(r->vis.positives.ptr, 0, sizeof(uint32_t)));
  return GBR_OK;
}

gbr_status gbr_vis_rows(gbr_renderer* r, uint64_t* rows, uint64_t* positives) {
  if (!r || !rows || !positives) return fail(GBR_ERR_ARGUMENT, "vis_rows: null argument");
  TRY(set_device(r));
  *rows = r->vis.rows;
  uint32_t p = 0;
  if (r->vis.positives.ptr) {
    HIP_TRY(hipStreamSynchronize(r->stream));
    HIP_TRY(hipMemcpy(&p, r->vis.positives.ptr, sizeof p, hipMemcpyDeviceToHost));
  }
  *positives = p;
  return GBR_OK;
}

gbr_status gbr_vis_fit(gbr_renderer* r, const gbr_vis_fit_params* params,
                       gbr_vis_fit_report* report) {
  if (!r || !params) return fail(GBR_ERR_ARGUMENT, "vis_fit: null argument");
  TRY(set_device(r));
  return r->vis.fit(*params, report, r->stream);
}

gbr_status gbr_vis_get_weights(gbr_renderer* r, float* out) {
  if (!r || !out) return fail(GBR_ERR_ARGUMENT, "vis_get_weights: null argument");
  TRY(set_device(r));
  return r->vis.get_weights(out, r->stream);
}

gbr_status gbr_vis_set_weights(gbr_renderer* r, const float* in) {
  if (!r || !in) return fail(GBR_ERR_ARGUMENT, "vis_set_weights: null argument");
  TRY(set_device(r));
  return r->vis.set_weights(in, r->stream);
}

gbr_status gbr_vis_loss_grad(gbr_renderer* r, const float* x, const uint8_t* labels,
                             int32_t rows, float pos_weight, float* loss, float* grad) {
  if (!r || !x || !labels || !loss || !grad)
    return fail(GBR_ERR_ARGUMENT, "vis_loss_grad: null argument");
  TRY(set_device(r));
  return r->vis.loss_grad(x, labels, rows, pos_weight, loss, grad, r->stream);
}

gbr_status gbr_vis_predict(gbr_renderer* r, const float* x, int32_t rows, float* logits) {
  if (!r || !x || !logits) return fail(GBR_ERR_ARGUMENT, "vis_predict: null argument");
  TRY(set_device(r));
  return r->vis.predict(x, rows, logits, r->stream);
}

gbr_status gbr_frame_timings(gbr_renderer* r, float* ms) {
  if (!r || !ms) return fail(GBR_ERR_ARGUMENT, "frame_timings: null argument");
  for (int i = 0; i < GBR_TIMING_SLOTS; ++i) ms[i] = 0.0f;
  if (!r->tun.timing) return fail(GBR_ERR_STATE, "frame_timings: tuning.timing is off");
  if (r->t_frames == 0) return fail(GBR_ERR_STATE, "frame_timings: no frame drawn since the last call");
  TRY(set_device(r));
  HIP_TRY(hipStreamSynchronize(r->stream));
  double sum[GBR_TIMING_SLOTS] = {};
  for (size_t f = 0; f < r->t_frames; ++f)
    for (int i = 0; i < GBR_TIMING_SLOTS; ++i) {
      if (!r->t_used[f][i]) continue;
      float v = 0.0f;
      HIP_TRY(hipEventElapsedTime(&v, r->t_pool[f][2 * i], r->t_pool[f][2 * i + 1]));
      sum[i] += v;
    }
  for (int i = 0; i < GBR_TIMING_SLOTS; ++i) ms[i] = (float)(sum[i] / (double)r->t_frames);
  r->t_frames = 0;
  return GBR_OK;
}

}  // extern "C"
