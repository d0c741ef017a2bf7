/*
 * gbr.h -- C ABI of the MI355X batch sensor renderer.
 *
 * Drop-in for habitat-sim's esp::gfx_batch::Renderer / RendererStandalone
 * path (src/esp/gfx_batch/Renderer.h, RendererStandalone.h,
 * DepthUnprojection.h).  Every entry point below names the reference method it
 * replaces (file:line relative to the habitat-sim source tree).  Entry points
 * marked "new" have no counterpart in the reference.
 *
 * Conventions
 *  - All functions return a gbr_status; 0 is success.  gbr_last_error()
 *    returns a thread-local message for the last failure.  No C++ exception
 *    crosses this boundary.
 *  - Matrices are 16 floats, column-major ([col*4 + row]), Magnum layout.
 *  - Images are bottom-up (first row = GL window y 0), like the reference's
 *    framebuffer reads, unless gbr_config.top_down is set for the per-scene
 *    planes.
 *  - A renderer owns one HIP stream; calls on one renderer must be serialised
 *    by the caller.  gbr_draw() is asynchronous; the *_into() readbacks
 *    synchronise the stream (the reference's "stalls the CPU" contract,
 *    RendererStandalone.h:174,196).
 */
#ifndef GBR_H
#define GBR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBR_ABI_VERSION 1

typedef int32_t gbr_status;
enum {
  GBR_OK = 0,
  GBR_ERR_ARGUMENT = 1,    /* null pointer, invalid flag or value */
  GBR_ERR_RANGE = 2,       /* scene / node / light index out of range */
  GBR_ERR_NOT_FOUND = 3,   /* unknown node hierarchy name */
  GBR_ERR_IO = 4,          /* file can't be opened or imported */
  GBR_ERR_SIZE = 5,        /* rectangle / image size mismatch */
  GBR_ERR_DEVICE = 6,      /* HIP runtime failure */
  GBR_ERR_OOM = 7,         /* device or host allocation failed */
  GBR_ERR_UNSUPPORTED = 8, /* feature not provided by this build */
  GBR_ERR_STATE = 9        /* conflicting state, e.g. duplicate name */
};

/* RendererFlag, Renderer.h:20-30 */
#define GBR_FLAG_NO_TEXTURES 0x1u
/* RendererStandaloneFlag, RendererStandalone.h */
#define GBR_STANDALONE_QUIET_LOG 0x1u
/* RendererFileFlag, Renderer.h:120-154 */
#define GBR_FILE_WHOLE 0x1u
#define GBR_FILE_GENERATE_MIPMAP 0x2u
/* RendererLightType, Renderer.h:170-173 */
#define GBR_LIGHT_DIRECTIONAL 0
#define GBR_LIGHT_POINT 1
/* largest max_light_count (the reference compiles its shaders for any count,
   Renderer.cpp:835-860) */
#define GBR_MAX_LIGHTS 64

/* Output planes produced by gbr_draw() (new; the reference has colour and
   window depth only, RendererStandalone.cpp:123-124) */
#define GBR_PLANE_COLOR 0x1u        /* RGBA8 */
#define GBR_PLANE_DEPTH 0x2u        /* window depth, Depth32F */
#define GBR_PLANE_LINEAR_DEPTH 0x4u /* unprojectDepth() result, far -> 0 */
#define GBR_PLANE_SEMANTIC 0x8u     /* uint32 semantic id, 0 = background */
#define GBR_PLANE_ALL 0xfu

/* pixel formats accepted by gbr_color_image_into() */
#define GBR_FORMAT_RGBA8 0
#define GBR_FORMAT_RGB8 1

/* where a destination pointer lives */
#define GBR_MEMORY_HOST 0
#define GBR_MEMORY_DEVICE 1

/* culling modes (new) */
#define GBR_CULL_NONE 0    /* the reference's method: every triangle set up */
#define GBR_CULL_FRUSTUM 1 /* per-chunk frustum test before setup (exact) */
#define GBR_CULL_HIZ 2     /* + two-phase depth-pyramid occlusion test; phase 1 =
                              chunks visible in the previous frame (exact) */
#define GBR_CULL_LEARNED 3 /* + phase 1 = chunks the visibility model proposes (exact) */

typedef struct gbr_renderer gbr_renderer;

/* RendererConfiguration (Renderer.h:51-113, defaults Renderer.cpp:58-64)
   + RendererStandaloneConfiguration (RendererStandalone.h) */
typedef struct gbr_config {
  int32_t tile_size[2];   /* default {128, 128} */
  int32_t tile_count[2];  /* default {1, 1}; scene count = product */
  uint32_t flags;         /* GBR_FLAG_* */
  uint32_t max_light_count; /* 0: every material flat; at most GBR_MAX_LIGHTS */
  float ambient_factor;   /* default 0.1 */
  int32_t device;         /* HIP device ordinal; -1 = current device
                             (gbr_get_config reports the ordinal in use) */
  uint32_t standalone_flags; /* GBR_STANDALONE_* */
  uint32_t planes;        /* GBR_PLANE_* produced by draw(); default ALL */
  uint32_t top_down;      /* new: per-scene planes stored top-down */
  uint32_t reserved[7];
} gbr_config;

/* SceneStats, Renderer.h:750-774 */
typedef struct gbr_scene_stats_info {
  uint64_t node_count;
  uint64_t draw_count;
  uint64_t draw_batch_count;
} gbr_scene_stats_info;

/* new: launch parameters, chosen by the autotuner */
typedef struct gbr_tuning {
  int32_t bin_size;           /* screen bin edge in pixels: 8, 16 or 32 */
  int32_t chunk_triangles;    /* triangles per setup work item: 1..1024 */
  float bin_capacity_factor;  /* per-bin entry capacity relative to tris/bins */
  int32_t min_bin_capacity;   /* lower bound of the per-bin capacity */
  int32_t cull_mode;          /* GBR_CULL_* */
  int32_t queue_grain;        /* setup work items claimed per queue fetch: 1..64 */
  int32_t timing;             /* 1: time each stage of gbr_draw() with HIP events */
  int32_t raster_mode;        /* 0: a lane per binned triangle, 1: a lane per
                                 (triangle, pixel) pair; same output */
  float proposal_threshold;   /* GBR_CULL_LEARNED: phase 1 = chunks whose logit
                                 exceeds this */
  int32_t lane_pixels;        /* raster_mode 0: triangles covering more than this
                                 many pixels of a bin are rasterised by a whole
                                 wave (the rectangle spread over its lanes), the
                                 rest by one lane each: 1..1024; default 64 */
  int32_t graph;              /* 1: gbr_draw() replays its kernel sequence as a
                                 captured HIP graph while nothing but cameras and
                                 transforms changed (timing and row collection
                                 draw directly) */
  int32_t proposal_history;   /* GBR_CULL_LEARNED: 1 = phase 1 also takes the
                                 chunks visible in the previous frame, and the
                                 model scores only the chunks that come into
                                 view (one that passed the frustum test in the
                                 previous frame without being visible waits
                                 for the depth-pyramid test): camera walks;
                                 0 = the model's proposal alone (default; the
                                 right choice when cameras jump) */
  int32_t heavy_entries;      /* raster_mode 0: bins that receive at least this many
                                 entries in a phase are rasterised first (0: bins in
                                 order); default 256 */
} gbr_tuning;

/* stages timed when gbr_tuning.timing is set (gbr_frame_timings) */
#define GBR_TIMING_NODES 0   /* absolute transforms, counter reset */
#define GBR_TIMING_CULL 1    /* chunk frustum test + proposal */
#define GBR_TIMING_SETUP1 2  /* triangle setup + binning, phase 1 */
#define GBR_TIMING_CLIP1 3   /* near/guard-band clipping, phase 1 */
#define GBR_TIMING_RASTER1 4 /* bin rasterisation, phase 1 */
#define GBR_TIMING_HIZ 5     /* depth pyramid + occlusion test */
#define GBR_TIMING_SETUP2 6
#define GBR_TIMING_CLIP2 7
#define GBR_TIMING_RASTER2 8
#define GBR_TIMING_RESOLVE 9 /* shading, depth, semantic planes */
#define GBR_TIMING_SLOTS 10

/* new: per-frame counters of the last gbr_draw() (read after a sync) */
typedef struct gbr_frame_stats_info {
  uint64_t work_items;        /* setup work items launched */
  uint64_t triangles;         /* triangles submitted */
  uint64_t bin_entries;       /* triangle-bin pairs written */
  uint64_t overflowed_bins;   /* bins that exceeded capacity (exact fallback) */
  uint64_t raster_items;      /* bins rasterised */
  uint64_t items_phase1;      /* work items set up in phase 1 (all, without culling) */
  uint64_t items_candidates;  /* work items tested against the depth pyramid */
  uint64_t items_phase2;      /* candidates set up in phase 2 */
  uint64_t triangles_setup;   /* triangles of the work items set up */
  uint64_t vertices_setup;    /* vertex-range entries of the work items set up */
  uint64_t triangles_phase2;  /* of which in phase 2 */
  uint64_t vertices_phase2;
  uint64_t raster_pixels;     /* pixel centres tested against binned triangles */
  uint64_t fragments;         /* of which inside their triangle */
  uint64_t triangles_clipped; /* triangles through the near/far/guard-band clipper */
  uint64_t clip_fan_triangles; /* front-facing triangles of their clipped fans */
  uint64_t items_scored;      /* GBR_CULL_LEARNED: work items scored by the model */
} gbr_frame_stats_info;

/* ---- parsed-file description consumed by gbr_add_file_data() (new) ---- */
typedef struct gbr_mesh_desc {
  uint32_t vertex_count;
  const float* positions;   /* [vertex_count*3] */
  const float* texcoords;   /* [vertex_count*2] or NULL; Y-up convention */
  const void* colors;       /* [vertex_count*4] float or uint8, or NULL */
  uint32_t colors_u8;       /* colors are uint8 (normalised) */
  uint32_t texcoords_half;  /* store texture coordinates as fp16 */
  uint32_t index_count;     /* triangle-list indices */
  const uint32_t* indices;
  uint32_t source_index_size; /* bytes per index in the source file */
  const float* normals;     /* [vertex_count*3] or NULL (+Z); used with lights */
} gbr_mesh_desc;

typedef struct gbr_texture_desc {
  uint32_t width, height, layers, levels;  /* level 0 size */
  const uint8_t* texels;    /* RGBA8, levels concatenated, each [layers][h][w] */
  uint32_t min_filter, mag_filter, wrap_s, wrap_t; /* GL enums */
  uint32_t compressed;      /* source was block-compressed: no mip generation */
} gbr_texture_desc;

typedef struct gbr_material_desc {
  uint32_t flat;            /* Trade::MaterialType::Flat */
  float base_color[4];
  int32_t texture;          /* file-local texture id or -1 */
  uint32_t layer;
  float texture_matrix[9];  /* column-major 3x3 */
} gbr_material_desc;

typedef struct gbr_node_desc {
  int32_t parent;           /* file-local node id or -1 */
  float transform[16];
  const char* name;         /* NULL if unnamed */
} gbr_node_desc;

typedef struct gbr_mesh_entry_desc {
  uint32_t node;
  uint32_t mesh;
  int32_t material;         /* material of the mesh primitive, -1 none */
  uint32_t has_view;        /* meshViewIndexOffset/Count/Material present */
  uint32_t view_index_offset_bytes;
  uint32_t view_index_count;
  int32_t view_material;
} gbr_mesh_entry_desc;

typedef struct gbr_file_desc {
  uint32_t mesh_count;
  const gbr_mesh_desc* meshes;
  uint32_t texture_count;
  const gbr_texture_desc* textures;
  uint32_t material_count;
  const gbr_material_desc* materials;
  uint32_t node_count;
  const gbr_node_desc* nodes;
  uint32_t mesh_entry_count;
  const gbr_mesh_entry_desc* mesh_entries;
  uint32_t has_scene;
} gbr_file_desc;

/* ---- library ------------------------------------------------------------ */
int32_t gbr_abi_version(void);
const char* gbr_last_error(void);
gbr_status gbr_config_init(gbr_config* config);           /* RendererConfiguration() defaults */
gbr_status gbr_tuning_init(gbr_tuning* tuning);

/* ---- lifecycle: Renderer(const RendererConfiguration&) Renderer.h:445-448,
   RendererStandalone ctor RendererStandalone.cpp:115-135, destroy() :735-737 */
gbr_status gbr_create(const gbr_config* config, gbr_renderer** out);
gbr_status gbr_destroy(gbr_renderer* r);

/* getters: flags() tileSize() tileCount() sceneCount() maxLightCount()
   Renderer.h:458-491, standaloneFlags() RendererStandalone.h */
gbr_status gbr_get_config(const gbr_renderer* r, gbr_config* out);
gbr_status gbr_scene_count(const gbr_renderer* r, uint64_t* out);

/* ---- assets: addFile() Renderer.h:517-556 / Renderer.cpp:354-867 ---------- */
/* Parses a glTF 2.0 / glb composite natively (MAGNUMX_mesh_views extras,
   KHR_texture_ktx, KHR_texture_transform, KHR_materials_unlit; KTX2 RGBA8/
   RGB8/RG8/R8 and BC1, PNG images; data URIs and glb chunks) with the
   conventions of the reference's Magnum importers, then applies addFile()'s
   logic as gbr_add_file_data().  A file that can't be read or parsed returns
   GBR_ERR_IO (the reference's `false`). */
gbr_status gbr_add_file(gbr_renderer* r, const char* filename, uint32_t file_flags,
                        const char* name);
/* new: addFile() on an already-parsed file; `filename` is the name used for
   GBR_FILE_WHOLE when `name` is NULL or empty.  GBR_ERR_IO / GBR_ERR_STATE
   correspond to the reference returning false. */
gbr_status gbr_add_file_data(gbr_renderer* r, const gbr_file_desc* file, uint32_t file_flags,
                             const char* name, const char* filename);
/* hasNodeHierarchy() Renderer.h:564 */
gbr_status gbr_has_node_hierarchy(const gbr_renderer* r, const char* name, int32_t* out);

/* ---- scenes ---------------------------------------------------------------- */
/* addNodeHierarchy(scene, name, bakeTransformation) Renderer.h:586-596;
   bake may be NULL (identity). */
gbr_status gbr_add_node_hierarchy(gbr_renderer* r, uint32_t scene, const char* name,
                                  const float* bake, uint64_t* out_node);
/* new: same, every draw of the hierarchy carries `semantic_id` */
gbr_status gbr_add_node_hierarchy_semantic(gbr_renderer* r, uint32_t scene, const char* name,
                                           const float* bake, uint32_t semantic_id,
                                           uint64_t* out_node);
/* new: addNodeHierarchy() `count` times into one scene, instance i from
   names[i], bakes[16*i] (NULL: identity for all) and semantics[i] (NULL: 0);
   top-level ids into out_nodes (may be NULL).  Stops at the first failure,
   keeping the instances added before it. */
gbr_status gbr_add_node_hierarchies(gbr_renderer* r, uint32_t scene, uint32_t count,
                                    const char* const* names, const float* bakes,
                                    const uint32_t* semantics, uint64_t* out_nodes);
/* new: sets the semantic id of every draw of the hierarchy whose top-level
   node is `node` (a gfx-replay instance's metadata, which the reference's
   BatchPlayerImplementation drops, BatchPlayerImplementation.cpp:51-91) */
gbr_status gbr_set_node_hierarchy_semantic(gbr_renderer* r, uint32_t scene, uint64_t node,
                                           uint32_t semantic_id);
/* addEmptyNode() Renderer.h:598 */
gbr_status gbr_add_empty_node(gbr_renderer* r, uint32_t scene, uint64_t* out_node);
/* addLight() Renderer.h:619-621, Renderer.cpp:990-1015: a light of
   GBR_LIGHT_* type placed by the scene node's absolute transform (point:
   its translation, directional: towards -backward), colour white, range
   infinite.  GBR_ERR_STATE when max_light_count is 0, GBR_ERR_RANGE for a
   node out of range.  A scene's lights beyond max_light_count don't shade. */
gbr_status gbr_add_light(gbr_renderer* r, uint32_t scene, uint64_t node, int32_t type,
                         uint64_t* out_light);
/* lightColors() Renderer.h:697-698, lightRanges() Renderer.h:713-714:
   mutable host views of the scene's lights (3 floats of colour, resp. one
   float of range, per light, `*out_stride` bytes apart), valid until the
   next addLight(), clearLights() or clear() of the scene; read at every
   gbr_draw() */
gbr_status gbr_light_colors(gbr_renderer* r, uint32_t scene, float** out_data, uint64_t* out_count,
                            uint64_t* out_stride);
gbr_status gbr_light_ranges(gbr_renderer* r, uint32_t scene, float** out_data, uint64_t* out_count,
                            uint64_t* out_stride);
/* clear() Renderer.h:631, clearLights() Renderer.h:640 */
gbr_status gbr_clear(gbr_renderer* r, uint32_t scene);
gbr_status gbr_clear_lights(gbr_renderer* r, uint32_t scene);
/* new: heterogeneous batches.  A scene renders into the bottom-left
   width x height pixels of its tile (1 <= width <= tile width, same for
   height; the projection maps onto that viewport) and the rest of the tile
   reads back as clear; by default the viewport is the whole tile.  A scene's
   sensor set (a subset of gbr_config.planes; 0 = all of them) limits the
   planes draw() writes for it; its other planes hold unspecified values.
   The reference's tiles are uniform (Renderer.h:51-113). */
gbr_status gbr_set_scene_viewport(gbr_renderer* r, uint32_t scene, int32_t width, int32_t height);
gbr_status gbr_scene_viewport(const gbr_renderer* r, uint32_t scene, int32_t* out2);
gbr_status gbr_set_scene_planes(gbr_renderer* r, uint32_t scene, uint32_t planes);

/* ---- cameras and transformations ----------------------------------------- */
/* camera() Renderer.h:647: combined projection*view */
gbr_status gbr_camera(const gbr_renderer* r, uint32_t scene, float* out16);
/* cameraDepthUnprojection() Renderer.h:653 */
gbr_status gbr_camera_depth_unprojection(const gbr_renderer* r, uint32_t scene, float* out2);
/* updateCamera(scene, projection, view) Renderer.h:665-667 */
gbr_status gbr_update_camera(gbr_renderer* r, uint32_t scene, const float* projection16,
                             const float* view16);
/* new: updateCamera() for `count` consecutive scenes from arrays of matrices */
gbr_status gbr_update_cameras(gbr_renderer* r, uint32_t first_scene, uint32_t count,
                              const float* projections16, const float* views16);
/* transformations(scene) Renderer.h:682-683: mutable host view of the local
   node matrices (16 floats each, contiguous), valid until the next structural
   change of the scene (add*, clear).  Obtaining the view marks the scene's
   transformations for upload at the next gbr_draw(). */
gbr_status gbr_transformations(gbr_renderer* r, uint32_t scene, float** out_data,
                               uint64_t* out_count);
/* new: mark transformations modified through a view obtained earlier */
gbr_status gbr_mark_transformations_dirty(gbr_renderer* r, uint32_t scene);
/* sceneStats() Renderer.h:730 */
gbr_status gbr_scene_stats(const gbr_renderer* r, uint32_t scene, gbr_scene_stats_info* out);

/* ---- rendering ------------------------------------------------------------- */
/* RendererStandalone::draw() RendererStandalone.cpp:155-159 (clear + draw of
   every scene), asynchronous on the renderer's stream */
gbr_status gbr_draw(gbr_renderer* r);
/* new: wait for the renderer's stream */
gbr_status gbr_synchronize(gbr_renderer* r);
/* new: the renderer's hipStream_t */
gbr_status gbr_stream(gbr_renderer* r, void** out_stream);

/* colorImageInto() RendererStandalone.cpp:168-180: rect = {x0, y0, x1, y1}
   in framebuffer pixels (tile grid, bottom-up); dst has rect size, rows
   `row_stride` bytes apart.  format GBR_FORMAT_RGBA8 or RGB8. */
gbr_status gbr_color_image_into(gbr_renderer* r, const int32_t* rect, void* dst,
                                int64_t row_stride, int32_t format, int32_t memory);
/* depthImageInto() RendererStandalone.cpp:189-201 (window depth) */
gbr_status gbr_depth_image_into(gbr_renderer* r, const int32_t* rect, float* dst,
                                int64_t row_stride, int32_t memory);
/* new: unprojected depth (depthImageInto() + unprojectDepth(), as
   BatchReplayRenderer::doRender() does, BatchReplayRenderer.cpp:139-148) */
gbr_status gbr_linear_depth_image_into(gbr_renderer* r, const int32_t* rect, float* dst,
                                       int64_t row_stride, int32_t memory);
/* new: semantic ids (classic R32UI convention, RenderTarget.cpp:78,190) */
gbr_status gbr_semantic_image_into(gbr_renderer* r, const int32_t* rect, uint32_t* dst,
                                   int64_t row_stride, int32_t memory);
/* colorCudaBufferDevicePointer() / depthCudaBufferDevicePointer()
   RendererStandalone.h:225-235: linearised framebuffer (tile grid, bottom-up)
   in device memory, valid until the next gbr_draw() / gbr_destroy() */
gbr_status gbr_color_device_pointer(gbr_renderer* r, const void** out);
gbr_status gbr_depth_device_pointer(gbr_renderer* r, const void** out);
/* new: per-scene planes, [scene][tile_h][tile_w] (RGBA8 / float / uint32),
   written in place by gbr_draw(); no copy */
gbr_status gbr_plane_device_pointer(gbr_renderer* r, uint32_t plane, const void** out);

/* ---- depth unprojection: DepthUnprojection.h:137-153 ----------------------- */
/* calculateDepthUnprojection() DepthUnprojection.cpp:98-101 */
gbr_status gbr_calculate_depth_unprojection(const float* projection16, float* out2);
/* unprojectDepth() DepthUnprojection.cpp:107-130 on host memory */
gbr_status gbr_unproject_depth(const float* unprojection2, float* depth, int64_t rows,
                               int64_t cols, int64_t row_stride);

/* ---- tuning and counters (new) --------------------------------------------- */
gbr_status gbr_get_tuning(const gbr_renderer* r, gbr_tuning* out);
gbr_status gbr_set_tuning(gbr_renderer* r, const gbr_tuning* tuning);
gbr_status gbr_frame_stats(gbr_renderer* r, gbr_frame_stats_info* out);
/* new: of the last frame's work items, those owning a visible pixel, and how
   many of them phase 1 set up (the proposal's recall is the ratio; 1 without
   culling and in frustum mode); synchronises the stream */
gbr_status gbr_frame_visibility(gbr_renderer* r, uint64_t* visible, uint64_t* visible_phase1);

/* new: autotuner.  Times gbr_draw() of the current scenes and cameras under
   candidate launch parameters, one parameter at a time (bin size, per-bin
   capacity factor, lane/workgroup raster split, setup claim size, raster
   mode, heavy-bin threshold, and with GBR_CULL_LEARNED the proposal
   threshold; each from the best values found so far; a candidate replaces
   the best only when it is more than 1% faster), and leaves the fastest set
   (every candidate renders the same pixels).  cull_mode,
   timing, graph, chunk_triangles and proposal_history are kept. */
typedef struct gbr_autotune_params {
  int32_t warmup; /* untimed draws per candidate (>= 1; absorbs rebuilds) */
  int32_t frames; /* timed draws per candidate (>= 1) */
  int32_t passes; /* rounds over the parameters (>= 1) */
} gbr_autotune_params;
typedef struct gbr_autotune_report {
  int32_t candidates;  /* candidates timed */
  float initial_ms;    /* GPU ms per frame with the tuning on entry */
  float best_ms;       /* ... with the tuning left set */
  gbr_tuning best;
} gbr_autotune_report;
gbr_status gbr_autotune_params_init(gbr_autotune_params* params);
gbr_status gbr_autotune(gbr_renderer* r, const gbr_autotune_params* params,
                        gbr_autotune_report* report);
/* ---- learned visibility proposal (new) -------------------------------------
   No counterpart in the reference, whose batch renderer submits every mesh
   of every scene (Renderer.cpp:1298-1351); the nearest precedent is the
   classic renderer's per-object frustum cull (RenderCamera.cpp:104-125).
   A 32 -> 256 -> 256 -> 1 MLP (BF16 MFMA, FP32 accumulation and master
   weights) scores each chunk that passes the frustum test from
   per-(camera, chunk) features; with GBR_CULL_LEARNED the chunks it scores
   above gbr_tuning.proposal_threshold form phase 1 of the two-phase
   depth-pyramid cull.  The proposal only orders work: output pixels never
   depend on it.  Training rows are (features, chunk visible) pairs taken
   from frames the renderer draws. */
#define GBR_VIS_FEATURES 32
#define GBR_VIS_HIDDEN 256
#define GBR_VIS_PARAMS 74497 /* W1[256][32] b1[256] W2[256][256] b2[256] w3[256] b3 */
/* feature sets of the 32 inputs */
#define GBR_VIS_FEATURES_POSE 0    /* default: camera position, sin/cos of yaw and
                                      pitch, the chunk's world box centre and extents
                                      relative to the camera (also in the camera
                                      frame), nothing from earlier frames */
#define GBR_VIS_FEATURES_HISTORY 1 /* ablation: visibility in earlier frames, the
                                      projected rectangle and depth, box and camera */

typedef struct gbr_vis_fit_params {
  int32_t steps;            /* optimiser steps */
  int32_t batch;            /* rows per step: multiple of 32, 32..65536 */
  float learning_rate;      /* Adam step size */
  float beta1, beta2, eps;  /* Adam */
  float pos_weight;         /* loss weight of rows labelled visible */
  int32_t reserved;
  uint64_t seed;            /* minibatch sampling */
} gbr_vis_fit_params;

typedef struct gbr_vis_fit_report {
  float loss_first;         /* mean weighted loss of the first step */
  float loss_last;          /* mean weighted loss over the last tenth of the steps */
  uint64_t rows;            /* training rows */
  uint64_t positives;       /* of which labelled visible */
} gbr_vis_fit_report;

gbr_status gbr_vis_fit_params_init(gbr_vis_fit_params* params);
/* random (He-uniform) weights from seed; keeps the training rows */
gbr_status gbr_vis_init(gbr_renderer* r, uint64_t seed);
/* gbr_draw() with the current cameras, then appends one training row per
   chunk that passed the frustum test strictly inside the clip volume */
gbr_status gbr_vis_collect(gbr_renderer* r, uint64_t* rows_added);
gbr_status gbr_vis_clear_rows(gbr_renderer* r);
/* selects the feature set (GBR_VIS_FEATURES_*) of collected rows and of the
   learned proposal; switching drops the training rows (the weights stay) */
gbr_status gbr_vis_set_feature_set(gbr_renderer* r, int32_t feature_set);
gbr_status gbr_vis_feature_set(const gbr_renderer* r, int32_t* out);
gbr_status gbr_vis_rows(gbr_renderer* r, uint64_t* rows, uint64_t* positives);
/* Adam on the weighted binary cross-entropy of the training rows */
gbr_status gbr_vis_fit(gbr_renderer* r, const gbr_vis_fit_params* params,
                       gbr_vis_fit_report* report);
gbr_status gbr_vis_get_weights(gbr_renderer* r, float* out /* GBR_VIS_PARAMS */);
gbr_status gbr_vis_set_weights(gbr_renderer* r, const float* in /* GBR_VIS_PARAMS */);
/* mean weighted loss and its gradient (GBR_VIS_PARAMS floats) for caller rows
   x[rows][GBR_VIS_FEATURES] (rounded to BF16) and labels; rows multiple of 32 */
gbr_status gbr_vis_loss_grad(gbr_renderer* r, const float* x, const uint8_t* labels,
                             int32_t rows, float pos_weight, float* loss, float* grad);
/* logits of caller rows x[rows][GBR_VIS_FEATURES] */
gbr_status gbr_vis_predict(gbr_renderer* r, const float* x, int32_t rows, float* logits);

/* new: milliseconds per GBR_TIMING_* stage, averaged over the gbr_draw()
   calls since the previous call (at most 4096; a stage that did not run
   counts 0); needs gbr_tuning.timing = 1; synchronises the stream */
gbr_status gbr_frame_timings(gbr_renderer* r, float* ms);

/* ---- Redwood depth noise: RedwoodNoiseModelGPUImpl (RedwoodNoiseModel.h:
   50-97, RedwoodNoiseModel.cu; Python RedwoodDepthNoiseModel,
   redwood_depth_noise_model.py:107-148) ------------------------------------ */
typedef struct gbr_redwood gbr_redwood;
/* the undistortion table (redwood-depth-dist-model.npy), [80][80][5] floats */
#define GBR_REDWOOD_MODEL_FLOATS (80 * 80 * 5)
/* RedwoodNoiseModelGPUImpl(model, gpuDeviceId, noiseMultiplier)
   (RedwoodNoiseModel.h:58-62): copies the table to `device` (-1: current).
   The reference seeds its per-thread curand states from rand(); here the
   deviates come from a Philox4x32-10 counter generator keyed by `seed`, so
   call k of a model reproduces from (seed, k). */
gbr_status gbr_redwood_create(const float* model, int32_t device, float noise_multiplier,
                              uint64_t seed, gbr_redwood** out);
gbr_status gbr_redwood_destroy(gbr_redwood* m);
/* simulateFromGPU(devDepth, rows, cols, devNoisyDepth) (RedwoodNoiseModel.h:
   78-81) for `images` consecutive rows x cols depth images in device memory
   (rows, cols >= 2), asynchronous on `stream` (a hipStream_t; NULL = the
   null stream); every call draws fresh deviates.  Depths >= 10 m give 0, and
   so do negative and NaN ones (the reference reads its table out of bounds
   for those). */
gbr_status gbr_redwood_simulate(gbr_redwood* m, const float* depth, int32_t images, int32_t rows,
                                int32_t cols, float* out, void* stream);
/* simulateFromCPU(depth, rows, cols) (RedwoodNoiseModel.h:70-72): host
   image in, host image out, synchronous */
gbr_status gbr_redwood_simulate_host(gbr_redwood* m, const float* depth, int32_t rows, int32_t cols,
                                     float* out);
/* new, no counterpart in the reference: gbr_redwood_simulate_host() for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  It stages the
   image in a fixed device array of the library (2 x
   GBR_REDWOOD_NOALLOC_MAX_PIXELS floats, 8 MiB per GPU, held until the
   process ends) and calls no allocator: rows x cols above
   GBR_REDWOOD_NOALLOC_MAX_PIXELS give GBR_ERR_SIZE.  Same output and
   generator sequence as gbr_redwood_simulate_host().  Every call on one GPU
   shares that array: calls that overlap on one GPU (several threads, or
   models on the same device) race and must be serialised by the caller. */
#define GBR_REDWOOD_NOALLOC_MAX_PIXELS (1 << 20)
gbr_status gbr_redwood_simulate_host_noalloc(gbr_redwood* m, const float* depth, int32_t rows,
                                             int32_t cols, float* out);
/* new: gbr_redwood_simulate with caller-given standard normal deviates, 3
   per pixel (y shift, x shift, disparity), in device memory */
gbr_status gbr_redwood_simulate_normals(gbr_redwood* m, const float* depth, int32_t images,
                                        int32_t rows, int32_t cols, const float* normals,
                                        float* out, void* stream);
/* new: generator calls so far (the next call's index) */
gbr_status gbr_redwood_calls(const gbr_redwood* m, uint64_t* out);
/* new: the generator's raw blocks: for pixels first .. first+n-1 of call
   `call`, Philox4x32-10 of counter (pixel lo, pixel hi, call lo, call hi)
   and key (seed lo, seed hi), 4 words each, into device memory */
gbr_status gbr_redwood_random_bits(gbr_redwood* m, uint64_t call, uint64_t first, uint64_t n,
                                   uint32_t* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GBR_H */
