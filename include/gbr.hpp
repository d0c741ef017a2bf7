/*
 * gbr.hpp -- header-only C++ classes shaped like habitat-sim's
 * esp::gfx_batch::Renderer / RendererStandalone over the C ABI of gbr.h.
 *
 * Same class, method and argument names as the reference
 * (src/esp/gfx_batch/Renderer.h:51-774, RendererStandalone.h:21-241,
 * DepthUnprojection.h:137-153), so C++ callers such as BatchReplayRenderer
 * (src/esp/sim/BatchReplayRenderer.cpp:22-197) change only their includes and
 * the small value types below, which stand in for the Magnum ones the
 * reference takes (Vector2i, Vector2, Matrix4, Range2Di, image views).
 *
 * Error behaviour follows the reference: addFile() returns false and prints
 * the message (Renderer.cpp:354-867); every other misuse is a
 * CORRADE_ASSERT there (message, then abort).  Here such a failure calls the
 * installed handler (setErrorHandler); the default prints
 * "gfx_batch: <message>" to stderr and aborts.  A handler may throw instead,
 * which the C ABI never does.
 *
 * The classes live in namespace gbr::gfx_batch; define
 * GBR_ESP_NAMESPACE_ALIAS before including to also get the reference's
 * spelling esp::gfx_batch.
 */
#ifndef GBR_HPP
#define GBR_HPP

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gbr.h"

namespace gbr {
namespace gfx_batch {

/* ---- value types standing in for Magnum's ---------------------------------- */

struct Vector2i {
  int x = 0, y = 0;
  Vector2i() = default;
  Vector2i(int x_, int y_) : x{x_}, y{y_} {}
  int product() const { return x * y; }
  bool operator==(const Vector2i& o) const { return x == o.x && y == o.y; }
  bool operator!=(const Vector2i& o) const { return !(*this == o); }
};

struct Vector2 {
  float x = 0.0f, y = 0.0f;
};

/* column-major 4x4 (Magnum layout: data[col*4 + row]); identity by default */
struct Matrix4 {
  float data[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  static Matrix4 fromColumnMajor(const float* m) {
    Matrix4 r;
    for (int i = 0; i < 16; ++i) r.data[i] = m[i];
    return r;
  }
  static Matrix4 translation(float x, float y, float z) {
    Matrix4 r;
    r.data[12] = x;
    r.data[13] = y;
    r.data[14] = z;
    return r;
  }
};

struct Color3 {
  float r = 0.0f, g = 0.0f, b = 0.0f;
};

/* Range2Di: min inclusive, max exclusive, framebuffer pixels */
struct Range2Di {
  Vector2i min, max;
  Range2Di() = default;
  Range2Di(Vector2i mn, Vector2i mx) : min{mn}, max{mx} {}
  Vector2i size() const { return {max.x - min.x, max.y - min.y}; }
};

/* Magnum::PixelFormat values the renderer reports or accepts */
enum class PixelFormat { RGBA8Unorm, RGB8Unorm, Depth32F, R32F, R32UI };

inline std::size_t pixelSize(PixelFormat f) { return f == PixelFormat::RGB8Unorm ? 3 : 4; }

/* MutableImageView2D: caller-owned pixels, rows `rowStride` bytes apart
   (0: tightly packed), bottom-up like GL */
struct MutableImageView2D {
  PixelFormat format;
  Vector2i size;
  void* data;
  std::ptrdiff_t rowStride;
  MutableImageView2D(PixelFormat f, Vector2i s, void* d, std::ptrdiff_t stride = 0)
      : format{f}, size{s}, data{d}, rowStride{stride ? stride : (std::ptrdiff_t)(s.x * pixelSize(f))} {}
};

/* Image2D: an owning image (colorImage(), depthImage()) */
struct Image2D {
  PixelFormat format;
  Vector2i size;
  std::vector<unsigned char> data;
  std::ptrdiff_t rowStride() const { return (std::ptrdiff_t)(size.x * pixelSize(format)); }
  template <class T>
  const T* pixels() const { return reinterpret_cast<const T*>(data.data()); }
};

/* StridedArrayView1D<T>: mutable view into renderer-owned memory */
template <class T>
struct StridedArrayView1D {
  T* base = nullptr;
  std::size_t count = 0;
  std::ptrdiff_t stride = sizeof(T);
  std::size_t size() const { return count; }
  T& operator[](std::size_t i) const {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (std::ptrdiff_t)i * stride);
  }
};

/* StridedArrayView2D<float>: rows x cols floats, rows `rowStride` bytes apart */
struct StridedArrayView2D {
  float* data;
  std::size_t rows, cols;
  std::ptrdiff_t rowStride;
};

/* ---- errors ---------------------------------------------------------------- */

typedef void (*ErrorHandler)(const char* message);

inline void defaultErrorHandler(const char* message) {
  std::fprintf(stderr, "gfx_batch: %s\n", message);
  std::abort();
}

inline ErrorHandler& errorHandlerSlot() {
  static ErrorHandler h = defaultErrorHandler;
  return h;
}

/* installs `h` (nullptr: the default) and returns the previous handler */
inline ErrorHandler setErrorHandler(ErrorHandler h) {
  ErrorHandler old = errorHandlerSlot();
  errorHandlerSlot() = h ? h : defaultErrorHandler;
  return old;
}

namespace detail {
inline void check(gbr_status st, const char* what) {
  if (st == GBR_OK) return;
  std::string msg = std::string(what) + ": " + gbr_last_error();
  errorHandlerSlot()(msg.c_str());
}
}  // namespace detail

/* ---- flags (Renderer.h:20-30,120-173; RendererStandalone.h:21-40) ----------- */

enum class RendererFlag : uint32_t { NoTextures = GBR_FLAG_NO_TEXTURES };
typedef uint32_t RendererFlags;
inline RendererFlags operator|(RendererFlag a, RendererFlag b) { return (uint32_t)a | (uint32_t)b; }

enum class RendererFileFlag : uint32_t { Whole = GBR_FILE_WHOLE, GenerateMipmap = GBR_FILE_GENERATE_MIPMAP };
struct RendererFileFlags {
  uint32_t bits = 0;
  RendererFileFlags() = default;
  RendererFileFlags(RendererFileFlag f) : bits{(uint32_t)f} {}
  explicit RendererFileFlags(uint32_t b) : bits{b} {}
};
inline RendererFileFlags operator|(RendererFileFlag a, RendererFileFlag b) {
  return RendererFileFlags{(uint32_t)a | (uint32_t)b};
}

enum class RendererLightType : int32_t { Directional = GBR_LIGHT_DIRECTIONAL, Point = GBR_LIGHT_POINT };

enum class RendererStandaloneFlag : uint32_t { QuietLog = GBR_STANDALONE_QUIET_LOG };
typedef uint32_t RendererStandaloneFlags;

/* ---- configurations --------------------------------------------------------- */

/* RendererConfiguration (Renderer.h:51-113, defaults Renderer.cpp:58-64) */
struct RendererConfiguration {
  gbr_config c;
  RendererConfiguration() { detail::check(gbr_config_init(&c), "RendererConfiguration"); }
  RendererConfiguration& setFlags(RendererFlags flags) {
    c.flags = flags;
    return *this;
  }
  RendererConfiguration& setTileSizeCount(const Vector2i& tileSize, const Vector2i& tileCount) {
    c.tile_size[0] = tileSize.x;
    c.tile_size[1] = tileSize.y;
    c.tile_count[0] = tileCount.x;
    c.tile_count[1] = tileCount.y;
    return *this;
  }
  RendererConfiguration& setMaxLightCount(unsigned count) {
    c.max_light_count = count;
    return *this;
  }
  RendererConfiguration& setAmbientFactor(float factor) {
    c.ambient_factor = factor;
    return *this;
  }
  /* new: output planes (GBR_PLANE_*) and their row order */
  RendererConfiguration& setPlanes(uint32_t planes, bool topDown = false) {
    c.planes = planes;
    c.top_down = topDown ? 1u : 0u;
    return *this;
  }
};

/* RendererStandaloneConfiguration (RendererStandalone.h:54-80) */
struct RendererStandaloneConfiguration {
  int32_t device = -1;
  RendererStandaloneFlags flags = 0;
  RendererStandaloneConfiguration& setCudaDevice(unsigned id) {
    device = (int32_t)id;
    return *this;
  }
  RendererStandaloneConfiguration& setFlags(RendererStandaloneFlags f) {
    flags = f;
    return *this;
  }
};

/* SceneStats (Renderer.h:750-774) */
struct SceneStats {
  std::size_t nodeCount;
  std::size_t drawCount;
  std::size_t drawBatchCount;
};

/* ---- Renderer (Renderer.h:437-743) ------------------------------------------ */

class Renderer {
 public:
  explicit Renderer(const RendererConfiguration& configuration)
      : Renderer{configuration, RendererStandaloneConfiguration{}} {}
  Renderer(const Renderer&) = delete;
  Renderer& operator=(const Renderer&) = delete;
  virtual ~Renderer() {
    if (r_) (void)gbr_destroy(r_);
  }

  RendererFlags flags() const { return config().flags; }
  Vector2i tileSize() const {
    const gbr_config c = config();
    return {c.tile_size[0], c.tile_size[1]};
  }
  Vector2i tileCount() const {
    const gbr_config c = config();
    return {c.tile_count[0], c.tile_count[1]};
  }
  std::size_t sceneCount() const {
    uint64_t n = 0;
    detail::check(gbr_scene_count(r_, &n), "sceneCount()");
    return (std::size_t)n;
  }
  unsigned maxLightCount() const { return config().max_light_count; }

  /* addFile(filename, flags, name) -> false + message when the file can't be
     imported or its names conflict (Renderer.h:517-556) */
  bool addFile(const std::string& filename, RendererFileFlags flags = {},
               const std::string& name = {}) {
    const gbr_status st = gbr_add_file(r_, filename.c_str(), flags.bits, name.empty() ? nullptr : name.c_str());
    if (st == GBR_OK) return true;
    if (st == GBR_ERR_IO || st == GBR_ERR_STATE || st == GBR_ERR_NOT_FOUND) {
      std::fprintf(stderr, "gfx_batch::Renderer::addFile(): %s\n", gbr_last_error());
      return false;
    }
    detail::check(st, "addFile()");
    return false;
  }
  bool hasNodeHierarchy(const std::string& name) const {
    int32_t out = 0;
    detail::check(gbr_has_node_hierarchy(r_, name.c_str(), &out), "hasNodeHierarchy()");
    return out != 0;
  }
  std::size_t addNodeHierarchy(unsigned sceneId, const std::string& name,
                               const Matrix4& bakeTransformation = {}) {
    uint64_t id = 0;
    detail::check(gbr_add_node_hierarchy(r_, sceneId, name.c_str(), bakeTransformation.data, &id),
                  "addNodeHierarchy()");
    return (std::size_t)id;
  }
  std::size_t addEmptyNode(unsigned sceneId) {
    uint64_t id = 0;
    detail::check(gbr_add_empty_node(r_, sceneId, &id), "addEmptyNode()");
    return (std::size_t)id;
  }
  std::size_t addLight(unsigned sceneId, std::size_t nodeId, RendererLightType type) {
    uint64_t id = 0;
    detail::check(gbr_add_light(r_, sceneId, nodeId, (int32_t)type, &id), "addLight()");
    return (std::size_t)id;
  }
  void clear(unsigned sceneId) { detail::check(gbr_clear(r_, sceneId), "clear()"); }
  void clearLights(unsigned sceneId) { detail::check(gbr_clear_lights(r_, sceneId), "clearLights()"); }

  Matrix4 camera(unsigned sceneId) const {
    Matrix4 m;
    detail::check(gbr_camera(r_, sceneId, m.data), "camera()");
    return m;
  }
  Vector2 cameraDepthUnprojection(unsigned sceneId) const {
    float v[2] = {0, 0};
    detail::check(gbr_camera_depth_unprojection(r_, sceneId, v), "cameraDepthUnprojection()");
    return {v[0], v[1]};
  }
  void updateCamera(unsigned sceneId, const Matrix4& projection, const Matrix4& view) {
    detail::check(gbr_update_camera(r_, sceneId, projection.data, view.data), "updateCamera()");
  }
  /* mutable view of the scene's local node transforms, read at the next
     draw() (Renderer.h:682-683) */
  StridedArrayView1D<Matrix4> transformations(unsigned sceneId) {
    float* data = nullptr;
    uint64_t count = 0;
    detail::check(gbr_transformations(r_, sceneId, &data, &count), "transformations()");
    StridedArrayView1D<Matrix4> v;
    v.base = reinterpret_cast<Matrix4*>(data);
    v.count = (std::size_t)count;
    v.stride = sizeof(Matrix4);
    return v;
  }
  StridedArrayView1D<Color3> lightColors(unsigned sceneId) {
    float* data = nullptr;
    uint64_t count = 0, stride = 0;
    detail::check(gbr_light_colors(r_, sceneId, &data, &count, &stride), "lightColors()");
    StridedArrayView1D<Color3> v;
    v.base = reinterpret_cast<Color3*>(data);
    v.count = (std::size_t)count;
    v.stride = (std::ptrdiff_t)stride;
    return v;
  }
  StridedArrayView1D<float> lightRanges(unsigned sceneId) {
    float* data = nullptr;
    uint64_t count = 0, stride = 0;
    detail::check(gbr_light_ranges(r_, sceneId, &data, &count, &stride), "lightRanges()");
    StridedArrayView1D<float> v;
    v.base = data;
    v.count = (std::size_t)count;
    v.stride = (std::ptrdiff_t)stride;
    return v;
  }
  SceneStats sceneStats(unsigned sceneId) const {
    gbr_scene_stats_info s{};
    detail::check(gbr_scene_stats(r_, sceneId, &s), "sceneStats()");
    return {(std::size_t)s.node_count, (std::size_t)s.draw_count, (std::size_t)s.draw_batch_count};
  }

  /* the C ABI handle, for the entry points without a reference counterpart
     (gbr_set_tuning, gbr_update_cameras, gbr_plane_device_pointer, ...) */
  gbr_renderer* handle() const { return r_; }

 protected:
  Renderer(const RendererConfiguration& configuration,
           const RendererStandaloneConfiguration& standalone) {
    gbr_config c = configuration.c;
    c.device = standalone.device;
    c.standalone_flags = standalone.flags;
    detail::check(gbr_create(&c, &r_), "Renderer");
  }
  gbr_config config() const {
    gbr_config c{};
    detail::check(gbr_get_config(r_, &c), "config");
    return c;
  }
  gbr_renderer* r_ = nullptr;
};

/* ---- RendererStandalone (RendererStandalone.h:120-241) ---------------------- */

class RendererStandalone : public Renderer {
 public:
  RendererStandalone(const RendererConfiguration& configuration,
                     const RendererStandaloneConfiguration& standaloneConfiguration)
      : Renderer{configuration, standaloneConfiguration} {}

  RendererStandaloneFlags standaloneFlags() const { return config().standalone_flags; }
  PixelFormat colorFramebufferFormat() const { return PixelFormat::RGBA8Unorm; }
  PixelFormat depthFramebufferFormat() const { return PixelFormat::Depth32F; }

  /* clear + draw of every scene, asynchronous (RendererStandalone.cpp:155-159) */
  void draw() { detail::check(gbr_draw(r_), "draw()"); }

  Image2D colorImage() { return whole(colorFramebufferFormat()); }
  /* RGBA8Unorm or RGB8Unorm into a caller view of exactly the rectangle's
     size; stalls until the frame is done (RendererStandalone.cpp:168-180) */
  void colorImageInto(const Range2Di& rectangle, const MutableImageView2D& image) {
    sizeCheck(rectangle, image, "colorImageInto()");
    int32_t fmt = GBR_FORMAT_RGBA8;
    if (image.format == PixelFormat::RGB8Unorm) fmt = GBR_FORMAT_RGB8;
    else if (image.format != PixelFormat::RGBA8Unorm)
      detail::check(GBR_ERR_ARGUMENT, "colorImageInto(): expected RGBA8Unorm or RGB8Unorm");
    const int32_t rect[4] = {rectangle.min.x, rectangle.min.y, rectangle.max.x, rectangle.max.y};
    detail::check(gbr_color_image_into(r_, rect, image.data, image.rowStride, fmt, GBR_MEMORY_HOST),
                  "colorImageInto()");
  }
  Image2D depthImage() { return whole(depthFramebufferFormat()); }
  /* window depth into a Depth32F (or R32F) view (RendererStandalone.cpp:189-201) */
  void depthImageInto(const Range2Di& rectangle, const MutableImageView2D& image) {
    sizeCheck(rectangle, image, "depthImageInto()");
    if (image.format != PixelFormat::Depth32F && image.format != PixelFormat::R32F)
      detail::check(GBR_ERR_ARGUMENT, "depthImageInto(): expected Depth32F");
    const int32_t rect[4] = {rectangle.min.x, rectangle.min.y, rectangle.max.x, rectangle.max.y};
    detail::check(gbr_depth_image_into(r_, rect, static_cast<float*>(image.data), image.rowStride,
                                       GBR_MEMORY_HOST),
                  "depthImageInto()");
  }
  /* linearised framebuffer in device memory, valid until the next draw()
     (RendererStandalone.h:225-235); the depth pointer is the depth buffer's,
     not the colour one the reference's Python binding returns
     (SimBindings.cpp:537-541) */
  const void* colorCudaBufferDevicePointer() {
    const void* p = nullptr;
    detail::check(gbr_color_device_pointer(r_, &p), "colorCudaBufferDevicePointer()");
    return p;
  }
  const void* depthCudaBufferDevicePointer() {
    const void* p = nullptr;
    detail::check(gbr_depth_device_pointer(r_, &p), "depthCudaBufferDevicePointer()");
    return p;
  }

 private:
  void sizeCheck(const Range2Di& rect, const MutableImageView2D& image, const char* what) {
    if (rect.size() != image.size) detail::check(GBR_ERR_SIZE, what);
  }
  Image2D whole(PixelFormat f) {
    const Vector2i size{tileSize().x * tileCount().x, tileSize().y * tileCount().y};
    Image2D img{f, size, std::vector<unsigned char>((std::size_t)size.product() * 4)};
    MutableImageView2D view{f, size, img.data.data()};
    if (f == PixelFormat::RGBA8Unorm) colorImageInto({{0, 0}, size}, view);
    else depthImageInto({{0, 0}, size}, view);
    return img;
  }
};

/* ---- depth unprojection (DepthUnprojection.h:137-153) ----------------------- */

inline Vector2 calculateDepthUnprojection(const Matrix4& projectionMatrix) {
  float v[2] = {0, 0};
  detail::check(gbr_calculate_depth_unprojection(projectionMatrix.data, v), "calculateDepthUnprojection()");
  return {v[0], v[1]};
}

inline void unprojectDepth(const Vector2& unprojection, const StridedArrayView2D& depth) {
  const float u[2] = {unprojection.x, unprojection.y};
  detail::check(gbr_unproject_depth(u, depth.data, (int64_t)depth.rows, (int64_t)depth.cols,
                                    (int64_t)depth.rowStride),
                "unprojectDepth()");
}

}  // namespace gfx_batch
}  // namespace gbr

#ifdef GBR_ESP_NAMESPACE_ALIAS
namespace esp {
namespace gfx_batch = ::gbr::gfx_batch;
}
#endif

#endif /* GBR_HPP */
