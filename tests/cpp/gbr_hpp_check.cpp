// Exercises include/gbr.hpp, the C++ classes shaped like habitat-sim's
// esp::gfx_batch::Renderer / RendererStandalone, through libgbr.so.
//
//   gbr_hpp_check host
//     host-only entry points (no GPU needed): configuration defaults
//     (Renderer.cpp:58-64), calculateDepthUnprojection / unprojectDepth on
//     DepthUnprojectionTest-style values (DepthUnprojection.cpp:98-130), and
//     the error path: constructing a renderer on a device that does not exist
//     reaches the error handler with a message instead of crashing.
//   gbr_hpp_check gpu <scenes dir> <out file>
//     GfxBatchRendererTest::singleMesh() (GfxBatchRendererTest.cpp:1521-1585)
//     written against the C++ classes: prints its checks and writes the
//     96x128 RGBA8 colour and Depth32F images (bottom-up) to <out file> for
//     the Python side to compare with the golden PNG and the Python API.
#define GBR_ESP_NAMESPACE_ALIAS
#include "gbr.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace gb = esp::gfx_batch;

static int failures = 0;

#define CHECK(cond)                                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

struct HandlerError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

static void throwing_handler(const char* message) { throw HandlerError(message); }

// orthographic projection of a size x 2 view, near 0.1, far 10, in the same
// float operation order as Matrix4::orthographicProjection
static gb::Matrix4 ortho(float w, float h, float n, float f) {
  gb::Matrix4 m;
  m.data[0] = 2.0f / w;
  m.data[5] = 2.0f / h;
  const float z = 2.0f / (n - f);
  m.data[10] = z;
  m.data[14] = n * z - 1.0f;
  return m;
}

static int host_checks() {
  gb::RendererConfiguration cfg;
  CHECK(cfg.c.tile_size[0] == 128 && cfg.c.tile_size[1] == 128);
  CHECK(cfg.c.tile_count[0] == 1 && cfg.c.tile_count[1] == 1);
  CHECK(cfg.c.max_light_count == 0);
  CHECK(cfg.c.flags == 0);
  CHECK(std::fabs(cfg.c.ambient_factor - 0.1f) < 1e-7f);
  cfg.setTileSizeCount({64, 32}, {3, 2}).setMaxLightCount(2).setAmbientFactor(0.25f)
      .setFlags((gb::RendererFlags)gb::RendererFlag::NoTextures);
  CHECK(cfg.c.tile_size[0] == 64 && cfg.c.tile_count[1] == 2 && cfg.c.max_light_count == 2);
  CHECK(cfg.c.flags == GBR_FLAG_NO_TEXTURES);

  // calculateDepthUnprojection: (P[2][2] - 1) / 2 and P[3][2] / 2
  const gb::Matrix4 P = ortho(2.0f, 2.0f, 0.1f, 10.0f);
  const gb::Vector2 u = gb::calculateDepthUnprojection(P);
  CHECK(u.x == (P.data[10] - 1.0f) * 0.5f);
  CHECK(u.y == P.data[14] * 0.5f);
  // unprojectDepth: z = b / (d + a), the far-plane value -> 0
  std::vector<float> depth = {0.0f, 0.5f, 1.0f, 0.25f, 0.75f, 1.0f};
  const gb::Vector2 pu{-0.5f, 2.0f};
  std::vector<float> expect(depth.size());
  for (std::size_t i = 0; i < depth.size(); ++i) {
    float z = pu.y / (depth[i] + pu.x);
    if (z == pu.y / (1.0f + pu.x)) z = 0.0f;
    expect[i] = z;
  }
  gb::unprojectDepth(pu, {depth.data(), 2, 3, (std::ptrdiff_t)(3 * sizeof(float))});
  for (std::size_t i = 0; i < depth.size(); ++i) CHECK(depth[i] == expect[i]);

  // misuse reaches the handler with the library's message
  gb::setErrorHandler(throwing_handler);
  bool raised = false;
  try {
    gb::RendererStandalone r{gb::RendererConfiguration{}, gb::RendererStandaloneConfiguration{}.setCudaDevice(4096)};
  } catch (const HandlerError& e) {
    raised = std::strlen(e.what()) > 0;
    std::printf("error path: %s\n", e.what());
  }
  CHECK(raised);
  gb::setErrorHandler(nullptr);
  std::printf("host checks: %d failures\n", failures);
  return failures ? 1 : 0;
}

static int gpu_single_mesh(const std::string& scenes, const std::string& out) {
  gb::setErrorHandler(throwing_handler);
  gb::RendererStandalone renderer{gb::RendererConfiguration{}.setTileSizeCount({128, 96}, {1, 1}),
                                  gb::RendererStandaloneConfiguration{}.setFlags(
                                      (gb::RendererStandaloneFlags)gb::RendererStandaloneFlag::QuietLog)};
  CHECK(renderer.tileSize() == gb::Vector2i(128, 96));
  CHECK(renderer.tileCount() == gb::Vector2i(1, 1));
  CHECK(renderer.sceneCount() == 1);
  CHECK(renderer.colorFramebufferFormat() == gb::PixelFormat::RGBA8Unorm);
  CHECK(renderer.depthFramebufferFormat() == gb::PixelFormat::Depth32F);
  CHECK(renderer.addFile(scenes + "/batch.gltf"));
  CHECK(!renderer.addFile(scenes + "/does-not-exist.gltf"));
  CHECK(renderer.hasNodeHierarchy("square"));
  CHECK(!renderer.hasNodeHierarchy("squares"));
  CHECK(!renderer.hasNodeHierarchy(""));
  // camera: orthographic 4:3 view of [-4/3, 4/3] x [-1, 1], one unit back
  const float w = (float)(2.0 * 4.0 / 3.0);
  renderer.updateCamera(0, ortho(w, 2.0f, 0.1f, 10.0f), gb::Matrix4::translation(0.0f, 0.0f, -1.0f));
  gb::Matrix4 bake;
  bake.data[0] = bake.data[5] = bake.data[10] = 0.4f;
  CHECK(renderer.addNodeHierarchy(0, "square", bake) == 0);
  const gb::SceneStats st = renderer.sceneStats(0);
  CHECK(st.nodeCount == 2 && st.drawCount == 1 && st.drawBatchCount == 1);
  auto t = renderer.transformations(0);
  CHECK(t.size() == 2);
  CHECK(std::memcmp(t[0].data, gb::Matrix4{}.data, sizeof(gb::Matrix4)) == 0);
  gb::Matrix4 two;
  two.data[0] = two.data[5] = two.data[10] = 2.0f;
  t[0] = two;
  renderer.draw();
  const gb::Image2D color = renderer.colorImage();
  const gb::Image2D depth = renderer.depthImage();
  CHECK(color.size == gb::Vector2i(128, 96));
  // Magnum images are bottom-up: pixels()[75][35] is row 75 from the bottom
  const uint8_t* c = color.pixels<uint8_t>() + (75 * 128 + 35) * 4;
  CHECK(c[0] == 0xcc && c[1] == 0xcc && c[2] == 0xcc && c[3] == 0xff);
  const uint8_t* c2 = color.pixels<uint8_t>() + (20 * 128 + 38) * 4;
  CHECK(c2[0] == 0x99 && c2[1] == 0x00 && c2[2] == 0x00 && c2[3] == 0xff);
  const float* d = depth.pixels<float>();
  CHECK(d[0] == 1.0f && d[95 * 128 + 127] == 1.0f);
  CHECK(std::fabs(d[64 * 128 + 48] - 0.0909091f) < 1e-6f);
  // colorImageInto() of a sub-rectangle into RGB8 with padded rows
  std::vector<uint8_t> rgb(10 * 40, 0);
  renderer.colorImageInto({{30, 70}, {40, 80}}, gb::MutableImageView2D{gb::PixelFormat::RGB8Unorm, {10, 10}, rgb.data(), 40});
  CHECK(rgb[(75 - 70) * 40 + (35 - 30) * 3] == 0xcc);
  bool size_error = false;
  try {
    renderer.colorImageInto({{0, 0}, {4, 4}}, gb::MutableImageView2D{gb::PixelFormat::RGBA8Unorm, {3, 4}, rgb.data()});
  } catch (const HandlerError&) {
    size_error = true;
  }
  CHECK(size_error);
  CHECK(renderer.colorCudaBufferDevicePointer() != nullptr);
  CHECK(renderer.depthCudaBufferDevicePointer() != nullptr);
  CHECK(renderer.depthCudaBufferDevicePointer() != renderer.colorCudaBufferDevicePointer());
  FILE* f = std::fopen(out.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "can't write %s\n", out.c_str());
    return 2;
  }
  std::fwrite(color.data.data(), 1, color.data.size(), f);
  std::fwrite(depth.data.data(), 1, depth.data.size(), f);
  std::fclose(f);
  std::printf("gpu single mesh: %d failures\n", failures);
  return failures ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "host") return host_checks();
  if (argc >= 4 && std::string(argv[1]) == "gpu") return gpu_single_mesh(argv[2], argv[3]);
  std::fprintf(stderr, "usage: %s host | gpu <scenes dir> <out file>\n", argv[0]);
  return 2;
}
