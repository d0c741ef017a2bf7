// Native import of batch-friendly glTF composites (glTF 2.0 / glb with
// MAGNUMX_mesh_views extras, KHR_texture_ktx, KHR_texture_transform,
// KHR_materials_unlit; KTX2 and PNG images) into the parsed-file description
// gbr_add_file_data() consumes.  The conventions are those of the reference's
// Magnum importers as Renderer::addFile() sees them (src/esp/gfx_batch/
// Renderer.cpp:354-867): images stored bottom row first, texture coordinate
// v flipped, texture transforms conjugated by the same flip, fans and strips
// as GL triangle order.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "gbr.h"

namespace gbr {
namespace imp {

struct Mesh {
  std::vector<float> positions;   // [V][3]
  std::vector<float> texcoords;   // [V][2] or empty
  std::vector<float> colors;      // [V][4] or empty
  std::vector<float> normals;     // [V][3] or empty
  std::vector<uint32_t> indices;  // triangle list
  uint32_t index_size = 4;        // bytes per index in the file
};

struct Texture {
  uint32_t width = 0, height = 0, layers = 1;
  std::vector<std::vector<uint8_t>> levels;  // RGBA8 [layers][h][w], bottom row first
  uint32_t min_filter = 9987, mag_filter = 9729, wrap_s = 10497, wrap_t = 10497;
  bool compressed = false;
};

struct Material {
  bool flat = false;
  float base_color[4] = {1, 1, 1, 1};
  int32_t texture = -1;
  uint32_t layer = 0;
  float texture_matrix[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // column-major
};

struct Node {
  int32_t parent = -1;
  float transform[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  bool has_name = false;
  std::string name;
};

struct Entry {
  uint32_t node = 0, mesh = 0;
  int32_t material = -1;
  bool has_view = false;
  uint32_t view_offset = 0, view_count = 0;
  int32_t view_material = -1;
};

struct File {
  std::vector<Mesh> meshes;
  std::vector<Texture> textures;
  std::vector<Material> materials;
  std::vector<Node> nodes;
  std::vector<Entry> entries;
  bool has_scene = true;

  // descriptors pointing into the vectors above (valid while *this lives)
  std::vector<gbr_mesh_desc> mesh_desc;
  std::vector<std::vector<uint8_t>> texel_blobs;
  std::vector<gbr_texture_desc> texture_desc;
  std::vector<gbr_material_desc> material_desc;
  std::vector<gbr_node_desc> node_desc;
  std::vector<gbr_mesh_entry_desc> entry_desc;
  gbr_file_desc desc() ;
};

// Reads and parses a .gltf/.glb file; on failure returns false with a
// message in err (addFile() returning false).
bool load_gltf(const std::string& path, File& out, std::string& err);

// Image decoders, bottom row first, RGBA8 (exposed for tests of the library).
bool decode_png(const uint8_t* data, size_t size, uint32_t& w, uint32_t& h,
                std::vector<uint8_t>& rgba, std::string& err);

}  // namespace imp
}  // namespace gbr
