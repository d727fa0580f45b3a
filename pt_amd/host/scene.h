// Host scene model: what the reference's SceneBuilder/Scene produce (src/scene/scene.h:13-66),
// the glTF subset loader that stands in for Assimp, and the packed buffers the device
// back end consumes (src/scene/scenebuffer.cpp, src/scene/lightsample.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mathx.h"
#include "pthip.h"

namespace pth {

// share/constants.h:6-19
enum : uint32_t {
    BASE_COLOUR_MAP_BIT = 1u << 0,
    EMISSION_MAP_BIT = 1u << 1,
    ROUGHNESS_MAP_BIT = 1u << 2,
    METALNESS_MAP_BIT = 1u << 3,
    ROUGHNESS_METALNESS_MAP_BIT = 1u << 4,
    SPECULAR_MAP_BIT = 1u << 5,
    SHEEN_MAP_BIT = 1u << 7,
    CLEARCOAT_MAP_BIT = 1u << 9,
    NORMAL_MAP_BIT = 1u << 12,
    MATERIAL_IS_TRANSPARENT_BIT = 1u << 13,
};
enum : uint32_t { LIGHT_TYPE_POINT = 1, LIGHT_TYPE_DIRECTIONAL = 2, LIGHT_TYPE_EMITTER = 3 };

struct Vertex {  // 32 B, src/scene/mesh.h:10-14
    float pos[3];
    float normal[3];
    float uv[2];
};
static_assert(sizeof(Vertex) == 32, "vertex layout");

struct Tri {
    uint32_t i[3];
};

struct Mesh {
    std::vector<Vertex> vertices;
    std::vector<Tri> tris;
};

struct Material {  // 96 B, src/scene/material.h:24-29
    float base_colour[4] = {0, 0, 0, 0};
    float emission[4] = {0, 0, 0, 0};
    float rough_metal_normal_bits[4] = {0, 0, 0, 0};
    float specular[4] = {0, 0, 0, 0};
    float sheen[4] = {0, 0, 0, 0};
    float cc_ccrgh_aniso[4] = {0, 0, 0, 0};

    uint32_t map_bits() const;
    void set_map_bits(uint32_t bits);
    bool is_emitter() const;  // src/scene/material.cpp:3-8
    bool is_transparent() const { return (map_bits() & MATERIAL_IS_TRANSPARENT_BIT) != 0; }
};
static_assert(sizeof(Material) == 96, "material layout");

struct Variant {  // ObjectVariant, src/scene/scene.h:18-23
    uint32_t mesh = 0;
    uint32_t material = 0;
    std::vector<M4> instances;
};

struct PointLight {
    V3 pos, col;
};
struct DirLight {
    V3 dir, col;
};

struct CameraSection {  // parse_camera defaults, src/scene/scenebuild.cpp:138-144
    V3 position{0, 0, 0};
    V3 lookdir{0, 0, -1};
    V3 up{0, 1, 0};
    float fov = 45.0f;
    float near_clip = 0.1f, far_clip = 1000.0f;
    float focus_dist = 0.0f, lens_radius = 0.0f;
    float exposure = 1.0f;
};

// A decoded texture (TextureData, src/scene/texturedata.cpp): 8-bit channels, rows top to
// bottom; srgb as the first material slot that used it decided (scene.cpp:42-72).
struct Texture {
    uint32_t width = 0, height = 0, channels = 0;
    bool srgb = false;
    std::vector<uint8_t> px;
    // TextureData::colour_at (texturedata.cpp:96-121): raw channels / 255, no sRGB decode.
    void colour_at(uint32_t x, uint32_t y, float out[4]) const;
    bool has_alpha_less_than_one() const;  // texturedata.cpp:138-154
};

struct Scene {
    std::vector<Texture> textures;
    std::vector<Variant> variants;
    std::vector<Mesh> meshes;
    std::vector<Material> materials;
    std::vector<PointLight> point_lights;
    std::vector<DirLight> dir_lights;
    V3 envcol{0, 0, 0};
    CameraSection camera;
    std::string resource_directory;
};

// Loads one glTF 2.0 file (.gltf with external or data-URI buffers, or .glb) into meshes,
// materials and variants the way Scene::from_file's Assimp import presents them
// (src/scene/scene.cpp:356-488): one mesh per primitive, one variant per mesh, instances
// from the node hierarchy.
Scene load_gltf(const std::string& path);

// SceneBuilder::read_scene_description (src/scene/scenebuild.cpp:279-365).
Scene load_scn(const std::string& path);

// Packed scene: SceneBuffer::build (src/scene/scenebuffer.cpp:65-168) + TableLightSampler.
struct PackedScene {
    std::vector<Vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<uint8_t> instances;  // 128 B records
    std::vector<Material> materials;
    std::vector<uint32_t> light_words;
    uint32_t light_count = 0;
    std::vector<pt_mesh_desc> meshes;
    std::vector<uint32_t> instance_mesh;
    float envcol[4] = {0, 0, 0, 0};
    std::vector<pt_texture_desc> textures;
    std::vector<uint8_t> texels;
    pt_scene_blob blob{};
    void bind_blob();
};

PackedScene pack_scene(const Scene& scene);

// The light sampler words alone (TableLightSampler ctor + copy, lightsample.cpp:89-221).
std::vector<uint32_t> build_light_words(const Scene& scene, const std::vector<uint32_t>& instance_offsets,
                                        uint32_t* light_count);

// AliasTable<T>::build (src/aliastable.h:21-67) over fixed-size values of `value_words`
// words each: bins of {p, q, alias, value...}.
std::vector<uint32_t> build_alias_table(const std::vector<float>& weights,
                                        const std::vector<uint32_t>& values, uint32_t value_words);

float from_srgb(float x);
float to_srgb(float x);

// Camera: basis of scenebuild.cpp:184-194, matrices of camera.h:16 / camera.cpp:21-25,
// uniform content of pathtrace.cpp:381-396.
void camera_block(const CameraSection& cam, float aspect, pt_camera* out);

}  // namespace pth
