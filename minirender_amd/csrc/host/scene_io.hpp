// Scene ingestion for the GPU path tracer's command line: the JSON v1 scene document
// (src/assets/scene_io.cpp:47-162), Wavefront OBJ meshes (src/assets/obj_loader.cpp:14-93), the
// object transforms (src/scene/scene.cpp:119-139) and the snapshot + SceneBvh flattening
// (scene.cpp:364-405, src/render/bvh.cpp:53-89) into the arrays of mrg_scene_desc.
//
// Base-color, normal-map and environment textures are decoded from PNG (image_io.hpp). Features of
// the document that the GPU integrator does not render (participating media, cloth and subsurface
// materials) are rejected with std::invalid_argument naming the feature, instead of being dropped
// silently.
#pragma once

#include <array>
#include <cstdint>
#include <filesystem>
#include <map>
#include <string>
#include <vector>

#include "image_io.hpp"
#include "mrg.h"

namespace mrg_host {

struct Corner {
    uint32_t             position = 0;
    std::array<float, 3> normal{0.0f, 0.0f, 0.0f};
    std::array<float, 2> uv{0.0f, 0.0f};
};

// MeshData (scene.hpp:33-60): shared positions, per-corner normals and UVs.
struct Mesh {
    std::vector<std::array<float, 3>>   positions;
    std::vector<std::array<Corner, 3>> triangles;
};

// Transform (scene.hpp:62-80): translation, XYZ Euler angles in degrees, scale.
struct Transform {
    std::array<float, 3> translation{0.0f, 0.0f, 0.0f};
    std::array<float, 3> rotation_degrees{0.0f, 0.0f, 0.0f};
    std::array<float, 3> scale{1.0f, 1.0f, 1.0f};
    void validate() const;
    // T * Rz * Ry * Rx * S in float, column-major m[column][row] (scene.cpp:126-133).
    std::array<std::array<float, 4>, 4> matrix() const;
    // transpose(inverse(upper 3x3 of matrix())) computed in double (scene.cpp:135-139).
    std::array<std::array<double, 3>, 3> normal_matrix() const;
};

struct Object {
    std::string id, name, mesh_id, material_id;
    bool        visible = true;
    Transform   transform;
};

// TextureAsset (scene.hpp:146-156) with its decoded level-0 image.
struct TextureAsset {
    std::string id;
    RgbImage    image;
    bool        srgb = true;  // ColorEncoding of the source
};

struct SceneDocument {
    mrg_camera                              camera{};
    std::array<float, 3>                    environment{0.0f, 0.0f, 0.0f};
    std::vector<std::string>                material_ids;
    std::vector<mrg_material>               materials;  // base_color_texture resolved against `textures`
    std::vector<std::string>                material_textures;  // base_color_texture id per material ("" = none)
    std::vector<std::string>                material_normal_maps;  // normal_texture id per material ("" = none)
    std::vector<TextureAsset>               textures;
    std::string                             environment_texture;  // EnvironmentMap (scene.hpp:188-193)
    float                                   environment_strength = 1.0f;
    float                                   environment_rotation_degrees = 0.0f;
    std::map<std::string, Mesh>             meshes;
    std::vector<std::string>                mesh_order;
    std::vector<Object>                     objects;
    std::vector<mrg_directional_light>      lights;
    void validate() const;
};

// Flattened snapshot: every visible object in document order, each triangle's corners in world
// space (instance-major, triangle-minor = the reference's primitive order).
struct FlatScene {
    std::vector<float>                 positions;  // [n][3][3]
    std::vector<float>                 normals;    // [n][3][3]
    std::vector<float>                 uvs;        // [n][3][2]
    std::vector<RgbImage>              textures;
    std::vector<mrg_texture>           texture_views;  // filled by desc()
    std::vector<uint32_t>              instance, triangle;
    std::vector<mrg_material>          materials;  // per instance
    std::vector<mrg_directional_light> lights;
    std::array<float, 3>               environment{0.0f, 0.0f, 0.0f};
    uint32_t                           environment_texture = 0;  // 1 + texture index, 0 = none
    float                              environment_strength = 1.0f, environment_rotation_degrees = 0.0f;
    mrg_camera                         camera{};
    mrg_scene_desc                     desc();
};

Mesh          load_obj(const std::filesystem::path& path);
SceneDocument load_scene(const std::filesystem::path& path);
FlatScene     flatten(const SceneDocument& document);

}  // namespace mrg_host
