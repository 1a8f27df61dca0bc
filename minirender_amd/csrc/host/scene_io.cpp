// Scene ingestion: JSON v1 document, OBJ meshes, transforms and flattening (see scene_io.hpp).
#include "scene_io.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iterator>
#include <limits>
#include <set>
#include <sstream>
#include <stdexcept>

#include "json.hpp"

namespace mrg_host {
namespace {

using V3 = std::array<float, 3>;

void require(bool ok, const std::string& message) {
    if (!ok) throw std::invalid_argument(message);
}

bool finite3(const V3& v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
V3   sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3   cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]};
}
float dot(const V3& a, const V3& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
float length(const V3& v) { return std::sqrt(dot(v, v)); }

std::string read_file(const std::filesystem::path& path, const char* what) {
    std::ifstream stream(path, std::ios::binary);
    if (!stream) throw std::runtime_error(std::string("Cannot open ") + what + ": " + path.generic_string());
    return std::string((std::istreambuf_iterator<char>(stream)), {});
}

// ------------------------------------------------------------------------------------------
// OBJ. The reference parses with tinyobjloader v2.0.0rc13 (triangulate = true). Restated here:
// "v x y z", "vn x y z", "vt u v" records; "f" corners v, v/vt, v//vn or v/vt/vn with 1-based or
// negative (relative) indices; other records (o, g, s, usemtl, mtllib, l, p) do not add triangles.
// A quad is split along its shorter diagonal; larger polygons are fanned from their first corner.

struct FaceIndex {
    int v = -1, vt = -1, vn = -1;
};

int resolve(long value, size_t count, size_t line) {
    long index = value > 0 ? value - 1 : long(count) + value;
    if (value == 0 || index < 0 || index >= long(count))
        throw std::runtime_error("OBJ parse error: index out of range on line " + std::to_string(line));
    return int(index);
}

FaceIndex parse_corner(const std::string& token, size_t nv, size_t nt, size_t nn, size_t line) {
    FaceIndex f;
    std::string parts[3];
    int         k = 0;
    for (char c : token) {
        if (c == '/') {
            if (++k > 2) throw std::runtime_error("OBJ parse error: bad face corner on line " + std::to_string(line));
        } else {
            parts[k] += c;
        }
    }
    auto num = [&](const std::string& s) {
        size_t used = 0;
        long   v    = 0;
        try {
            v = std::stol(s, &used);
        } catch (const std::exception&) {
            throw std::runtime_error("OBJ parse error: bad index on line " + std::to_string(line));
        }
        if (used != s.size()) throw std::runtime_error("OBJ parse error: bad index on line " + std::to_string(line));
        return v;
    };
    if (parts[0].empty()) throw std::runtime_error("OBJ parse error: face corner without position on line " + std::to_string(line));
    f.v = resolve(num(parts[0]), nv, line);
    if (!parts[1].empty()) f.vt = resolve(num(parts[1]), nt, line);
    if (!parts[2].empty()) f.vn = resolve(num(parts[2]), nn, line);
    return f;
}

}  // namespace

Mesh load_obj(const std::filesystem::path& path) {
    const std::string text = read_file(path, "OBJ");
    std::vector<float> vertices, normals, texcoords;
    std::vector<std::array<FaceIndex, 3>> faces;
    std::istringstream in(text);
    in.imbue(std::locale::classic());
    std::string raw;
    size_t      line_no = 0;
    while (std::getline(in, raw)) {
        ++line_no;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        std::istringstream ls(raw);
        ls.imbue(std::locale::classic());
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        auto floats = [&](std::vector<float>& out, int n, int required) {
            int got = 0;
            for (; got < n; ++got) {
                float v;
                if (!(ls >> v)) break;
                out.push_back(v);
            }
            if (got < required) throw std::runtime_error("OBJ parse error: short record on line " + std::to_string(line_no));
            for (; got < n; ++got) out.push_back(0.0f);
        };
        if (key == "v") {
            floats(vertices, 3, 3);
        } else if (key == "vn") {
            floats(normals, 3, 3);
        } else if (key == "vt") {
            floats(texcoords, 2, 1);
        } else if (key == "f") {
            std::vector<FaceIndex> poly;
            std::string            token;
            while (ls >> token) poly.push_back(parse_corner(token, vertices.size() / 3, texcoords.size() / 2, normals.size() / 3, line_no));
            if (poly.size() < 3) throw std::runtime_error("OBJ parse error: face with fewer than 3 corners on line " + std::to_string(line_no));
            if (poly.size() == 3) {
                faces.push_back({poly[0], poly[1], poly[2]});
            } else if (poly.size() == 4) {
                auto p   = [&](const FaceIndex& f) { return V3{vertices[f.v * 3], vertices[f.v * 3 + 1], vertices[f.v * 3 + 2]}; };
                auto d02 = sub(p(poly[2]), p(poly[0]));
                auto d13 = sub(p(poly[3]), p(poly[1]));
                if (dot(d02, d02) <= dot(d13, d13)) {
                    faces.push_back({poly[0], poly[1], poly[2]});
                    faces.push_back({poly[0], poly[2], poly[3]});
                } else {
                    faces.push_back({poly[0], poly[1], poly[3]});
                    faces.push_back({poly[1], poly[2], poly[3]});
                }
            } else {
                for (size_t k = 1; k + 1 < poly.size(); ++k) faces.push_back({poly[0], poly[k], poly[k + 1]});
            }
        }
    }
    if (vertices.size() / 3 > std::numeric_limits<uint32_t>::max())
        throw std::runtime_error("OBJ exceeds supported position index range");
    Mesh mesh;
    for (size_t i = 0; i + 2 < vertices.size(); i += 3) mesh.positions.push_back({vertices[i], vertices[i + 1], vertices[i + 2]});
    // obj_loader.cpp:38-88: corner normals normalized (zero rejected), missing normals from the face
    for (const auto& face : faces) {
        std::array<Corner, 3> tri;
        std::array<bool, 3>   needs_normal{};
        for (int k = 0; k < 3; ++k) {
            Corner& c  = tri[k];
            c.position = uint32_t(face[k].v);
            if (face[k].vn >= 0) {
                const size_t n = size_t(face[k].vn) * 3;
                c.normal       = {normals[n], normals[n + 1], normals[n + 2]};
                const float l  = length(c.normal);
                if (!(l > 1e-10f)) throw std::runtime_error("OBJ normal is zero");
                c.normal = {c.normal[0] / l, c.normal[1] / l, c.normal[2] / l};
            } else {
                needs_normal[k] = true;
            }
            if (face[k].vt >= 0) c.uv = {texcoords[size_t(face[k].vt) * 2], texcoords[size_t(face[k].vt) * 2 + 1]};
        }
        const V3 a = mesh.positions[tri[0].position], b = mesh.positions[tri[1].position], c = mesh.positions[tri[2].position];
        const V3 n = cross(sub(b, a), sub(c, a));
        const float l = length(n);
        if (!(l > 1e-10f)) throw std::runtime_error("OBJ has a degenerate triangle");
        const float inv = 1.0f / std::sqrt(dot(n, n));
        for (int k = 0; k < 3; ++k)
            if (needs_normal[k]) tri[k].normal = {n[0] * inv, n[1] * inv, n[2] * inv};
        mesh.triangles.push_back(tri);
    }
    // MeshData::validate (scene.cpp:38-66)
    require(!mesh.positions.empty() && !mesh.triangles.empty(), "Mesh must contain positions and triangles");
    for (const auto& p : mesh.positions) require(finite3(p), "Mesh position must be finite");
    for (const auto& t : mesh.triangles) {
        for (const auto& c : t) {
            require(finite3(c.normal) && dot(c.normal, c.normal) > 1e-12f, "Mesh normal must be finite and nonzero");
            require(std::isfinite(c.uv[0]) && std::isfinite(c.uv[1]), "Mesh UV must be finite");
        }
        const V3 area = cross(sub(mesh.positions[t[1].position], mesh.positions[t[0].position]),
                              sub(mesh.positions[t[2].position], mesh.positions[t[0].position]));
        require(dot(area, area) > 1e-20f, "Degenerate mesh triangle");
    }
    return mesh;
}

// ------------------------------------------------------------------------------------------
// Transforms. glm 1.0.3 (the reference's math library) restated for the operations used by
// Transform::matrix: translate(I, t) puts t in column 3; rotate(m, a, axis) multiplies m by the
// axis-angle matrix built from c = cos(a), s = sin(a), temp = (1 - c) * axis (glm's column order);
// scale(m, s) scales columns 0..2. Angles are degrees * 0.01745329251994329576923690768489 in
// float. Bitwise agreement with glm also depends on the platform's cosf/sinf ("parity unpinned").

namespace {

using M4 = std::array<std::array<float, 4>, 4>;  // m[column][row]

M4 rotate(const M4& m, float angle, const V3& axis_in) {
    const float c = std::cos(angle), s = std::sin(angle);
    const float inv  = 1.0f / std::sqrt(dot(axis_in, axis_in));
    const V3    axis = {axis_in[0] * inv, axis_in[1] * inv, axis_in[2] * inv};
    const V3    temp = {(1.0f - c) * axis[0], (1.0f - c) * axis[1], (1.0f - c) * axis[2]};
    float r[3][3];
    r[0][0] = c + temp[0] * axis[0];
    r[0][1] = temp[0] * axis[1] + s * axis[2];
    r[0][2] = temp[0] * axis[2] - s * axis[1];
    r[1][0] = temp[1] * axis[0] - s * axis[2];
    r[1][1] = c + temp[1] * axis[1];
    r[1][2] = temp[1] * axis[2] + s * axis[0];
    r[2][0] = temp[2] * axis[0] + s * axis[1];
    r[2][1] = temp[2] * axis[1] - s * axis[0];
    r[2][2] = c + temp[2] * axis[2];
    M4 out;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 4; ++row)
            out[col][row] = (m[0][row] * r[col][0] + m[1][row] * r[col][1]) + m[2][row] * r[col][2];
    out[3] = m[3];
    return out;
}

}  // namespace

void Transform::validate() const {
    require(finite3(translation) && finite3(rotation_degrees) && finite3(scale), "Transform must be finite");
    require(std::abs(scale[0]) > 1e-6f && std::abs(scale[1]) > 1e-6f && std::abs(scale[2]) > 1e-6f,
            "Transform scale must be invertible");
}

std::array<std::array<float, 4>, 4> Transform::matrix() const {
    validate();
    M4 m{};
    for (int i = 0; i < 4; ++i) m[i][i] = 1.0f;
    m[3] = {translation[0], translation[1], translation[2], 1.0f};
    const float k = 0.01745329251994329576923690768489f;
    m = rotate(m, rotation_degrees[2] * k, {0.0f, 0.0f, 1.0f});
    m = rotate(m, rotation_degrees[1] * k, {0.0f, 1.0f, 0.0f});
    m = rotate(m, rotation_degrees[0] * k, {1.0f, 0.0f, 0.0f});
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 4; ++row) m[col][row] = m[col][row] * scale[col];
    return m;
}

std::array<std::array<double, 3>, 3> Transform::normal_matrix() const {
    const M4 f = matrix();
    double   m[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) m[c][r] = double(f[c][r]);
    // inverse by the adjugate (glm's mat3 inverse), then transpose
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                       m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    const double inv_det = 1.0 / det;
    double       inv[3][3];
    inv[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * inv_det;
    inv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * inv_det;
    inv[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * inv_det;
    inv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * inv_det;
    inv[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * inv_det;
    inv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * inv_det;
    inv[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * inv_det;
    inv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * inv_det;
    inv[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * inv_det;
    std::array<std::array<double, 3>, 3> out;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out[c][r] = double(float(inv[r][c]));  // transpose, stored as float
    return out;
}

// ------------------------------------------------------------------------------------------
// JSON v1 document (scene_io.cpp:47-162).

namespace {

V3 vector3(const Json& j) {
    if (!j.is_array() || j.size() != 3) throw std::invalid_argument("Expected a three-component vector");
    return {float(j[0].number()), float(j[1].number()), float(j[2].number())};
}

float number_or(const Json& obj, const char* key, float fallback) {
    const Json* v = obj.find(key);
    return v ? float(v->number()) : fallback;
}

bool bool_or(const Json& obj, const char* key, bool fallback) {
    const Json* v = obj.find(key);
    return v ? v->boolean() : fallback;
}

std::string string_or(const Json& obj, const char* key) {
    const Json* v = obj.find(key);
    return v ? v->string() : std::string();
}

uint32_t texture_filter(const std::string& name) {
    // parse_texture_filter (texture.cpp:139-151)
    if (name.empty() || name == "trilinear") return MRG_FILTER_TRILINEAR;
    if (name == "nearest") return MRG_FILTER_NEAREST;
    if (name == "bilinear") return MRG_FILTER_BILINEAR;
    throw std::invalid_argument("Unknown texture filter: " + name);
}

uint32_t material_kind(const std::string& name) {
    // scene.cpp:219-239; the GPU integrator renders the first four
    if (name == "diffuse") return 0;
    if (name == "conductor") return 1;
    if (name == "mirror") return 2;
    if (name == "dielectric") return 3;
    if (name == "cloth" || name == "subsurface")
        throw std::invalid_argument("Material kind " + name + " is not supported by the GPU integrator");
    throw std::invalid_argument("Unsupported material kind: " + name);
}

}  // namespace

SceneDocument load_scene(const std::filesystem::path& path) {
    const std::string text = read_file(path, "scene");
    try {
        const Json json = Json::parse(text);
        if (!json.at("version").is_integer() || json.at("version").number() != 1)
            throw std::invalid_argument("Unsupported scene version");
        SceneDocument doc;
        const auto    dir = std::filesystem::absolute(path).parent_path();
        const Json&   cam = json.at("camera");
        const V3      pos = vector3(cam.at("position")), tgt = vector3(cam.at("target")), up = vector3(cam.at("up"));
        std::copy(pos.begin(), pos.end(), doc.camera.position);
        std::copy(tgt.begin(), tgt.end(), doc.camera.target);
        std::copy(up.begin(), up.end(), doc.camera.up);
        doc.camera.vertical_fov_degrees = float(cam.at("vertical_fov").number());
        doc.camera.near_plane           = float(cam.at("near").number());
        doc.camera.far_plane            = float(cam.at("far").number());
        doc.camera.aperture_radius      = number_or(cam, "aperture_radius", 0.0f);
        doc.camera.focus_distance       = number_or(cam, "focus_distance", 4.0f);
        doc.environment                 = vector3(json.at("environment"));
        if (const Json* media = json.find("media"); media && media->size() != 0)
            throw std::invalid_argument("Participating media are not supported by the GPU integrator");
        if (const Json* env = json.find("environment_map")) {
            doc.environment_texture          = string_or(*env, "texture");
            doc.environment_strength         = number_or(*env, "strength", 1.0f);
            doc.environment_rotation_degrees = number_or(*env, "rotation_degrees", 0.0f);
        }
        for (const Json& v : json.at("materials").array()) {
            mrg_material m{};
            const std::string id = v.at("id").string();
            m.kind               = material_kind(v.at("kind").string());
            const V3 color       = vector3(v.at("base_color"));
            const V3 emission    = v.contains("emission") ? vector3(v.at("emission")) : V3{0.0f, 0.0f, 0.0f};
            std::copy(color.begin(), color.end(), m.base_color);
            std::copy(emission.begin(), emission.end(), m.emission);
            m.roughness    = number_or(v, "roughness", 0.35f);
            m.ior          = number_or(v, "ior", 1.5f);
            m.double_sided = bool_or(v, "double_sided", true) ? 1u : 0u;
            if (v.contains("cloth") || v.contains("subsurface"))
                throw std::invalid_argument("Material " + id + ": cloth / subsurface data is not supported by the GPU integrator");
            doc.material_textures.push_back(string_or(v, "base_color_texture"));
            doc.material_normal_maps.push_back(string_or(v, "normal_texture"));
            m.normal_scale  = number_or(v, "normal_scale", 1.0f);
            m.normal_flip_y = bool_or(v, "normal_flip_y", false) ? 1u : 0u;
            m.compensate_energy = bool_or(v, "compensate_energy", false) ? 1u : 0u;
            m.texture_filter    = texture_filter(string_or(v, "texture_filter"));
            doc.material_ids.push_back(id);
            doc.materials.push_back(m);
        }
        for (const Json& v : json.at("meshes").array()) {
            const std::string id = v.at("id").string();
            if (doc.meshes.count(id)) throw std::invalid_argument("Duplicate Mesh id: " + id);
            const std::string source = v.at("source").string();
            doc.meshes.emplace(id, load_obj(dir / std::filesystem::u8path(source)));
            doc.mesh_order.push_back(id);
        }
        if (const Json* textures = json.find("textures")) {
            for (const Json& v : textures->array()) {
                TextureAsset t;
                t.id                       = v.at("id").string();
                const std::string encoding = v.contains("encoding") ? v.at("encoding").string() : std::string("srgb");
                if (encoding != "srgb" && encoding != "linear") throw std::invalid_argument("Unknown texture encoding");
                t.image = read_texture(dir / std::filesystem::u8path(v.at("source").string()), encoding == "srgb");
                t.srgb  = encoding == "srgb";
                doc.textures.push_back(std::move(t));
            }
        }
        for (size_t i = 0; i < doc.materials.size(); ++i) {
            auto find = [&](const std::string& tex) {
                size_t k = 0;
                while (k < doc.textures.size() && doc.textures[k].id != tex) ++k;
                if (k == doc.textures.size()) throw std::invalid_argument("Material references missing texture: " + tex);
                return k;
            };
            if (!doc.material_textures[i].empty()) doc.materials[i].base_color_texture = uint32_t(find(doc.material_textures[i]) + 1);
            if (!doc.material_normal_maps[i].empty()) {
                // scene.cpp:295-300
                const size_t k = find(doc.material_normal_maps[i]);
                if (doc.textures[k].srgb) throw std::invalid_argument("Normal map must reference a linear texture");
                doc.materials[i].normal_texture = uint32_t(k + 1);
            }
        }
        for (const Json& v : json.at("objects").array()) {
            Object o;
            o.id          = v.at("id").string();
            o.name        = v.contains("name") ? v.at("name").string() : o.id;
            o.mesh_id     = v.at("mesh").string();
            o.material_id = v.at("material").string();
            o.visible     = bool_or(v, "visible", true);
            const Json& t = v.at("transform");
            o.transform.translation      = vector3(t.at("translation"));
            o.transform.rotation_degrees = vector3(t.at("rotation"));
            o.transform.scale            = vector3(t.at("scale"));
            if (v.contains("cloth_simulation"))
                throw std::invalid_argument("Object " + o.id + ": cloth simulation is not supported by the GPU integrator");
            doc.objects.push_back(o);
        }
        for (const Json& v : json.at("directional_lights").array()) {
            mrg_directional_light l{};
            const V3 d = vector3(v.at("direction")), c = vector3(v.at("color"));
            std::copy(d.begin(), d.end(), l.direction);
            std::copy(c.begin(), c.end(), l.color);
            l.intensity = float(v.at("intensity").number());
            v.at("id").string();
            doc.lights.push_back(l);
        }
        doc.validate();
        return doc;
    } catch (const std::exception& error) {
        throw std::runtime_error("Scene " + path.generic_string() + ": " + error.what());
    }
}

void SceneDocument::validate() const {
    // Camera::validate (scene.cpp:141-157)
    const V3 pos = {camera.position[0], camera.position[1], camera.position[2]};
    const V3 tgt = {camera.target[0], camera.target[1], camera.target[2]};
    const V3 up  = {camera.up[0], camera.up[1], camera.up[2]};
    require(finite3(pos) && finite3(tgt) && finite3(up), "Camera vectors must be finite");
    require(length(sub(tgt, pos)) > 1e-6f, "Camera position and target must differ");
    require(length(cross(sub(tgt, pos), up)) > 1e-6f, "Camera up must not be parallel to view direction");
    require(std::isfinite(camera.vertical_fov_degrees) && camera.vertical_fov_degrees > 0 && camera.vertical_fov_degrees < 179,
            "Camera FOV must be between 0 and 179 degrees");
    require(std::isfinite(camera.near_plane) && std::isfinite(camera.far_plane) && camera.near_plane > 0 &&
                camera.far_plane > camera.near_plane,
            "Camera clipping planes must satisfy 0 < near < far");
    require(std::isfinite(camera.aperture_radius) && camera.aperture_radius >= 0 && std::isfinite(camera.focus_distance) &&
                camera.focus_distance > 0,
            "Camera lens requires nonnegative aperture and positive focus distance");
    require(finite3(environment) && environment[0] >= 0 && environment[1] >= 0 && environment[2] >= 0,
            "Environment must be finite and nonnegative");
    std::set<std::string> texture_ids;
    for (const TextureAsset& t : textures) require(texture_ids.insert(t.id).second, "Duplicate Texture id: " + t.id);
    require(environment_texture.empty() || texture_ids.count(environment_texture) != 0,
            "Environment references missing texture: " + environment_texture);
    require(std::isfinite(environment_strength) && environment_strength >= 0 && std::isfinite(environment_rotation_degrees),
            "Environment strength and rotation must be finite; strength must be nonnegative");
    // Material::validate (scene.cpp:241-266), unique ids
    std::set<std::string> ids;
    for (size_t i = 0; i < materials.size(); ++i) {
        require(ids.insert(material_ids[i]).second, "Duplicate Material id: " + material_ids[i]);
        const mrg_material& m = materials[i];
        for (int k = 0; k < 3; ++k) {
            require(std::isfinite(m.base_color[k]) && m.base_color[k] >= 0 && m.base_color[k] <= 1, "Material base color must be in [0,1]");
            require(std::isfinite(m.emission[k]) && m.emission[k] >= 0, "Material emission must be nonnegative");
        }
        require(std::isfinite(m.roughness) && m.roughness >= 0 && m.roughness <= 1, "Roughness must be in [0,1]");
        require(std::isfinite(m.ior) && m.ior > 0, "IOR must be positive");
    }
    std::set<std::string> object_ids;
    for (const Object& o : objects) {
        require(object_ids.insert(o.id).second, "Duplicate Object id: " + o.id);
        require(meshes.count(o.mesh_id) != 0, "Object references missing mesh: " + o.mesh_id);
        require(ids.count(o.material_id) != 0, "Object references missing material: " + o.material_id);
        o.transform.validate();
    }
    for (const mrg_directional_light& l : lights) {
        const V3 d = {l.direction[0], l.direction[1], l.direction[2]};
        const V3 c = {l.color[0], l.color[1], l.color[2]};
        require(finite3(d) && length(d) > 1e-6f, "Light direction must be nonzero");
        require(finite3(c) && c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && std::isfinite(l.intensity) && l.intensity >= 0,
                "Light color/intensity must be nonnegative");
    }
}

// ------------------------------------------------------------------------------------------
// make_snapshot (scene.cpp:364-405) + the SceneBvh flattening (bvh.cpp:53-89): visible objects in
// document order; world position = vec3(M * vec4(p, 1)) in float (glm's mat4 * vec4 sums the
// column products pairwise: (c0 x + c1 y) + (c2 z + c3 w)); per-corner world normal =
// normalize(N * n) in double, stored as float; mirrored = det(M3) < 0 in double.

FlatScene flatten(const SceneDocument& doc) {
    doc.validate();
    FlatScene flat;
    flat.camera      = doc.camera;
    flat.environment = doc.environment;
    flat.lights      = doc.lights;
    for (const TextureAsset& t : doc.textures) flat.textures.push_back(t.image);
    for (size_t k = 0; k < doc.textures.size(); ++k)
        if (!doc.environment_texture.empty() && doc.textures[k].id == doc.environment_texture) flat.environment_texture = uint32_t(k + 1);
    flat.environment_strength         = doc.environment_strength;
    flat.environment_rotation_degrees = doc.environment_rotation_degrees;
    uint32_t instance = 0;
    for (const Object& o : doc.objects) {
        if (!o.visible) continue;
        const Mesh& mesh = doc.meshes.at(o.mesh_id);
        const auto  m    = o.transform.matrix();
        const auto  nm   = o.transform.normal_matrix();
        const size_t mi  = size_t(std::find(doc.material_ids.begin(), doc.material_ids.end(), o.material_id) - doc.material_ids.begin());
        mrg_material material = doc.materials[mi];
        double       m3[3][3];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) m3[c][r] = double(m[c][r]);
        const double det = m3[0][0] * (m3[1][1] * m3[2][2] - m3[2][1] * m3[1][2]) - m3[1][0] * (m3[0][1] * m3[2][2] - m3[2][1] * m3[0][2]) +
                           m3[2][0] * (m3[0][1] * m3[1][2] - m3[1][1] * m3[0][2]);
        material.mirrored = det < 0 ? 1u : 0u;
        flat.materials.push_back(material);
        for (uint32_t t = 0; t < mesh.triangles.size(); ++t) {
            for (const Corner& c : mesh.triangles[t]) {
                const auto& p = mesh.positions[c.position];
                for (int r = 0; r < 3; ++r) {
                    const float w = (m[0][r] * p[0] + m[1][r] * p[1]) + (m[2][r] * p[2] + m[3][r] * 1.0f);
                    if (!std::isfinite(w)) throw std::invalid_argument("Transformed triangle position is not finite");
                    flat.positions.push_back(w);
                }
                double n[3];
                for (int r = 0; r < 3; ++r)
                    n[r] = (nm[0][r] * double(c.normal[0]) + nm[1][r] * double(c.normal[1])) + nm[2][r] * double(c.normal[2]);
                const double inv = 1.0 / std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
                for (int r = 0; r < 3; ++r) flat.normals.push_back(float(n[r] * inv));
                flat.uvs.push_back(c.uv[0]);
                flat.uvs.push_back(c.uv[1]);
            }
            flat.instance.push_back(instance);
            flat.triangle.push_back(t);
        }
        ++instance;
    }
    return flat;
}

mrg_scene_desc FlatScene::desc() {
    texture_views.clear();
    for (const RgbImage& t : textures) texture_views.push_back(mrg_texture{t.width, t.height, t.rgb.data()});
    mrg_scene_desc d{};
    d.uvs           = uvs.empty() ? nullptr : uvs.data();
    d.texture_count = uint32_t(texture_views.size());
    d.textures      = texture_views.empty() ? nullptr : texture_views.data();
    d.environment_texture          = environment_texture;
    d.environment_strength         = environment_strength;
    d.environment_rotation_degrees = environment_rotation_degrees;
    d.triangle_count    = instance.size();
    d.positions         = positions.data();
    d.normals           = normals.data();
    d.instance          = instance.data();
    d.triangle          = triangle.data();
    d.instance_count    = uint32_t(materials.size());
    d.materials         = materials.data();
    d.environment[0]    = environment[0];
    d.environment[1]    = environment[1];
    d.environment[2]    = environment[2];
    d.directional_count = uint32_t(lights.size());
    d.directional       = lights.data();
    return d;
}

}  // namespace mrg_host
