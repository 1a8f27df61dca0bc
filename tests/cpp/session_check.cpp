// Drives libmrg through the C++ session header the way the reference's host code would: uploads a
// flattened scene, runs SceneBvh::intersect on a ray set, and renders one job through
// minirender::gpu::RenderSession (start, pause, resume, wait, progress, preview, result).
//
//   session_check <input.bin> <output.bin>
//
// input (little endian): u32 magic 'MRGS', u64 n, u32 instances, u32 directional, f32 env[3],
//   mrg_camera, u32 width, height, spp, u64 seed, u32 mode, depth, rr, rr_depth, u64 rays,
//   f32 positions[n*9], f32 normals[n*9], u32 instance[n], u32 triangle[n],
//   mrg_material[instances], mrg_directional_light[directional], f32 rays[rays*6]
// output: u32 magic 'MRGO', u32 state, u32 pause_accepted, u32 resume_accepted, u64 recorded,
//   u32 preview_w, preview_h, f32 image[h*w*3], mrg_hit[rays]
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "mrg_session.hpp"

namespace {

template <class T>
void get(std::ifstream& in, T& v) {
    in.read(reinterpret_cast<char*>(&v), sizeof v);
}
template <class T>
void get_vec(std::ifstream& in, std::vector<T>& v, size_t n) {
    v.resize(n);
    in.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T)));
}
template <class T>
void put(std::ofstream& out, const T& v) {
    out.write(reinterpret_cast<const char*>(&v), sizeof v);
}

}  // namespace

int main(int argc, char** argv) {
    namespace gpu = minirender::gpu;
    if (argc != 3) {
        std::cerr << "usage: session_check <input.bin> <output.bin>\n";
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    uint32_t      magic = 0, instances = 0, directional = 0, width = 0, height = 0, spp = 0, mode = 0, depth = 0, rr = 0,
             rr_depth = 0;
    uint64_t   n = 0, seed = 0, nrays = 0;
    float      env[3];
    mrg_camera camera{};
    get(in, magic);
    if (magic != 0x5347524du) {
        std::cerr << "bad input\n";
        return 2;
    }
    get(in, n);
    get(in, instances);
    get(in, directional);
    in.read(reinterpret_cast<char*>(env), sizeof env);
    get(in, camera);
    get(in, width);
    get(in, height);
    get(in, spp);
    get(in, seed);
    get(in, mode);
    get(in, depth);
    get(in, rr);
    get(in, rr_depth);
    get(in, nrays);
    std::vector<float>                 positions, normals, rays;
    std::vector<uint32_t>              instance, triangle;
    std::vector<mrg_material>          materials;
    std::vector<mrg_directional_light> lights;
    get_vec(in, positions, n * 9);
    get_vec(in, normals, n * 9);
    get_vec(in, instance, n);
    get_vec(in, triangle, n);
    get_vec(in, materials, instances);
    get_vec(in, lights, directional);
    get_vec(in, rays, nrays * 6);
    if (!in) {
        std::cerr << "truncated input\n";
        return 2;
    }
    mrg_scene_desc desc{};
    desc.triangle_count    = n;
    desc.positions         = positions.data();
    desc.normals           = normals.data();
    desc.instance          = instance.data();
    desc.triangle          = triangle.data();
    desc.instance_count    = instances;
    desc.materials         = materials.data();
    desc.environment[0]    = env[0];
    desc.environment[1]    = env[1];
    desc.environment[2]    = env[2];
    desc.directional_count = directional;
    desc.directional       = lights.data();

    try {
        gpu::Device device(0);
        auto        scene = std::make_shared<gpu::SceneBvh>(device, desc);
        const auto  hits  = scene->intersect(rays);

        gpu::RenderSession session(device);
        // an idle session: terminal, no result, empty preview
        if (!gpu::render_terminal(session.progress().state) || session.result()) return 3;
        bool rejected = false;
        try {
            session.start(nullptr, camera, {width, height, spp, seed});
        } catch (const std::invalid_argument&) {
            rejected = true;
        }
        if (!rejected) return 4;
        session.start(scene, camera, {width, height, spp, seed}, {mode, depth, rr != 0, rr_depth}, {4});
        const bool paused  = session.pause();
        const bool resumed = session.resume();
        session.wait();
        const gpu::RenderProgress p = session.progress();
        if (p.state != gpu::RenderState::completed) {
            std::cerr << "job ended " << gpu::render_state_name(p.state) << ": " << p.error << "\n";
            return 5;
        }
        const auto image   = session.result();
        const auto preview = session.preview(8);
        if (!image || !preview.image || preview.recorded_samples != p.recorded_samples) return 6;

        std::ofstream out(argv[2], std::ios::binary);
        put(out, uint32_t(0x4f47524du));
        put(out, uint32_t(p.state));
        put(out, uint32_t(paused));
        put(out, uint32_t(resumed));
        put(out, p.recorded_samples);
        put(out, preview.image->width);
        put(out, preview.image->height);
        out.write(reinterpret_cast<const char*>(image->rgb.data()), std::streamsize(image->rgb.size() * sizeof(float)));
        out.write(reinterpret_cast<const char*>(hits.data()), std::streamsize(hits.size() * sizeof(mrg_hit)));
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
