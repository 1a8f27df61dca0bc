// Command line of the GPU path tracer: the reference's `--render` entry point
// (src/app/render_cli.cpp:46-181, dispatched from cli.cpp:33-35) with `--integrator gpu`:
//
//   mrg_render --render scene.json output-prefix [--width N --height N --spp N --depth N
//              --mode mis|bsdf|light --seed N --exposure EV --no-rr --rr-depth N --threads N
//              --tile N --batch N --integrator gpu]
//
// writes output-prefix.exr (linear float RGB) and output-prefix.png (display image), printing
// progress lines to stderr and the job summary to stdout. --threads and --tile are validated as
// the reference does and have no GPU meaning; --batch is the progressive pass size.
//
// Diagnostics used by the tests (no GPU needed):
//   mrg_render --scene-info scene.json [flat.bin]   triangle / instance counts, optional dump
//   mrg_render --obj mesh.obj mesh.bin              OBJ loader output
//   mrg_render --encode image.bin output-prefix [--exposure EV]   EXR + PNG of a raw image
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

#include "image_io.hpp"
#include "mrg_session.hpp"
#include "scene_io.hpp"

namespace {

namespace gpu = minirender::gpu;

// Strict numeric option values: the whole token must be the number, in the C locale, with no
// leading blank or sign for counts. A malformed value is an input error naming the token.
[[noreturn]] void bad_number(std::string_view text) {
    throw std::invalid_argument("Invalid numeric render option: " + std::string(text));
}

// Unsigned decimal count: digits only, no overflow of T.
template <typename T>
T parse_count(std::string_view text) {
    static_assert(std::is_unsigned_v<T>, "counts are unsigned");
    if (text.empty()) bad_number(text);
    T value = 0;
    for (const char c : text) {
        if (c < '0' || c > '9') bad_number(text);
        const T digit = T(c - '0');
        if (value > (std::numeric_limits<T>::max() - digit) / 10) bad_number(text);
        value = T(value * 10 + digit);
    }
    return value;
}

// Decimal real: [+-] digits [. digits] [(e|E) [+-] digits], at least one mantissa digit; the scan
// accepts exactly that grammar (no blanks, inf, nan or hexadecimal forms) and strtod converts it.
double parse_real(std::string_view text) {
    size_t k      = 0;
    auto   digits = [&] {
        const size_t from = k;
        while (k < text.size() && text[k] >= '0' && text[k] <= '9') ++k;
        return k - from;
    };
    if (k < text.size() && (text[k] == '+' || text[k] == '-')) ++k;
    size_t mantissa = digits();
    if (k < text.size() && text[k] == '.') {
        ++k;
        mantissa += digits();
    }
    if (mantissa == 0) bad_number(text);
    if (k < text.size() && (text[k] == 'e' || text[k] == 'E')) {
        ++k;
        if (k < text.size() && (text[k] == '+' || text[k] == '-')) ++k;
        if (digits() == 0) bad_number(text);
    }
    if (k != text.size()) bad_number(text);
    const std::string token(text);
    return std::strtod(token.c_str(), nullptr);
}

template <class T>
void put(std::ofstream& out, const T& v) {
    out.write(reinterpret_cast<const char*>(&v), sizeof v);
}
template <class T>
void put_vec(std::ofstream& out, const std::vector<T>& v) {
    out.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

// u64 n, u32 instances, u32 directional, f32 env[3], mrg_camera, then positions, normals,
// instance, triangle, mrg_material[instances], mrg_directional_light[directional], uvs [n][3][2],
// u32 textures, then per texture u32 width, u32 height, f32 rgb[height][width][3], then
// u32 environment texture (1 + index, 0 = none), f32 strength, f32 rotation degrees
void dump_flat(const mrg_host::FlatScene& f, const std::string& path) {
    std::ofstream out(path, std::ios::binary);
    out.exceptions(std::ios::badbit | std::ios::failbit);
    put(out, uint64_t(f.instance.size()));
    put(out, uint32_t(f.materials.size()));
    put(out, uint32_t(f.lights.size()));
    put(out, f.environment);
    put(out, f.camera);
    put_vec(out, f.positions);
    put_vec(out, f.normals);
    put_vec(out, f.instance);
    put_vec(out, f.triangle);
    put_vec(out, f.materials);
    put_vec(out, f.lights);
    put_vec(out, f.uvs);
    put(out, uint32_t(f.textures.size()));
    for (const auto& t : f.textures) {
        put(out, t.width);
        put(out, t.height);
        put_vec(out, t.rgb);
    }
    put(out, f.environment_texture);
    put(out, f.environment_strength);
    put(out, f.environment_rotation_degrees);
}

int scene_info(int argc, char** argv) {
    if (argc < 1) throw std::invalid_argument("Usage: --scene-info scene.json [flat.bin]");
    const auto doc  = mrg_host::load_scene(std::filesystem::u8path(argv[0]));
    const auto flat = mrg_host::flatten(doc);
    std::cout << "triangles: " << flat.instance.size() << "\ninstances: " << flat.materials.size()
              << "\nmeshes: " << doc.meshes.size() << "\ndirectional lights: " << flat.lights.size() << '\n';
    if (argc >= 2) dump_flat(flat, argv[1]);
    return 0;
}

// u32 width, u32 height, f32 rgb[height][width][3]: the decoded linear texture
int dump_texture(int argc, char** argv) {
    if (argc < 3) throw std::invalid_argument("Usage: --texture image.png srgb|linear texture.bin");
    const std::string encoding = argv[1];
    if (encoding != "srgb" && encoding != "linear") throw std::invalid_argument("Unknown texture encoding");
    const auto    img = mrg_host::read_texture(std::filesystem::u8path(argv[0]), encoding == "srgb");
    std::ofstream out(argv[2], std::ios::binary);
    out.exceptions(std::ios::badbit | std::ios::failbit);
    put(out, img.width);
    put(out, img.height);
    put_vec(out, img.rgb);
    return 0;
}

// u32 positions, u32 triangles, f32 positions[3n], then per corner: u32 position, f32 normal[3], f32 uv[2]
int dump_obj(int argc, char** argv) {
    if (argc < 2) throw std::invalid_argument("Usage: --obj mesh.obj mesh.bin");
    const auto    mesh = mrg_host::load_obj(std::filesystem::u8path(argv[0]));
    std::ofstream out(argv[1], std::ios::binary);
    out.exceptions(std::ios::badbit | std::ios::failbit);
    put(out, uint32_t(mesh.positions.size()));
    put(out, uint32_t(mesh.triangles.size()));
    for (const auto& p : mesh.positions) put(out, p);
    for (const auto& t : mesh.triangles)
        for (const auto& c : t) {
            put(out, c.position);
            put(out, c.normal);
            put(out, c.uv);
        }
    return 0;
}

// Output files of a prefix: "<prefix>.exr" (linear float) and "<prefix>.png" (display image). A
// prefix that already names one of the two (".png" / ".exr") is taken without that extension; the
// directory part is created when missing.
void write_outputs(const mrg_host::RgbImage& image, const std::filesystem::path& prefix, double exposure) {
    std::string base = prefix.u8string();
    const bool named = base.size() > 4 && (base.compare(base.size() - 4, 4, ".png") == 0 || base.compare(base.size() - 4, 4, ".exr") == 0);
    if (named) base.resize(base.size() - 4);
    const std::filesystem::path stem = std::filesystem::u8path(base);
    if (stem.has_parent_path()) std::filesystem::create_directories(stem.parent_path());
    mrg_host::write_exr(image, std::filesystem::u8path(base + ".exr"));
    mrg_host::write_png(mrg_host::display_image(image, exposure), std::filesystem::u8path(base + ".png"));
}

int encode(int argc, char** argv) {
    if (argc < 2) throw std::invalid_argument("Usage: --encode image.bin output-prefix [--exposure EV]");
    double exposure = 0;
    if (argc >= 4 && std::string_view(argv[2]) == "--exposure") exposure = parse_real(argv[3]);
    std::ifstream in(argv[0], std::ios::binary);
    mrg_host::RgbImage image;
    in.read(reinterpret_cast<char*>(&image.width), 4);
    in.read(reinterpret_cast<char*>(&image.height), 4);
    if (!in || image.width == 0 || image.height == 0 || uint64_t(image.width) * image.height > (uint64_t(1) << 28))
        throw std::invalid_argument("bad image file");
    image.rgb.resize(size_t(image.width) * image.height * 3);
    in.read(reinterpret_cast<char*>(image.rgb.data()), std::streamsize(image.rgb.size() * 4));
    if (!in) throw std::invalid_argument("truncated image file");
    write_outputs(image, std::filesystem::u8path(argv[1]), exposure);
    return 0;
}

int render(int argc, char** argv) {
    if (argc < 2)
        throw std::invalid_argument(
            "Usage: --render scene.json output-prefix [--width N --height N --spp N --depth N "
            "--mode mis|bsdf|light --seed N --exposure EV --no-rr --rr-depth N --threads N --tile "
            "N --batch N --integrator gpu]");
    uint64_t width = 256, height = 256, seed = 1;
    uint32_t spp = 64, depth = 8, rr_depth = 3, threads = 0, tile = 16, batch = 4, mode = MRG_MODE_MIS;
    bool     rr       = true;
    double   exposure = 0;
    // Every option with a value, and what its value sets. --no-rr is the one switch without a value.
    struct Setting {
        std::string_view                        flag;
        std::function<void(std::string_view)> set;
    };
    const Setting settings[] = {
        {"--width", [&](std::string_view v) { width = parse_count<uint64_t>(v); }},
        {"--height", [&](std::string_view v) { height = parse_count<uint64_t>(v); }},
        {"--spp", [&](std::string_view v) { spp = parse_count<uint32_t>(v); }},
        {"--depth", [&](std::string_view v) { depth = parse_count<uint32_t>(v); }},
        {"--rr-depth", [&](std::string_view v) { rr_depth = parse_count<uint32_t>(v); }},
        {"--threads", [&](std::string_view v) { threads = parse_count<uint32_t>(v); }},
        {"--tile", [&](std::string_view v) { tile = parse_count<uint32_t>(v); }},
        {"--batch", [&](std::string_view v) { batch = parse_count<uint32_t>(v); }},
        {"--seed", [&](std::string_view v) { seed = parse_count<uint64_t>(v); }},
        {"--exposure", [&](std::string_view v) { exposure = parse_real(v); }},
        {"--mode",
         [&](std::string_view v) {
             static const std::pair<std::string_view, uint32_t> modes[] = {
                 {"bsdf", MRG_MODE_BSDF}, {"light", MRG_MODE_LIGHT}, {"mis", MRG_MODE_MIS}};
             const auto* m = std::find_if(std::begin(modes), std::end(modes), [&](const auto& e) { return e.first == v; });
             if (m == std::end(modes)) throw std::invalid_argument("Sampling mode must be bsdf, light or mis");
             mode = m->second;
         }},
        {"--integrator",
         [&](std::string_view v) {
             if (v == "gpu") return;
             if (v == "path" || v == "sppm")
                 throw std::invalid_argument("The CPU integrators (path, sppm) are not part of this build; use --integrator gpu");
             throw std::invalid_argument("Integrator must be gpu");
         }},
    };
    std::vector<std::string_view> args(argv + 2, argv + argc);
    for (size_t k = 0; k < args.size(); ++k) {
        if (args[k] == "--no-rr") {
            rr = false;
            continue;
        }
        if (k + 1 == args.size()) throw std::invalid_argument("Render option is missing its value: " + std::string(args[k]));
        const auto* s = std::find_if(std::begin(settings), std::end(settings), [&](const Setting& e) { return e.flag == args[k]; });
        if (s == std::end(settings)) throw std::invalid_argument("Unknown render option: " + std::string(args[k]));
        s->set(args[++k]);
    }
    // RenderSettings / IntegratorSettings / SessionSettings::validate
    if (width == 0 || height == 0 || width > 16384 || height > 16384 || width * height > 64u * 1024u * 1024u || spp == 0)
        throw std::invalid_argument("Render dimensions or sample count exceed the image budget");
    if (depth > 128 || rr_depth > 128) throw std::invalid_argument("Path and roulette depths must not exceed 128");
    if (threads > 32 || tile == 0 || tile > 128 || batch == 0 || batch > 1024)
        throw std::invalid_argument("Session requires <=32 threads, tile size in [1,128], batch in [1,1024]");
    if (!std::isfinite(exposure) || std::abs(exposure) > 64) throw std::invalid_argument("Exposure must be finite and in [-64,64]");

    auto       flat  = mrg_host::flatten(mrg_host::load_scene(std::filesystem::u8path(argv[0])));
    const auto start = std::chrono::steady_clock::now();
    gpu::Device device(0);
    auto        scene = std::make_shared<gpu::SceneBvh>(device, flat.desc());
    gpu::RenderSession session(device);
    session.start(scene, flat.camera, {uint32_t(width), uint32_t(height), spp, seed}, {mode, depth, rr, rr_depth}, {batch});
    while (!session.wait_for(std::chrono::milliseconds(200))) {
        const auto p = session.progress();
        std::cerr << "progress: " << gpu::render_state_name(p.state) << ' ' << p.recorded_samples << '/' << p.total_samples << '\n';
    }
    session.wait();
    const auto progress = session.progress();
    if (progress.state != gpu::RenderState::completed)
        throw std::runtime_error(progress.error.empty() ? "Render did not complete" : progress.error);
    const auto         result = session.result();
    mrg_host::RgbImage image{result->width, result->height, result->rgb};
    const auto         finished = std::chrono::steady_clock::now();
    write_outputs(image, std::filesystem::u8path(argv[1]), exposure);
    static const char* modes[] = {"bsdf", "light", "mis"};
    const auto&        st      = progress.statistics;
    std::cout << "GPU path render complete\nmode: " << modes[mode] << "\nresolution: " << width << 'x' << height
              << "\nspp: " << spp << "\ntriangles: " << flat.instance.size() << "\npaths: " << st.paths
              << "\nrecorded samples: " << progress.recorded_samples << "\nextension rays: " << st.extension_rays
              << "\nshadow rays: " << st.shadow_rays << "\nscattering events: " << st.scattering_events
              << "\nroulette terminations: " << st.roulette_terminations
              << "\nelapsed ms: " << std::chrono::duration<double, std::milli>(finished - start).count()
              << "\nEXR and PNG saved.\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 2) {
            const std::string_view mode(argv[1]);
            if (mode == "--render") return render(argc - 2, argv + 2);
            if (mode == "--scene-info") return scene_info(argc - 2, argv + 2);
            if (mode == "--obj") return dump_obj(argc - 2, argv + 2);
            if (mode == "--texture") return dump_texture(argc - 2, argv + 2);
            if (mode == "--encode") return encode(argc - 2, argv + 2);
        }
        std::cerr << "Usage: mrg_render --render scene.json output-prefix [options] | --scene-info scene.json [flat.bin] | "
                     "--obj mesh.obj mesh.bin | --encode image.bin output-prefix [--exposure EV]\n";
        return 2;
    } catch (const std::exception& error) {
        std::cerr << error.what() << '\n';
        return 1;
    }
}
