// Headless command-line front end: the reference's `pt --headless ...` (src/main.cpp:16-118)
// with PathTraceDispatcher::start (src/pathtrace/ptdispatch.cpp:42-95) driving the MI355X
// back end through its C ABI.
//
//   pt --headless -w W -h H -s SPP -b BOUNCES -o out.png [-f png|hdr] scene.scn
//
// Options, defaults and error messages follow src/main.cpp and src/argparse.cpp: long and
// short names, attached short values (-w1024), `--` ends options; spp 32, bounces 2,
// 400x400, png. png: the accumulation through the ACES tone mapper to 8-bit RGBA
// (pt_read_tonemapped, as PathTraceDispatcher::put_result + write_png do); hdr: the
// accumulation as Radiance RGBE (write_hdr). Without --headless the reference opens a
// window; this build has no interactive front end and reports that.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pthip.h"
#include "pthost.h"

namespace {

// The option table: every option the reference's CLI accepts, in its help order
// (src/main.cpp:18-26, with --help first as CommandLineParser's constructor adds it).
enum class Kind { FLAG, TEXT, INT, REAL };

struct OptSpec {
    const char* name;  // "--width"
    char letter;       // 'w' for -w, 0: none
    Kind kind;
    const char* help;
};

constexpr OptSpec kOptions[] = {
    {"--help", 0, Kind::FLAG, "Displays help."},
    {"--width", 'w', Kind::INT, "The width of the rendered image."},
    {"--height", 'h', Kind::INT, "The height of the rendered image."},
    {"--headless", 0, Kind::FLAG, "Run without window and user interface."},
    {"--samples", 's', Kind::INT, "Number of samples to take."},
    {"--bounces", 'b', Kind::INT, "Maximum number of ray bounces."},
    {"--out", 'o', Kind::TEXT, "Specifies in which file to place the output."},
    {"--format", 'f', Kind::TEXT, "The image format of the output file."},
    {"--time", 't', Kind::REAL, "The maximum rendering time."},
};

const OptSpec* by_name(const std::string& tok)
{
    for (const OptSpec& o : kOptions)
        if (tok == o.name)
            return &o;
    return nullptr;
}

const OptSpec* by_letter(char c)
{
    for (const OptSpec& o : kOptions)
        if (o.letter != 0 && o.letter == c)
            return &o;
    return nullptr;
}

// The command line as scanned: raw option values by long name, then the operands.
struct CommandLine {
    std::map<std::string, std::string> values;
    std::vector<std::string> operands;

    bool has(const char* name) const { return values.count(name) != 0; }
    std::optional<std::string> text(const char* name) const
    {
        auto it = values.find(name);
        return it == values.end() ? std::nullopt : std::optional<std::string>(it->second);
    }
};

// Scans argv once, left to right. A token is one of: "--" (every later token is an operand),
// "--name" (a long option, its value the next token unless it is a flag), "-x..." (a short
// option: the value attached, else the next token; a flag takes nothing), or an operand. The
// failures are the reference's messages (src/argparse.cpp:61-105).
CommandLine scan(int argc, char** argv)
{
    CommandLine cl;
    bool operands_only = false;
    int i = 1;
    auto next_value = [&](const std::string& shown) -> std::string {
        if (i + 1 >= argc)
            throw std::runtime_error("Option " + shown + " has no value.");
        return argv[++i];
    };
    for (; i < argc; i++) {
        const std::string tok = argv[i];
        const bool dashed = !operands_only && tok.size() >= 2 && tok[0] == '-';
        if (tok == "--") {
            operands_only = true;
        } else if (dashed && tok[1] == '-') {
            const OptSpec* o = by_name(tok);
            if (!o)
                throw std::runtime_error("Unknown option " + tok + ".");
            cl.values[o->name] = o->kind == Kind::FLAG ? std::string("true") : next_value(tok);
        } else if (dashed) {
            const std::string head = tok.substr(0, 2);
            const OptSpec* o = by_letter(tok[1]);
            if (!o)
                throw std::runtime_error("Unknown option " + head + ".");
            const std::string attached = tok.substr(2);
            if (o->kind == Kind::FLAG) {
                if (!attached.empty())
                    throw std::runtime_error("Switch option " + head + " has a value.");
                cl.values[o->name] = "true";
            } else {
                cl.values[o->name] = attached.empty() ? next_value(tok) : attached;
            }
        } else {
            cl.operands.push_back(tok);
        }
    }
    return cl;
}

// Number syntax the reference accepts: an optional sign, digits, and for decimals an optional
// '.' with further digits (src/argparse.cpp:22-32).
bool number_syntax(const std::string& v, bool fraction)
{
    size_t k = (!v.empty() && (v[0] == '+' || v[0] == '-')) ? 1 : 0;
    const size_t digits_at = k;
    while (k < v.size() && v[k] >= '0' && v[k] <= '9')
        k++;
    if (k == digits_at)
        return false;
    if (fraction && k < v.size() && v[k] == '.') {
        k++;
        while (k < v.size() && v[k] >= '0' && v[k] <= '9')
            k++;
    }
    return k == v.size();
}

std::optional<int> int_value(const CommandLine& cl, const char* name)
{
    const auto v = cl.text(name);
    if (!v)
        return std::nullopt;
    if (!number_syntax(*v, false))
        throw std::runtime_error(std::string("The value of option ") + name + " was not an integer: " + *v);
    return std::stoi(*v);
}

std::optional<float> real_value(const CommandLine& cl, const char* name)
{
    const auto v = cl.text(name);
    if (!v)
        return std::nullopt;
    if (!number_syntax(*v, true))
        throw std::runtime_error(std::string("The value of option ") + name + " was not a decimal number: " + *v);
    return std::stof(*v);
}

void print_help(const char* prog)
{
    std::cout << "Usage: " << prog << " [options] path-to-scene\nOptions:";
    for (const OptSpec& o : kOptions) {
        std::cout << "\n  " << o.name;
        if (o.letter)
            std::cout << ", -" << o.letter;
        std::cout << "\t" << o.help;
    }
    std::cout << std::endl;
}

}  // namespace

int main(int argc, char** argv)
{
    CommandLine cl;
    int samples, bounces, width, height;
    std::optional<std::string> out;
    std::string format;
    bool headless;
    try {
        cl = scan(argc, argv);
        if (cl.has("--help")) {
            print_help(argv[0]);
            return 0;
        }
        if (cl.operands.empty()) {
            std::cerr << "Error: Expected path to scene." << std::endl;
            print_help(argv[0]);
            return 1;
        }
        if (cl.operands.size() > 1) {
            std::cerr << "Error: Unexpected number of arguments." << std::endl;
            print_help(argv[0]);
            return 1;
        }
        headless = cl.has("--headless");
        samples = int_value(cl, "--samples").value_or(32);
        if (samples < 1) {
            std::cerr << "Error: Invalid sample count: " << samples << "." << std::endl;
            return 1;
        }
        bounces = int_value(cl, "--bounces").value_or(2);
        if (bounces < 0) {
            std::cerr << "Error: Invalid number of ray bounces: " << bounces << "." << std::endl;
            return 1;
        }
        out = cl.text("--out");
        if (headless && !out) {
            std::cerr << "Error: No output file was given (required for headless execution)." << std::endl;
            return 1;
        }
        width = int_value(cl, "--width").value_or(400);
        height = int_value(cl, "--height").value_or(400);
        if (width < 1 || height < 1) {
            std::cerr << "Error: Invalid dimensions: (" << width << ", " << height << ")." << std::endl;
            return 1;
        }
        format = cl.text("--format").value_or("png");
        if (format != "png" && format != "PNG" && format != "hdr" && format != "HDR") {
            std::cerr << "Error: Output format is invalid: " << format << std::endl;
            return 1;
        }
        (void)real_value(cl, "--time");  // checked, unused (as in the reference)
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        print_help(argv[0]);
        return 1;
    }
    if (!headless) {
        std::cerr << "Error: this build has no interactive front end; use --headless." << std::endl;
        return 1;
    }
    const bool png = format == "png" || format == "PNG";

    pth_scene* scene = nullptr;
    if (pth_scene_load(cl.operands[0].c_str(), &scene) != PT_OK) {
        std::cerr << "Error: " << pth_last_error() << std::endl;
        return 1;
    }
    pth_scene_info info{};
    pth_scene_info_get(scene, &info);
    std::cout << "Scene has:\n"
              << "    " << info.object_variants << " meshes\n"
              << "    " << info.materials << " materials\n"
              << "    " << info.point_lights << " point lights\n"
              << "    " << info.directional_lights << " directional lights\n"
              << "    " << info.emissive_instances << " instances with emission\n"
              << "    0 textures\n"
              << "    " << info.instances << " instances\n"
              << "    " << info.unique_vertices << " unique vertices\n"
              << "    " << info.unique_triangles << " unique triangles\n"
              << "    " << info.total_triangles << " total triangles\n"
              << std::endl;

    pth_camera_params cp{};
    pt_camera cam{};
    pth_scene_camera_params(scene, uint32_t(width), uint32_t(height), &cp);
    pth_camera_block(&cp, &cam);

    pt_ctx* ctx = nullptr;
    if (pt_create(pth_scene_blob(scene), uint32_t(width), uint32_t(height), &ctx) != PT_OK) {
        std::cerr << "Error: " << pt_last_error(nullptr) << std::endl;
        pth_scene_free(scene);
        return 1;
    }
    int rc = pt_set_camera(ctx, &cam);
    if (rc == PT_OK)
        rc = pt_set_max_bounces(ctx, uint32_t(bounces));
    // Timing window of ptdispatch.cpp:54-65: from the first submit to the readback (the tone
    // mapping of the png path included, as put_result runs it inside the window).
    std::vector<float> img(png ? 0 : size_t(width) * height * 4);
    std::vector<uint8_t> img8(png ? size_t(width) * height * 4 : 0);
    const auto t0 = std::chrono::high_resolution_clock::now();
    if (rc == PT_OK)
        rc = pt_render(ctx, uint32_t(samples));
    if (rc == PT_OK)
        rc = png ? pt_read_tonemapped(ctx, img8.data(), size_t(width) * 4)
                 : pt_read_accumulation(ctx, img.data(), size_t(width) * 16);
    const float secs = std::chrono::duration<float>(std::chrono::high_resolution_clock::now() - t0).count();
    if (rc != PT_OK) {
        std::cerr << "Error: " << pt_last_error(ctx) << std::endl;
        pt_destroy(ctx);
        pth_scene_free(scene);
        return 1;
    }
    std::cout << "Rendering took " << secs << " seconds with " << samples << " samples taken per pixel.\n";
    std::cout << samples / secs << " samples per second.\n";
    pt_stats st{};
    if (pt_get_stats(ctx, &st) == PT_OK)
        std::cout << (st.extension_rays + st.shadow_rays) / secs * 1e-6 << " Mrays per second ("
                  << st.extension_rays << " extension + " << st.shadow_rays << " shadow rays).\n";
    std::cout << "Writing image to " << *out << std::endl;
    const int wrc = png ? pth_write_png(out->c_str(), img8.data(), uint32_t(width), uint32_t(height), size_t(width) * 4)
                        : pth_write_hdr(out->c_str(), img.data(), uint32_t(width), uint32_t(height), size_t(width) * 16);
    if (wrc != PT_OK) {
        std::cerr << "Error: " << pth_last_error() << std::endl;
        rc = 1;
    }
    pt_destroy(ctx);
    pth_scene_free(scene);
    return rc == PT_OK ? 0 : 1;
}
