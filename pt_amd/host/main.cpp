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
#include <regex>
#include <string>
#include <vector>

#include "pthip.h"
#include "pthost.h"

namespace {

enum class OptType { STRING, SWITCH, INTEGER, DECIMAL };

struct Option {
    std::string name, short_name;
    OptType type;
    std::string description;
    std::optional<std::string> value;
};

struct Parser {
    std::vector<Option> opts;
    std::map<std::string, size_t> by_name, by_short;

    void add(const std::string& n, const std::string& s, OptType t, const std::string& d)
    {
        by_name[n] = opts.size();
        if (!s.empty())
            by_short[s] = opts.size();
        opts.push_back({n, s, t, d, std::nullopt});
    }

    void parse(int argc, char** argv, std::vector<std::string>& rest)
    {
        bool ended = false;
        for (int i = 1; i < argc; i++) {
            const std::string tok = argv[i];
            if (!ended && tok.size() >= 3 && tok[0] == '-' && tok[1] == '-') {
                if (!by_name.count(tok))
                    throw std::runtime_error("Unknown option " + tok + ".");
                Option& o = opts[by_name[tok]];
                if (o.type == OptType::SWITCH) {
                    o.value = "true";
                } else {
                    if (++i >= argc)
                        throw std::runtime_error("Option " + tok + " has no value.");
                    o.value = std::string(argv[i]);
                }
            } else if (!ended && tok.size() >= 2 && tok[0] == '-' && tok[1] != '-') {
                const std::string s = tok.substr(0, 2);
                if (!by_short.count(s))
                    throw std::runtime_error("Unknown option " + s + ".");
                Option& o = opts[by_short[s]];
                if (o.type == OptType::SWITCH) {
                    if (tok.size() != 2)
                        throw std::runtime_error("Switch option " + s + " has a value.");
                    o.value = "true";
                } else if (tok.size() == 2) {
                    if (++i >= argc)
                        throw std::runtime_error("Option " + tok + " has no value.");
                    o.value = std::string(argv[i]);
                } else {
                    o.value = tok.substr(2);
                }
            } else if (tok == "--") {
                ended = true;
            } else {
                rest.push_back(tok);
            }
        }
    }

    const Option& get(const std::string& n) const { return opts.at(by_name.at(n)); }
    bool flag(const std::string& n) const { return get(n).value == std::string("true"); }
    std::optional<std::string> str(const std::string& n) const { return get(n).value; }
    std::optional<int> integer(const std::string& n) const
    {
        const Option& o = get(n);
        if (!o.value)
            return std::nullopt;
        if (!std::regex_match(*o.value, std::regex("[\\-\\+]?[0-9]+")))
            throw std::runtime_error("The value of option " + n + " was not an integer: " + *o.value);
        return std::stoi(*o.value);
    }
    std::optional<float> decimal(const std::string& n) const
    {
        const Option& o = get(n);
        if (!o.value)
            return std::nullopt;
        if (!std::regex_match(*o.value, std::regex("[\\-\\+]?[0-9]+(\\.[0-9]*)?")))
            throw std::runtime_error("The value of option " + n + " was not a decimal number: " + *o.value);
        return std::stof(*o.value);
    }
    std::string help() const
    {
        std::string h = "Options:";
        for (const Option& o : opts)
            h += "\n  " + o.name + (o.short_name.empty() ? "" : ", " + o.short_name) + "\t" + o.description;
        return h;
    }
};

void print_help(const char* prog, const Parser& p)
{
    std::cout << "Usage: " << prog << " [options] path-to-scene\n" << p.help() << std::endl;
}

}  // namespace

int main(int argc, char** argv)
{
    Parser parser;
    parser.add("--help", "", OptType::SWITCH, "Displays help.");
    parser.add("--width", "-w", OptType::INTEGER, "The width of the rendered image.");
    parser.add("--height", "-h", OptType::INTEGER, "The height of the rendered image.");
    parser.add("--headless", "", OptType::SWITCH, "Run without window and user interface.");
    parser.add("--samples", "-s", OptType::INTEGER, "Number of samples to take.");
    parser.add("--bounces", "-b", OptType::INTEGER, "Maximum number of ray bounces.");
    parser.add("--out", "-o", OptType::STRING, "Specifies in which file to place the output.");
    parser.add("--format", "-f", OptType::STRING, "The image format of the output file.");
    parser.add("--time", "-t", OptType::DECIMAL, "The maximum rendering time.");

    std::vector<std::string> args;
    int samples, bounces, width, height;
    std::optional<std::string> out;
    std::string format;
    bool headless;
    try {
        parser.parse(argc, argv, args);
        if (parser.flag("--help")) {
            print_help(argv[0], parser);
            return 0;
        }
        if (args.empty()) {
            std::cerr << "Error: Expected path to scene." << std::endl;
            print_help(argv[0], parser);
            return 1;
        }
        if (args.size() > 1) {
            std::cerr << "Error: Unexpected number of arguments." << std::endl;
            print_help(argv[0], parser);
            return 1;
        }
        headless = parser.flag("--headless");
        samples = parser.integer("--samples").value_or(32);
        if (samples < 1) {
            std::cerr << "Error: Invalid sample count: " << samples << "." << std::endl;
            return 1;
        }
        bounces = parser.integer("--bounces").value_or(2);
        if (bounces < 0) {
            std::cerr << "Error: Invalid number of ray bounces: " << bounces << "." << std::endl;
            return 1;
        }
        out = parser.str("--out");
        if (headless && !out) {
            std::cerr << "Error: No output file was given (required for headless execution)." << std::endl;
            return 1;
        }
        width = parser.integer("--width").value_or(400);
        height = parser.integer("--height").value_or(400);
        if (width < 1 || height < 1) {
            std::cerr << "Error: Invalid dimensions: (" << width << ", " << height << ")." << std::endl;
            return 1;
        }
        format = parser.str("--format").value_or("png");
        if (format != "png" && format != "PNG" && format != "hdr" && format != "HDR") {
            std::cerr << "Error: Output format is invalid: " << format << std::endl;
            return 1;
        }
        (void)parser.decimal("--time");  // parsed, unused (as in the reference)
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        print_help(argv[0], parser);
        return 1;
    }
    if (!headless) {
        std::cerr << "Error: this build has no interactive front end; use --headless." << std::endl;
        return 1;
    }
    const bool png = format == "png" || format == "PNG";

    pth_scene* scene = nullptr;
    if (pth_scene_load(args[0].c_str(), &scene) != PT_OK) {
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
