// C ABI of the host scene surface (include/pthost.h).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "hdrio.h"
#include "png.h"
#include "pthost.h"
#include "scene.h"

struct pth_scene {
    pth::Scene scene;
    pth::PackedScene packed;
};

static thread_local std::string g_error;

template <typename F>
static int guarded(F&& f)
{
    try {
        f();
        return PT_OK;
    } catch (const std::bad_alloc&) {
        g_error = "out of host memory";
        return PT_ENOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return PT_EINVAL;
    }
}

extern "C" {

int pth_scene_load(const char* scn_path, pth_scene** out)
{
    if (!scn_path || !out) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    *out = nullptr;
    return guarded([&] {
        pth_scene* s = new pth_scene;
        try {
            s->scene = pth::load_scn(scn_path);
            s->packed = pth::pack_scene(s->scene);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

const pt_scene_blob* pth_scene_blob(const pth_scene* s) { return s ? &s->packed.blob : nullptr; }

int pth_scene_info_get(const pth_scene* s, pth_scene_info* out)
{
    if (!s || !out) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    pth_scene_info info{};
    const pth::Scene& sc = s->scene;
    info.object_variants = uint32_t(sc.variants.size());
    info.materials = uint32_t(sc.materials.size());
    info.point_lights = uint32_t(sc.point_lights.size());
    info.directional_lights = uint32_t(sc.dir_lights.size());
    for (const pth::Mesh& m : sc.meshes) {
        info.unique_vertices += m.vertices.size();
        info.unique_triangles += m.tris.size();
    }
    for (const pth::Variant& v : sc.variants) {
        info.instances += uint32_t(v.instances.size());
        info.total_triangles += uint64_t(v.instances.size()) * sc.meshes[v.mesh].tris.size();
        if (sc.materials[v.material].is_emitter())
            info.emissive_instances += uint32_t(v.instances.size());
    }
    info.light_count = s->packed.light_count;
    *out = info;
    return PT_OK;
}

int pth_scene_camera_params(const pth_scene* s, uint32_t width, uint32_t height, pth_camera_params* out)
{
    if (!s || !out || width == 0 || height == 0) {
        g_error = "invalid argument";
        return PT_EINVAL;
    }
    const pth::CameraSection& c = s->scene.camera;
    pth_camera_params p{};
    for (int i = 0; i < 3; i++) {
        p.position[i] = c.position[i];
        p.lookdir[i] = c.lookdir[i];
        p.up[i] = c.up[i];
    }
    p.fov_deg = c.fov;
    p.near_clip = c.near_clip;
    p.far_clip = c.far_clip;
    p.focus_dist = c.focus_dist;
    p.lens_radius = c.lens_radius;
    p.exposure = c.exposure;
    p.aspect = float(width) / float(height);  // main.cpp:98
    *out = p;
    return PT_OK;
}

void pth_scene_free(pth_scene* s) { delete s; }

int pth_camera_block(const pth_camera_params* p, pt_camera* out)
{
    if (!p || !out) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    return guarded([&] {
        pth::CameraSection c;
        c.position = pth::V3(p->position[0], p->position[1], p->position[2]);
        c.lookdir = pth::V3(p->lookdir[0], p->lookdir[1], p->lookdir[2]);
        c.up = pth::V3(p->up[0], p->up[1], p->up[2]);
        c.fov = p->fov_deg;
        c.near_clip = p->near_clip;
        c.far_clip = p->far_clip;
        c.focus_dist = p->focus_dist;
        c.lens_radius = p->lens_radius;
        c.exposure = p->exposure;
        pth::camera_block(c, p->aspect, out);
    });
}

int pth_write_hdr(const char* path, const float* rgba, uint32_t width, uint32_t height, size_t row_pitch)
{
    if (!path || !rgba) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    return guarded([&] { pth::write_hdr_file(path, rgba, width, height, row_pitch); });
}

int pth_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height, size_t row_pitch)
{
    if (!path || !rgba8) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    return guarded([&] {
        if (width == 0 || height == 0 || row_pitch < size_t(width) * 4)
            throw std::runtime_error("invalid image size");
        const std::vector<uint8_t> png = pth::png_encode_rgba8(rgba8, width, height, row_pitch);
        FILE* f = std::fopen(path, "wb");
        if (!f)
            throw std::runtime_error("Could not write PNG file.");
        const size_t wrote = std::fwrite(png.data(), 1, png.size(), f);
        const bool ok = std::fclose(f) == 0 && wrote == png.size();
        if (!ok)
            throw std::runtime_error("Could not write PNG file.");
    });
}

int pth_read_hdr(const char* path, float** out, uint32_t* width, uint32_t* height)
{
    if (!path || !out || !width || !height) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    return guarded([&] {
        std::vector<float> img = pth::read_hdr_file(path, *width, *height);
        float* buf = (float*)std::malloc(img.size() * sizeof(float));
        if (!buf)
            throw std::bad_alloc();
        memcpy(buf, img.data(), img.size() * sizeof(float));
        *out = buf;
    });
}

int pth_decode_image(const uint8_t* data, size_t size, uint32_t* width, uint32_t* height, uint32_t* channels,
                     uint8_t** px)
{
    if (!data || !width || !height || !channels || !px) {
        g_error = "null argument";
        return PT_EINVAL;
    }
    return guarded([&] {
        pth::Image img = pth::load_texture_image(data, size);
        uint8_t* buf = (uint8_t*)std::malloc(std::max<size_t>(img.px.size(), 1));
        if (!buf)
            throw std::bad_alloc();
        memcpy(buf, img.px.data(), img.px.size());
        *width = img.width;
        *height = img.height;
        *channels = img.channels;
        *px = buf;
    });
}

double pth_meansq(const float* img, const float* ref, uint32_t width, uint32_t height)
{
    return pth::meansq(img, ref, width, height);
}

int pth_alias_table(const float* weights, uint32_t n, const uint32_t* values, uint32_t value_words,
                    uint32_t* out_words)
{
    if (!weights || n == 0 || (value_words && !values) || !out_words) {
        g_error = "invalid argument";
        return PT_EINVAL;
    }
    return guarded([&] {
        std::vector<float> w(weights, weights + n);
        std::vector<uint32_t> v(values, values + size_t(n) * value_words);
        std::vector<uint32_t> words = pth::build_alias_table(w, v, value_words);
        memcpy(out_words, words.data(), words.size() * 4);
    });
}

float pth_from_srgb(float x) { return pth::from_srgb(x); }
float pth_to_srgb(float x) { return pth::to_srgb(x); }

void pth_free(void* p) { std::free(p); }

const char* pth_last_error(void) { return g_error.c_str(); }

}  // extern "C"
