/*
 * pthost.h - C ABI of the host-side scene surface (libpthost.so).
 *
 * The reference keeps these parts on the host and so does this build: the .scn scene
 * description, glTF loading, the packed scene buffer, the light alias tables, the camera
 * matrices and the .hdr writer. They produce the pt_scene_blob / pt_camera that the device
 * back end (pthip.h) consumes, so the scene packing stays host code as in the reference.
 *
 * Reference interfaces replaced, one per entry point, are cited below. Every call returns
 * PT_OK or a negative PT_E* status; pth_last_error() gives the message (thread-local).
 */
#ifndef PTHOST_H_INCLUDED
#define PTHOST_H_INCLUDED

#include "pthip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pth_scene pth_scene;

/* Camera parameters of a .scn camera section (src/scene/scenebuild.cpp:132-198) plus the
 * aspect ratio main() sets (src/main.cpp:98). */
typedef struct pth_camera_params {
    float position[3];
    float lookdir[3];
    float up[3];
    float fov_deg;
    float near_clip;
    float far_clip;
    float focus_dist;
    float lens_radius;
    float exposure;
    float aspect;
} pth_camera_params;

typedef struct pth_scene_info {
    uint32_t object_variants; /* "meshes" of Scene::scene_details (src/scene/scene.cpp:490-518) */
    uint32_t materials;
    uint32_t point_lights;
    uint32_t directional_lights;
    uint32_t emissive_instances;
    uint32_t instances;
    uint64_t unique_vertices;
    uint64_t unique_triangles;
    uint64_t total_triangles;
    uint32_t light_count; /* bins of the light alias table */
} pth_scene_info;

/* SceneBuilder::read_scene_description (src/scene/scenebuild.cpp:279-365) with glTF loading
 * (Scene::from_file, src/scene/scene.cpp:533-568), followed by SceneBuffer::build
 * (src/scene/scenebuffer.cpp:65-168) and TableLightSampler (src/scene/lightsample.cpp:89-178). */
int pth_scene_load(const char* scn_path, pth_scene** out);

/* The packed scene; valid until pth_scene_free. */
const pt_scene_blob* pth_scene_blob(const pth_scene* scene);

/* Scene::scene_details counts (src/scene/scene.cpp:490-518). */
int pth_scene_info_get(const pth_scene* scene, pth_scene_info* out);

/* The scene's camera section, with aspect = width/height as main() sets it
 * (src/main.cpp:98). */
int pth_scene_camera_params(const pth_scene* scene, uint32_t width, uint32_t height,
                            pth_camera_params* out);

void pth_scene_free(pth_scene* scene);

/* Camera basis (scenebuild.cpp:184-194) -> Camera::view_matrix / projection_matrix
 * (camera.h:16, camera.cpp:21-25) -> PathTracer::set_camera (pathtrace.cpp:381-396). */
int pth_camera_block(const pth_camera_params* params, pt_camera* out);

/* write_hdr (src/io/imagewrite.cpp:19-37): RGBA32F rows -> Radiance RGBE file with the
 * byte stream stb_image_write 1.16 produces (alpha dropped, RLE scanlines). */
int pth_write_hdr(const char* path, const float* rgba, uint32_t width, uint32_t height,
                  size_t row_pitch);

/* write_png (src/io/imagewrite.cpp:44-67): 8-bit RGBA rows (row_pitch bytes apart, e.g. from
 * pt_read_tonemapped) -> PNG file. The reference encodes with stb_image_write; this is another
 * valid PNG of the same pixels (per-row filters, fixed-code DEFLATE). */
int pth_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height, size_t row_pitch);

/* Reads a Radiance .hdr (RGBE, flat or RLE) into width*height*4 floats (alpha = 1), with the
 * RGBE decode stb_image 2.30 uses; *out is malloc'd, release with pth_free. */
int pth_read_hdr(const char* path, float** out, uint32_t* width, uint32_t* height);

/* util/meansq/meansq.cpp:30-41: sum over RGB of (img - ref)^2 / (3*W*H - 1). */
double pth_meansq(const float* img_rgba, const float* ref_rgba, uint32_t width, uint32_t height);

/* AliasTable<T>::build (src/aliastable.h:21-67) + copy (aliastable.h:73-77): bins of
 * {f32 p, f32 q, u32 alias, value_words value words} for n weights, written to out_words
 * (n * (3 + value_words) words). */
int pth_alias_table(const float* weights, uint32_t n, const uint32_t* values, uint32_t value_words,
                    uint32_t* out_words);

/* Texture image decoding as TextureData does it (src/scene/texturedata.cpp:30-46 with
 * stb_image 2.30): PNG or JPEG (baseline, progressive); 1, 2 or 4 channels of 8 bits (three
 * widened to four with alpha 255; colour JPEGs are RGBA, gray ones one channel). *px is
 * malloc'd (width*height*channels bytes), release with pth_free. */
int pth_decode_image(const uint8_t* data, size_t size, uint32_t* width, uint32_t* height, uint32_t* channels,
                     uint8_t** px);

/* from_srgb / to_srgb (src/colour.cpp:3-15). */
float pth_from_srgb(float x);
float pth_to_srgb(float x);

void pth_free(void* p);

const char* pth_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
