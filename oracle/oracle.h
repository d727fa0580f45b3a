/*
 * oracle.h - CPU oracle of the path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the project's plain C++ restatement of the reference's
 * GPU path tracer (shaders/pathtrace/*.hlsl*), used by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg as the checker. The product (pt_amd/) never links, loads
 * or calls it.
 *
 * It consumes the same packed scene (pt_scene_blob) and camera block (pt_camera) as the
 * device back end and builds its own acceleration structure (binned-SAH BVH2 per mesh plus
 * a BVH2 over instances). Ray/triangle tests are the watertight test of Woop, Benthin and
 * Wald (JCGT 2013) evaluated in double precision; shading is fp32 as in the shaders.
 */
#ifndef PT_ORACLE_H_INCLUDED
#define PT_ORACLE_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#include "pthip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

typedef struct oracle_counts {
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t paths;
} oracle_counts;

int oracle_scene_create(const pt_scene_blob* blob, oracle_scene** out);
void oracle_scene_free(oracle_scene* s);

/* Renders samples [sample_index, sample_index + spp) of every pixel into accum (RGBA32F,
 * width*height*4), exactly as the raygen shader does across 1-spp dispatches
 * (pathtrace.rg.hlsl:9-67). If rows_step > 1 only rows y % rows_step == 0 are traced (a
 * bounded sample for timing). threads <= 0 uses all OpenMP threads. */
int oracle_render(const oracle_scene* s, const pt_camera* cam, uint32_t width, uint32_t height,
                  uint32_t sample_index, uint32_t spp, uint32_t max_bounces, float* accum,
                  uint32_t rows_step, int threads, oracle_counts* counts);

/* Estimator variants of oracle_render_ex, for the oracle's own checks against closed forms and
 * equivalent estimators (the reference build always samples lights and plays Russian roulette):
 *   ORACLE_NO_NEE   SAMPLE_LIGHTS undefined (pathtrace.ch.hlsl:8): no light sampling, emission
 *                   added unweighted at every hit (pathtrace.ch.hlsl:266-268);
 *   ORACLE_NO_RR    no Russian roulette (pathtrace.rg.hlsl:53).
 * Planted mistakes, which the checks must detect:
 *   ORACLE_PLANT_NO_DIST2  light sampling's area-to-solid-angle conversion without d^2;
 *   ORACLE_PLANT_MIS_SWAP  the emitter-hit balance weight with the light pdf on top;
 *   ORACLE_PLANT_RR_Q      surviving paths divided by q instead of 1 - q. */
#define ORACLE_NO_NEE 1u
#define ORACLE_NO_RR 2u
#define ORACLE_PLANT_NO_DIST2 0x100u
#define ORACLE_PLANT_MIS_SWAP 0x200u
#define ORACLE_PLANT_RR_Q 0x400u
int oracle_render_ex(const oracle_scene* s, const pt_camera* cam, uint32_t width, uint32_t height,
                     uint32_t sample_index, uint32_t spp, uint32_t max_bounces, float* accum,
                     uint32_t rows_step, int threads, oracle_counts* counts, uint32_t flags);

/* light_pdf (lightsample.hlsli:36-50): area probability of the emitter triangle prim of the
 * light bin emitter_index (light table p times triangle table p). */
float oracle_light_pdf(const oracle_scene* s, uint32_t emitter_index, uint32_t prim);

/* evaluate_brdf of n (i, o) pairs with m = normalize(i + o) (params as oracle_eval_brdf). */
int oracle_eval_brdf_many(const float* params, size_t n, const float* i, const float* o, float* rgb_out);

/* Primary hits of one sample index (pathtrace.rg.hlsl:18-35). */
int oracle_primary(const oracle_scene* s, const pt_camera* cam, uint32_t width, uint32_t height,
                   uint32_t sample_index, pt_hit* hits, pt_ray* rays_out);

/* Closest (any_hit == 0) or first (any_hit != 0) hits of arbitrary rays. */
int oracle_trace_rays(const oracle_scene* s, const pt_ray* rays, uint32_t count, pt_hit* hits, int any_hit);

/* Sampler restatement (sample.hlsli:94-190): the first n values drawn by a sequence of
 * sample_1d (pattern bit 0) / sample_2d (pattern bit 1) calls after
 * sampler_start_new(pixel, sample_index, seed). pattern[k] is 1 or 2. */
int oracle_sampler_draw(uint32_t px, uint32_t py, uint32_t sample_index, uint32_t seed, const uint8_t* pattern,
                        uint32_t n_calls, float* out);

/* BSDF restatement (pathtrace.ch.hlsl:79-157). params: base rgb, alpha xy, metalness,
 * specular rgb, sheen rgb, clearcoat, clearcoat_alpha (14 floats). */
int oracle_eval_brdf(const float* params, const float* i, const float* o, const float* m, float* rgb_out,
                     float* pdf_out);
int oracle_sample_direction(const float* o, const float* alpha, float metalness, const float* u3, float* i_out,
                            float* m_out, float* w_out, float* pdf_out);

/* Camera ray restatement (sample.hlsli:258-280). */
int oracle_camera_ray(const pt_camera* cam, uint32_t px, uint32_t py, uint32_t width, uint32_t height,
                      const float* u_pixel, const float* u_lens, float* origin, float* dir);

/* Light sampling restatement (lightsample.hlsli:13-108): returns emission rgb, light_dir,
 * light_dist, pdf for a shading point and three uniforms. */
int oracle_sample_light(const oracle_scene* s, const float* point, const float* u3, float* emission,
                        float* light_dir, float* light_dist, float* pdf);

/* The PNG path's ACES tone mapping (tonemap.cs.hlsl:30-66, ptdispatch.cpp:29) and write_png's
 * 8-bit conversion (imagewrite.cpp:39-42): n RGBA float pixels -> n RGBA8 pixels. */
int oracle_tonemap(const float* rgba, size_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif

#endif
