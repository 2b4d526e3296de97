/*
 * oracle.h - TEST INFRASTRUCTURE ONLY.
 *
 * C interface of the scalar CPU restatement of VulkanHybridRaytracer's hybrid
 * pass (shadowRaygen.rgen, hybridRaygen.rgen, hybridHit.rchit, miss.rmiss,
 * shadow.rmiss, helpers.glsl, random.glsl) together with a CPU LBVH build and
 * traversal that define the bit-exact result of the GPU library. It is the
 * checker of the GPU path and the CPU baseline of the benchmark; nothing in
 * the product links or calls it.
 *
 * Inputs use the POD layouts of include/hr/hr.h. Images are host float arrays
 * with the same layouts as the GPU buffers.
 */
#ifndef HR_ORACLE_H
#define HR_ORACLE_H

#include <stdint.h>
#include "../include/hr/hr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

typedef struct or_counters {
    uint64_t primary_rays, shadow_rays, reflection_rays, hit_shadow_rays, occluded_rays;
    uint64_t node_visits, tri_tests;
} or_counters;

/* random.glsl / PCG32 known-answer hooks */
uint32_t or_tea(uint32_t v0, uint32_t v1);
uint32_t or_lcg(uint32_t* state);
float or_rnd(uint32_t* state);
void or_pcg32_seq(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out);
float or_u01_stream(uint32_t pixel_seed, uint32_t purpose, uint32_t index, uint32_t k);
uint32_t or_morton3(uint32_t x, uint32_t y, uint32_t z);
/* treelet restructuring in scenes created afterwards (1, the default: as the
 * GPU library builds; 0: the plain Karras tree, for checks of that stage) */
void or_set_treelets(int on);

/* math hooks: out must hold n values */
void or_atan2(const float* y, const float* x, uint32_t n, float* out);
void or_asin(const float* x, uint32_t n, float* out);
void or_sincos2pi(const float* u, uint32_t n, float* s, float* c);
/* helpers.glsl scalar BRDF terms (for closed-form checks) */
float or_distribution_ggx(const float* N, const float* H, float roughness);
float or_geometry_smith(const float* N, const float* V, const float* L, float roughness);
void or_fresnel_schlick(float cos_theta, const float* F0, float* out3);
void or_sample_disk(const hr_light* light, const float* pos, const float* L,
                    float r1, float r2, float* dir_out);
/* sampling and glass hooks (for closed-form and distribution checks):
 * sampleGGX half-vector around unit N; sampleRect direction of stratum s of S
 * from pos towards the rectangle light centred at centre; the glass direction
 * of hybridRaygen.rgen:131-144 (V towards the eye) and of hybridHit.rchit:
 * 158-173 (rd = incoming ray direction). */
void or_sample_ggx(const float* N, float roughness, float r1, float r2, float* out3);
void or_sample_rect(const float* centre, const hr_light_shape* shape, const float* pos, uint32_t s, uint32_t S,
                    float r1, float r2, float* out3);
void or_glass_raygen(const float* N, const float* V, float ior, float* out3);
void or_glass_hit(const float* N, const float* rd, float ior, float* out3);
/* adaptive shadow sampling: the cell (cx of nx columns, cy of ny rows) of
 * sample s of a nested sequence of S, the stopping rule and the next stage */
void or_nested_cell(uint32_t s, uint32_t S, uint32_t* cx, uint32_t* nx, uint32_t* cy, uint32_t* ny);
int or_adaptive_stop(uint32_t k, uint32_t N, uint32_t S);
/* the ray or_trace_shadows traces for adaptive sample sI of light li at pixel
 * (x, y): 0 = none (factor set without rays), 1 = sampled light, 2 = hard light */
int or_adaptive_shadow_ray(const or_scene* s, uint32_t W, uint32_t H, const float* position, const float* normal,
                           const hr_light* lights, const hr_light_shape* shapes, uint32_t nl, const hr_camera* cam,
                           const hr_trace_params* params, uint32_t x, uint32_t y, uint32_t li, uint32_t sI,
                           hr_ray* ray);
uint32_t or_adaptive_next(uint32_t N, uint32_t S);
/* TESTS ONLY: plant one deliberate mistake (see PLANT_* in hr_oracle.cpp;
 * 0 restores the faithful oracle) to show that a check catches it. */
void or_set_plant(int plant);
/* Watertight ray/triangle test: returns 1 on a hit with t in [tmin, tmax]. */
int or_intersect_tri(const float* org, const float* dir, float tmin, float tmax,
                     const float* v0, const float* v1, const float* v2,
                     float* t, float* u, float* v);

/* scene: same inputs as hr_scene_build */
int or_scene_build(const hr_mesh* meshes, uint32_t mesh_count,
                   const hr_primitive* prims, uint32_t prim_count,
                   const hr_material* materials, uint32_t material_count,
                   uint32_t leaf_max, or_scene** out);
int or_scene_refit(or_scene* s, const float* models);
int or_scene_set_env(or_scene* s, const hr_env* env);
/* material textures, as hr_scene_set_textures */
int or_scene_set_textures(or_scene* s, const hr_texture* textures, uint32_t count);
void or_scene_destroy(or_scene* s);
uint32_t or_scene_tri_count(const or_scene* s);
uint32_t or_scene_node_count(const or_scene* s);
uint32_t or_scene_depth(const or_scene* s);
/* which: as hr_scene_export */
int or_scene_export(const or_scene* s, uint32_t which, void* dst, uint64_t bytes);

/* ray queries; brute != 0 tests every triangle instead of walking the BVH */
void or_trace_rays(const or_scene* s, const hr_ray* rays, uint32_t n, int any_hit, int brute,
                   hr_hit* hits, or_counters* cnt, uint32_t threads);

/* passes over rows y = row_start, row_start + row_step, ... < H.
 * G-buffer planes are float4[H][W] host arrays (only the listed rows are
 * written / read). */
void or_gbuffer_primary(const or_scene* s, const hr_camera* cam, uint32_t W, uint32_t H,
                        float* position, float* normal, float* albedo, float* material,
                        float* emissive, uint32_t row_start, uint32_t row_step,
                        or_counters* cnt, uint32_t threads);
void or_trace_shadows(const or_scene* s, uint32_t W, uint32_t H,
                      const float* position, const float* normal,
                      const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                      const hr_camera* cam, const hr_trace_params* params, float* shadow,
                      uint32_t row_start, uint32_t row_step, or_counters* cnt, uint32_t threads);
void or_shade_hybrid(const or_scene* s, uint32_t W, uint32_t H,
                     const float* position, const float* normal, const float* albedo,
                     const float* material, const float* emissive,
                     const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                     const hr_camera* cam, const hr_trace_params* params, const float* shadow,
                     float* lit, float* refl, uint32_t row_start, uint32_t row_step,
                     or_counters* cnt, uint32_t threads);
/* display pass (postFragment.frag) over n float4 pixels; out / rgba8 may be NULL */
void or_post_gamma(const float* in, uint64_t n, float* out, uint8_t* rgba8);
/* pure ray tracing mode (raygen.rgen, closesthit.rchit): accum float4[H][W] */
void or_render_rt(const or_scene* s, const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                  const hr_camera* cam, const hr_trace_params* params, uint32_t W, uint32_t H, float* accum,
                  uint32_t row_start, uint32_t row_step, or_counters* cnt, uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
