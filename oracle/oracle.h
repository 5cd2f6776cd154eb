// CPU oracle C API -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load liboracle.so, and only as the checker / the timed CPU baseline.  The
// product (drawlab_amd/) never links or calls it.
//
// Where the reference's two backends disagree, the oracle (and the HIP
// backend it checks) take these decisions:
//  * sampling streams: one pcg32 per (pixel, sample), seeded with
//    initstate = pixel + (seed << 32), initseq = sample index; draw order
//    pixel 2D, aperture 2D, then per bounce light 2D (+1D when the picked
//    emitter is the environment map), BSDF 2D, Russian roulette 1D.
//    (the reference seeds one pcg32 per 32x32 block:
//    src/tracer/sampler/independent.cpp:30-35)
//  * closest-hit ties: lowest global primitive index (mesh order, then
//    face order); the reference octree keeps the last equal hit
//    (src/tracer/accel/octree.cpp:63-70), available as DLO_ACCEL_OCTREE.
//    Its unpadded slab test also rejects some hits of rays aimed exactly at
//    triangle edges; renders use DLO_ACCEL_OCTREE_ROBUST.
//  * maxDepth / rrDepth: path-length cap (-1 = none, as the reference) and
//    the bounce index from which Russian roulette runs (4 = reference
//    `depth > 3`, src/tracer/integrator/path.cpp:90).
//  * a miss adds the environment map with MIS (src/optix/cuda/miss.cu:11-49);
//    the CPU path has no environment emitter (envmap.cpp:21-33 throws).
//  * point lights include the receiver cosine and are delta (MIS weight 1),
//    as include/optix/light/point.h:15-31 and hitgroup_pgs.cu:99.
//  * the environment importance table uses the same row orientation as the
//    lookup (the reference builds it from the vertically flipped bitmap,
//    envmap.cpp:200-233 vs texture/bitmap.cpp:86).
//  * a path whose throughput becomes exactly zero stops (all later terms
//    are products with zero).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DLO_BSDF_DIFFUSE = 0,     // albedo[3]
    DLO_BSDF_MIRROR = 1,      // -
    DLO_BSDF_DIELECTRIC = 2,  // intIOR, extIOR
    DLO_BSDF_MICROFACET = 3,  // alpha, intIOR, extIOR, kd[3]
    DLO_BSDF_ANISO_GGX = 4,   // pd[3], ps[3], ax, ay
    DLO_BSDF_CONDUCTOR = 5,   // alpha, eta[3], k[3]
};
enum { DLO_EMITTER_AREA = 0, DLO_EMITTER_POINT = 1, DLO_EMITTER_ENV = 2 };
enum { DLO_CAMERA_VIRTUAL = 0, DLO_CAMERA_PERSPECTIVE = 1, DLO_CAMERA_OPENCV = 2 };
enum { DLO_FILTER_GAUSSIAN = 0, DLO_FILTER_BOX = 1, DLO_FILTER_TENT = 2, DLO_FILTER_MITCHELL = 3 };
// BRUTE: every triangle, lowest index on equal t (the semantic definition).
// OCTREE: the reference octree exactly (last equal hit, unpadded slabs).
// OCTREE_LOWEST: reference octree traversal with the lowest-index rule.
// OCTREE_ROBUST: the reference octree with conservative box tests and the
//   lowest-index rule; equal to BRUTE (checked by the tests).
enum { DLO_ACCEL_BRUTE = 0, DLO_ACCEL_OCTREE = 1, DLO_ACCEL_OCTREE_LOWEST = 2, DLO_ACCEL_OCTREE_ROBUST = 3 };

void* dlo_scene_new(void);
void dlo_scene_free(void* scene);
int dlo_scene_add_bsdf(void* scene, int type, const float* params, int n);
int dlo_scene_add_texture(void* scene, const float* rgb, int w, int h);
void dlo_scene_set_bsdf_textures(void* scene, int bsdf, const int* tex5, int object_space);
void dlo_texture_sample(const float* rgb, int w, int h, const float* uv, uint32_t n, float* out);
int dlo_scene_add_mesh(void* scene, const float* V, uint32_t nv, const uint32_t* F, uint32_t nf, const float* N,
                       const float* UV, int bsdf, const float* radiance);
int dlo_scene_add_point_light(void* scene, const float* position, const float* intensity);
int dlo_scene_add_envmap(void* scene, const float* rgb, int w, int h, float scale, int visual);
void dlo_scene_set_camera_virtual(void* scene, const float* eye, const float* lookat, const float* up, float fov,
                                  int width, int height);
void dlo_scene_set_camera_perspective(void* scene, const float* sample_to_camera, const float* camera_to_world,
                                      int width, int height, float near_clip, float far_clip);
void dlo_scene_set_camera_basis(void* scene, const float* eye_u_v_w, int width, int height);
void dlo_scene_set_filter_table(void* scene, float radius, const float* table33);
void dlo_scene_get_camera(void* scene, float* out12);
void dlo_scene_set_filter(void* scene, int type, float radius, float p0, float p1);
void dlo_scene_get_filter_table(void* scene, float* out33);
void dlo_scene_set_background(void* scene, const float* rgb);
/* OpencvCamera: camera_to_world = inverse extrinsic (row-major), eye,
   intrinsics (cx, cy, fx, fy). */
void dlo_scene_set_camera_opencv(void* scene, const float* camera_to_world, const float* eye,
                                 const float* intrinsics, int width, int height);
/* Integrator and viewer options: normals != 0 renders |shading normal| of
   the first hit (normal.cpp:9-19); hide_lights != 0 makes area-light
   surfaces transparent to path rays (hitgroup_pgs.cu:55-64); eps is the ray
   offset (the CPU path's Epsilon 1e-4 by default). */
void dlo_scene_set_options(void* scene, int normals, int hide_lights, float eps);
int dlo_scene_build(void* scene);
uint32_t dlo_scene_num_prims(void* scene);

void dlo_trace_closest(void* scene, int mode, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                       float* v);
void dlo_trace_shadow(void* scene, int mode, const float* rays, uint32_t n, uint8_t* occluded);
void dlo_camera_rays(void* scene, const float* pos, uint32_t n, float* rays);
void dlo_render_samples(void* scene, const int* xys, uint32_t n, uint32_t seed, int max_depth, int rr_depth,
                        int accel, float* out);
int dlo_render(void* scene, int spp, uint32_t seed, int max_depth, int rr_depth, int accel, int threads,
               int max_blocks, float* rgb_out, double* stats);

void dlo_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out);
void dlo_pcg32_float(uint64_t initstate, uint64_t initseq, uint32_t n, float* out);
void dlo_warp(int which, const float* in, uint32_t n, float a0, float a1, float* out);
void dlo_optics(int which, const float* in, uint32_t n, float* out);
void dlo_bsdf_query(void* scene, int bsdf, const float* in, uint32_t n, float* out);
// Scene::sampleEmitterDirection + pdfEmitterDirection at shading points.
// in: p(3) shading normal(3) uniforms(3) per query (the third uniform is
// the environment map's extra draw); out: spectrum(3) d(3) dist pdf
// pdf_of_direction emitter delta = 11 floats.
void dlo_emitter_query(void* scene, const float* in, uint32_t n, float* out);
void dlo_env_tables(void* scene, float* q, uint32_t* alias, float* pdf);
uint32_t dlo_spiral_blocks(int W, int H, int* out);

// Planted mistakes, used only by the tests that show each pin of the oracle
// fails when the oracle is wrong.  0 (none) everywhere else.  Environment
// tables take the setting when the map is added.
enum {
    DLO_MUT_NONE = 0,
    DLO_MUT_GGX_PDF_NO_WHZ = 1,         // GGX lobe pdf D/(4 wh.wo) without wh.z
    DLO_MUT_DIELECTRIC_SWAP_IOR = 2,    // intIOR and extIOR exchanged
    DLO_MUT_DIFFUSE_UNIFORM_PDF = 3,    // diffuse pdf 1/(2 pi) for a cosine sampler
    DLO_MUT_EMITTER_HIT_NO_MIS = 4,     // emitter hits weighted 1 beside MIS-weighted NEE
    DLO_MUT_AREA_PDF_NO_DIST2 = 5,      // area light solid-angle pdf without dist^2
    DLO_MUT_CONDUCTOR_SWAP_ETA_K = 6,   // conductor Fresnel with eta and k exchanged
    DLO_MUT_RR_NO_DIVIDE = 7,           // Russian roulette survivors not divided by p
    DLO_MUT_MICROFACET_G_WI_TWICE = 8,  // Beckmann G = G1(wi) G1(wi)
    DLO_MUT_POINT_INV_DIST = 9,         // point light I cos / d instead of / d^2
    DLO_MUT_ENV_PDF_TIMES_AREA = 10,    // environment pdf times the pixel's solid angle
    DLO_MUT_DIELECTRIC_NO_FRESNEL = 11, // dielectric always refracts (no Fresnel choice)
};
void dlo_set_mutation(int m);

#ifdef __cplusplus
}
#endif
