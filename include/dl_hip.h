/*
 * dl_hip.h -- C ABI of the DrawLab MI355X (gfx950) rendering backend.
 *
 * A third backend beside the reference's CPU (TBB) and OptiX ones.  Host
 * code (the C++ plugin layer in drawlab_amd/csrc/host, or any FFI caller)
 * exports the scene as plain descriptors and drives HIP through these
 * entry points.  Conventions: every call returns a dl_status; on failure
 * dl_hip_last_error(ctx) describes it.  Host memory is caller-owned and only
 * read during the call; device memory is owned by the context.  A context is
 * used by one host thread at a time.
 *
 * Entry points and the reference interfaces they replace:
 *   dl_hip_create / dl_hip_destroy
 *       DeviceContext ctor/dtor        src/optix/host/device_context.cpp:29-120
 *   dl_hip_upload_scene
 *       OptixRenderer::initOptixScene  src/tracer/backend/optix_renderer.cpp:25-82
 *       (addMesh/addLight/addMaterial/updateCamera, include/optix/host/optix_scene.h:16-76)
 *   dl_hip_build_accel
 *       Accel::build / OCTree::build   include/tracer/accel.h:13-58,
 *                                      src/tracer/accel/octree.cpp:135-151
 *       OptixAccel::build              src/optix/host/optix_accel.cpp:26-152
 *   dl_hip_trace_closest / dl_hip_trace_shadow
 *       Accel::rayIntersect(ray, its, shadowRay)  include/tracer/accel.h:50-57,
 *                                      src/tracer/accel/octree.cpp:153-212
 *   dl_hip_render
 *       CPURenderer::renderAsync       src/tracer/backend/cpu_renderer.cpp:57-135
 *       OptixScene::render + launch    src/optix/host/optix_scene.cpp:125-140,
 *                                      src/optix/host/device_context.cpp:303-309
 *   dl_hip_render_paths
 *       PathIntegrator::Li per camera sample (cpu_renderer.cpp:36-51,
 *       src/tracer/integrator/path.cpp:13-100), for per-path parity checks.
 */
#ifndef DL_HIP_H
#define DL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dl_status {
    DL_OK = 0,
    DL_ERR_INVALID = 1, /* bad argument or descriptor */
    DL_ERR_HIP = 2,     /* a HIP runtime call failed */
    DL_ERR_OOM = 3,     /* device allocation failed */
    DL_ERR_STATE = 4,   /* call out of order (e.g. render before upload) */
    DL_ERR_NO_DEVICE = 5
} dl_status;

/* BSDF types and the meaning of dl_material_desc.params (reference plugin names) */
enum {
    DL_BSDF_DIFFUSE = 0,    /* "diffuse":    albedo[3]                      diffuse.cpp:16-19  */
    DL_BSDF_MIRROR = 1,     /* "mirror":     -                              mirror.cpp         */
    DL_BSDF_DIELECTRIC = 2, /* "dielectric": intIOR, extIOR                 dielectric.cpp:26-32 */
    DL_BSDF_MICROFACET = 3, /* "microfacet": alpha, intIOR, extIOR, kd[3]   microfacet.cpp:13-33 */
    DL_BSDF_ANISO_GGX = 4,  /* "aniso_ggx":  pd[3], ps[3], ax, ay           aniso_ggx.cpp:12-20 */
    DL_BSDF_CONDUCTOR = 5,  /* "conductor":  alpha, eta[3], k[3] (GGX, complex-IOR Fresnel;
                               no reference counterpart) */
    DL_BSDF_COUNT = 6
};

enum { DL_LIGHT_AREA = 0, DL_LIGHT_POINT = 1, DL_LIGHT_ENV = 2 };
enum { DL_CAMERA_VIRTUAL = 0, DL_CAMERA_PERSPECTIVE = 1, DL_CAMERA_OPENCV = 2 };

typedef struct dl_mesh_desc {
    const float* positions;    /* 3 * num_vertices, world space */
    const uint32_t* indices;   /* 3 * num_triangles */
    const float* normals;      /* 3 * num_vertices, or NULL */
    const float* texcoords;    /* 2 * num_vertices, or NULL */
    const float* tangents;     /* 3 * num_vertices, or NULL (Mesh::activate, mesh.cpp:30-44) */
    uint32_t num_vertices;
    uint32_t num_triangles;
    int32_t material;          /* index into dl_scene_desc.materials */
    int32_t light;             /* index into dl_scene_desc.lights, -1 if not an emitter */
} dl_mesh_desc;

/* Texture slots of a material (texture children of the reference plugins):
   diffuse: DL_TEX_ALBEDO ("albedo", diffuse.cpp:69-75);
   aniso_ggx: DL_TEX_PD, DL_TEX_PS, DL_TEX_AXAY, DL_TEX_NORMAL, DL_TEX_TANGENT
   ("pd", "ps", "axay", "normal", "tangent", aniso_ggx.cpp:43-58). */
enum {
    DL_TEX_ALBEDO = 0,
    DL_TEX_PD = 0,
    DL_TEX_PS = 1,
    DL_TEX_AXAY = 2,
    DL_TEX_NORMAL = 3,
    DL_TEX_TANGENT = 4,
    DL_TEX_SLOTS = 5
};

typedef struct dl_material_desc {
    int32_t type;              /* DL_BSDF_* */
    float params[16];
    int32_t tex[DL_TEX_SLOTS]; /* 1 + index into dl_scene_desc.textures; 0 = the constant in params */
    int32_t maps_object_space; /* aniso_ggx normal / tangent maps: 0 = tangent space (the reference
                                  default is_tangent_space = true), 1 = object space */
} dl_material_desc;

/* Bitmap texture (BitmapTexture, src/tracer/texture/bitmap.cpp), sampled as
   the reference's GPU path samples it (src/optix/host/cuda_texture.cpp:
   normalized coordinates, bilinear filter, wrap addressing; the texel
   centres at ((i + 1/2) / width, (j + 1/2) / height)), with binary32
   interpolation weights.  rgb: height x width x 3 floats, row 0 at v = 0 --
   BitmapTexture::createCUDATexture uploads float images (EXR, HDR) flipped
   upside down and 8-bit images as stored, read as value / 255. */
typedef struct dl_texture_desc {
    int32_t width, height;
    const float* rgb;
} dl_texture_desc;

typedef struct dl_light_desc {
    int32_t type;              /* DL_LIGHT_* */
    int32_t mesh;              /* area: the emitting mesh */
    float radiance[3];         /* area: radiance; point: intensity */
    float position[3];         /* point */
    const float* env_rgb;      /* env: height x width x 3, row 0 = +y (theta = 0) */
    int32_t env_width, env_height;
    float env_scale;
    int32_t env_visual;
} dl_light_desc;

/* DL_CAMERA_VIRTUAL: VirtualCamera (src/tracer/camera/virtual.cpp:16-65)
   after its constructor -- eye, U, V, W as the reference computes them.
   DL_CAMERA_PERSPECTIVE: PerspectiveCamera (src/tracer/camera/perspective.cpp:
   21-95) after activate() -- the sample-to-camera and camera-to-world
   matrices (row-major), 1/size and the clip distances.
   DL_CAMERA_OPENCV: OpencvCamera (src/tracer/camera/opencv.cpp:9-24, rays as
   include/optix/camera/opencv.h:24-47 casts them) -- camera_to_world = the
   inverse of the extrinsic matrix (row-major), eye = its translation,
   intrinsics = (cx, cy, fx, fy) in pixels; film row 0 is image row 0. */
typedef struct dl_camera_desc {
    int32_t type;              /* DL_CAMERA_* */
    int32_t width, height;
    float eye[3], U[3], V[3], W[3];
    float sample_to_camera[16];
    float camera_to_world[16];
    float inv_size[2];
    float near_clip, far_clip;
    float intrinsics[4];       /* DL_CAMERA_OPENCV: cx, cy, fx, fy */
} dl_camera_desc;

/* ImageBlock filter tabulation (src/core/bitmap/block.cpp:12-27):
   table[i] = filter(radius * i / 32) for i < 32, table[32] = 0.
   dl_hip_upload_scene accepts radii in (0, 32] with border and lookup as
   dl_film_filter sets them (DL_ERR_INVALID "bad reconstruction filter"). */
typedef struct dl_film_desc {
    float radius;
    int32_t border;            /* ceil(radius - 0.5) */
    float lookup;              /* 32 / radius */
    float table[33];
} dl_film_desc;

typedef struct dl_scene_desc {
    const dl_mesh_desc* meshes;
    uint32_t num_meshes;
    const dl_material_desc* materials;
    uint32_t num_materials;
    const dl_light_desc* lights;  /* in Scene::m_emitters order (scene.cpp:36-81) */
    uint32_t num_lights;
    dl_camera_desc camera;
    dl_film_desc film;
    float bg_color[3];
    const dl_texture_desc* textures;  /* referenced by dl_material_desc.tex */
    uint32_t num_textures;
} dl_scene_desc;

typedef struct dl_accel_stats {
    uint32_t num_nodes;
    uint32_t num_leaves;
    uint32_t max_depth;
    uint32_t num_prims;
    double build_ms;
    double sah_cost;
    uint64_t node_bytes;       /* S_node * num_nodes */
    uint64_t tri_bytes;        /* S_tri * num_prims */
} dl_accel_stats;

/* DL_BUILD_GPU: PLOC clustering on the device (SAH-like quality);
   DL_BUILD_HOST_SAH: binned SAH on the host (object-median splits below
   depth 40, leaves of at most 255 triangles); DL_BUILD_GPU_LBVH: Karras LBVH.
   A GPU-built tree deeper than 64 levels (pathological geometry) is rebuilt
   by the host builder within the same call. */
enum { DL_BUILD_GPU = 0, DL_BUILD_HOST_SAH = 1, DL_BUILD_GPU_LBVH = 2 };

typedef struct dl_render_params {
    int32_t spp;               /* samples per pixel (the sampler's sampleCount) */
    uint32_t seed;             /* stream key: initstate = pixel + (seed << 32) */
    int32_t max_depth;         /* -1 = unlimited (reference behaviour) */
    int32_t rr_depth;          /* RR when bounce index >= rr_depth; 4 = reference */
    int32_t spp_per_pass;      /* samples per pixel per wavefront pass (0 = auto: at most 2^27 paths;
                                  at most 2^29 paths when given) */
    int32_t first_sample;      /* index of the first sample (progressive rendering) */
    int32_t flags;             /* DL_RENDER_* */
    int32_t guide_steps;       /* DL_RENDER_GUIDED training: Adam steps after each pass (0 = 16 until
                                  the network has taken 256 steps, then 4) */
} dl_render_params;

enum {
    DL_RENDER_COUNT_TRAVERSAL = 1, /* count BVH node visits / triangle tests of closest-hit rays */
    DL_RENDER_ACCUMULATE = 2,      /* add to the film of the previous call instead of clearing it
                                      (progressive rendering, src/optix/cuda/integrator/path.cu:86-90) */
    DL_RENDER_GUIDED = 4,          /* sample directions from the learned proposal too (one-sample MIS
                                      with BSDF sampling) and train it online after every pass */
    DL_RENDER_GUIDE_FROZEN = 8,    /* with DL_RENDER_GUIDED: use the proposal without training it */
    DL_RENDER_NORMALS = 16         /* the "normals" integrator (src/tracer/integrator/normal.cpp:9-19):
                                      per camera sample the component-wise absolute value of the
                                      shading normal at the first hit, 0 on a miss */
};

typedef struct dl_render_stats {
    double paths;              /* camera samples traced */
    double closest_rays;       /* closest-hit queries issued */
    double shadow_rays;        /* shadow queries issued */
    double invalid_samples;    /* NaN / Inf / negative samples dropped (block.cpp:82-88) */
    double ms_total;           /* wall time of the render loop (hipEvents) */
    double ms_closest;         /* summed kernel time per stage */
    double ms_shadow;
    double ms_shade;
    double ms_film;
    double ms_other;
    double launches_closest;
    double launches_shadow;
    double nodes_closest;      /* with DL_RENDER_COUNT_TRAVERSAL: node visits of closest-hit rays */
    double tris_closest;       /* with DL_RENDER_COUNT_TRAVERSAL: triangle tests of closest-hit rays */
    double ms_guide;           /* with DL_RENDER_GUIDED: proposal inference kernels */
    double ms_train;           /* with DL_RENDER_GUIDED: online training (batches, gradients, Adam) */
    double train_steps;        /* with DL_RENDER_GUIDED: Adam steps taken during the render */
    double nodes_shadow;       /* with DL_RENDER_COUNT_TRAVERSAL: node visits of shadow rays */
    double tris_shadow;        /* with DL_RENDER_COUNT_TRAVERSAL: triangle tests of shadow rays */
} dl_render_stats;

/* ---- context ---------------------------------------------------------- */
dl_status dl_hip_create(int device, void** ctx);
dl_status dl_hip_destroy(void* ctx);
const char* dl_hip_last_error(void* ctx);
const char* dl_hip_device_name(void* ctx);

/* ---- scene and acceleration structure ---------------------------------- */
dl_status dl_hip_upload_scene(void* ctx, const dl_scene_desc* scene);
dl_status dl_hip_build_accel(void* ctx, int builder, dl_accel_stats* stats);

/* ---- traversal on device-resident batches ------------------------------
 * rays: 8 planes of n floats (ox, oy, oz, dx, dy, dz, tmin, tmax);
 * hits: prim (global primitive index, 0xffffffff on a miss), t, u, v.
 * Pointers are device pointers (dl_hip_malloc); the call is asynchronous on
 * the context's stream unless synchronised with dl_hip_sync. */
dl_status dl_hip_trace_closest(void* ctx, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v);
dl_status dl_hip_trace_shadow(void* ctx, const float* rays, uint32_t n, uint8_t* occluded);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dl_hip_trace_closest with the traversal's scratch (the
 * stack overflow of a fixed persistent grid of 1024 blocks and its work
 * counter) in a static per-device array of 92.3 MB, held until the process
 * ends, instead of the context's buffers.  The call allocates nothing;
 * inputs and outputs are the caller's device pointers, as for
 * dl_hip_trace_closest, and the results are identical.  n is at most 2^31
 * (DL_ERR_INVALID beyond).  Every call on one device uses that device's one
 * array: calls that overlap on one GPU (several contexts or threads) race.
 * Returns after the kernel has finished. */
dl_status dl_hip_trace_closest_noalloc(void* ctx, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                                       float* v);
/* Same with counters: per-ray node visits and triangle tests (uint32 each),
 * the traversal work of the algorithmic byte count (bench, CPU replay). */
dl_status dl_hip_trace_closest_counted(void* ctx, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                                       float* v, uint32_t* node_visits, uint32_t* tri_tests);
dl_status dl_hip_trace_shadow_counted(void* ctx, const float* rays, uint32_t n, uint8_t* occluded,
                                      uint32_t* node_visits, uint32_t* tri_tests);

/* ---- rendering --------------------------------------------------------- */
/* Renders spp samples per pixel; writes W*H*3 floats (row-major, row 0 =
 * top) to rgb_host if non-NULL. */
dl_status dl_hip_render(void* ctx, const dl_render_params* params, float* rgb_host, dl_render_stats* stats);
/* Radiance of every path of one sample index: out_host = W*H*5 floats
 * (sample position x, y, radiance r, g, b) for sample `sample`. */
dl_status dl_hip_render_paths(void* ctx, const dl_render_params* params, uint32_t sample, float* out_host);

/* ---- device memory and timing helpers ---------------------------------- */
dl_status dl_hip_malloc(void* ctx, size_t bytes, void** dptr);
dl_status dl_hip_free(void* ctx, void* dptr);
dl_status dl_hip_memcpy_h2d(void* ctx, void* dst, const void* src, size_t bytes);
dl_status dl_hip_memcpy_d2h(void* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel on the context stream
 * (asynchronous; both pointers 16-byte aligned).  Its known byte count also
 * calibrates hardware traffic counters.  (No reference counterpart: the
 * reference's CUDABuffer copies host<->device only,
 * include/optix/host/cuda_buffer.h:46-57.) */
dl_status dl_hip_memcpy_d2d(void* ctx, void* dst, const void* src, size_t bytes);
dl_status dl_hip_sync(void* ctx);
/* Kernel time of the last dl_hip_trace_* call (hipEvents on the context
 * stream, milliseconds). */
dl_status dl_hip_last_kernel_ms(void* ctx, double* ms);

/* Device memory held by the context (every allocation it made and has not
   freed): now, and the peak since the previous call (or since creation).
   (No reference counterpart.) */
dl_status dl_hip_memory_usage(void* ctx, uint64_t* current_bytes, uint64_t* peak_bytes);

/* Replaces the camera (and output size) of the uploaded scene; the film of
   the next render starts empty.  Replaces OptixScene::updateCamera / resize
   (src/optix/host/optix_scene.cpp:43-52). */
dl_status dl_hip_set_camera(void* ctx, const dl_camera_desc* camera);

/* ---- shared passes -------------------------------------------------------
   One image of a pass that several images share (dl_hip_render_batch): its
   scene, its camera, the samples it adds in this pass and the film slot
   that accumulates them.  Each image's samples and film are exactly those
   of a dl_hip_render of it (its scene current) with the same pass sizes, so
   its result is the same bit for bit. */
enum { DL_BATCH_MAX = 32 };
typedef struct dl_batch_item {
    dl_camera_desc camera;     /* the image's camera and size */
    int32_t spp;               /* samples per pixel added in this pass */
    int32_t first_sample;      /* index of the first of them */
    uint32_t seed;             /* stream key, as dl_render_params.seed */
    int32_t film;              /* film slot, 0 .. DL_BATCH_MAX - 1 (distinct within a pass) */
    int32_t accumulate;        /* nonzero: add to the slot's film (same image size); zero: start it empty */
    float* rgb_out;            /* host buffer of width * height * 3: the resolved image after this pass, or NULL */
    uint64_t scene;            /* key of a resident scene (dl_hip_scene_keep), or 0: the current scene */
} dl_batch_item;
/* Renders one wavefront pass holding the paths of n images (1..DL_BATCH_MAX)
   of the current scene or of resident scenes: the paths of every image
   share the pass's queues and kernels (camera rays, traversal, material-
   sorted shading, shadow rays), then each image's film gathers its own
   samples with its scene's filter.  Images of different scenes are rendered
   from one device layout holding every resident scene (built on first use,
   about as many device bytes as those scenes' geometry, BVHs, lights and
   textures, and kept until a resident scene changes or dl_hip_trim), each
   path carrying its scene's index.  The scenes of one pass must agree in
   hide_lights and epsilon (dl_view_options).  `shared` gives max_depth,
   rr_depth and flags (DL_RENDER_COUNT_TRAVERSAL and DL_RENDER_NORMALS;
   DL_RENDER_GUIDED is rejected); its other fields are ignored.  At most
   2^27 paths (sum of width * height * spp).  The film slots belong to the
   context and persist across calls and scene changes until dl_hip_trim.
   The current scene stays current.  (No reference counterpart: the render
   service packs jobs into shared passes with it.) */
dl_status dl_hip_render_batch(void* ctx, const dl_render_params* shared, const dl_batch_item* items, uint32_t n,
                              dl_render_stats* stats);

/* Viewer settings changed between renders without re-uploading the scene,
   as the OptiX renderer's context holds them (include/tracer/backend/
   context.h:12-14, applied by OptixRenderer::updateOptixScene,
   src/tracer/backend/optix_renderer.cpp:84-102). */
typedef struct dl_view_options {
    float bg_color[3];         /* colour of camera rays that miss when no visible environment map */
    float epsilon;             /* ray offset: tmin of camera, continuation and shadow rays and the
                                  shadow rays' end short of the light; 0 = 1e-4 (the CPU path's
                                  Epsilon, the default) */
    int32_t hide_lights;       /* nonzero: surfaces of area lights are invisible to camera rays and
                                  transparent to every path ray (hitgroup_pgs.cu:55-64): camera rays
                                  pass through without their emission; later rays collect it (MIS-
                                  weighted) and pass through; neither counts as a bounce nor draws
                                  random numbers.  Shadow rays are still blocked by them. */
} dl_view_options;
/* Sets the options (NULL restores the defaults: the uploaded scene's
   bg_color, epsilon 0, lights shown); they hold until changed or until
   another scene is uploaded or selected. */
dl_status dl_hip_set_view_options(void* ctx, const dl_view_options* opts);

/* Frees the context's scenes (current and resident), acceleration
   structures and render state (the context keeps only its fixed buffers;
   upload a scene again to render).  (No reference counterpart; used by the
   render service under memory pressure.) */
dl_status dl_hip_trim(void* ctx);

/* Copies the acceleration structure the traversal kernels read -- 4-wide
   nodes and pre-gathered triangles, layouts in drawlab_amd/csrc/bvh_layout.h
   -- to host buffers of stats.node_bytes and stats.tri_bytes bytes (either
   may be NULL).  (No reference counterpart: for CPU replays of the
   traversal that recompute its node visits and triangle tests.) */
dl_status dl_hip_accel_export(void* ctx, void* nodes, void* tris);

/* Resident scenes (no reference counterpart; for callers that render
   several scenes in turn, such as the render service).
   dl_hip_scene_keep tags the current scene (uploaded and built) with a
   nonzero key: when another scene is uploaded or selected, it stays in
   device memory instead of being freed.  dl_hip_scene_select makes the
   scene kept under `key` current again -- geometry, BVH, materials,
   lights, environment, camera and film as they were -- without upload or
   build (DL_ERR_STATE when no scene is kept under that key).
   dl_hip_scene_drop frees the scene kept under `key` (0: every kept scene
   that is not current; the current scene too when it is the one named).
   dl_hip_scene_bytes gives the device bytes the scene kept under `key` (0:
   the current scene) holds: its geometry, materials, lights, environment,
   textures and BVH (DL_ERR_STATE when there is no such scene). */
dl_status dl_hip_scene_keep(void* ctx, uint64_t key);
dl_status dl_hip_scene_select(void* ctx, uint64_t key);
dl_status dl_hip_scene_drop(void* ctx, uint64_t key);
dl_status dl_hip_scene_bytes(void* ctx, uint64_t key, uint64_t* bytes);

/* ---- learned directional proposal ----------------------------------------
 * A mixture of 8 von Mises-Fisher lobes per shading point predicted by a
 * 96-64-64-64-64-48 ReLU MLP (bfloat16 weights and activations, float32
 * accumulation on the matrix cores), trained online (KL divergence from the
 * incident radiance, Adam).  Rendering with DL_RENDER_GUIDED samples
 * directions from it in one-sample MIS with BSDF sampling (50/50).  Inputs
 * and outputs below are device arrays (dl_hip_malloc): x is n x 96 features
 * (layout of oracle/guide.py), out n x 48 raw outputs, mix n x 40 (w, kappa,
 * mu per lobe), dirs n x 3 unit directions, weights n training weights.
 * (No reference counterpart: the reference has no path guiding.) */
/* Resets the network: params (21,808 floats, host) or seeded initial weights. */
dl_status dl_hip_guide_init(void* ctx, uint64_t seed, float learning_rate, const float* params);
/* Copies the float32 parameters and the accumulated gradient (host arrays of
   21,808 floats; either may be NULL). */
dl_status dl_hip_guide_params(void* ctx, float* params, float* grads);
dl_status dl_hip_guide_eval(void* ctx, const float* x, uint32_t n, float* out, float* mix);
/* One step: loss = sum(-w log q(d)) / sum(w); apply = 0 accumulates the
   gradient without updating (for inspection). */
dl_status dl_hip_guide_train_step(void* ctx, const float* x, const float* dirs, const float* weights, uint32_t n,
                                  int apply, float* loss);

/* ---- render service ----------------------------------------------------
 * Many render jobs on one device: scenes are registered once (their arrays
 * copied), jobs name a scene and render parameters and run on `lanes`
 * worker contexts, each with its own HIP stream, so kernels of independent
 * jobs overlap on the hardware queues.  A lane keeps its last scene (BVH
 * included) resident for the next job of that scene.  Admission keeps the
 * device bytes reserved by all lanes within budget_bytes; a job that could
 * never fit fails.  Every job's image is bit-identical to dl_hip_render of
 * the same scene and parameters (a plain image does not depend on the pass
 * size; a guided one, whose network trains between passes, is identical for
 * the same spp_per_pass: a job with spp_per_pass 0 renders in passes of at
 * most 2^26 paths).
 * No reference counterpart: the reference renders one scene per process
 * (src/tracer/main.cpp:34-48); its closest analogue is the progressive
 * accumulation of src/optix/host/param_buffer.cpp:87-95. */
enum { DL_POLICY_FIFO = 0, DL_POLICY_SJF = 1 };  /* shortest job (pixels * spp) first */
enum { DL_JOB_QUEUED = 0, DL_JOB_RUNNING = 1, DL_JOB_DONE = 2, DL_JOB_FAILED = 3, DL_JOB_CANCELLED = 4 };

typedef struct dl_service_config {
    int32_t device;
    int32_t lanes;             /* worker contexts (>= 1) */
    int32_t policy;            /* DL_POLICY_* */
    int32_t builder;           /* DL_BUILD_* for scene uploads */
    uint64_t budget_bytes;     /* device memory all lanes together may reserve */
    int32_t max_batch;         /* jobs a lane packs into shared passes: 0 = DL_BATCH_MAX, 1 = one at a time */
    int32_t single_scene_passes; /* nonzero: only jobs of one scene share a pass; zero: jobs of every scene
                                  resident on the lane do (one layout holding those scenes, reserved
                                  as about their bytes again) */
} dl_service_config;

typedef struct dl_job_info {
    int32_t state;             /* DL_JOB_* */
    int32_t lane;              /* lane that ran it, -1 if none */
    double t_submit, t_start, t_finish;  /* seconds since the service started */
    double paths;              /* camera samples traced */
    double ms_gpu;             /* sum of the render calls' event time */
    uint64_t reserved_bytes;   /* admission estimate for the job's lane */
    uint64_t peak_bytes;       /* the lane context's device peak while it ran */
} dl_job_info;

typedef struct dl_service_stats {
    uint64_t jobs_done, jobs_failed, jobs_cancelled;
    uint64_t budget_bytes;
    uint64_t peak_reserved_bytes;  /* largest sum of lane reservations */
    uint64_t estimate_misses;      /* jobs whose lane peak exceeded its reservation */
    uint64_t scene_uploads;        /* uploads + BVH builds performed */
    uint64_t scene_selects;        /* jobs that switched a lane to a scene it kept resident */
    double paths;
    uint64_t batched_jobs;         /* jobs that joined a running group of shared passes */
    uint64_t shared_passes;        /* passes that held more than one job */
    uint64_t cross_scene_passes;   /* passes that held jobs of more than one scene */
} dl_service_stats;

dl_status dl_hip_service_start(const dl_service_config* cfg, void** svc);
/* Copies the scene's arrays; the descriptor may be freed afterwards. */
dl_status dl_hip_service_add_scene(void* svc, const dl_scene_desc* scene, uint64_t* scene_id);
/* camera: NULL renders the scene's camera; otherwise this job's camera
   (e.g. another output size). */
dl_status dl_hip_service_submit(void* svc, uint64_t scene_id, const dl_camera_desc* camera,
                                const dl_render_params* params, uint64_t* job_id);
dl_status dl_hip_service_poll(void* svc, uint64_t job_id, dl_job_info* info);
/* Blocks until the job is done, failed or cancelled. */
dl_status dl_hip_service_wait(void* svc, uint64_t job_id, dl_job_info* info);
/* W*H*3 floats of a finished job (row 0 = top, W x H of the job's camera);
   the image is then released. */
dl_status dl_hip_service_result(void* svc, uint64_t job_id, float* rgb_out);
dl_status dl_hip_service_cancel(void* svc, uint64_t job_id);
dl_status dl_hip_service_stats(void* svc, dl_service_stats* stats);
/* Cancels queued jobs, lets running ones finish their pass, frees all. */
dl_status dl_hip_service_stop(void* svc);
const char* dl_hip_service_last_error(void* svc);

/* ---- host-side helpers shared with the plugin layer -------------------- */
/* VirtualCamera constructor arithmetic (virtual.cpp:16-42). */
void dl_camera_virtual(const float eye[3], const float lookat[3], const float up[3], float fov, int width, int height,
                       dl_camera_desc* out);
/* PerspectiveCamera constructor + activate (perspective.cpp:21-70): fov is
   the horizontal field of view in degrees; to_world is row-major (NULL =
   identity).  Returns DL_ERR_INVALID if the projection is not invertible. */
dl_status dl_camera_perspective(const float to_world[16], float fov, float near_clip, float far_clip, int width,
                                int height, dl_camera_desc* out);
/* OpencvCamera constructor (opencv.cpp:9-24): extrinsic is the row-major
   world-to-camera matrix; its inverse, the eye (the inverse's translation)
   and the intrinsics (pixels) fill `out`.  Returns DL_ERR_INVALID if the
   extrinsic matrix is not invertible or a focal length is zero. */
dl_status dl_camera_opencv(const float extrinsic[16], float cx, float cy, float fx, float fy, int width, int height,
                           dl_camera_desc* out);
/* ReconstructionFilter + ImageBlock tabulation (rfilter.cpp:13-98,
 * block.cpp:12-27).  type: 0 gaussian(radius, stddev) 1 box 2 tent
 * 3 mitchell(radius, B, C). */
void dl_film_filter(int type, float radius, float p0, float p1, dl_film_desc* out);
/* Mesh::activate tangent computation (mesh.cpp:30-44, 144-177). */
void dl_mesh_tangents(const float* positions, uint32_t nv, const uint32_t* indices, uint32_t nf,
                      const float* texcoords, float* tangents_out);

#ifdef __cplusplus
}
#endif
#endif /* DL_HIP_H */
