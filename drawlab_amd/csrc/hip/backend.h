// Internal state of a dl_hip context (one device, one stream).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <map>
#include <unordered_map>
#include <vector>

#include "../../../include/dl_hip.h"
#include "../bvh_layout.h"

namespace dl {

constexpr int kTraceBlockHost = 128;  // threads per traversal block (trace.h kTraceBlock)
constexpr int kTraceBlocksPerCu = 16;  // at most: 8 waves per SIMD

// Device copy of the scene in the layouts the kernels read.  All arrays are
// indexed by global primitive id (mesh order, then face order) or global
// vertex id.
struct DevScene {
    // geometry
    float* pos = nullptr;        // 3 * nv   (global vertex positions)
    float* nrm = nullptr;        // 3 * nv   (zeros for meshes without normals)
    float* tan = nullptr;        // 3 * nv   (zeros for meshes without tangents)
    uint32_t* idx = nullptr;     // 3 * nprims (global vertex indices)
    uint32_t* prim_mesh = nullptr;  // nprims -> mesh index
    uint8_t* prim_info = nullptr;   // nprims -> material type | kPrimLight (one byte: 2 MB per 2M triangles)
    float* uv = nullptr;         // 2 * nv   (texture coordinates; zeros for meshes without)
    // bitmap textures: texels (rgb, 0) of all textures, per texture
    // (first texel, width, height, 0)
    float4* texels = nullptr;
    int4* tex_info = nullptr;
    // per mesh records (MeshRec)
    void* meshes = nullptr;
    // materials (dl_material_desc) and lights (LightRec)
    void* materials = nullptr;
    void* lights = nullptr;
    float* light_cdf = nullptr;  // concatenated area CDFs (dpdf.h)
    uint32_t* light_guide = nullptr;  // per emitter CDF: bucket index of its lower_bound (cdf_guide)
    // environment map
    float4* env_rgb = nullptr;   // w * h texels (rgb, 0)
    float* env_q = nullptr;
    uint32_t* env_alias = nullptr;
    float* env_pdf = nullptr;
    // acceleration structure
    float4* nodes = nullptr;
    float4* tris = nullptr;
};

// Per-mesh record read by the shading kernels.
// prim_info flag: the primitive's mesh is an area light.
constexpr uint32_t kPrimLight = 8u;
static_assert(DL_BSDF_COUNT <= (int)kPrimLight, "material types fit below the light flag");

struct MeshRec {
    uint32_t prim_offset;
    uint32_t num_prims;
    int32_t material;
    int32_t light;
    int32_t has_normals;
    int32_t has_tangents;
    int32_t has_texcoords;
    float area_norm;   // 1 / total area (DiscretePDF::getNormalization)
    uint32_t cdf_offset;  // into light_cdf (num_prims + 1 entries), emitters only
    uint32_t guide_offset;  // into light_guide (2^guide_bits + 1 entries), emitters only
    uint32_t guide_bits;
    int32_t pad;
};

// Bucket index of an area CDF c[0..n] (c[n] = 1): entry g, for g = 0 ..
// 2^bits, is lower_bound(c, g / 2^bits), so the lower_bound of any x in
// [g, g + 1) / 2^bits lies in [guide[g], guide[g + 1]] -- a search of that
// range finds the same index as a search of the whole CDF.
inline uint32_t cdf_guide_bits(uint32_t n) {
    uint32_t b = 0;
    while (b < 12 && (1u << b) < n) b++;
    return b;
}

struct LightRec {
    int32_t type;
    int32_t mesh;
    float radiance[3];
    float position[3];
    int32_t env_w, env_h;
    float env_scale;
    int32_t env_visual;
    uint32_t env_base;  // first texel of the environment in the env_* arrays
    int32_t pad[3];
};

// Lights, background and BVH root of one scene of a layout that holds
// several (a pass shared by images of different scenes); paths carry the
// index of their scene's record.
struct SceneRec {
    int32_t light0;      // first light of the scene in the light array
    int32_t num_lights;
    int32_t env_light;   // its environment light (index into the light array), or -1
    int32_t root;        // its BVH root node
    float bg[3];
    int32_t pad;
};

// Everything a context holds for one scene (geometry, BVH, materials,
// lights, environment, camera, film).  The current scene lives in the Ctx
// fields of the same names; scenes kept resident for a later
// dl_hip_scene_select live in Ctx::stash as SceneStates.
struct SceneState {
    bool has_scene = false;
    bool has_accel = false;
    uint32_t nprims = 0, nverts = 0;
    std::vector<float> tri_verts;
    std::vector<MeshRec> mesh_recs;
    std::vector<dl_material_desc> mats;
    std::vector<LightRec> light_recs;
    int env_light = -1;
    dl_camera_desc cam{};
    dl_film_desc film{};
    float bg[3] = {0, 0, 0};
    bool view_set = false;  // view options given (dl_hip_set_view_options)
    dl_view_options view{};
    float bbox_lo[3] = {0, 0, 0}, bbox_inv[3] = {1, 1, 1};
    DevScene ds;
    dl_accel_stats accel{};
    std::vector<void*> scene_allocs;
    uint32_t num_textures = 0;
    uint64_t num_texels = 0;
};

// What a pass over several scenes reads of one of them, current or resident.
struct SceneParts {
    uint32_t nprims = 0, nverts = 0;
    const std::vector<MeshRec>* meshes = nullptr;
    const std::vector<dl_material_desc>* mats = nullptr;
    const std::vector<LightRec>* lights = nullptr;
    int env_light = -1;
    const float* bg = nullptr;  // background (the view options' when set)
    bool hide = false;          // view options: hidden lights
    float eps = 0.0f;           // ray offset
    const dl_film_desc* film = nullptr;
    const DevScene* ds = nullptr;
    const dl_accel_stats* accel = nullptr;
    uint32_t num_textures = 0;
    uint64_t num_texels = 0;
};

// Device copy of several resident scenes in one layout (their geometry,
// materials, lights, environments, textures and BVHs concatenated, every
// index rebased), so that the paths of images of different scenes share one
// wavefront pass: each path carries the index of its scene's record
// (SceneRec: its lights, environment, background and BVH root).  Built from
// the scenes' device arrays; a member that changes (rebuilt, replaced,
// dropped) invalidates it.
struct SceneUnion {
    std::vector<uint64_t> keys;  // member scenes in record order (0: the current scene when it has no key)
    std::vector<int32_t> light0, root;  // per member: first light, BVH root node
    DevScene ds;
    SceneRec* recs = nullptr;  // device, one per member
    int32_t* roots = nullptr;  // device, one per member
    std::vector<SceneRec> host_recs;  // source of the last upload of recs (outlives its asynchronous copy)
    std::vector<void*> allocs;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    char devname[256] = {0};
    int num_cus = 0;
    double last_kernel_ms = 0.0;
    bool bvh_too_deep = false;  // the last GPU build's tree exceeded kMaxBinaryDepth
    // traversal stack overflow (entries kLdsStack..kMaxStack of every lane of
    // a persistent traversal grid)
    int* stack_ovf = nullptr;
    uint32_t trace_blocks = 0;      // blocks the stack overflow holds (kTraceBlocksPerCu per CU)
    uint32_t trace_grid = 0;        // blocks of a persistent traversal launch: all resident at once
    uint32_t trace_grid_count = 0;  // the same for the counting variants
    uint32_t* work = nullptr;  // work counters of the batch traversal kernels

    // scene
    bool has_scene = false;
    bool has_accel = false;
    uint32_t nprims = 0, nverts = 0;
    std::vector<float> tri_verts;  // 9 floats per prim (host copy for builders)
    std::vector<MeshRec> mesh_recs;
    std::vector<dl_material_desc> mats;
    std::vector<LightRec> light_recs;
    int env_light = -1;
    dl_camera_desc cam{};
    dl_film_desc film{};
    float bg[3] = {0, 0, 0};
    bool view_set = false;  // view options given (dl_hip_set_view_options); they belong to the scene
    dl_view_options view{};
    float bbox_lo[3] = {0, 0, 0}, bbox_inv[3] = {1, 1, 1};  // scene bounds: lo, 1 / extent per axis
    DevScene ds;
    dl_accel_stats accel{};
    std::vector<void*> scene_allocs;
    uint32_t num_textures = 0;
    uint64_t num_texels = 0;
    uint64_t scene_key = 0;                 // nonzero: keep the current scene resident when replaced
    std::map<uint64_t, SceneState> stash;   // resident scenes by key
    SceneUnion* uni = nullptr;              // resident scenes in one layout (passes over several scenes)

    // render scratch (reused across calls, grown on demand)
    std::vector<void*> render_allocs;
    size_t render_capacity = 0;
    void* rs = nullptr;  // RenderState*
    // films of shared-pass images (dl_hip_render_batch), by slot: float4
    // accumulators and resolved RGB of w x h pixels (room for cap)
    struct FilmSlot {
        float* film = nullptr;
        float* rgb = nullptr;
        uint32_t w = 0, h = 0, cap = 0;
    };
    std::vector<FilmSlot> film_slots;
    void* guide = nullptr;  // GuideState* (learned directional proposal)

    // device memory held by this context (every allocation goes through
    // dmalloc / dfree): current and peak bytes
    std::unordered_map<void*, size_t> dev_allocs;
    size_t dev_bytes = 0, dev_peak = 0;
};

// Tracked device allocation of a context.
inline hipError_t dmalloc(Ctx* c, void** p, size_t bytes) {
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        c->dev_allocs[*p] = bytes;
        c->dev_bytes += bytes;
        if (c->dev_bytes > c->dev_peak) c->dev_peak = c->dev_bytes;
    }
    return e;
}
inline hipError_t dfree(Ctx* c, void* p) {
    if (!p) return hipSuccess;
    auto it = c->dev_allocs.find(p);
    if (it != c->dev_allocs.end()) {
        c->dev_bytes -= it->second;
        c->dev_allocs.erase(it);
    }
    return hipFree(p);
}

dl_status set_error(Ctx* c, dl_status s, const std::string& msg);
dl_status hip_fail(Ctx* c, hipError_t e, const char* what);

#define DL_HIP_CHECK(ctx, call)                                    \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return ::dl::hip_fail((ctx), e_, #call); \
    } while (0)

// trace.hip launchers
dl_status launch_trace_closest(Ctx* c, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v,
                               uint32_t* cn, uint32_t* ct);
dl_status launch_trace_shadow(Ctx* c, const float* rays, uint32_t n, uint8_t* occ, uint32_t* cn = nullptr,
                              uint32_t* ct = nullptr);
dl_status noalloc_scratch_init(Ctx* c);
int render_trace_blocks_per_cu(bool counting);
int batch_trace_blocks_per_cu(bool counting);
dl_status launch_trace_closest_noalloc(Ctx* c, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                                       float* v);

// integrator.hip
dl_status render(Ctx* c, const dl_render_params* p, float* rgb_host, dl_render_stats* stats, int paths_sample,
                 float* paths_out);
void free_render_state(Ctx* c);
dl_status render_batch(Ctx* c, const dl_render_params* shared, const dl_batch_item* items, uint32_t n,
                       dl_render_stats* stats);
void free_film_slots(Ctx* c);
void invalidate_film(Ctx* c);  // the next render clears the film even with DL_RENDER_ACCUMULATE

// guide.hip
struct GuideState;
struct Net;
void free_guide(Ctx* c);
dl_status guide_init(Ctx* c, uint64_t seed, float lr, const float* params);
// bf16 weights / biases of the current network (guide must be initialised)
void guide_net_ptrs(Ctx* c, const uint16_t** w, const uint16_t** wt, const float** bias);
// one training step on R records; rad: final path radiance (float4 per path)
dl_status guide_train_records(Ctx* c, const float* records, uint32_t R, const float4* rad, int steps);
int guide_step_count(Ctx* c);  // Adam steps taken since guide_init

// scene_union.hip
// The scene of key k (0: the current scene) as a pass reads it; false if
// there is none.
bool scene_parts(Ctx* c, uint64_t key, SceneParts& out);
// A union holding every scene of `keys` (normalised: 0 stands for the
// current scene when it has no key); reused while it holds them all,
// otherwise rebuilt over every scene resident on the context.
dl_status scene_union_for(Ctx* c, const std::vector<uint64_t>& keys, SceneUnion** out);
void free_scene_union(Ctx* c);

// bvh_gpu.hip
dl_status build_bvh_gpu(Ctx* c, dl_accel_stats* stats, int method);
// Binary nodes of any builder -> the 4-wide nodes the traversal reads.
dl_status collapse_bvh4(Ctx* c, uint32_t num_binary, dl_accel_stats* stats);

}  // namespace dl
