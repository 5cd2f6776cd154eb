// C ABI entry points (include/dl_hip.h): context, scene upload, acceleration
// structure, traversal batches, device memory helpers.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../host/bvh_build.h"
#include "../host/scene_prep.h"
#include "backend.h"

#include <algorithm>

namespace dl {

dl_status set_error(Ctx* c, dl_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

dl_status hip_fail(Ctx* c, hipError_t e, const char* what) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return set_error(c, e == hipErrorOutOfMemory ? DL_ERR_OOM : DL_ERR_HIP, buf);
}

namespace {

template <typename T>
dl_status upload(Ctx* c, T** dst, const T* src, size_t count) {
    *dst = nullptr;
    if (count == 0) return DL_OK;
    void* p = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, &p, count * sizeof(T)));
    c->scene_allocs.push_back(p);
    if (src) DL_HIP_CHECK(c, hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    *dst = (T*)p;
    return DL_OK;
}

// Exchanges the context's current scene with s.
void swap_scene(Ctx* c, SceneState& s) {
    std::swap(c->has_scene, s.has_scene);
    std::swap(c->has_accel, s.has_accel);
    std::swap(c->nprims, s.nprims);
    std::swap(c->nverts, s.nverts);
    std::swap(c->tri_verts, s.tri_verts);
    std::swap(c->mesh_recs, s.mesh_recs);
    std::swap(c->mats, s.mats);
    std::swap(c->light_recs, s.light_recs);
    std::swap(c->env_light, s.env_light);
    std::swap(c->cam, s.cam);
    std::swap(c->film, s.film);
    std::swap(c->bg, s.bg);
    std::swap(c->view_set, s.view_set);
    std::swap(c->view, s.view);
    std::swap(c->bbox_lo, s.bbox_lo);
    std::swap(c->bbox_inv, s.bbox_inv);
    std::swap(c->ds, s.ds);
    std::swap(c->accel, s.accel);
    std::swap(c->scene_allocs, s.scene_allocs);
    std::swap(c->num_textures, s.num_textures);
    std::swap(c->num_texels, s.num_texels);
}

void free_scene(Ctx* c) {
    free_scene_union(c);
    if (c->stream) hipStreamSynchronize(c->stream);
    // the acceleration structure is allocated by the builders, outside
    // scene_allocs
    if (c->ds.nodes) dfree(c, c->ds.nodes);
    if (c->ds.tris) dfree(c, c->ds.tris);
    for (void* p : c->scene_allocs) dfree(c, p);
    SceneState empty;
    swap_scene(c, empty);  // resets every scene field
    c->scene_key = 0;
}

// Frees one resident (stashed) scene.
void free_stashed(Ctx* c, SceneState& s) {
    free_scene_union(c);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (s.ds.nodes) dfree(c, s.ds.nodes);
    if (s.ds.tris) dfree(c, s.ds.tris);
    for (void* p : s.scene_allocs) dfree(c, p);
    s = SceneState();
}

// Camera descriptor checks shared by upload and set_camera.
dl_status check_camera(Ctx* c, const dl_camera_desc& cam) {
    if (cam.width <= 0 || cam.height <= 0) return set_error(c, DL_ERR_INVALID, "bad camera size");
    // one sample per pixel must fit the largest pass (2^27 paths); pixel
    // indices are 32-bit throughout the integrator and film
    if ((uint64_t)cam.width * (uint64_t)cam.height > (1ull << 27))
        return set_error(c, DL_ERR_INVALID, "camera has more than 2^27 pixels");
    if (cam.type != DL_CAMERA_VIRTUAL && cam.type != DL_CAMERA_PERSPECTIVE && cam.type != DL_CAMERA_OPENCV)
        return set_error(c, DL_ERR_INVALID, "unknown camera type");
    if (cam.type == DL_CAMERA_OPENCV) {
        for (int k = 0; k < 4; k++)
            if (!std::isfinite(cam.intrinsics[k])) return set_error(c, DL_ERR_INVALID, "opencv intrinsics must be finite");
        if (cam.intrinsics[2] == 0.0f || cam.intrinsics[3] == 0.0f)
            return set_error(c, DL_ERR_INVALID, "opencv focal lengths must be nonzero");
    }
    return DL_OK;
}

// Releases the current scene before another becomes current: kept resident
// under its key when it has one, freed otherwise.
void release_current(Ctx* c) {
    invalidate_film(c);
    if (c->scene_key != 0 && c->has_scene) {
        SceneState s;
        swap_scene(c, s);
        c->stash[c->scene_key] = std::move(s);
        c->scene_key = 0;
    } else {
        free_scene(c);
    }
}

void free_all_scenes(Ctx* c) {
    free_scene(c);
    for (auto& kv : c->stash) free_stashed(c, kv.second);
    c->stash.clear();
}

#define DL_TRY(expr)                     \
    do {                                 \
        dl_status s_ = (expr);           \
        if (s_ != DL_OK) return s_;      \
    } while (0)

// Plain streaming copy: 16 B per lane per iteration, grid-stride; the tail
// (bytes % 16) is copied byte-wise by the first lanes.
__global__ __launch_bounds__(256) void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16,
                                              uint8_t* __restrict__ dtail, const uint8_t* __restrict__ stail,
                                              uint32_t ntail) {
    const size_t G = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = i0; i < n16; i += G) dst[i] = src[i];
    if (i0 < ntail) dtail[i0] = stail[i0];
}

}  // namespace
}  // namespace dl

using namespace dl;

extern "C" {

dl_status dl_hip_create(int device, void** out) {
    if (!out) return DL_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return DL_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return DL_ERR_INVALID;
    Ctx* c = new Ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return DL_ERR_HIP;
    }
    snprintf(c->devname, sizeof(c->devname), "%s (%s)", prop.name, prop.gcnArchName);
    c->num_cus = prop.multiProcessorCount;
    // persistent traversal grids: every block resident at once (at most 16
    // per CU, 8 waves per SIMD), as each kernel's registers allow
    const uint32_t cus = (uint32_t)std::max(1, c->num_cus);
    c->trace_blocks = cus * (uint32_t)kTraceBlocksPerCu;
    c->trace_grid = cus * (uint32_t)std::min(render_trace_blocks_per_cu(false), batch_trace_blocks_per_cu(false));
    c->trace_grid_count = cus * (uint32_t)std::min(render_trace_blocks_per_cu(true), batch_trace_blocks_per_cu(true));
    e = dmalloc(c, (void**)&c->stack_ovf, sizeof(int) * (size_t)c->trace_blocks * kTraceBlockHost * (kMaxStack - kLdsStack));
    if (e == hipSuccess) e = dmalloc(c, (void**)&c->work, 4 * sizeof(uint32_t));
    if (e != hipSuccess) {
        dl_hip_destroy(c);
        return DL_ERR_OOM;
    }
    // the device address of the allocation-free entry point's scratch, once
    // per device, with the device current
    if (noalloc_scratch_init(c) != DL_OK) {
        dl_hip_destroy(c);
        return DL_ERR_HIP;
    }
    *out = c;
    return DL_OK;
}

dl_status dl_hip_destroy(void* h) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    hipSetDevice(c->device);
    free_render_state(c);
    free_film_slots(c);
    free_guide(c);
    free_all_scenes(c);
    if (c->stack_ovf) dfree(c, c->stack_ovf);
    if (c->work) dfree(c, c->work);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return DL_OK;
}

const char* dl_hip_last_error(void* h) { return h ? ((Ctx*)h)->err.c_str() : "null context"; }
const char* dl_hip_device_name(void* h) { return h ? ((Ctx*)h)->devname : ""; }

dl_status dl_hip_upload_scene(void* h, const dl_scene_desc* sd) {
    Ctx* c = (Ctx*)h;
    if (!c || !sd) return DL_ERR_INVALID;
    hipSetDevice(c->device);
    free_scene_union(c);  // its member "current scene without a key" is about to change
    release_current(c);
    if (sd->num_meshes == 0) return set_error(c, DL_ERR_INVALID, "scene has no meshes");
    DL_TRY(check_camera(c, sd->camera));
    {
        // the reconstruction filter as dl_film_filter makes it: the film
        // gathers size their footprints from the radius and border, and
        // index the table with |offset| * lookup <= radius * lookup <= 32
        const dl_film_desc& f = sd->film;
        bool ok = std::isfinite(f.radius) && f.radius > 0.0f && f.radius <= 32.0f &&
                  f.border == (int32_t)std::ceil(f.radius - 0.5f) && std::isfinite(f.lookup) && f.lookup > 0.0f &&
                  f.lookup * f.radius <= 32.001f;
        for (int k = 0; k < 33 && ok; k++) ok = std::isfinite(f.table[k]);
        if (!ok) return set_error(c, DL_ERR_INVALID, "bad reconstruction filter");
    }
    for (uint32_t i = 0; i < sd->num_materials; i++) {
        if (sd->materials[i].type < 0 || sd->materials[i].type >= DL_BSDF_COUNT)
            return set_error(c, DL_ERR_INVALID, "unknown material type");
        for (int k = 0; k < DL_TEX_SLOTS; k++)
            if (sd->materials[i].tex[k] < 0 || (uint32_t)sd->materials[i].tex[k] > sd->num_textures)
                return set_error(c, DL_ERR_INVALID, "material texture index out of range");
    }
    uint64_t texels = 0;
    for (uint32_t t = 0; t < sd->num_textures; t++) {
        const dl_texture_desc& td = sd->textures[t];
        if (!td.rgb || td.width <= 0 || td.height <= 0 || (uint64_t)td.width * (uint64_t)td.height > (1ull << 28))
            return set_error(c, DL_ERR_INVALID, "invalid texture");
        texels += (uint64_t)td.width * (uint64_t)td.height;
    }
    if (texels >= 0x7fffffffull) return set_error(c, DL_ERR_INVALID, "textures too large");
    uint64_t nv = 0, nf = 0;
    for (uint32_t m = 0; m < sd->num_meshes; m++) {
        const dl_mesh_desc& md = sd->meshes[m];
        if (!md.positions || !md.indices) return set_error(c, DL_ERR_INVALID, "mesh without positions/indices");
        if (md.material < 0 || (uint32_t)md.material >= sd->num_materials)
            return set_error(c, DL_ERR_INVALID, "mesh material index out of range");
        if (md.light >= (int32_t)sd->num_lights) return set_error(c, DL_ERR_INVALID, "mesh light index out of range");
        for (uint64_t k = 0; k < 3ull * md.num_triangles; k++)
            if (md.indices[k] >= md.num_vertices) return set_error(c, DL_ERR_INVALID, "vertex index out of range");
        nv += md.num_vertices;
        nf += md.num_triangles;
    }
    if (nf >= 0x7fffffffull || nv >= 0x7fffffffull) return set_error(c, DL_ERR_INVALID, "scene too large");
    c->nprims = (uint32_t)nf;
    c->nverts = (uint32_t)nv;

    std::vector<float> pos(3 * nv), nrm(3 * nv, 0.0f), tan(3 * nv, 0.0f), uvs(2 * nv, 0.0f);
    std::vector<uint32_t> idx(3 * nf), prim_mesh(nf);
    c->tri_verts.assign(9 * nf, 0.0f);
    c->mesh_recs.assign(sd->num_meshes, MeshRec{});
    std::vector<float> cdf_all;
    std::vector<uint32_t> guide_all;
    uint64_t vo = 0, fo = 0;
    for (uint32_t m = 0; m < sd->num_meshes; m++) {
        const dl_mesh_desc& md = sd->meshes[m];
        MeshRec& r = c->mesh_recs[m];
        r.prim_offset = (uint32_t)fo;
        r.num_prims = md.num_triangles;
        r.material = md.material;
        r.light = md.light;
        r.has_normals = md.normals != nullptr;
        r.has_texcoords = md.texcoords != nullptr;
        r.has_tangents = md.tangents != nullptr || md.texcoords != nullptr;
        std::memcpy(&pos[3 * vo], md.positions, sizeof(float) * 3 * md.num_vertices);
        if (md.normals) std::memcpy(&nrm[3 * vo], md.normals, sizeof(float) * 3 * md.num_vertices);
        if (md.texcoords) std::memcpy(&uvs[2 * vo], md.texcoords, sizeof(float) * 2 * md.num_vertices);
        if (md.tangents)
            std::memcpy(&tan[3 * vo], md.tangents, sizeof(float) * 3 * md.num_vertices);
        else if (md.texcoords)
            dl_mesh_tangents(md.positions, md.num_vertices, md.indices, md.num_triangles, md.texcoords, &tan[3 * vo]);
        for (uint32_t f = 0; f < md.num_triangles; f++) {
            prim_mesh[fo + f] = m;
            for (int k = 0; k < 3; k++) {
                uint32_t vi = md.indices[3 * f + k];
                idx[3 * (fo + f) + k] = (uint32_t)(vo + vi);
                std::memcpy(&c->tri_verts[9 * (fo + f) + 3 * k], md.positions + 3 * (size_t)vi, 3 * sizeof(float));
            }
        }
        if (md.light >= 0) {
            std::vector<float> cdf;
            r.area_norm = area_cdf(md.positions, md.indices, md.num_triangles, cdf);
            r.cdf_offset = (uint32_t)cdf_all.size();
            cdf_all.insert(cdf_all.end(), cdf.begin(), cdf.end());
            // bucket index of the CDF; a CDF that rounding left non-monotone
            // (a degenerate last triangle) gets one bucket spanning every
            // entry, so its searches stay those of the whole CDF
            const bool monotone = std::is_sorted(cdf.begin(), cdf.end());
            r.guide_bits = monotone ? cdf_guide_bits(md.num_triangles) : 0;
            r.guide_offset = (uint32_t)guide_all.size();
            const uint32_t G = 1u << r.guide_bits;
            for (uint32_t g = 0; g <= G; g++) {
                const float key = (float)g / (float)G;  // exact: G is a power of two
                guide_all.push_back(!monotone ? (g == 0 ? 0u : md.num_triangles)
                                              : (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), key) -
                                                           cdf.begin()));
            }
        }
        vo += md.num_vertices;
        fo += md.num_triangles;
    }
    // scene bounds (network inputs use positions normalised to them)
    {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t i = 0; i < c->tri_verts.size(); i += 3)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], c->tri_verts[i + k]);
                hi[k] = std::max(hi[k], c->tri_verts[i + k]);
            }
        for (int k = 0; k < 3; k++) {
            c->bbox_lo[k] = lo[k];
            c->bbox_inv[k] = hi[k] > lo[k] ? 1.0f / (hi[k] - lo[k]) : 1.0f;
        }
    }
    c->mats.assign(sd->materials, sd->materials + sd->num_materials);
    c->light_recs.assign(sd->num_lights, LightRec{});
    c->env_light = -1;
    std::vector<float4> env_rgb;
    std::vector<float> env_q, env_pdf;
    std::vector<uint32_t> env_alias;
    for (uint32_t l = 0; l < sd->num_lights; l++) {
        const dl_light_desc& ld = sd->lights[l];
        LightRec& r = c->light_recs[l];
        r.type = ld.type;
        r.mesh = ld.mesh;
        for (int k = 0; k < 3; k++) {
            r.radiance[k] = ld.radiance[k];
            r.position[k] = ld.position[k];
        }
        if (ld.type == DL_LIGHT_AREA) {
            if (ld.mesh < 0 || (uint32_t)ld.mesh >= sd->num_meshes || sd->meshes[ld.mesh].light != (int32_t)l)
                return set_error(c, DL_ERR_INVALID, "area light must reference the mesh that references it");
        } else if (ld.type == DL_LIGHT_ENV) {
            if (c->env_light >= 0) return set_error(c, DL_ERR_INVALID, "There is more than one envmap in the scene!");
            if (!ld.env_rgb || ld.env_width <= 0 || ld.env_height <= 0)
                return set_error(c, DL_ERR_INVALID, "envmap without pixels");
            c->env_light = (int)l;
            r.env_w = ld.env_width;
            r.env_h = ld.env_height;
            r.env_scale = ld.env_scale;
            r.env_visual = ld.env_visual;
            const size_t n = (size_t)ld.env_width * ld.env_height;
            env_rgb.resize(n);
            for (size_t i = 0; i < n; i++)
                env_rgb[i] = make_float4(ld.env_rgb[3 * i], ld.env_rgb[3 * i + 1], ld.env_rgb[3 * i + 2], 0.0f);
            env_tables(ld.env_rgb, ld.env_width, ld.env_height, env_q, env_alias, env_pdf);
        } else if (ld.type != DL_LIGHT_POINT) {
            return set_error(c, DL_ERR_INVALID, "unknown light type");
        }
    }
    c->cam = sd->camera;
    c->film = sd->film;
    for (int k = 0; k < 3; k++) c->bg[k] = sd->bg_color[k];

    DevScene& ds = c->ds;
    DL_TRY(upload(c, &ds.pos, pos.data(), pos.size()));
    DL_TRY(upload(c, &ds.nrm, nrm.data(), nrm.size()));
    DL_TRY(upload(c, &ds.tan, tan.data(), tan.size()));
    DL_TRY(upload(c, &ds.idx, idx.data(), idx.size()));
    DL_TRY(upload(c, &ds.prim_mesh, prim_mesh.data(), prim_mesh.size()));
    // per primitive: its material's type and whether its mesh is a light, so
    // the path kernels classify a hit with one load
    std::vector<uint8_t> prim_info(nf);
    for (const MeshRec& r : c->mesh_recs) {
        const uint8_t info = (uint8_t)((uint32_t)c->mats[r.material].type | (r.light >= 0 ? kPrimLight : 0u));
        std::fill(prim_info.begin() + r.prim_offset, prim_info.begin() + r.prim_offset + r.num_prims, info);
    }
    DL_TRY(upload(c, &ds.prim_info, prim_info.data(), prim_info.size()));
    MeshRec* dm = nullptr;
    DL_TRY(upload(c, &dm, c->mesh_recs.data(), c->mesh_recs.size()));
    ds.meshes = dm;
    dl_material_desc* dmat = nullptr;
    DL_TRY(upload(c, &dmat, c->mats.data(), c->mats.size()));
    ds.materials = dmat;
    LightRec* dl_ = nullptr;
    DL_TRY(upload(c, &dl_, c->light_recs.data(), c->light_recs.size()));
    ds.lights = dl_;
    DL_TRY(upload(c, &ds.light_cdf, cdf_all.data(), cdf_all.size()));
    DL_TRY(upload(c, &ds.light_guide, guide_all.data(), guide_all.size()));
    DL_TRY(upload(c, &ds.env_rgb, env_rgb.data(), env_rgb.size()));
    DL_TRY(upload(c, &ds.env_q, env_q.data(), env_q.size()));
    DL_TRY(upload(c, &ds.env_alias, env_alias.data(), env_alias.size()));
    DL_TRY(upload(c, &ds.env_pdf, env_pdf.data(), env_pdf.size()));
    DL_TRY(upload(c, &ds.uv, uvs.data(), uvs.size()));
    {
        std::vector<float4> tx((size_t)texels);
        std::vector<int4> info(sd->num_textures);
        size_t o = 0;
        for (uint32_t t = 0; t < sd->num_textures; t++) {
            const dl_texture_desc& td = sd->textures[t];
            const size_t n = (size_t)td.width * (size_t)td.height;
            for (size_t i = 0; i < n; i++) tx[o + i] = make_float4(td.rgb[3 * i], td.rgb[3 * i + 1], td.rgb[3 * i + 2], 0.f);
            info[t] = make_int4((int)o, td.width, td.height, 0);
            o += n;
        }
        c->num_textures = sd->num_textures;
        c->num_texels = o;
        DL_TRY(upload(c, &ds.texels, tx.data(), tx.size()));
        DL_TRY(upload(c, &ds.tex_info, info.data(), info.size()));
        DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));  // tx / info leave scope
    }
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->has_scene = true;
    return DL_OK;
}

dl_status dl_hip_build_accel(void* h, int builder, dl_accel_stats* stats) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (!c->has_scene) return set_error(c, DL_ERR_STATE, "build_accel before upload_scene");
    hipSetDevice(c->device);
    free_scene_union(c);  // it may hold this scene's previous BVH
    if (c->ds.nodes) {
        dfree(c, c->ds.nodes);
        dfree(c, c->ds.tris);
        c->ds.nodes = nullptr;
        c->ds.tris = nullptr;
    }
    dl_accel_stats st{};
    if (builder != DL_BUILD_GPU && builder != DL_BUILD_GPU_LBVH && builder != DL_BUILD_HOST_SAH)
        return set_error(c, DL_ERR_INVALID, "unknown builder");
    bool host = builder == DL_BUILD_HOST_SAH;
    if (!host) {
        c->bvh_too_deep = false;
        const dl_status s = build_bvh_gpu(c, &st, builder);
        if (s != DL_OK && !c->bvh_too_deep) return s;
        if (s != DL_OK) {
            // a tree deeper than the traversal stack allows (pathological
            // geometry): rebuilt by the host builder, whose depth is bounded
            dfree(c, c->ds.nodes);
            dfree(c, c->ds.tris);
            c->ds.nodes = nullptr;
            c->ds.tris = nullptr;
            st = dl_accel_stats{};
            host = true;
            c->err.clear();
        }
    }
    if (host) {
        auto t0 = std::chrono::steady_clock::now();
        HostBvh bvh;
        build_bvh_host(c->tri_verts.data(), c->nprims, bvh);
        auto t1 = std::chrono::steady_clock::now();
        void* pn = nullptr;
        void* pt = nullptr;
        DL_HIP_CHECK(c, dmalloc(c, &pn, bvh.nodes.size() * sizeof(BvhNodeH)));
        DL_HIP_CHECK(c, dmalloc(c, &pt, std::max<size_t>(1, bvh.tris.size()) * sizeof(BvhTriH)));
        DL_HIP_CHECK(c, hipMemcpy(pn, bvh.nodes.data(), bvh.nodes.size() * sizeof(BvhNodeH), hipMemcpyHostToDevice));
        if (!bvh.tris.empty())
            DL_HIP_CHECK(c, hipMemcpy(pt, bvh.tris.data(), bvh.tris.size() * sizeof(BvhTriH), hipMemcpyHostToDevice));
        c->ds.nodes = (float4*)pn;
        c->ds.tris = (float4*)pt;
        st.num_nodes = (uint32_t)bvh.nodes.size();
        st.num_leaves = bvh.num_leaves;
        st.max_depth = bvh.max_depth;
        st.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    {
        const auto t0 = std::chrono::steady_clock::now();
        DL_TRY(collapse_bvh4(c, st.num_nodes, &st));
        st.build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    st.num_prims = c->nprims;
    st.node_bytes = (uint64_t)st.num_nodes * kNode4Bytes;
    st.tri_bytes = (uint64_t)c->nprims * kTriBytes;
    c->accel = st;
    c->has_accel = true;
    if (stats) *stats = st;
    return DL_OK;
}

dl_status dl_hip_trace_closest(void* h, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v) {
    Ctx* c = (Ctx*)h;
    if (!c || !rays || !prim || !t || !u || !v) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "trace before build_accel");
    return launch_trace_closest(c, rays, n, prim, t, u, v, nullptr, nullptr);
}

dl_status dl_hip_trace_closest_noalloc(void* h, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                                       float* v) {
    Ctx* c = (Ctx*)h;
    if (!c || !rays || !prim || !t || !u || !v) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "trace before build_accel");
    return launch_trace_closest_noalloc(c, rays, n, prim, t, u, v);
}

dl_status dl_hip_trace_closest_counted(void* h, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                                       float* v, uint32_t* nodes, uint32_t* tris) {
    Ctx* c = (Ctx*)h;
    if (!c || !rays || !prim || !t || !u || !v || !nodes || !tris) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "trace before build_accel");
    return launch_trace_closest(c, rays, n, prim, t, u, v, nodes, tris);
}

dl_status dl_hip_trace_shadow(void* h, const float* rays, uint32_t n, uint8_t* occ) {
    Ctx* c = (Ctx*)h;
    if (!c || !rays || !occ) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "trace before build_accel");
    return launch_trace_shadow(c, rays, n, occ);
}

dl_status dl_hip_trace_shadow_counted(void* h, const float* rays, uint32_t n, uint8_t* occ, uint32_t* nodes,
                                      uint32_t* tris) {
    Ctx* c = (Ctx*)h;
    if (!c || !rays || !occ || !nodes || !tris) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "trace before build_accel");
    return launch_trace_shadow(c, rays, n, occ, nodes, tris);
}

dl_status dl_hip_render(void* h, const dl_render_params* p, float* rgb_host, dl_render_stats* stats) {
    Ctx* c = (Ctx*)h;
    if (!c || !p) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "render before build_accel");
    return render(c, p, rgb_host, stats, -1, nullptr);
}

dl_status dl_hip_render_batch(void* h, const dl_render_params* shared, const dl_batch_item* items, uint32_t n,
                              dl_render_stats* stats) {
    Ctx* c = (Ctx*)h;
    if (!c || !shared || !items) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "render before build_accel");
    if (n == 0 || n > DL_BATCH_MAX) return set_error(c, DL_ERR_INVALID, "a shared pass holds 1 to DL_BATCH_MAX images");
    bool used[DL_BATCH_MAX] = {};
    for (uint32_t i = 0; i < n; i++) {
        DL_TRY(check_camera(c, items[i].camera));
        if (items[i].spp <= 0) return set_error(c, DL_ERR_INVALID, "spp must be positive");
        if (items[i].film < 0 || items[i].film >= DL_BATCH_MAX)
            return set_error(c, DL_ERR_INVALID, "film slot out of range");
        if (used[items[i].film]) return set_error(c, DL_ERR_INVALID, "two images of a pass share a film slot");
        used[items[i].film] = true;
    }
    return render_batch(c, shared, items, n, stats);
}

dl_status dl_hip_render_paths(void* h, const dl_render_params* p, uint32_t sample, float* out) {
    Ctx* c = (Ctx*)h;
    if (!c || !p || !out) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "render before build_accel");
    return render(c, p, nullptr, nullptr, (int)sample, out);
}

dl_status dl_hip_malloc(void* h, size_t bytes, void** dptr) {
    Ctx* c = (Ctx*)h;
    if (!c || !dptr) return DL_ERR_INVALID;
    hipSetDevice(c->device);
    DL_HIP_CHECK(c, dmalloc(c, dptr, bytes));
    return DL_OK;
}
dl_status dl_hip_free(void* h, void* p) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    DL_HIP_CHECK(c, dfree(c, p));
    return DL_OK;
}
dl_status dl_hip_memcpy_h2d(void* h, void* dst, const void* src, size_t bytes) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    DL_HIP_CHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return DL_OK;
}
dl_status dl_hip_memcpy_d2h(void* h, void* dst, const void* src, size_t bytes) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    DL_HIP_CHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return DL_OK;
}
dl_status dl_hip_memcpy_d2d(void* h, void* dst, const void* src, size_t bytes) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (bytes == 0) return DL_OK;
    if (!dst || !src) return set_error(c, DL_ERR_INVALID, "null pointer");
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) != 0) return set_error(c, DL_ERR_INVALID, "pointers must be 16-byte aligned");
    (void)hipSetDevice(c->device);
    const size_t n16 = bytes / 16;
    const uint32_t ntail = (uint32_t)(bytes % 16);
    const size_t want = (n16 + 255) / 256;
    const int grid = (int)std::max<size_t>(1, std::min<size_t>(want, (size_t)std::max(1, c->num_cus) * 32));
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    k_copy<<<grid, 256, 0, c->stream>>>((uint4*)dst, (const uint4*)src, n16, (uint8_t*)dst + n16 * 16,
                                        (const uint8_t*)src + n16 * 16, ntail);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    return DL_OK;
}
dl_status dl_hip_memory_usage(void* h, uint64_t* current, uint64_t* peak) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (current) *current = c->dev_bytes;
    if (peak) *peak = c->dev_peak;
    c->dev_peak = c->dev_bytes;  // the next query reports the peak since this one
    return DL_OK;
}
dl_status dl_hip_set_camera(void* h, const dl_camera_desc* cam) {
    Ctx* c = (Ctx*)h;
    if (!c || !cam) return DL_ERR_INVALID;
    if (!c->has_scene) return set_error(c, DL_ERR_STATE, "set_camera before upload_scene");
    DL_TRY(check_camera(c, *cam));
    c->cam = *cam;
    invalidate_film(c);
    return DL_OK;
}
dl_status dl_hip_set_view_options(void* h, const dl_view_options* o) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (!c->has_scene) return set_error(c, DL_ERR_STATE, "set_view_options before upload_scene");
    if (!o) {
        c->view_set = false;
        c->view = dl_view_options{};
        invalidate_film(c);
        return DL_OK;
    }
    if (!(o->epsilon >= 0.0f && o->epsilon < 1.0f))
        return set_error(c, DL_ERR_INVALID, "epsilon must be in [0, 1)");
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(o->bg_color[k])) return set_error(c, DL_ERR_INVALID, "bg_color must be finite");
    c->view_set = true;
    c->view = *o;
    invalidate_film(c);
    return DL_OK;
}
dl_status dl_hip_trim(void* h) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    (void)hipSetDevice(c->device);
    free_render_state(c);
    free_film_slots(c);
    free_guide(c);
    free_all_scenes(c);
    return DL_OK;
}
dl_status dl_hip_accel_export(void* h, void* nodes, void* tris) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "accel_export before build_accel");
    (void)hipSetDevice(c->device);
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (nodes) DL_HIP_CHECK(c, hipMemcpy(nodes, c->ds.nodes, c->accel.node_bytes, hipMemcpyDeviceToHost));
    if (tris) DL_HIP_CHECK(c, hipMemcpy(tris, c->ds.tris, c->accel.tri_bytes, hipMemcpyDeviceToHost));
    return DL_OK;
}
dl_status dl_hip_scene_keep(void* h, uint64_t key) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (key == 0) return set_error(c, DL_ERR_INVALID, "scene key 0 is reserved");
    if (!c->has_accel) return set_error(c, DL_ERR_STATE, "scene_keep before build_accel");
    (void)hipSetDevice(c->device);
    free_scene_union(c);  // members are named by key
    auto it = c->stash.find(key);
    if (it != c->stash.end()) {  // an older scene under the same key is replaced
        free_stashed(c, it->second);
        c->stash.erase(it);
    }
    c->scene_key = key;
    return DL_OK;
}
dl_status dl_hip_scene_select(void* h, uint64_t key) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    if (key == 0) return set_error(c, DL_ERR_INVALID, "scene key 0 is reserved");
    (void)hipSetDevice(c->device);
    if (c->scene_key == key && c->has_scene) return DL_OK;
    auto it = c->stash.find(key);
    if (it == c->stash.end()) return set_error(c, DL_ERR_STATE, "no resident scene with this key");
    SceneState s = std::move(it->second);
    c->stash.erase(it);
    release_current(c);
    swap_scene(c, s);
    c->scene_key = key;
    return DL_OK;
}
dl_status dl_hip_scene_drop(void* h, uint64_t key) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    (void)hipSetDevice(c->device);
    for (auto it = c->stash.begin(); it != c->stash.end();) {
        if (key == 0 || it->first == key) {
            free_stashed(c, it->second);
            it = c->stash.erase(it);
        } else {
            ++it;
        }
    }
    if (key != 0 && c->scene_key == key) free_scene(c);
    return DL_OK;
}
dl_status dl_hip_scene_bytes(void* h, uint64_t key, uint64_t* bytes) {
    Ctx* c = (Ctx*)h;
    if (!c || !bytes) return DL_ERR_INVALID;
    const bool current = c->has_scene && (key == 0 || key == c->scene_key);
    const auto it = c->stash.find(key);
    if (!current && (key == 0 || it == c->stash.end())) return set_error(c, DL_ERR_STATE, "no such scene");
    const std::vector<void*>& allocs = current ? c->scene_allocs : it->second.scene_allocs;
    const DevScene& ds = current ? c->ds : it->second.ds;
    auto size_of = [&](const void* p) -> uint64_t {
        const auto a = c->dev_allocs.find(const_cast<void*>(p));
        return a == c->dev_allocs.end() ? 0 : a->second;
    };
    uint64_t sum = size_of(ds.nodes) + size_of(ds.tris);
    for (void* p : allocs) sum += size_of(p);
    *bytes = sum;
    return DL_OK;
}
dl_status dl_hip_sync(void* h) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return DL_OK;
}
dl_status dl_hip_last_kernel_ms(void* h, double* ms) {
    Ctx* c = (Ctx*)h;
    if (!c || !ms) return DL_ERR_INVALID;
    *ms = c->last_kernel_ms;
    return DL_OK;
}

}  // extern "C"
