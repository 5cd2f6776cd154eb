// Several resident scenes in one device layout, for wavefront passes whose
// paths belong to images of different scenes (dl_hip_render_batch items that
// name different resident scenes).
//
// Every per-scene array is concatenated in member order and every index it
// holds is rebased on the way: vertex indices by the vertices before the
// scene, primitive -> mesh by the meshes before it, mesh records (first
// primitive, material, light, CDF and CDF-index offsets), materials (texture slots), lights
// (mesh, first environment texel) and texture records (first texel) on the
// host, and on the device the 4-wide BVH nodes (interior child refs by the
// nodes before the scene, leaf refs by the triangles before it) and the
// triangles' primitive ids.  Traversal of a ray then starts at its scene's
// root node and never leaves that scene's subtree, and a hit's primitive id
// leads, through the rebased records, to the same mesh, material, texture and
// light as in the scene's own layout -- so a path's arithmetic is unchanged
// and its image is bit-identical to a render of its scene alone.
#include <algorithm>
#include <cstring>
#include <vector>

#include "backend.h"
#include "dl_math.h"

namespace dl {

namespace {

constexpr int kUnionBlock = 256;

__global__ __launch_bounds__(kUnionBlock) void k_rebase_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                            size_t n, uint32_t add) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i] + add;
}

// 4-wide nodes: record 3 holds the four child refs, record 2's third word
// the four leaf counts (one byte each).  An interior ref (>= 0) moves by the
// nodes before the scene, a leaf ref (~first triangle, count > 0) by the
// triangles before it; an empty slot (~0, count 0) stays.
__global__ __launch_bounds__(kUnionBlock) void k_rebase_nodes(int4* __restrict__ dst, const int4* __restrict__ src,
                                                              size_t n, int node0, int tri0) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int4 r0 = src[4 * i], r1 = src[4 * i + 1], r2 = src[4 * i + 2];
        int4 r3 = src[4 * i + 3];
        const uint32_t cnt = (uint32_t)r2.z;
        int* ref = &r3.x;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t cj = (cnt >> (8 * j)) & 0xffu;
            if (ref[j] >= 0)
                ref[j] += node0;
            else if (cj > 0)
                ref[j] = ~(~ref[j] + tri0);
        }
        dst[4 * i] = r0;
        dst[4 * i + 1] = r1;
        dst[4 * i + 2] = r2;
        dst[4 * i + 3] = r3;
    }
}

// Triangles: the primitive id (bits of record 2's second word) moves by the
// primitives before the scene.
__global__ __launch_bounds__(kUnionBlock) void k_rebase_tris(int4* __restrict__ dst, const int4* __restrict__ src,
                                                             size_t n, int prim0) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int4 t2 = src[3 * i + 2];
        t2.y += prim0;
        dst[3 * i] = src[3 * i];
        dst[3 * i + 1] = src[3 * i + 1];
        dst[3 * i + 2] = t2;
    }
}

int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + kUnionBlock - 1) / kUnionBlock, 8192)); }

template <typename T>
dl_status union_alloc(Ctx* c, SceneUnion* u, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return DL_OK;
    void* q = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, &q, count * sizeof(T)));
    u->allocs.push_back(q);
    *p = (T*)q;
    return DL_OK;
}

template <typename T>
dl_status copy_at(Ctx* c, T* dst, size_t at, const T* src, size_t count) {
    if (count == 0) return DL_OK;
    DL_HIP_CHECK(c, hipMemcpyAsync(dst + at, src, count * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    return DL_OK;
}

#define DL_TRY(expr)                \
    do {                            \
        dl_status s_ = (expr);      \
        if (s_ != DL_OK) return s_; \
    } while (0)

// Entries of each per-scene array (the sizes the scene's upload gave them).
struct Extent {
    size_t verts, prims, meshes, mats, lights, cdf, guide, env, textures, texels, nodes, tris;
};

Extent extent_of(const SceneParts& s) {
    Extent e{};
    e.verts = s.nverts;
    e.prims = s.nprims;
    e.meshes = s.meshes->size();
    e.mats = s.mats->size();
    e.lights = s.lights->size();
    for (const MeshRec& m : *s.meshes)
        if (m.light >= 0) {
            e.cdf = std::max<size_t>(e.cdf, (size_t)m.cdf_offset + m.num_prims + 1);
            e.guide = std::max<size_t>(e.guide, (size_t)m.guide_offset + (1u << m.guide_bits) + 1);
        }
    for (const LightRec& l : *s.lights)
        if (l.type == DL_LIGHT_ENV) e.env = (size_t)l.env_w * (size_t)l.env_h;
    e.textures = s.num_textures;
    e.texels = s.num_texels;
    e.nodes = s.accel->node_bytes / kNode4Bytes;
    e.tris = s.accel->tri_bytes / kTriBytes;
    return e;
}

dl_status build_union(Ctx* c, SceneUnion* u, const std::vector<uint64_t>& members,
                      const std::vector<SceneParts>& parts, const std::vector<Extent>& ext, const Extent& tot);

}  // namespace

bool scene_parts(Ctx* c, uint64_t key, SceneParts& out) {
    auto fill = [&](uint32_t nprims, uint32_t nverts, const std::vector<MeshRec>& meshes,
                    const std::vector<dl_material_desc>& mats, const std::vector<LightRec>& lights, int env_light,
                    const float* bg, bool view_set, const dl_view_options& view, const dl_film_desc& film,
                    const DevScene& ds, const dl_accel_stats& accel, uint32_t ntex, uint64_t ntexels) {
        out.nprims = nprims;
        out.nverts = nverts;
        out.meshes = &meshes;
        out.mats = &mats;
        out.lights = &lights;
        out.env_light = env_light;
        out.bg = view_set ? view.bg_color : bg;
        out.hide = view_set && view.hide_lights != 0;
        out.eps = view_set && view.epsilon > 0.0f ? view.epsilon : kEpsilon;
        out.film = &film;
        out.ds = &ds;
        out.accel = &accel;
        out.num_textures = ntex;
        out.num_texels = ntexels;
    };
    if (c->has_accel && (key == 0 || key == c->scene_key)) {
        fill(c->nprims, c->nverts, c->mesh_recs, c->mats, c->light_recs, c->env_light, c->bg, c->view_set, c->view,
             c->film, c->ds, c->accel, c->num_textures, c->num_texels);
        return true;
    }
    auto it = c->stash.find(key);
    if (key == 0 || it == c->stash.end() || !it->second.has_accel) return false;
    const SceneState& s = it->second;
    fill(s.nprims, s.nverts, s.mesh_recs, s.mats, s.light_recs, s.env_light, s.bg, s.view_set, s.view, s.film, s.ds,
         s.accel, s.num_textures, s.num_texels);
    return true;
}

void free_scene_union(Ctx* c) {
    if (!c->uni) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->uni->allocs) (void)dfree(c, p);
    delete c->uni;
    c->uni = nullptr;
}

dl_status scene_union_for(Ctx* c, const std::vector<uint64_t>& keys, SceneUnion** out) {
    *out = nullptr;
    if (c->uni) {
        bool all = true;
        for (uint64_t k : keys)
            all = all && std::find(c->uni->keys.begin(), c->uni->keys.end(), k) != c->uni->keys.end();
        if (all) {
            *out = c->uni;
            return DL_OK;
        }
        free_scene_union(c);
    }
    // members: every resident scene (the current one first)
    std::vector<uint64_t> members;
    if (c->has_accel) members.push_back(c->scene_key);
    for (const auto& kv : c->stash)
        if (kv.second.has_accel && kv.first != c->scene_key) members.push_back(kv.first);
    for (uint64_t k : keys)
        if (std::find(members.begin(), members.end(), k) == members.end())
            return set_error(c, DL_ERR_STATE, "no resident scene with this key");
    if (members.size() > 255) return set_error(c, DL_ERR_INVALID, "more than 255 resident scenes");
    std::vector<SceneParts> parts(members.size());
    std::vector<Extent> ext(members.size());
    Extent tot{};
    for (size_t m = 0; m < members.size(); m++) {
        if (!scene_parts(c, members[m], parts[m])) return set_error(c, DL_ERR_STATE, "resident scene without a BVH");
        ext[m] = extent_of(parts[m]);
        const Extent& e = ext[m];
        tot.verts += e.verts;
        tot.prims += e.prims;
        tot.meshes += e.meshes;
        tot.mats += e.mats;
        tot.lights += e.lights;
        tot.cdf += e.cdf;
        tot.guide += e.guide;
        tot.env += e.env;
        tot.textures += e.textures;
        tot.texels += e.texels;
        tot.nodes += e.nodes;
        tot.tris += e.tris;
    }
    if (tot.prims >= 0x7fffffffull || tot.verts >= 0x7fffffffull || tot.nodes >= 0x7fffffffull ||
        tot.texels >= 0x7fffffffull || tot.env >= 0xffffffffull)
        return set_error(c, DL_ERR_INVALID, "resident scenes too large for one layout");

    c->uni = new SceneUnion();
    const dl_status st = build_union(c, c->uni, members, parts, ext, tot);
    if (st != DL_OK) {
        free_scene_union(c);
        return st;
    }
    *out = c->uni;
    return DL_OK;
}

namespace {

dl_status build_union(Ctx* c, SceneUnion* u, const std::vector<uint64_t>& members,
                      const std::vector<SceneParts>& parts, const std::vector<Extent>& ext, const Extent& tot) {
    u->keys = members;
    DevScene& d = u->ds;
    MeshRec* d_meshes = nullptr;
    dl_material_desc* d_mats = nullptr;
    LightRec* d_lights = nullptr;
    DL_TRY(union_alloc(c, u, &d.pos, 3 * tot.verts));
    DL_TRY(union_alloc(c, u, &d.nrm, 3 * tot.verts));
    DL_TRY(union_alloc(c, u, &d.tan, 3 * tot.verts));
    DL_TRY(union_alloc(c, u, &d.uv, 2 * tot.verts));
    DL_TRY(union_alloc(c, u, &d.idx, 3 * tot.prims));
    DL_TRY(union_alloc(c, u, &d.prim_mesh, tot.prims));
    DL_TRY(union_alloc(c, u, &d.prim_info, tot.prims));
    DL_TRY(union_alloc(c, u, &d_meshes, tot.meshes));
    DL_TRY(union_alloc(c, u, &d_mats, tot.mats));
    DL_TRY(union_alloc(c, u, &d_lights, tot.lights));
    DL_TRY(union_alloc(c, u, &d.light_cdf, tot.cdf));
    DL_TRY(union_alloc(c, u, &d.light_guide, tot.guide));
    DL_TRY(union_alloc(c, u, &d.env_rgb, tot.env));
    DL_TRY(union_alloc(c, u, &d.env_q, tot.env));
    DL_TRY(union_alloc(c, u, &d.env_alias, tot.env));
    DL_TRY(union_alloc(c, u, &d.env_pdf, tot.env));
    DL_TRY(union_alloc(c, u, &d.texels, tot.texels));
    DL_TRY(union_alloc(c, u, &d.tex_info, tot.textures));
    DL_TRY(union_alloc(c, u, &d.nodes, 4 * tot.nodes));
    DL_TRY(union_alloc(c, u, &d.tris, 3 * tot.tris));
    DL_TRY(union_alloc(c, u, &u->recs, members.size()));
    DL_TRY(union_alloc(c, u, &u->roots, members.size()));
    d.meshes = d_meshes;
    d.materials = d_mats;
    d.lights = d_lights;

    // host-side records, rebased
    std::vector<MeshRec> meshes;
    std::vector<dl_material_desc> mats;
    std::vector<LightRec> lights;
    std::vector<int4> tex_info(tot.textures);
    Extent o{};
    for (size_t m = 0; m < members.size(); m++) {
        const SceneParts& s = parts[m];
        const Extent& e = ext[m];
        for (MeshRec r : *s.meshes) {
            r.prim_offset += (uint32_t)o.prims;
            r.material += (int32_t)o.mats;
            if (r.light >= 0) r.light += (int32_t)o.lights;
            r.cdf_offset += (uint32_t)o.cdf;
            r.guide_offset += (uint32_t)o.guide;  // its entries index the mesh's own CDF
            meshes.push_back(r);
        }
        for (dl_material_desc md : *s.mats) {
            for (int k = 0; k < DL_TEX_SLOTS; k++)
                if (md.tex[k] > 0) md.tex[k] += (int32_t)o.textures;
            mats.push_back(md);
        }
        for (LightRec l : *s.lights) {
            if (l.mesh >= 0) l.mesh += (int32_t)o.meshes;
            l.env_base += (uint32_t)o.env;
            lights.push_back(l);
        }
        if (e.textures > 0) {
            DL_HIP_CHECK(c, hipMemcpyAsync(&tex_info[o.textures], s.ds->tex_info, e.textures * sizeof(int4),
                                           hipMemcpyDeviceToHost, c->stream));
        }
        u->light0.push_back((int32_t)o.lights);
        u->root.push_back((int32_t)o.nodes);
        // device arrays: plain copies and rebased copies
        const DevScene& sd = *s.ds;
        DL_TRY(copy_at(c, d.pos, 3 * o.verts, sd.pos, 3 * e.verts));
        DL_TRY(copy_at(c, d.nrm, 3 * o.verts, sd.nrm, 3 * e.verts));
        DL_TRY(copy_at(c, d.tan, 3 * o.verts, sd.tan, 3 * e.verts));
        DL_TRY(copy_at(c, d.uv, 2 * o.verts, sd.uv, 2 * e.verts));
        DL_TRY(copy_at(c, d.prim_info, o.prims, sd.prim_info, e.prims));
        DL_TRY(copy_at(c, d.light_cdf, o.cdf, sd.light_cdf, e.cdf));
        DL_TRY(copy_at(c, d.light_guide, o.guide, sd.light_guide, e.guide));
        DL_TRY(copy_at(c, d.env_rgb, o.env, sd.env_rgb, e.env));
        DL_TRY(copy_at(c, d.env_q, o.env, sd.env_q, e.env));
        DL_TRY(copy_at(c, d.env_alias, o.env, sd.env_alias, e.env));  // alias entries are within the environment
        DL_TRY(copy_at(c, d.env_pdf, o.env, sd.env_pdf, e.env));
        DL_TRY(copy_at(c, d.texels, o.texels, sd.texels, e.texels));
        if (e.prims > 0) {
            k_rebase_u32<<<grid_for(3 * e.prims), kUnionBlock, 0, c->stream>>>(d.idx + 3 * o.prims, sd.idx,
                                                                                3 * e.prims, (uint32_t)o.verts);
            k_rebase_u32<<<grid_for(e.prims), kUnionBlock, 0, c->stream>>>(d.prim_mesh + o.prims, sd.prim_mesh,
                                                                            e.prims, (uint32_t)o.meshes);
        }
        if (e.nodes > 0)
            k_rebase_nodes<<<grid_for(e.nodes), kUnionBlock, 0, c->stream>>>(
                reinterpret_cast<int4*>(d.nodes) + 4 * o.nodes, reinterpret_cast<const int4*>(sd.nodes), e.nodes,
                (int)o.nodes, (int)o.tris);
        if (e.tris > 0)
            k_rebase_tris<<<grid_for(e.tris), kUnionBlock, 0, c->stream>>>(
                reinterpret_cast<int4*>(d.tris) + 3 * o.tris, reinterpret_cast<const int4*>(sd.tris), e.tris,
                (int)o.prims);
        DL_HIP_CHECK(c, hipGetLastError());
        o.verts += e.verts;
        o.prims += e.prims;
        o.meshes += e.meshes;
        o.mats += e.mats;
        o.lights += e.lights;
        o.cdf += e.cdf;
        o.guide += e.guide;
        o.env += e.env;
        o.textures += e.textures;
        o.texels += e.texels;
        o.nodes += e.nodes;
        o.tris += e.tris;
    }
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));  // tex_info downloads
    o = Extent{};
    for (size_t m = 0; m < members.size(); m++) {
        for (size_t t = 0; t < ext[m].textures; t++) tex_info[o.textures + t].x += (int)o.texels;
        o.textures += ext[m].textures;
        o.texels += ext[m].texels;
    }
    auto up = [&](auto* dst, const auto& v) -> dl_status {
        if (v.empty()) return DL_OK;
        DL_HIP_CHECK(c, hipMemcpyAsync(dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, c->stream));
        return DL_OK;
    };
    DL_TRY(up(d_meshes, meshes));
    DL_TRY(up(d_mats, mats));
    DL_TRY(up(d_lights, lights));
    DL_TRY(up(d.tex_info, tex_info));
    DL_TRY(up(u->roots, u->root));
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));  // the host vectors leave scope
    return DL_OK;
}

}  // namespace

}  // namespace dl
