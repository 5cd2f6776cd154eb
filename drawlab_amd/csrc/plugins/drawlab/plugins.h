// Plugin classes of the scene description (the reference's src/tracer/*
// plugins, registered under the same XML `type` names with the same
// property names and defaults).  They hold the parameters the HIP backend
// needs and export them as dl_hip.h descriptors; radiance is computed on the
// GPU only.
#pragma once
#include <memory>
#include <utility>
#include <vector>

#include "dl_hip.h"
#include "drawlab/core.h"

namespace drawlab {

class BSDF;
class Emitter;
class Camera;
class Integrator;
class Sampler;
class ReconstructionFilter;

// ---- Texture (include/tracer/texture.h, texture/bitmap.cpp) ----------------
// A bitmap texture, loaded when constructed: rgb rows as uploaded (row 0 at
// v = 0).
class Texture : public Object {
public:
    EClassType getClassType() const override { return ETexture; }
    int width = 0, height = 0;
    std::vector<float> rgb;
};

// ---- BSDF (include/tracer/bsdf.h:46-111) -----------------------------------
class BSDF : public Object {
public:
    EClassType getClassType() const override { return EBSDF; }
    // Material descriptor of the C ABI (type + parameter block); texture
    // slots are filled in by exportScene from `textures`.
    virtual dl_material_desc descriptor() const = 0;
    // (DL_TEX_* slot, texture) of the textured inputs
    std::vector<std::pair<int, const Texture*>> textures;
    bool maps_object_space = false;
};

// ---- Mesh (include/tracer/mesh.h) ------------------------------------------
class Mesh : public Object {
public:
    EClassType getClassType() const override { return EMesh; }
    void addChild(Object* obj) override;
    void activate() override;
    std::string toString() const override;

    uint32_t getVertexCount() const { return (uint32_t)(m_V.size() / 3); }
    uint32_t getTriangleCount() const { return (uint32_t)(m_F.size() / 3); }
    bool hasVertexNormal() const { return !m_N.empty(); }
    bool hasTexCoord() const { return !m_UV.empty(); }
    bool isEmitter() const { return m_emitter != nullptr; }
    BSDF* getBSDF() const { return m_bsdf; }
    Emitter* getEmitter() const { return m_emitter; }
    const std::vector<float>& positions() const { return m_V; }
    const std::vector<uint32_t>& indices() const { return m_F; }
    const std::vector<float>& normals() const { return m_N; }
    const std::vector<float>& texcoords() const { return m_UV; }

protected:
    std::vector<float> m_V, m_N, m_UV;
    std::vector<uint32_t> m_F;
    std::string m_name;
    BSDF* m_bsdf = nullptr;
    Emitter* m_emitter = nullptr;
};

// ---- Emitter (include/tracer/emitter.h:7-118) -------------------------------
class Emitter : public Object {
public:
    EClassType getClassType() const override { return EEmitter; }
    virtual bool isEnvironmentEmitter() const { return false; }
    void setMesh(Mesh* m) { m_mesh = m; }
    Mesh* getMesh() const { return m_mesh; }
    // Light descriptor; `mesh` is the index of the emitting mesh (area lights).
    virtual dl_light_desc descriptor(int mesh) const = 0;

protected:
    Mesh* m_mesh = nullptr;
};

// ---- ReconstructionFilter (include/tracer/rfilter.h) -----------------------
class ReconstructionFilter : public Object {
public:
    EClassType getClassType() const override { return EReconstructionFilter; }
    virtual dl_film_desc descriptor() const = 0;
};

// ---- Camera (include/tracer/camera.h:19-64) ----------------------------------
class Camera : public Object {
public:
    EClassType getClassType() const override { return ECamera; }
    void addChild(Object* obj) override;
    void activate() override;
    virtual dl_camera_desc descriptor() const = 0;
    const ReconstructionFilter* getReconstructionFilter() const { return m_rfilter; }
    int width() const { return m_w; }
    int height() const { return m_h; }

protected:
    int m_w = 1280, m_h = 720;
    ReconstructionFilter* m_rfilter = nullptr;
    std::unique_ptr<ReconstructionFilter> m_owned_rfilter;
};

// ---- Integrator / Sampler ----------------------------------------------------
class Integrator : public Object {
public:
    EClassType getClassType() const override { return EIntegrator; }
    // DL_RENDER_* flags of the integrator on the HIP backend, or -1 when
    // it has no HIP counterpart.
    virtual int hipFlags() const { return -1; }
    // Path-tracing parameters: maxDepth (-1 = unlimited) and rrDepth.
    virtual bool isPathTracer() const { return false; }
    virtual int maxDepth() const { return -1; }
    virtual int rrDepth() const { return 4; }
};

class Sampler : public Object {
public:
    EClassType getClassType() const override { return ESampler; }
    virtual int getSampleCount() const = 0;
    virtual uint32_t getSeed() const { return 0; }
};

// ---- Scene (include/tracer/scene.h, src/tracer/scene/scene.cpp) -------------
class Scene : public Object {
public:
    explicit Scene(const PropertyList& props);
    ~Scene() override;
    EClassType getClassType() const override { return EScene; }
    void addChild(Object* obj) override;
    void activate() override;
    std::string toString() const override;

    const std::vector<Mesh*>& getMeshes() const { return m_meshes; }
    const std::vector<Emitter*>& getEmitters() const { return m_emitters; }
    const std::vector<BSDF*>& getBSDFs() const { return m_bsdfs; }
    const Camera* getCamera() const { return m_camera; }
    const Integrator* getIntegrator() const { return m_integrator; }
    const Sampler* getSampler() const { return m_sampler; }
    Color3f getBackgroundColor() const { return m_bg_color; }
    int getMeshBsdfIdx(int mesh) const { return m_mesh_bsdf_idx[mesh]; }
    int getMeshLightIdx(int mesh) const { return m_mesh_light_idx[mesh]; }
    int getEnvironmentEmitterIdx() const { return m_envmap_idx; }

private:
    std::vector<Mesh*> m_meshes;
    std::vector<Emitter*> m_emitters;
    std::vector<BSDF*> m_bsdfs;
    std::vector<int> m_mesh_bsdf_idx, m_mesh_light_idx;
    std::vector<Object*> m_owned;
    Camera* m_camera = nullptr;
    Integrator* m_integrator = nullptr;
    Sampler* m_sampler = nullptr;
    int m_envmap_idx = -1;
    Color3f m_bg_color;
};

// ---- scene export -----------------------------------------------------------
// A dl_scene_desc that points into the Scene's arrays (valid while the Scene
// lives), with the render parameters of its integrator and sampler.
struct ExportedScene {
    std::vector<dl_texture_desc> textures;
    std::vector<dl_mesh_desc> meshes;
    std::vector<dl_material_desc> materials;
    std::vector<dl_light_desc> lights;
    dl_scene_desc desc{};
    dl_render_params params{};
};
// Throws drawlab::Exception for what the HIP backend cannot render.
void exportScene(const Scene& scene, ExportedScene& out);

// ---- images -----------------------------------------------------------------
struct Image {
    int width = 0, height = 0;
    std::vector<float> rgb;  // row-major, row 0 = top
};
// Wavefront OBJ corners de-indexed per distinct (v, t, n), before any
// transform: positions (3 per corner), normals and texcoords (empty unless
// every corner has one), triangles (3 corner indices each).
struct ObjData {
    std::vector<float> P, N, UV;
    std::vector<uint32_t> F;
};
// Reads an OBJ file (through the binary cache in $DRAWLAB_MESH_CACHE when
// that names a directory; *from_cache tells whether the cache answered).
ObjData loadOBJ(const std::string& filename, bool* from_cache = nullptr);

// Radiance .hdr (RGBE, run-length encoded or flat), .pfm, uncompressed .exr.
Image loadImage(const std::string& filename);
// Texture rows as the reference uploads them (row 0 at v = 0, 8-bit
// samples / 255): .png / .exr / .pfm rows bottom up, .hdr top down.
Image loadTextureImage(const std::string& filename);
void savePFM(const std::string& filename, const Image& img);
// 8-bit sRGB PNG, as Bitmap::savePNG writes it.
void savePNG(const std::string& filename, const Image& img);
void saveEXR(const std::string& filename, const Image& img);

}  // namespace drawlab
