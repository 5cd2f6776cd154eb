// Plugins of the scene description, registered under the reference's XML
// type names with its property names and defaults (each cites the reference
// constructor it mirrors).  They carry parameters only; the HIP backend
// evaluates them.
#include <cmath>
#include <fstream>
#include <sstream>
#include <set>
#include <unordered_map>

#include "drawlab/plugins.h"

namespace drawlab {

// ============================================================================
// Mesh  (src/tracer/mesh/mesh.cpp:15-88)
// ============================================================================
void Mesh::addChild(Object* obj) {
    switch (obj->getClassType()) {
        case EBSDF:
            if (m_bsdf) throw Exception("Mesh: tried to register multiple BSDF instances!");
            m_bsdf = static_cast<BSDF*>(obj);
            break;
        case EEmitter: {
            Emitter* e = static_cast<Emitter*>(obj);
            if (m_emitter) throw Exception("Mesh: tried to register multiple Emitter instances!");
            e->setMesh(this);
            m_emitter = e;
        } break;
        default: throw Exception("Mesh::addChild(<%s>) is not supported!", classTypeName(obj->getClassType()));
    }
}

void Mesh::activate() {
    if (!m_bsdf) {
        // no material assigned: a diffuse BRDF with the default albedo
        m_bsdf = static_cast<BSDF*>(ObjectFactory::createInstance("diffuse", PropertyList()));
    }
    if (m_F.empty()) throw Exception("Mesh \"%s\" has no triangles", m_name);
    const uint32_t nv = getVertexCount();
    for (uint32_t i : m_F)
        if (i >= nv) throw Exception("Mesh \"%s\": vertex index %i out of range", m_name, (int)i);
}

std::string Mesh::toString() const {
    return format("Mesh[name = \"%s\", vertexCount = %i, triangleCount = %i]", m_name, (int)getVertexCount(),
                  (int)getTriangleCount());
}

// Wavefront OBJ (src/tracer/mesh/obj.cpp:17-87): positions, texcoords and
// normals de-indexed per distinct (v, vt, vn) corner in order of first use,
// polygons fanned into triangles (obj_io.cpp), positions mapped by toWorld
// as points and normals by its inverse transpose.
class WavefrontOBJ : public Mesh {
public:
    explicit WavefrontOBJ(const PropertyList& props) {
        const std::string filename = resolvePath(props.getString("filename"));
        const Transform trafo = props.getTransform("toWorld", Transform());
        ObjData d = loadOBJ(filename);
        const size_t nc = d.P.size() / 3;
        m_V.resize(3 * nc);
        for (size_t i = 0; i < nc; i++) {
            const Point3f p = trafo.point(Point3f(d.P[3 * i], d.P[3 * i + 1], d.P[3 * i + 2]));
            m_V[3 * i] = p[0];
            m_V[3 * i + 1] = p[1];
            m_V[3 * i + 2] = p[2];
        }
        if (!d.N.empty()) {
            m_N.resize(3 * nc);
            for (size_t i = 0; i < nc; i++) {
                const Normal3f n = trafo.normal(Normal3f(d.N[3 * i], d.N[3 * i + 1], d.N[3 * i + 2]));
                m_N[3 * i] = n[0];
                m_N[3 * i + 1] = n[1];
                m_N[3 * i + 2] = n[2];
            }
        }
        m_UV = std::move(d.UV);
        m_F = std::move(d.F);
        m_name = filename;
    }
};
DRAWLAB_REGISTER_CLASS(WavefrontOBJ, "obj");

// Rectangle (src/tracer/mesh/rectangle.cpp:9-41): corners pos, pos+v0,
// pos+v1, pos+v0+v1; triangles (0,1,2), (3,2,1); normal normalize(v0 x v1).
class Rectangle : public Mesh {
public:
    explicit Rectangle(const PropertyList& props) {
        const Vector3f pos = props.getVector("pos");
        const Vector3f v0 = props.getVector("v0");
        const Vector3f v1 = props.getVector("v1");
        const Normal3f n = normalized(cross(v0, v1));
        const Vector3f p[4] = {pos, pos + v0, pos + v1, (pos + v0) + v1};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 3; j++) {
                m_V.push_back(p[i][j]);
                m_N.push_back(n[j]);
            }
        m_UV = {1, 0, 1, 1, 0, 0, 0, 1};
        m_F = {0, 1, 2, 3, 2, 1};
        m_name = "rectangle";
    }
};
DRAWLAB_REGISTER_CLASS(Rectangle, "rectangle");

// ============================================================================
// Textures (src/tracer/texture/bitmap.cpp, texture.cpp): "filename" loaded
// when constructed.  The reference's GPU path samples the bitmap at the hit's
// texture coordinates and does not apply the uoffset / voffset / uscale /
// vscale properties of Texture2D (texture.cpp:15-28, used by the CPU eval
// only, which throws); they are parsed and likewise unused.
// ============================================================================
class BitmapTexture : public Texture {
public:
    explicit BitmapTexture(const PropertyList& props) : m_filename(props.getString("filename")) {
        if (props.getString("coordinates", "uv") != "uv") throw Exception("Only UV coordinates are supported at the moment!");
        (void)props.getFloat("uoffset", 0.0f);
        (void)props.getFloat("voffset", 0.0f);
        const float s = props.getFloat("uvscale", 1.0f);
        (void)props.getFloat("uscale", s);
        (void)props.getFloat("vscale", s);
        Image img = loadTextureImage(resolvePath(m_filename));
        width = img.width;
        height = img.height;
        rgb = std::move(img.rgb);
    }
    std::string toString() const override { return format("BitmapTexture[%s, %ix%i]", m_filename, width, height); }

private:
    std::string m_filename;
};
DRAWLAB_REGISTER_CLASS(BitmapTexture, "bitmap");

// ============================================================================
// BSDFs
// ============================================================================
static void put3(float* p, const Color3f& c) {
    p[0] = c[0];
    p[1] = c[1];
    p[2] = c[2];
}

// Texture children by name -> DL_TEX_* slot (a later child of the same
// name replaces an earlier one, as the reference's member assignment does).
class TexturableBSDF : public BSDF {
protected:
    void bind(int slot, const Object* obj) {
        const Texture* t = static_cast<const Texture*>(obj);
        for (auto& s : textures)
            if (s.first == slot) {
                s.second = t;
                return;
            }
        textures.emplace_back(slot, t);
    }
};

// Diffuse (src/tracer/bsdf/diffuse.cpp:14-24, 69-76): albedo, default 0.5;
// a texture child named "albedo" replaces it.
class Diffuse : public TexturableBSDF {
public:
    explicit Diffuse(const PropertyList& p) : m_albedo(p.getColor("albedo", Color3f(0.5f))) {}
    void addChild(Object* obj) override {
        if (obj->getClassType() == ETexture && obj->getObjectName() == "albedo")
            bind(DL_TEX_ALBEDO, obj);
        else
            warn("The child(%s) of diffuse is not used!", obj->getObjectName());
    }
    dl_material_desc descriptor() const override {
        dl_material_desc d{};
        d.type = DL_BSDF_DIFFUSE;
        put3(d.params, m_albedo);
        return d;
    }
    std::string toString() const override { return format("Diffuse[albedo = %s]", m_albedo); }

private:
    Color3f m_albedo;
};
DRAWLAB_REGISTER_CLASS(Diffuse, "diffuse");

// Mirror (src/tracer/bsdf/mirror.cpp).
class Mirror : public BSDF {
public:
    explicit Mirror(const PropertyList&) {}
    dl_material_desc descriptor() const override {
        dl_material_desc d{};
        d.type = DL_BSDF_MIRROR;
        return d;
    }
    std::string toString() const override { return "Mirror[]"; }
};
DRAWLAB_REGISTER_CLASS(Mirror, "mirror");

// Dielectric (src/tracer/bsdf/dielectric.cpp:21-29): intIOR 1.5046, extIOR 1.000277.
class Dielectric : public BSDF {
public:
    explicit Dielectric(const PropertyList& p)
        : m_int(p.getFloat("intIOR", 1.5046f)), m_ext(p.getFloat("extIOR", 1.000277f)) {}
    dl_material_desc descriptor() const override {
        dl_material_desc d{};
        d.type = DL_BSDF_DIELECTRIC;
        d.params[0] = m_int;
        d.params[1] = m_ext;
        return d;
    }
    std::string toString() const override { return format("Dielectric[intIOR = %f, extIOR = %f]", m_int, m_ext); }

private:
    float m_int, m_ext;
};
DRAWLAB_REGISTER_CLASS(Dielectric, "dielectric");

// Microfacet (src/tracer/bsdf/microfacet.cpp:11-33): alpha 0.1, intIOR,
// extIOR, kd 0.5 (Beckmann + dielectric Fresnel + diffuse base).
class Microfacet : public BSDF {
public:
    explicit Microfacet(const PropertyList& p)
        : m_alpha(p.getFloat("alpha", 0.1f)),
          m_int(p.getFloat("intIOR", 1.5046f)),
          m_ext(p.getFloat("extIOR", 1.000277f)),
          m_kd(p.getColor("kd", Color3f(0.5f))) {}
    dl_material_desc descriptor() const override {
        dl_material_desc d{};
        d.type = DL_BSDF_MICROFACET;
        d.params[0] = m_alpha;
        d.params[1] = m_int;
        d.params[2] = m_ext;
        put3(d.params + 3, m_kd);
        return d;
    }
    std::string toString() const override { return format("Microfacet[alpha = %f, kd = %s]", m_alpha, m_kd); }

private:
    float m_alpha, m_int, m_ext;
    Color3f m_kd;
};
DRAWLAB_REGISTER_CLASS(Microfacet, "microfacet");

// AnisoGGX (src/tracer/bsdf/aniso_ggx.cpp:10-58): pd, ps, axay (x, y used),
// default 0.5 each; texture children "pd", "ps", "axay", "normal",
// "tangent"; is_tangent_space (default true) for the normal / tangent maps.
class AnisoGGX : public TexturableBSDF {
public:
    explicit AnisoGGX(const PropertyList& p)
        : m_pd(p.getColor("pd", Color3f(0.5f))),
          m_ps(p.getColor("ps", Color3f(0.5f))),
          m_axay(p.getColor("axay", Color3f(0.5f))) {
        maps_object_space = !p.getBoolean("is_tangent_space", true);
    }
    void addChild(Object* obj) override {
        if (obj->getClassType() != ETexture) {
            warn("The child(%s) of aniso_ggx is not used!", obj->getObjectName());
            return;
        }
        const std::string& n = obj->getObjectName();
        if (n == "pd") bind(DL_TEX_PD, obj);
        else if (n == "ps") bind(DL_TEX_PS, obj);
        else if (n == "axay") bind(DL_TEX_AXAY, obj);
        else if (n == "normal") bind(DL_TEX_NORMAL, obj);
        else if (n == "tangent") bind(DL_TEX_TANGENT, obj);
    }
    dl_material_desc descriptor() const override {
        dl_material_desc d{};
        d.type = DL_BSDF_ANISO_GGX;
        put3(d.params, m_pd);
        put3(d.params + 3, m_ps);
        d.params[6] = m_axay[0];
        d.params[7] = m_axay[1];
        return d;
    }
    std::string toString() const override { return format("AnisoGGX[pd = %s, ps = %s, axay = %s]", m_pd, m_ps, m_axay); }

private:
    Color3f m_pd, m_ps, m_axay;
};
DRAWLAB_REGISTER_CLASS(AnisoGGX, "aniso_ggx");

// Conductor: isotropic GGX with the complex-IOR Fresnel term (no reference
// counterpart; an extension with its own type name).  alpha 0.1, eta, k
// default to copper.
class Conductor : public BSDF {
public:
    explicit Conductor(const PropertyList& p)
        : m_alpha(p.getFloat("alpha", 0.1f)),
          m_eta(p.getColor("eta", Color3f(0.2f, 0.92f, 1.1f))),
          m_k(p.getColor("k", Color3f(3.912f, 2.452f, 2.142f))) {}
    dl_material_desc descriptor() const override {
        dl_material_desc d{};
        d.type = DL_BSDF_CONDUCTOR;
        d.params[0] = m_alpha;
        put3(d.params + 1, m_eta);
        put3(d.params + 4, m_k);
        return d;
    }
    std::string toString() const override { return format("Conductor[alpha = %f, eta = %s, k = %s]", m_alpha, m_eta, m_k); }

private:
    float m_alpha;
    Color3f m_eta, m_k;
};
DRAWLAB_REGISTER_CLASS(Conductor, "conductor");

// ============================================================================
// Emitters
// ============================================================================
// AreaLight (src/tracer/emitter/area.cpp:6-10): radiance (required).
class AreaLight : public Emitter {
public:
    explicit AreaLight(const PropertyList& p) : m_radiance(p.getColor("radiance")) {}
    dl_light_desc descriptor(int mesh) const override {
        dl_light_desc d{};
        d.type = DL_LIGHT_AREA;
        d.mesh = mesh;
        put3(d.radiance, m_radiance);
        return d;
    }
    std::string toString() const override { return format("AreaLight[radiance = %s]", m_radiance); }

private:
    Color3f m_radiance;
};
DRAWLAB_REGISTER_CLASS(AreaLight, "area");

// PointLight (src/tracer/emitter/point.cpp:11-15): intensity, position (required).
class PointLight : public Emitter {
public:
    explicit PointLight(const PropertyList& p) : m_intensity(p.getColor("intensity")), m_position(p.getPoint("position")) {}
    dl_light_desc descriptor(int) const override {
        dl_light_desc d{};
        d.type = DL_LIGHT_POINT;
        d.mesh = -1;
        put3(d.radiance, m_intensity);
        put3(d.position, m_position);
        return d;
    }
    std::string toString() const override { return format("PointLight[intensity = %s, position = %s]", m_intensity, m_position); }

private:
    Color3f m_intensity;
    Point3f m_position;
};
DRAWLAB_REGISTER_CLASS(PointLight, "point");

// EnvironmentMap (src/tracer/emitter/envmap.cpp:13-19): filename, visual
// (default true), scale (default 1).  Row 0 of the lat-long table is the
// direction +y, in the orientation the reference's texture upload produces:
// the top row of a Radiance .hdr file, the bottom row of an OpenEXR file
// (texture/bitmap.cpp:86 flips every float bitmap; only .hdr was already
// flipped on load, bitmap.cpp:109).
class EnvironmentMap : public Emitter {
public:
    explicit EnvironmentMap(const PropertyList& p)
        : m_filename(p.getString("filename")), m_visual(p.getBoolean("visual", true)), m_scale(p.getFloat("scale", 1.f)) {
        Image img = loadImage(resolvePath(m_filename));
        m_w = img.width;
        m_h = img.height;
        m_rgb = std::move(img.rgb);
        if (toLower(m_filename).size() >= 4 && endsWith(toLower(m_filename), ".exr")) {
            for (int y = 0; y < m_h / 2; y++)
                for (int i = 0; i < 3 * m_w; i++)
                    std::swap(m_rgb[(size_t)y * 3 * m_w + i], m_rgb[(size_t)(m_h - 1 - y) * 3 * m_w + i]);
        }
    }
    bool isEnvironmentEmitter() const override { return true; }
    dl_light_desc descriptor(int) const override {
        dl_light_desc d{};
        d.type = DL_LIGHT_ENV;
        d.mesh = -1;
        d.env_rgb = m_rgb.data();
        d.env_width = m_w;
        d.env_height = m_h;
        d.env_scale = m_scale;
        d.env_visual = m_visual ? 1 : 0;
        return d;
    }
    std::string toString() const override { return format("EnvironmentMap[%s, %ix%i]", m_filename, m_w, m_h); }

private:
    std::string m_filename;
    bool m_visual;
    float m_scale;
    int m_w = 0, m_h = 0;
    std::vector<float> m_rgb;
};
DRAWLAB_REGISTER_CLASS(EnvironmentMap, "envmap");

// ============================================================================
// Reconstruction filters (src/tracer/rfilter/rfilter.cpp:13-98)
// ============================================================================
class FilterPlugin : public ReconstructionFilter {
public:
    FilterPlugin(int type, float radius, float p0, float p1, const char* name)
        : m_type(type), m_radius(radius), m_p0(p0), m_p1(p1), m_name(name) {}
    dl_film_desc descriptor() const override {
        dl_film_desc d;
        dl_film_filter(m_type, m_radius, m_p0, m_p1, &d);
        return d;
    }
    std::string toString() const override { return format("%s[radius = %f]", m_name, m_radius); }

private:
    int m_type;
    float m_radius, m_p0, m_p1;
    const char* m_name;
};
struct GaussianFilter : FilterPlugin {
    explicit GaussianFilter(const PropertyList& p)
        : FilterPlugin(0, p.getFloat("radius", 2.0f), p.getFloat("stddev", 0.5f), 0.f, "GaussianFilter") {}
};
struct BoxFilter : FilterPlugin {
    explicit BoxFilter(const PropertyList&) : FilterPlugin(1, 0.5f, 0.f, 0.f, "BoxFilter") {}
};
struct TentFilter : FilterPlugin {
    explicit TentFilter(const PropertyList&) : FilterPlugin(2, 1.0f, 0.f, 0.f, "TentFilter") {}
};
struct MitchellNetravaliFilter : FilterPlugin {
    explicit MitchellNetravaliFilter(const PropertyList& p)
        : FilterPlugin(3, p.getFloat("radius", 2.0f), p.getFloat("B", 1.0f / 3.0f), p.getFloat("C", 1.0f / 3.0f),
                       "MitchellNetravaliFilter") {}
};
DRAWLAB_REGISTER_CLASS(GaussianFilter, "gaussian");
DRAWLAB_REGISTER_CLASS(MitchellNetravaliFilter, "mitchell");
DRAWLAB_REGISTER_CLASS(TentFilter, "tent");
DRAWLAB_REGISTER_CLASS(BoxFilter, "box");

// ============================================================================
// Cameras
// ============================================================================
void Camera::addChild(Object* obj) {
    if (obj->getClassType() != EReconstructionFilter)
        throw Exception("Camera::addChild(<%s>) is not supported!", classTypeName(obj->getClassType()));
    if (m_rfilter) throw Exception("Camera: tried to register multiple reconstruction filters!");
    m_rfilter = static_cast<ReconstructionFilter*>(obj);
    m_owned_rfilter.reset(m_rfilter);
}
void Camera::activate() {
    // no filter given: the default Gaussian (virtual.cpp:44-50)
    if (!m_rfilter) {
        m_owned_rfilter.reset(
            static_cast<ReconstructionFilter*>(ObjectFactory::createInstance("gaussian", PropertyList())));
        m_rfilter = m_owned_rfilter.get();
    }
}

// VirtualCamera (src/tracer/camera/virtual.cpp:14-28): eye, lookat, up
// (required), fov 30 (vertical, degrees), width 1280, height 720.
class VirtualCamera : public Camera {
public:
    explicit VirtualCamera(const PropertyList& p) {
        m_w = p.getInteger("width", 1280);
        m_h = p.getInteger("height", 720);
        m_fov = p.getFloat("fov", 30.0f);
        m_eye = p.getVector("eye");
        m_lookat = p.getVector("lookat");
        m_up = p.getVector("up");
    }
    dl_camera_desc descriptor() const override {
        dl_camera_desc d;
        dl_camera_virtual(m_eye.v, m_lookat.v, m_up.v, m_fov, m_w, m_h, &d);
        return d;
    }
    std::string toString() const override {
        return format("VirtualCamera[eye = %s, lookat = %s, fov = %f, size = %ix%i]", m_eye, m_lookat, m_fov, m_w, m_h);
    }

private:
    float m_fov;
    Vector3f m_eye, m_lookat, m_up;
};
DRAWLAB_REGISTER_CLASS(VirtualCamera, "virtual");

// PerspectiveCamera (src/tracer/camera/perspective.cpp:21-38): width 1280,
// height 720, toWorld, fov 30 (horizontal), nearClip 1e-4, farClip 1e4.
class PerspectiveCamera : public Camera {
public:
    explicit PerspectiveCamera(const PropertyList& p) {
        m_w = p.getInteger("width", 1280);
        m_h = p.getInteger("height", 720);
        m_to_world = p.getTransform("toWorld", Transform());
        m_fov = p.getFloat("fov", 30.0f);
        m_near = p.getFloat("nearClip", 1e-4f);
        m_far = p.getFloat("farClip", 1e4f);
    }
    dl_camera_desc descriptor() const override {
        dl_camera_desc d;
        float tw[16];
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) tw[4 * j + i] = m_to_world.matrix().m[j][i];
        if (dl_camera_perspective(tw, m_fov, m_near, m_far, m_w, m_h, &d) != DL_OK)
            throw Exception("PerspectiveCamera: the projection is not invertible");
        return d;
    }
    std::string toString() const override { return format("PerspectiveCamera[fov = %f, size = %ix%i]", m_fov, m_w, m_h); }

private:
    Transform m_to_world;
    float m_fov, m_near, m_far;
};
DRAWLAB_REGISTER_CLASS(PerspectiveCamera, "perspective");

// OpencvCamera (src/tracer/camera/opencv.cpp:9-24): cx, cy, fx, fy
// (pixels) and the extrinsic (world-to-camera) transform, all required;
// width 1280, height 720.  Rays as the reference's GPU path casts them
// (include/optix/camera/opencv.h:24-47; its CPU sampleRay throws).
class OpencvCamera : public Camera {
public:
    explicit OpencvCamera(const PropertyList& p) {
        m_cx = p.getFloat("cx");
        m_cy = p.getFloat("cy");
        m_fx = p.getFloat("fx");
        m_fy = p.getFloat("fy");
        m_ext = p.getTransform("extrinsic");
        m_w = p.getInteger("width", 1280);
        m_h = p.getInteger("height", 720);
    }
    dl_camera_desc descriptor() const override {
        dl_camera_desc d;
        float ext[16];
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) ext[4 * j + i] = m_ext.matrix().m[j][i];
        if (dl_camera_opencv(ext, m_cx, m_cy, m_fx, m_fy, m_w, m_h, &d) != DL_OK)
            throw Exception("OpencvCamera: the extrinsic matrix is not invertible or a focal length is zero");
        return d;
    }
    std::string toString() const override {
        return format("OpencvCamera[cxcy = %f %f, fxfy = %f %f, size = %ix%i]", m_cx, m_cy, m_fx, m_fy, m_w, m_h);
    }

private:
    float m_cx, m_cy, m_fx, m_fy;
    Transform m_ext;
};
DRAWLAB_REGISTER_CLASS(OpencvCamera, "opencv");

// ============================================================================
// Integrators and sampler
// ============================================================================
// PathIntegrator (src/tracer/integrator/path.cpp): unidirectional path
// tracing with MIS.  maxDepth (-1 = unlimited, the reference) and rrDepth
// (bounce index where Russian roulette starts; 4 = the reference's
// `depth > 3`) are extensions with reference-preserving defaults.
class PathIntegrator : public Integrator {
public:
    explicit PathIntegrator(const PropertyList& p)
        : m_max_depth(p.getInteger("maxDepth", -1)), m_rr_depth(p.getInteger("rrDepth", 4)) {}
    int hipFlags() const override { return 0; }
    bool isPathTracer() const override { return true; }
    int maxDepth() const override { return m_max_depth; }
    int rrDepth() const override { return m_rr_depth; }
    std::string toString() const override { return format("PathIntegrator[maxDepth = %i, rrDepth = %i]", m_max_depth, m_rr_depth); }

private:
    int m_max_depth, m_rr_depth;
};
DRAWLAB_REGISTER_CLASS(PathIntegrator, "path");

// Integrators of the reference that the hip backend does not render.
class OtherIntegrator : public Integrator {
public:
    OtherIntegrator(const PropertyList&, const char* name) : m_name(name) {}
    std::string toString() const override { return m_name; }

private:
    std::string m_name;
};
// NormalIntegrator (src/tracer/integrator/normal.cpp:9-19): |shading
// normal| of the first hit, rendered by DL_RENDER_NORMALS.
struct NormalIntegrator : OtherIntegrator {
    explicit NormalIntegrator(const PropertyList& p) : OtherIntegrator(p, "NormalIntegrator[]") {}
    int hipFlags() const override { return DL_RENDER_NORMALS; }
};
struct WhittedIntegrator : OtherIntegrator {
    explicit WhittedIntegrator(const PropertyList& p) : OtherIntegrator(p, "whitted") {}
};
DRAWLAB_REGISTER_CLASS(NormalIntegrator, "normals");
DRAWLAB_REGISTER_CLASS(WhittedIntegrator, "whitted");

// Independent (src/tracer/sampler/independent.cpp:16-19): sampleCount 1.
// `seed` keys the per-(pixel, sample) pcg32 streams (extension, default 0).
class Independent : public Sampler {
public:
    explicit Independent(const PropertyList& p)
        : m_count(p.getInteger("sampleCount", 1)), m_seed((uint32_t)p.getInteger("seed", 0)) {}
    int getSampleCount() const override { return m_count; }
    uint32_t getSeed() const override { return m_seed; }
    std::string toString() const override { return format("Independent[sampleCount = %i]", m_count); }

private:
    int m_count;
    uint32_t m_seed;
};
DRAWLAB_REGISTER_CLASS(Independent, "independent");

// ============================================================================
// Scene  (src/tracer/scene/scene.cpp:6-107)
// ============================================================================
Scene::Scene(const PropertyList& p) {
    m_bg_color = p.getColor("bg_color", Color3f(0.f));
    (void)p.getFloat("epsilon", 1e-3f);
}
// The scene deletes every object it holds once: what it was given directly
// and the materials of its meshes (shared through <ref> or not).
Scene::~Scene() {
    std::set<Object*> all(m_owned.begin(), m_owned.end());
    all.insert(m_bsdfs.begin(), m_bsdfs.end());
    for (Object* o : all) delete o;
}

void Scene::activate() {
    if (!m_integrator) throw Exception("No integrator was specified!");
    if (!m_camera) throw Exception("No camera was specified!");
    if (!m_sampler) {
        m_sampler = static_cast<Sampler*>(ObjectFactory::createInstance("independent", PropertyList()));
        m_owned.push_back(m_sampler);
    }
}

void Scene::addChild(Object* obj) {
    switch (obj->getClassType()) {
        case EMesh: {
            Mesh* mesh = static_cast<Mesh*>(obj);
            m_meshes.push_back(mesh);
            m_owned.push_back(mesh);
            if (mesh->isEmitter()) {
                m_mesh_light_idx.push_back((int)m_emitters.size());
                m_emitters.push_back(mesh->getEmitter());
                m_owned.push_back(mesh->getEmitter());
            } else {
                m_mesh_light_idx.push_back(-1);
            }
            // one material per distinct BSDF object, in order of first use
            int idx = -1;
            for (size_t i = 0; i < m_bsdfs.size(); i++)
                if (m_bsdfs[i] == mesh->getBSDF()) idx = (int)i;
            if (idx < 0) {
                idx = (int)m_bsdfs.size();
                m_bsdfs.push_back(mesh->getBSDF());
            }
            m_mesh_bsdf_idx.push_back(idx);
        } break;
        case EEmitter: {
            Emitter* e = static_cast<Emitter*>(obj);
            if (e->isEnvironmentEmitter()) {
                if (m_envmap_idx >= 0) throw Exception("There is more than one envmap in the scene!");
                m_envmap_idx = (int)m_emitters.size();
            }
            m_emitters.push_back(e);
            m_owned.push_back(e);
        } break;
        case ESampler:
            if (m_sampler) throw Exception("There can only be one sampler per scene!");
            m_sampler = static_cast<Sampler*>(obj);
            m_owned.push_back(obj);
            break;
        case ECamera:
            if (m_camera) throw Exception("There can only be one camera per scene!");
            m_camera = static_cast<Camera*>(obj);
            m_owned.push_back(obj);
            break;
        case EIntegrator:
            if (m_integrator) throw Exception("There can only be one integrator per scene!");
            m_integrator = static_cast<Integrator*>(obj);
            m_owned.push_back(obj);
            break;
        case EBSDF:
            // declared at scene level for <ref>: owned by the pool, not a
            // material of the scene until a mesh uses it
            m_owned.push_back(obj);
            break;
        case ETexture: m_owned.push_back(obj); break;
        default: throw Exception("Scene::addChild(<%s>) is not supported!", classTypeName(obj->getClassType()));
    }
}

DRAWLAB_REGISTER_CLASS(Scene, "scene");

std::string Scene::toString() const {
    std::string s = "Scene[\n  integrator = " + m_integrator->toString() + ",\n  sampler = " + m_sampler->toString() +
                    ",\n  camera = " + m_camera->toString() + ",\n  meshes = {\n";
    for (const Mesh* m : m_meshes) s += "    " + m->toString() + "\n";
    s += "  }\n  emitters = {\n";
    for (const Emitter* e : m_emitters) s += "    " + e->toString() + "\n";
    return s + "  }\n]";
}

// ============================================================================
// export to the C ABI  (mirrors OptixRenderer::initOptixScene,
// src/tracer/backend/optix_renderer.cpp:25-82)
// ============================================================================
void exportScene(const Scene& scene, ExportedScene& out) {
    const Integrator* integ = scene.getIntegrator();
    if (integ->hipFlags() < 0)
        throw Exception("integrator '%s' is not supported by the hip backend (only 'path' and 'normals')",
                        integ->toString());
    if (scene.getEmitters().empty()) throw Exception("There is not emitter in the scene.");
    out.meshes.clear();
    out.materials.clear();
    out.lights.clear();
    out.textures.clear();
    std::vector<const Texture*> tex_objs;
    for (const BSDF* b : scene.getBSDFs()) {
        dl_material_desc d = b->descriptor();
        // each distinct texture once, in order of first use
        // (BSDF::createOptixMaterial, e.g. aniso_ggx.cpp:79-138)
        for (const auto& s : b->textures) {
            size_t k = 0;
            while (k < tex_objs.size() && tex_objs[k] != s.second) k++;
            if (k == tex_objs.size()) {
                tex_objs.push_back(s.second);
                dl_texture_desc td{};
                td.width = s.second->width;
                td.height = s.second->height;
                td.rgb = s.second->rgb.data();
                out.textures.push_back(td);
            }
            d.tex[s.first] = (int32_t)k + 1;
        }
        d.maps_object_space = b->maps_object_space ? 1 : 0;
        out.materials.push_back(d);
    }
    const auto& meshes = scene.getMeshes();
    for (size_t i = 0; i < meshes.size(); i++) {
        const Mesh* m = meshes[i];
        dl_mesh_desc d{};
        d.positions = m->positions().data();
        d.indices = m->indices().data();
        d.normals = m->hasVertexNormal() ? m->normals().data() : nullptr;
        d.texcoords = m->hasTexCoord() ? m->texcoords().data() : nullptr;
        d.tangents = nullptr;
        d.num_vertices = m->getVertexCount();
        d.num_triangles = m->getTriangleCount();
        d.material = scene.getMeshBsdfIdx((int)i);
        d.light = scene.getMeshLightIdx((int)i);
        out.meshes.push_back(d);
    }
    for (const Emitter* e : scene.getEmitters()) {
        int mesh = -1;
        for (size_t i = 0; i < meshes.size(); i++)
            if (meshes[i] == e->getMesh()) mesh = (int)i;
        out.lights.push_back(e->descriptor(mesh));
    }
    dl_scene_desc& s = out.desc;
    s = dl_scene_desc{};
    s.meshes = out.meshes.data();
    s.num_meshes = (uint32_t)out.meshes.size();
    s.materials = out.materials.data();
    s.num_materials = (uint32_t)out.materials.size();
    s.lights = out.lights.data();
    s.num_lights = (uint32_t)out.lights.size();
    s.textures = out.textures.data();
    s.num_textures = (uint32_t)out.textures.size();
    s.camera = scene.getCamera()->descriptor();
    s.film = scene.getCamera()->getReconstructionFilter()->descriptor();
    put3(s.bg_color, scene.getBackgroundColor());
    dl_render_params& p = out.params;
    p = dl_render_params{};
    p.spp = scene.getSampler()->getSampleCount();
    p.seed = scene.getSampler()->getSeed();
    p.max_depth = integ->maxDepth();
    p.rr_depth = integ->rrDepth();
    p.flags = integ->hipFlags();
}

}  // namespace drawlab
