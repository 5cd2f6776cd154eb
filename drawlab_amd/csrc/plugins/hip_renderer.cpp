// HipRenderer: renders a loaded Scene with the HIP backend through the C ABI
// -- the third backend beside CPURenderer and OptixRenderer
// (src/tracer/backend/cpu_renderer.cpp:57-135,
// src/tracer/backend/optix_renderer.cpp:25-82) -- and the C API of the plugin
// layer (drawlab_host.h).
#include <cstring>
#include <memory>

#include "drawlab/plugins.h"
#include "drawlab_host.h"

namespace drawlab {

namespace {
thread_local std::string g_error;

struct Handle {
    std::unique_ptr<Object> root;
    ExportedScene ex;
    bool exported = false;
    Scene* scene() const { return static_cast<Scene*>(root.get()); }
};

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

void ensure_export(Handle* h) {
    if (!h->exported) {
        exportScene(*h->scene(), h->ex);
        h->exported = true;
    }
}

// One render of the exported scene on one device.  Errors of the backend
// come back as drawlab::Exception carrying dl_hip_last_error's message.
void render_scene(const ExportedScene& ex, int device, int builder, int spp, float* rgb, dl_render_stats* stats) {
    void* ctx = nullptr;
    if (dl_hip_create(device, &ctx) != DL_OK) {
        const std::string msg = ctx ? dl_hip_last_error(ctx) : "no HIP device";
        if (ctx) dl_hip_destroy(ctx);
        throw Exception("hip backend: %s", msg);
    }
    std::unique_ptr<void, dl_status (*)(void*)> guard(ctx, dl_hip_destroy);
    auto check = [&](dl_status s) {
        if (s != DL_OK) throw Exception("%s", dl_hip_last_error(ctx));
    };
    check(dl_hip_upload_scene(ctx, &ex.desc));
    check(dl_hip_build_accel(ctx, builder, nullptr));
    dl_render_params p = ex.params;
    if (spp > 0) p.spp = spp;
    check(dl_hip_render(ctx, &p, rgb, stats));
}
}  // namespace

// CPURenderer::renderAsync's contract (cpu_renderer.cpp:57-135): render the
// scene, then write <scene file without extension>.exr and .png (and .pfm).
void renderToFile(const Scene& scene, const std::string& scene_file, int device, int builder, int spp) {
    ExportedScene ex;
    exportScene(scene, ex);
    const int W = ex.desc.camera.width, H = ex.desc.camera.height;
    Image img;
    img.width = W;
    img.height = H;
    img.rgb.resize((size_t)W * H * 3);
    dl_render_stats st{};
    render_scene(ex, device, builder, spp, img.rgb.data(), &st);
    std::string out = scene_file;
    const size_t dot = out.find_last_of('.');
    if (dot != std::string::npos && out.find('/', dot) == std::string::npos) out.erase(dot);
    saveEXR(out + ".exr", img);
    savePNG(out + ".png", img);
    savePFM(out + ".pfm", img);
    std::printf("Render done: %dx%d, %.0f paths in %.1f ms (%.1f Mpaths/s); wrote %s.exr\n", W, H, st.paths,
                st.ms_total, st.paths / (st.ms_total * 1e3), out.c_str());
}

}  // namespace drawlab

using namespace drawlab;

extern "C" {

const char* dlx_last_error(void) { return g_error.c_str(); }

int dlx_load_xml(const char* path, void** scene) {
    return guarded([&] {
        std::unique_ptr<Handle> h(new Handle);
        h->root.reset(loadFromXML(path));
        if (h->root->getClassType() != Object::EScene) throw Exception("The root element of \"%s\" is not a <scene>", path);
        *scene = h.release();
    });
}

void dlx_free(void* scene) { delete static_cast<Handle*>(scene); }

int dlx_export(void* scene, const dl_scene_desc** desc, dl_render_params* params) {
    return guarded([&] {
        Handle* h = static_cast<Handle*>(scene);
        ensure_export(h);
        if (desc) *desc = &h->ex.desc;
        if (params) *params = h->ex.params;
    });
}

int dlx_describe(void* scene, char* buf, size_t n) {
    return guarded([&] {
        const std::string s = static_cast<Handle*>(scene)->scene()->toString();
        if (n) {
            std::strncpy(buf, s.c_str(), n - 1);
            buf[n - 1] = 0;
        }
    });
}

int dlx_render(void* scene, int device, int builder, int spp, float* rgb_out, dl_render_stats* stats) {
    return guarded([&] {
        Handle* h = static_cast<Handle*>(scene);
        ensure_export(h);
        render_scene(h->ex, device, builder, spp, rgb_out, stats);
    });
}

int dlx_save_image(const char* path, const float* rgb, int width, int height) {
    return guarded([&] {
        Image img;
        img.width = width;
        img.height = height;
        img.rgb.assign(rgb, rgb + (size_t)width * height * 3);
        if (endsWith(toLower(path), ".pfm"))
            savePFM(path, img);
        else if (endsWith(toLower(path), ".exr"))
            saveEXR(path, img);
        else if (endsWith(toLower(path), ".png"))
            savePNG(path, img);
        else
            throw Exception("Unsupported output format \"%s\" (expected .exr, .png or .pfm)", path);
    });
}

int dlx_load_image(const char* path, int* width, int* height, float* rgb) {
    return guarded([&] {
        const Image img = loadImage(path);
        *width = img.width;
        *height = img.height;
        if (rgb) std::memcpy(rgb, img.rgb.data(), img.rgb.size() * sizeof(float));
    });
}

}  // extern "C"
