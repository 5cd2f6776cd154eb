// drawlab: command-line renderer (src/tracer/main.cpp:14-61) with the HIP
// backend.  Usage mirrors the reference:
//   drawlab -s scene.xml [-b hip] [-t threads] [--device N] [--builder gpu|host_sah|lbvh] [--spp N]
// The scene's directory is searched for relative resource paths; the image
// is written beside the scene as <name>.exr, <name>.png and <name>.pfm.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "drawlab/plugins.h"

namespace drawlab {
void renderToFile(const Scene& scene, const std::string& scene_file, int device, int builder, int spp);
}

using namespace drawlab;

static void usage() {
    std::printf(
        "usage: drawlab --scene=string [options] ...\n"
        "options:\n"
        "  -s, --scene      Scene xml file (string)\n"
        "  -t, --thread     Thread num used in cpu backend (int [=4]; unused by hip)\n"
        "  -b, --backend    Backend:[hip] (string [=hip])\n"
        "      --device     HIP device index (int [=0])\n"
        "      --builder    BVH builder:[gpu, host_sah, lbvh] (string [=gpu])\n"
        "      --spp        Override the sampler's sampleCount (int)\n"
        "  -?, --help       print this message\n");
}

int main(int argc, char** argv) {
    std::string scene, backend = "hip", builder = "gpu";
    int device = 0, spp = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i], v;
        const size_t eq = a.find('=');
        if (eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
        }
        auto value = [&]() -> std::string {
            if (!v.empty()) return v;
            if (i + 1 >= argc) {
                std::fprintf(stderr, "option needs a value: %s\n", a.c_str());
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "-s" || a == "--scene")
            scene = value();
        else if (a == "-b" || a == "--backend")
            backend = value();
        else if (a == "-t" || a == "--thread")
            (void)value();
        else if (a == "--device")
            device = std::atoi(value().c_str());
        else if (a == "--builder")
            builder = value();
        else if (a == "--spp")
            spp = std::atoi(value().c_str());
        else if (a == "-?" || a == "--help") {
            usage();
            return 0;
        } else if (a == "--gui") {
            std::fprintf(stderr, "Fatal error: --gui is not supported by the hip renderer\n");
            return 1;
        } else {
            std::fprintf(stderr, "unknown option: %s\n", a.c_str());
            usage();
            return 1;
        }
    }
    if (scene.empty()) {
        usage();
        return 1;
    }
    if (backend != "hip") {
        std::fprintf(stderr, "Fatal error: unknown backend:  %s (the renderer here is 'hip')\n", backend.c_str());
        return 1;
    }
    const int b = builder == "gpu" ? DL_BUILD_GPU : builder == "host_sah" ? DL_BUILD_HOST_SAH
                                                  : builder == "lbvh" ? DL_BUILD_GPU_LBVH : -1;
    if (b < 0) {
        std::fprintf(stderr, "Fatal error: unknown builder: %s\n", builder.c_str());
        return 1;
    }
    try {
        if (!endsWith(toLower(scene), ".xml"))
            throw Exception("unknown file: \"%s\", expected an extension of type .xml", scene);
        std::unique_ptr<Object> root(loadFromXML(scene));
        if (root->getClassType() != Object::EScene) throw Exception("the root element is not a <scene>");
        renderToFile(*static_cast<Scene*>(root.get()), scene, device, b, spp);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Fatal error: %s\n", e.what());
        return -1;
    }
    return 0;
}
