/*
 * drawlab_host.h -- C API of the host plugin layer (libdrawlab_host.so):
 * loads a DrawLab scene XML file through the reference's plugin model
 * (loadFromXML, src/core/parser/parser.cpp:19-387, and the plugins under
 * src/tracer/), exports it as the dl_hip.h descriptors, and renders it with
 * the HIP backend the way the reference's renderers do
 * (CPURenderer::renderAsync, src/tracer/backend/cpu_renderer.cpp:57-135).
 * Every call returns 0 on success; dlx_last_error() describes a failure
 * (the message of the drawlab::Exception the reference would throw).
 */
#ifndef DRAWLAB_HOST_H
#define DRAWLAB_HOST_H

#include "dl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Parse a scene file; *scene receives a handle owning the object tree.
   Replaces loadFromXML (parser.cpp:19) + the Scene root (main.cpp:32). */
int dlx_load_xml(const char* path, void** scene);
void dlx_free(void* scene);
const char* dlx_last_error(void);

/* Descriptors of a loaded scene (valid until dlx_free) and its render
   parameters (sampler sampleCount/seed, integrator maxDepth/rrDepth).
   Replaces OptixRenderer::initOptixScene's export (optix_renderer.cpp:25-82). */
int dlx_export(void* scene, const dl_scene_desc** desc, dl_render_params* params);

/* Scene::toString (scene.cpp:109-137) into buf (truncated to n bytes). */
int dlx_describe(void* scene, char* buf, size_t n);

/* Render on HIP device `device` with BVH builder `builder` (DL_BUILD_*):
   W*H*3 floats into rgb_out (row 0 = top).  spp <= 0 uses the sampler's
   sampleCount.  Replaces CPURenderer::renderAsync's render loop. */
int dlx_render(void* scene, int device, int builder, int spp, float* rgb_out, dl_render_stats* stats);

/* Images: .exr (uncompressed) / .png (8-bit sRGB, Bitmap::savePNG) / .pfm
   written, .exr / .pfm / .hdr read.
   dlx_load_image: call with rgb = NULL to get the size, then with a buffer
   of w*h*3 floats.  Replaces Bitmap::saveEXR / Bitmap(filename)
   (src/core/bitmap/bitmap.cpp:33-214). */
int dlx_save_image(const char* path, const float* rgb, int width, int height);
int dlx_load_image(const char* path, int* width, int* height, float* rgb);

#ifdef __cplusplus
}
#endif
#endif
