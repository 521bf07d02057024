// hr_cli.cpp -- the headless renderer `happyray`: Main.cpp's "-nw" path over
// the C ABI.
//
// Reference behaviour followed:
//   main                             src/Main.cpp:42-83
//   SDLGLApplication::init           src/Application/SDLGLApplication.cpp:54-81
//   CUDAApplication::initScene       src/Application/CUDAApplication.cu:63-71
//   SDLGLApplication::setRenderMode  src/Application/SDLGLApplication.cpp:637-650
//   SDLGLApplication::dumpFrames     src/Application/SDLGLApplication.cpp:627-635
//   SDLGLApplication::renderFrame    src/Application/SDLGLApplication.cpp:776-805
//   SDLGLApplication::writeScreenShot src/Application/SDLGLApplication.cpp:120-143
//
// Usage: happyray [scene.cfg] -nw, then the render mode (0-3) and the samples
// per pixel on standard input.  Frame 0 of the scene is built (two-level grid,
// densities 1.2 / 2.2, RTEngine.cu:77-78) and rendered spp times with image
// ids 0..spp-1 into a running mean, written gamma-corrected to
// screenshot0.png.  There is no window: without -nw the program says so and
// exits.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "happyray.h"

namespace {

int fail(const char* what)
{
    std::cerr << what << ": " << hr_last_error() << "\n";
    return 1;
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc > 3) {
        std::cerr << "Invalid arguments. Use no argumets, <scene.cfg>, or -nw.\n";
        return 1;
    }
    std::string config = "scene.cfg";
    bool no_window = false;
    for (int i = 1; i < argc; ++i) {
        if (std::string(argv[i]) == "-nw") no_window = true;
        else config = argv[i];
    }
    if (!no_window) {
        std::cerr << "happyray renders without a window only: run it with -nw\n";
        return 1;
    }

    hr_scene* scene = nullptr;
    if (hr_scene_load(config.c_str(), &scene)) return fail("loading the scene");
    hr_scene_info info;
    if (hr_scene_get_info(scene, &info)) return fail("scene info");
    hr_options opt = { 0, 0x1234u, 0x5678u };
    hr_ctx* ctx = nullptr;
    if (hr_create(&ctx, &opt)) return fail("creating the context");
    if (hr_scene_upload(ctx, scene)) return fail("uploading the scene");
    // initScene uploads key frame 0 twice with coefficient 1
    if (hr_set_frame(ctx, 0, 0, 1.f)) return fail("selecting frame 0");
    // a scene with an instances file renders through the instanced hierarchy,
    // as the reference's TLGRIDHIERARCHY build does when there is more than one
    // instance (RTEngine.cu:263-264); that build renders ambient occlusion only
    const bool hierarchy = info.num_instances > 1;
    if (hr_set_grid(ctx, hierarchy ? HR_GRID_HIERARCHY : HR_GRID_TWO_LEVEL, 1.2f, 2.2f)) return fail("grid settings");
    if (hr_build(ctx)) return fail("building the grid");
    int res_x = 512, res_y = 512;
    if (info.has_camera) {
        res_x = info.cam_res_x;
        res_y = info.cam_res_y;
    } else {
        // CameraManager defaults (src/Application/CameraManager.hpp:54-75)
        const float pos[3] = { 0.f, 0.f, 0.f }, dir[3] = { 0.f, 0.f, 1.f }, up[3] = { 0.f, -1.f, 0.f };
        if (hr_set_camera(ctx, pos, dir, up, 66.f, res_x, res_y)) return fail("camera");
    }

    std::cout << "available render modes:\n ";
    std::cout << "DEFAULT          : 0\n ";
    std::cout << "PATH_TRACE_FAST  : 1\n ";
    std::cout << "PATH_TRACE_SIMPLE: 2\n ";
    std::cout << "AMBIENT_OCCLUSION: 3\n ";
    std::cout << "select render mode: ";
    int mode = 0;
    std::cin >> mode;
    mode = mode < 0 ? 0 : (mode > 3 ? 3 : mode);
    if (hierarchy && mode != 3) {
        std::cout << "(the instanced hierarchy renders ambient occlusion)\n";
        mode = 3;
    }
    std::cout << "samples per pixel: ";
    int spp = 1;
    std::cin >> spp;
    spp = spp < 1 ? 1 : spp;
    std::cout << "\n";

    hr_grid_info gi;
    if (hr_get_grid_info(ctx, &gi)) return fail("grid info");
    const float dx = gi.bmax[0] - gi.bmin[0], dy = gi.bmax[1] - gi.bmin[1], dz = gi.bmax[2] - gi.bmin[2];
    hr_render_params p;
    p.jitter = 1;
    p.ao_samples = 4;                                    // NUMAMBIENTOCCLUSIONSAMPLES (AOIntegrator.h:45)
    p.ao_dist = std::sqrt(dx * dx + dy * dy + dz * dz) * 0.05f;   // RTEngine.cu:259
    p.pt_depth = 64;
    p.collect_stats = 0;
    const hr_mode modes[4] = { HR_MODE_SIMPLE, HR_MODE_PATH_MULTI, HR_MODE_PATH, HR_MODE_AO };
    std::vector<float> rgb(3 * (size_t)res_x * (size_t)res_y);
    for (int i = 0; i < spp; ++i)
        if (hr_render(ctx, modes[mode], i, &p, i + 1 == spp ? rgb.data() : nullptr)) return fail("rendering");
    hr_timing tm;
    if (hr_get_timing(ctx, &tm) == HR_OK) std::cout << "render time of the last sample: " << tm.render_ms << " ms\n";
    const char* out = "screenshot0.png";
    if (hr_write_screenshot(out, rgb.data(), res_x, res_y)) return fail("writing the screenshot");
    std::cout << "Wrote " << out << "\n";
    hr_destroy(ctx);
    hr_scene_free(scene);
    return 0;
}
