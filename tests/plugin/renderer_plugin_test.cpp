// renderer_plugin_test.cpp -- drives the HIP wavefront tracer through the reference's renderer
// plugin interface (include/renderers/renderer.h:14-35) the way RenderSession does
// (render_session.h:59-117): a std::unique_ptr<Renderer>, renderAsync / pollRender with the
// RenderRecord's events, getSpp, reset, getRadianceBuffer, and a renderer switch on one
// DeviceGlobals.  The plugin's images are compared bit for bit with a renderer driven
// directly through the C ABI on the same scene and seeds.
//
// Exit status: 0 all checks passed, 1 a check failed, 77 no HIP device (nothing was run).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../gpu-pathed_amd/csrc/renderer.h"

using namespace rays;

namespace {

int failures = 0;

void expect(bool ok, const std::string &what) {
    printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
    fflush(stdout);
    if (!ok) failures++;
}

void must(int status) {
    if (status != PT_OK) throw std::runtime_error(pt_last_error());
}

// The scene-side state of a libpathed scene: camera and materials (what RenderSession builds).
Scene sceneOf(const pt_scene *s, int w, int h) {
    pt_scene_info info;
    must(pt_scene_get_info(s, &info));
    const int64_t nm = info.n_materials;
    std::vector<int> type(nm);
    std::vector<float> albedo(3 * nm), emit(3 * nm), ior(nm), alpha(nm);
    must(pt_scene_export(s, nullptr, nullptr, nullptr, nullptr, nullptr, type.data(), albedo.data(), emit.data(), ior.data(), alpha.data(),
                         nullptr));
    std::vector<MaterialParams> mats(nm);
    for (int64_t i = 0; i < nm; i++) {
        mats[i].type = (MaterialType)type[i];
        mats[i].albedo = Vec3(albedo[3 * i], albedo[3 * i + 1], albedo[3 * i + 2]);
        mats[i].emit = Vec3(emit[3 * i], emit[3 * i + 1], emit[3 * i + 2]);
        mats[i].ior = ior[i];
        mats[i].alpha = alpha[i];
    }
    const Camera cam(Vec3(info.camera_origin[0], info.camera_origin[1], info.camera_origin[2]),
                     Vec3(info.camera_target[0], info.camera_target[1], info.camera_target[2]),
                     Vec3(info.camera_up[0], info.camera_up[1], info.camera_up[2]), info.camera_vfov, Resolution{w, h}, info.camera_flip != 0);
    return Scene(cam, mats);
}

// One pass through the interface, polled to completion as the session's paint loop does.
float pass(Renderer &r, int spp, HeadlessPbo &pbo, const Scene &scene, const DeviceGlobals &g) {
    const RenderRecord rec = r.renderAsync(spp, &pbo, scene, g);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::seconds(60);
    while (!r.pollRender(&pbo, rec)) {
        if (std::chrono::steady_clock::now() > deadline) throw std::runtime_error("pollRender did not report the pass finished within 60 s");
    }
    float ms = -1.f;
    if (hipEventElapsedTime(&ms, rec.beginEvent, rec.endEvent) != hipSuccess) ms = -1.f;
    return ms;
}

}  // namespace

int main() {
    int ndev = 0;
    pt_device_count(&ndev);
    if (ndev == 0) {
        printf("no HIP device: nothing run\n");
        return 77;
    }
    const int W = 64, H = 48, spp = 4, passes = 4;
    try {
        pt_scene *world = nullptr;
        must(pt_scene_create(&world));
        must(pt_scene_preset(world, 0, 0, 0));  // the Cornell box
        Scene scene = sceneOf(world, W, H);
        scene.setMaxDepth(5);

        DeviceGlobals globals;
        {
            std::unique_ptr<Renderer> early = makeRenderer(RendererType::CUDA);
            bool threw = false;
            try {
                HeadlessPbo p;
                early->renderAsync(1, &p, scene, globals);
            } catch (const std::runtime_error &) {
                threw = true;
            }
            expect(threw, "renderAsync before init / mallocWorld throws");
        }
        globals.mallocWorld(world, W, H);
        globals.copyCamera(scene.getCamera());
        globals.updateMaterials(scene);

        // RenderSession holds its renderer through the base class (render_session.h:123)
        std::unique_ptr<Renderer> renderer = makeRenderer(RendererType::CUDA);
        renderer->init(W, H, scene);
        expect(renderer->getSpp() == 0, "getSpp is 0 after init");
        HeadlessPbo pbo;
        pbo.id = 7;
        bool timesOk = true, sppOk = true;
        for (int i = 0; i < passes; i++) {
            const float ms = pass(*renderer, spp, pbo, scene, globals);
            timesOk = timesOk && ms > 0.f;
            sppOk = sppOk && renderer->getSpp() == spp * (i + 1);
        }
        expect(timesOk, "RenderRecord events give a positive pass time (hipEventElapsedTime)");
        expect(sppOk, "getSpp counts the samples of every finished pass");
        const std::vector<float> img = renderer->getRadianceBuffer();
        double mean = 0.0;
        bool finite = img.size() == 3 * (size_t)W * H;
        for (float v : img) {
            finite = finite && isfinite(v);
            mean += v;
        }
        mean /= img.empty() ? 1 : (double)img.size();
        expect(finite && mean > 0.05, "getRadianceBuffer: 3*W*H finite values, mean " + std::to_string(mean));

        // the same render driven directly through the C ABI: identical bytes
        pt_renderer *direct = nullptr;
        must(pt_renderer_create(W, H, &direct));
        must(pt_renderer_set_scene(direct, world));
        must(pt_renderer_set_type(direct, PT_RENDERER_CUDA));
        must(pt_renderer_set_max_depth(direct, 5));
        must(pt_renderer_set_nee(direct, 1));
        for (int i = 0; i < passes; i++) {
            must(pt_renderer_render_async(direct, spp));
            must(pt_renderer_synchronize(direct));
        }
        std::vector<float> want(3 * (size_t)W * H);
        std::vector<uint8_t> wantRgba(4 * (size_t)W * H);
        must(pt_renderer_get_radiance(direct, want.data()));
        must(pt_renderer_get_rgba8(direct, wantRgba.data()));
        expect(memcmp(want.data(), img.data(), want.size() * sizeof(float)) == 0, "radiance equals the C-ABI renderer's bit for bit");
        expect(pbo.rgba == wantRgba, "the pixel buffer holds the pass's uchar4 sRGB framebuffer");

        // reset: the next pass starts the accumulation again (path_tracer.cu:307-310)
        renderer->reset();
        pass(*renderer, 2, pbo, scene, globals);
        expect(renderer->getSpp() == 2, "after reset() the next pass restarts the count");

        // an edit through the device globals (the session's subscriber) plus a reset
        scene.setColor(0, Vec3(0.9f, 0.1f, 0.1f));
        globals.updateMaterial(0, scene.getMaterial(0));
        renderer->reset();
        pass(*renderer, 4, pbo, scene, globals);
        const std::vector<float> edited = renderer->getRadianceBuffer();
        expect(renderer->getSpp() == 4 && memcmp(edited.data(), img.data(), img.size() * sizeof(float)) != 0,
               "a material edit changes the image and restarts accumulation");

        // the renderer switch of render_session.h:103-114: a new renderer on the same globals
        renderer = makeRenderer(RendererType::Normals);
        renderer->init(W, H, scene);
        pass(*renderer, 1, pbo, scene, globals);
        const std::vector<float> nrm = renderer->getRadianceBuffer();
        // g_buffer.cu: colour = (n + 1) / 2 of the unit normal at the hit, black on a miss
        bool unit = true;
        int hits = 0;
        for (size_t p = 0; p < (size_t)W * H; p++) {
            const float *c = &nrm[3 * p];
            if (c[0] == 0.f && c[1] == 0.f && c[2] == 0.f) continue;
            const float n[3] = {2.f * c[0] - 1.f, 2.f * c[1] - 1.f, 2.f * c[2] - 1.f};
            unit = unit && fabsf(sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) - 1.f) < 1e-3f;
            hits++;
        }
        expect(renderer->getSpp() == 1 && unit && hits > W * H / 2,
               "a Normals renderer on the same device globals renders the normals view (" + std::to_string(hits) + " hits)");

        bool mismatch = false;
        try {
            DeviceGlobals other;
            other.mallocWorld(world, W / 2, H);
            std::unique_ptr<Renderer> r2 = makeRenderer(RendererType::Optix);
            r2->init(W, H, scene);
            r2->renderAsync(1, nullptr, scene, other);
        } catch (const std::invalid_argument &) {
            mismatch = true;
        }
        expect(mismatch, "renderAsync on device globals of another image size throws");

        pt_renderer_destroy(direct);
        renderer.reset();
        pt_scene_destroy(world);
    } catch (const std::exception &e) {
        printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    printf("%s\n", failures ? "FAILED" : "PASSED");
    return failures ? 1 : 0;
}
