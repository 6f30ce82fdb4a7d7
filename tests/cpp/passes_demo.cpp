// Engine-side use of include/vx/passes.hpp in the order WorldRenderer::DrawWorld
// runs the reference's passes (WorldRenderer.cpp:239-274): rebuild, primary hit,
// ambient visibility, shading, temporal accumulation, reflection.
//
// passes_demo                      synthetic world, one rebuild
// passes_demo MODS PREFAB OUT      spawns MODS/PREFAB (a .pf) into the reference's
//                                  524x188x524 volume, runs one OnUpdate and
//                                  writes the volume bytes to OUT
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "vx/passes.hpp"

static int spawn_prefab(vx::Context& ctx, const char* mods, const char* prefab, const char* out) {
    vx::PrefabAsset pf(mods, prefab);
    std::map<int32_t, std::unique_ptr<vx::VoxAsset>> owned;
    std::map<int32_t, const vx::VoxAsset*> models;
    for (const vx_prefab_entity& e : pf.Entities()) {
        if (e.model >= 0 && !owned.count(e.model)) {
            owned[e.model] = std::make_unique<vx::VoxAsset>(ctx, pf.Files()[e.model]);
            models[e.model] = owned[e.model].get();
        }
    }
    vx::ShadowVoxSystem shadowVox(ctx);
    const std::vector<vx::Entity> ids = shadowVox.Spawn(pf, models);
    // the geometry pass's draw list for the same spawn
    std::vector<const vx_model*> byFile(pf.Files().size(), nullptr);
    for (const auto& kv : owned) byFile[kv.first] = kv.second->GetImage();
    const std::vector<vx_vox_cmd> draws = pf.VoxCmds(byFile);
    shadowVox.OnUpdate(0.016f);
    int32_t d[3], levels = 0;
    vx::check(vx_volume_dims(shadowVox.GetVolumeImage(), d, &levels));
    std::vector<uint8_t> bytes((size_t)d[0] * d[1] * d[2]);
    vx::check(vx_volume_download(shadowVox.GetVolumeImage(), bytes.data(), bytes.size(), nullptr));
    FILE* f = std::fopen(out, "wb");
    if (!f || std::fwrite(bytes.data(), 1, bytes.size(), f) != bytes.size()) {
        std::printf("cannot write %s\n", out);
        if (f) std::fclose(f);
        return 1;
    }
    std::fclose(f);
    std::printf("spawned %zu entities, %zu instances, %zu lights, %zu draws, %zu dirty regions\n",
                pf.Entities().size(), ids.size(), pf.Lights().size(), draws.size(),
                shadowVox.LastDirtyRegions().size());
    return 0;
}

// ViewData of a pinhole camera at eye (world units) looking at target:
// glm::lookAtRH and glm::perspectiveRH_NO with the reference's NEAR/FAR
// (Common.frag:12-13) and fov 0.8 (EditorCamera.h:11), column-major.
static vx_view make_view(const double eye[3], const double target[3], int w, int h) {
    vx_view v;
    std::memset(&v, 0, sizeof(v));
    double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double s[3] = {f[1] * 0.0 - f[2] * 1.0, f[2] * 0.0 - f[0] * 0.0, f[0] * 1.0 - f[1] * 0.0};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double V[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0, s[2], u[2], -f[2], 0,
                    -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]), -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
                    f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2], 1};
    const double nearp = 0.1, farp = 4096.0, t = std::tan(0.4), aspect = (double)w / h;
    double P[16] = {0};
    P[0] = 1.0 / (aspect * t);
    P[5] = 1.0 / t;
    P[10] = -(farp + nearp) / (farp - nearp);
    P[11] = -1.0;
    P[14] = -(2.0 * farp * nearp) / (farp - nearp);
    // inverse of the rigid view: transpose the rotation, rotate the translation back
    double IV[16] = {V[0], V[4], V[8], 0, V[1], V[5], V[9], 0, V[2], V[6], V[10], 0, eye[0], eye[1], eye[2], 1};
    double IP[16] = {0};
    IP[0] = 1.0 / P[0];
    IP[5] = 1.0 / P[5];
    IP[11] = 1.0 / P[14];
    IP[14] = -1.0;
    IP[15] = P[10] / P[14];
    for (int k = 0; k < 16; ++k) {
        v.last_view[k] = v.view[k] = (float)V[k];
        v.inv_view[k] = (float)IV[k];
        v.proj[k] = (float)P[k];
        v.inv_proj[k] = (float)IP[k];
    }
    v.res[0] = (float)w; v.res[1] = (float)h;
    v.ires[0] = 1.0f / (float)w; v.ires[1] = 1.0f / (float)h;
    for (int k = 0; k < 3; ++k) v.cam_pos[k] = (float)eye[k];
    v.width = w;
    v.height = h;
    return v;
}

template <class T>
static T* dev_alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw vx::Error(VX_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) throw vx::Error(VX_ERR_DEVICE, "hipMemset failed");
    return static_cast<T*>(p);
}

// One frame of WorldRenderer::DrawWorld's passes over the synthetic world
// (WorldRenderer.cpp:224-274): geometry, ambient shadow + AO rays and the
// shading, TAA of the light buffer, then the reflection pass reading it.
static int frame(vx::Context& ctx, vx::VoxAsset& world, vx::ShadowVoxSystem& shadowVox) {
    const int W = 96, H = 64;
    const size_t npx = (size_t)W * H;
    const double eye[3] = {-0.5, 6.7, -0.5}, target[3] = {3.2, 1.0, 3.2};
    vx_view view = make_view(eye, target, W, H);
    vx_gbuf gb{dev_alloc<float>(3 * npx), dev_alloc<float>(3 * npx), dev_alloc<uint8_t>(npx)};
    float* depth = dev_alloc<float>(npx);
    vx_trace_out vis{dev_alloc<uint8_t>(npx), dev_alloc<float>(npx), dev_alloc<float>(npx), nullptr};
    float* rgb = dev_alloc<float>(3 * npx);
    float* hist[2] = {dev_alloc<float>(4 * npx), dev_alloc<float>(4 * npx)};
    float* refl = dev_alloc<float>(4 * npx);
    float* pal = dev_alloc<float>(256 * 6);
    std::vector<float> hp(256 * 6);
    for (int m = 1; m < 256; ++m) {
        hp[6 * m + 0] = (float)(m % 7) / 7.0f;
        hp[6 * m + 1] = (float)(m % 5) / 5.0f;
        hp[6 * m + 2] = (float)(m % 3) / 3.0f;
        hp[6 * m + 3] = 0.05f + 0.9f * (float)(m % 11) / 11.0f;
        hp[6 * m + 4] = (m % 4 == 0) ? 1.0f : 0.0f;
    }
    if (hipMemcpy(pal, hp.data(), hp.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return 1;
    vx::GeometryVoxelPipeline::Get().Use(ctx, nullptr, view, world, gb);
    vx::check(vx_gbuf_depth(ctx.get(), &view, &gb, 0, depth, nullptr));
    vx_trace_params tp{VX_MARCH_SPARSE, VX_MARCH_SUPERSPARSE, VX_MARCH_SPARSE, 1, 1, 0, 128.0f, 128.0f, 256.0f,
                       {0.4242641f, 0.5656854f, 0.7071068f}, 1u, 0, 1, 0};
    vx::LightAmbientPipeline::Get().Use(ctx, nullptr, view, gb, shadowVox.GetVolumeImage(), pal, tp, vis);
    float* occl = dev_alloc<float>(npx);
    vx::LightAmbientPipeline::Get().UseOcclusion(ctx, nullptr, view, gb, depth, 1u, occl);
    const float sun[6] = {0.4242641f, 0.5656854f, 0.7071068f, 0.45f, 0.45f, 0.4f};
    vx::check(vx_shade_forward(ctx.get(), &view, &gb, &vis, occl, pal, sun, rgb, nullptr));
    vx_taa_io taa{rgb, hist[0], nullptr, depth, gb.normal, gb.material, pal, hist[1]};
    vx::LightTAAPipeline::Get().Use(ctx, nullptr, view, taa, 1u);
    vx_skybox sky{nullptr, 0, 0};
    vx_reflection_params rp{VX_MARCH_SPARSE, 256.0f, 1u, 0};
    vx::LightReflectionPipeline::Get().Use(ctx, nullptr, view, hist[1], gb, depth, shadowVox.GetVolumeImage(), sky,
                                           pal, rp, refl);
    std::vector<uint8_t> mat(npx);
    std::vector<float> out(4 * npx);
    if (hipMemcpy(mat.data(), gb.material, npx, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(out.data(), refl, out.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return 1;
    size_t hit = 0;
    double sum = 0.0;
    for (size_t i = 0; i < npx; ++i) {
        hit += mat[i] != 0;
        sum += (double)out[4 * i] + out[4 * i + 1] + out[4 * i + 2];
    }
    std::printf("frame: %zu hit pixels, reflection sum %.4f\n", hit, sum);
    for (void* p : {(void*)gb.pos, (void*)gb.normal, (void*)gb.material, (void*)depth, (void*)vis.shadow,
                    (void*)vis.ao, (void*)vis.spec, (void*)rgb, (void*)hist[0], (void*)hist[1], (void*)refl,
                    (void*)pal, (void*)occl})
        (void)hipFree(p);
    return hit > npx / 4 && sum > 0.0 ? 0 : 1;
}

int main(int argc, char** argv) {
    try {
        vx::Context ctx(0);
        if (argc == 4) return spawn_prefab(ctx, argv[1], argv[2], argv[3]);
        vx::VoxAsset world(ctx, 64, 64, 64);
        vx_scene_desc scene{7u, 0.3f, 10, 2, 8, 0};
        world.Generate(scene);
        vx::ShadowVoxSystem shadowVox(ctx, 32, 32, 32, 2);
        const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        shadowVox.CreateEntity(&world, I);
        shadowVox.OnUpdate(0.016f);
        (void)shadowVox.LastDirtyRegions();
        if (frame(ctx, world, shadowVox)) return 1;
        std::printf("passes.hpp demo: context, asset, rebuild, frame OK\n");
        return 0;
    } catch (const vx::Error& e) {
        std::printf("vx error %d: %s\n", e.code, e.what());
        return 1;
    }
}
