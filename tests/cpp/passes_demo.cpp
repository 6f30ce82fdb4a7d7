// Engine-side use of include/vx/passes.hpp in the order WorldRenderer::DrawWorld
// runs the reference's passes (WorldRenderer.cpp:239-274): rebuild, primary hit,
// ambient + reflection visibility, shading, point and spot lights.
//
// passes_demo                      synthetic world, one rebuild
// passes_demo MODS PREFAB OUT      spawns MODS/PREFAB (a .pf) into the reference's
//                                  524x188x524 volume, runs one OnUpdate and
//                                  writes the volume bytes to OUT
#include <cstdio>
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
        std::printf("passes.hpp demo: context, asset, rebuild OK\n");
        return 0;
    } catch (const vx::Error& e) {
        std::printf("vx error %d: %s\n", e.code, e.what());
        return 1;
    }
}
