// Engine-side use of include/vx/passes.hpp in the order WorldRenderer::DrawWorld
// runs the reference's passes (WorldRenderer.cpp:239-274): rebuild, primary hit,
// ambient + reflection visibility, shading, point and spot lights.
#include <cstdio>
#include <vector>

#include "vx/passes.hpp"

int main() {
    try {
        vx::Context ctx(0);
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
