"""The C++ pass wrappers (include/vx/passes.hpp) compile against the C ABI and
link against the built library, with the point/spot light pipelines
instantiated the way the reference's DrawLight callbacks use them."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "voxelengine_amd")

LIGHTS = r'''
#include "vx/passes.hpp"
void draw_lights(vx::Context& ctx, void* cmd, const vx_view& view, const vx_gbuf& gb, vx_volume* vol,
                 const float* pal, vx_light* lights_dev, float* rgb) {
    vx_light_params lp{VX_MARCH_SPARSE, VX_MARCH_SUPERSPARSE, 0u, 0};
    float pos[3] = {1, 2, 3}, col[3] = {1, 1, 1}, dir[3] = {0, 0, 1};
    vx::LightPointPipeline::Get().Use(ctx, cmd, view, gb, vol, pal, lp, lights_dev, rgb,
        [&](vx::LightPointPipeline& P) { P.DrawLight(pos, 5.0f, col, 1.0f); });
    vx::LightSpotPipeline::Get().Use(ctx, cmd, view, gb, vol, pal, lp, lights_dev, rgb,
        [&](vx::LightSpotPipeline& P) { P.DrawLight(pos, 5.0f, col, 1.0f, dir, 0.5f, 1.0f); });
}
'''


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    pytest.skip("hipcc not found")


def test_passes_header_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "libvx_amd.so")):
        pytest.skip("library not built")
    src = tmp_path / "lights.cpp"
    src.write_text(LIGHTS)
    demo = os.path.join(ROOT, "tests", "cpp", "passes_demo.cpp")
    out = tmp_path / "demo"
    cmd = [_hipcc(), "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), demo, str(src),
           "-L", LIB_DIR, "-lvx_amd", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert out.exists()
