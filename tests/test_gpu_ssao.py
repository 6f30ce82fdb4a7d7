"""Screen-space occlusion (vx_ssao: calculateOcclusion, LightAmbient.frag:
59-109) bit for bit against the oracle, and the shading that takes it as its
ambient factor (LightAmbient.frag:211)."""
import numpy as np
import pytest
import torch

from helpers import assert_bitwise, device_scene, oracle_occ

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,frame", [((160, 112), 0), ((97, 61), 3)])
def test_ssao_and_ambient_match_oracle(ctx, oracle, size, frame):
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, 96, seed=41, nboxes=60, box=(3, 14))
    W, H = size
    vd = views.scene_view(model.dims, W, H, frame=frame)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, vd, g)
    depth = torch.zeros(W * H, dtype=torch.float32, device="cuda")
    api.gbuf_depth(ctx, vd, g, depth)
    s = torch.zeros(W * H, dtype=torch.float32, device="cuda")
    api.ssao(ctx, vd, g, depth, s, seed=17)
    torch.cuda.synchronize()
    nrm, mat = g.normal.cpu().numpy(), g.material.cpu().numpy()
    ref = oracle.ssao(vd, nrm, mat, depth.cpu().numpy(), seed=17)
    assert_bitwise(s.cpu().numpy(), ref, "ssao")
    hit = mat > 0
    assert (ref[hit] < 1).mean() > 0.05 and (ref[~hit] == 1).all()
    # the ambient pass shading with the occlusion factor
    pal = configs.make_palette()
    params = configs.trace_params(n_ao=4, ao_maxt=16.0, n_spec=0, march="reference")
    vis = api.Visibility(W, H)
    pal_d = torch.from_numpy(pal).cuda()
    api.trace(ctx, occ, vd, g, pal_d, params, vis)
    light = configs.default_light()
    rgb = torch.zeros((W * H, 3), dtype=torch.float32, device="cuda")
    api.shade_forward(ctx, vd, g, vis, pal_d, light, rgb, ssao=s)
    torch.cuda.synchronize()
    want = oracle.shade_f32(vd, g.pos.cpu().numpy(), nrm, mat, vis.shadow.cpu().numpy(), vis.ao.cpu().numpy(), pal,
                            light, ssao=ref)
    assert_bitwise(rgb.cpu().numpy(), want, "shaded")
