"""Real-data fixtures: the ModernHouse and FarmHouse prefabs voxelized into the
reference's 524x188x524 ShadowVox volume (tests/golden/prefabs.npz, written by
tests/golden/make_fixtures.py from the reference's .pf/.v assets)."""
import hashlib
import os

import numpy as np
import pytest

from helpers import IDENTITY

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "prefabs.npz")


def load(name):
    z = np.load(GOLDEN)
    mi = z[f"{name}_inst_model"]
    models = [z[f"{name}_model{i}"] for i in range(int(mi.max()) + 1)]
    frames = []
    for f in range(2):
        frames.append(dict(sha=z[f"{name}_frame{f}_sha256"].tobytes().hex(),
                           popcount=int(z[f"{name}_frame{f}_popcount"]),
                           regions=[tuple(int(v) for v in r) for r in z[f"{name}_frame{f}_regions"]]))
    return (models, mi, z[f"{name}_inst_matrix"], z[f"{name}_inst_moved"], z[f"{name}_inst_pivot"], frames,
            tuple(int(v) for v in z["volume_dims"]))


def frame_pairs(mat, moved, frame):
    if frame == 0:
        return [(IDENTITY, m) for m in mat]
    return list(zip(mat, moved))


@pytest.mark.parametrize("name", ["ModernHouse", "FarmHouse"])
def test_oracle_reproduces_prefab_fixture(oracle, name):
    models, mi, mat, moved, piv, frames, dims = load(name)
    vol = np.zeros(dims[::-1], np.uint8)
    for f in range(2):
        insts = [oracle.make_instance(models[m], p, c, pv) for m, (p, c), pv in
                 zip(mi, frame_pairs(mat, moved, f), piv)]
        regs = oracle.rebuild(vol, insts)
        assert hashlib.sha256(vol.tobytes()).hexdigest() == frames[f]["sha"]
        assert int(np.unpackbits(vol).sum()) == frames[f]["popcount"]
        assert [tuple(r) for r in regs.tolist()] == frames[f]["regions"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ModernHouse", "FarmHouse"])
def test_gpu_rebuild_reproduces_prefab_fixture(ctx, name):
    import torch
    from voxelengine_amd import api
    models, mi, mat, moved, piv, frames, dims = load(name)
    dev = [torch.from_numpy(np.ascontiguousarray(m).reshape(-1)).cuda() for m in models]
    vol = api.Volume(ctx, *dims, nlevels=2)
    vol.clear()
    for f in range(2):
        insts = [api.make_instance(dev[m].data_ptr(), models[m].shape[::-1], p, c, pv) for m, (p, c), pv in
                 zip(mi, frame_pairs(mat, moved, f), piv)]
        regs = vol.rebuild(insts)
        host = vol.download()
        assert hashlib.sha256(host.tobytes()).hexdigest() == frames[f]["sha"]
        assert regs == frames[f]["regions"]
        vol.build_levels(regs)
