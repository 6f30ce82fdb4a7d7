"""Real scenes through the loaders into the rebuild, on the device: .v files
read by vx_model_load, prefabs spawned by vx_prefab_load, the volume rebuilt
in the reference's first-frame order and compared with the fixtures and the
oracle."""
import hashlib

import numpy as np
import pytest

from golden_mods import GOLDEN, write_mod_folder
from helpers import IDENTITY

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods(tmp_path_factory):
    return write_mod_folder(str(tmp_path_factory.mktemp("gpu_assets")))


@pytest.fixture(scope="module")
def z():
    return np.load(GOLDEN)


def test_model_load_uploads_and_builds_mips(ctx, mods, z, oracle):
    from voxelengine_amd import api
    P = api.Prefab(mods, "default/FarmHouse.pf")
    models = P.load_models(ctx)
    files = [str(f) for f in z["FarmHouse_model_files"]]
    for fi, m in models.items():
        k = files.index(P.files[fi].rsplit("/", 1)[1])
        ref = z[f"FarmHouse_model{k}"]
        assert m.dims == ref.shape[::-1]
        mips = oracle.build_mips(ref)
        for lvl in range(3):
            np.testing.assert_array_equal(m.download(lvl), mips[lvl])


def _volume_after(ctx, P, models, order, dims):
    from voxelengine_amd import api
    vol = api.Volume(ctx, *dims, nlevels=2)
    vol.clear()
    regs = vol.rebuild(P.instances(models, order=order))
    return vol, regs


@pytest.mark.parametrize("order_kind", ["reference", "file"])
def test_farmhouse_rebuild_from_loaded_prefab(ctx, mods, z, order_kind):
    from voxelengine_amd import api, _lib as L
    dims = tuple(int(v) for v in z["volume_dims"])
    P = api.Prefab(mods, "default/FarmHouse.pf")
    models = P.load_models(ctx)
    if order_kind == "reference":
        order, key = None, "FarmHouse_ref_frame0"
    else:
        order, key = [i for i, e in enumerate(P.entities) if e.flags & L.ENT_VOX], "FarmHouse_frame0"
    vol, regs = _volume_after(ctx, P, models, order, dims)
    host = vol.download()
    assert hashlib.sha256(host.tobytes()).hexdigest() == z[f"{key}_sha256"].tobytes().hex()
    assert int(np.unpackbits(host).sum()) == int(z[f"{key}_popcount"])
    assert regs == [tuple(int(v) for v in r) for r in z[f"{key}_regions"]]


def test_modernhouse_with_nested_player_matches_oracle(ctx, mods, z, oracle):
    """ModernHouse spawns the player_ik prefab (rotated, scaled bones with
    pivots) as a nested instance; the rebuild of the whole spawn in the
    reference order matches the oracle on the same instances."""
    from voxelengine_amd import api
    dims = tuple(int(v) for v in z["volume_dims"])
    P = api.Prefab(mods, "default/ModernHouse.pf")
    models = P.load_models(ctx)
    vol, regs = _volume_after(ctx, P, models, None, dims)
    host_models = {fi: m.download(0) for fi, m in models.items()}
    ref = np.zeros(dims[::-1], np.uint8)
    oi = [oracle.make_instance(host_models[P.entities[i].model], IDENTITY, P.world(i),
                               np.array(P.entities[i].pivot, np.float32))
          for i in P.rebuild_order() if P.entities[i].model >= 0]
    assert len(oi) > 6    # the house's own models plus the player's bones
    rr = oracle.rebuild(ref, oi)
    np.testing.assert_array_equal(vol.download(), ref)
    assert regs == [tuple(r) for r in rr.tolist()]


def test_cpp_shadowvox_spawn_matches_reference_order(mods, z, tmp_path):
    """passes.hpp: PrefabAsset + ShadowVoxSystem::Spawn + OnUpdate (C++) give
    the reference-order FarmHouse volume."""
    import os
    import subprocess
    demo = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "passes_demo")
    if not os.path.exists(demo):
        pytest.skip("tests/cpp/passes_demo is not built (build() builds it)")
    out = tmp_path / "farmhouse.bin"
    r = subprocess.run([demo, mods, "default/FarmHouse.pf", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "67 instances, 1 lights, 67 draws" in r.stdout, r.stdout
    assert hashlib.sha256(out.read_bytes()).hexdigest() == z["FarmHouse_ref_frame0_sha256"].tobytes().hex()
