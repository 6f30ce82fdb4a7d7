"""Shared test helpers: scenes built by the product on the device and mirrored
for the oracle, bitwise comparisons, ray sets."""
from __future__ import annotations

import numpy as np

IDENTITY = np.eye(4, dtype=np.float32).reshape(16)


def bits(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    if a.dtype == np.float64:
        return a.view(np.uint64)
    return a


def assert_bitwise(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ba, bb = bits(a), bits(b)
    bad = np.nonzero(ba.reshape(-1) != bb.reshape(-1))[0]
    assert bad.size == 0, f"{what}: {bad.size} of {ba.size} differ; first at {bad[:5]}: " \
                          f"{a.reshape(-1)[bad[:5]]} vs {b.reshape(-1)[bad[:5]]}"


def compare_ray_results(got, ref, what=""):
    for f in ("hit", "term", "steps", "fetches", "voxel", "t", "hitpos", "normal"):
        assert_bitwise(got[f], ref[f], f"{what}.{f}")


def random_rays(rng, n, lo, hi, maxt=(1.0, 300.0)):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = rng.uniform(*maxt, size=(n, 1))
    return np.concatenate([o, d.astype(np.float32), t.astype(np.float32)], axis=1).astype(np.float32)


def edge_rays(S):
    """Rays that exercise the reference's special cases: axis-aligned and signed-zero
    directions, integer (voxel-boundary) origins, exact diagonal ties, tiny and
    non-finite components, out-of-volume origins, small and huge maxt."""
    F = 2 * S
    rays = []
    c = F / 2.0
    for ax in range(3):
        for sgn in (1.0, -1.0):
            d = [0.0, 0.0, 0.0]
            d[ax] = sgn
            rays.append([c + 0.25, c + 0.5, c + 0.75] + d + [200.0])
            rays.append([c, c, c] + d + [200.0])
            d2 = list(d)
            d2[(ax + 1) % 3] = -0.0
            rays.append([c + 0.3, c + 0.6, c + 0.1] + d2 + [100.0])
    for dd in ([1, 1, 1], [1, -1, 1], [-1, -1, -1], [1, 1, 0], [0, 1, -1]):
        v = np.array(dd, np.float64)
        v = v / np.linalg.norm(v)
        rays.append([c, c, c] + list(v) + [300.0])
        rays.append([c + 0.5, c + 0.5, c + 0.5] + list(v) + [300.0])
        rays.append([1.0, 2.0, 3.0] + list(v) + [300.0])
    for tiny in (1e-30, 1e-40, -1e-40, 1e-45):
        rays.append([c + 0.1, c + 0.2, c + 0.3, 0.6, 0.8, tiny, 400.0])
    rays.append([c, c, c, float("inf"), 1.0, 0.5, 100.0])
    rays.append([c, c, c, float("nan"), 0.5, 0.5, 100.0])
    rays.append([-5.0, c, c, 1.0, 0.0, 0.0, 500.0])
    rays.append([-5.0, -5.0, -5.0, 0.5, 0.6, 0.62, 1000.0])
    rays.append([F + 3.0, c, c, -1.0, 0.01, 0.02, 1000.0])
    rays.append([F + 1.5, F + 1.5, F + 1.5, 1.0, 1.0, 1.0, 10.0])
    rays.append([c + 0.1, c + 0.2, c + 0.3, 0.3, 0.4, 0.5, 0.5])
    rays.append([c + 0.1, c + 0.2, c + 0.3, 0.3, 0.4, 0.5, 1e9])
    rays.append([c + 0.1, c + 0.2, c + 0.3, 0.3, 0.4, 0.5, 0.0])
    rays.append([c + 0.1, c + 0.2, c + 0.3, 0.0, 0.0, 0.0, 50.0])
    rays.append([1e9, c, c, -1.0, 0.1, 0.1, 50.0])
    return np.array(rays, dtype=np.float32)


def device_scene(api, ctx, N, seed=5, fill=0.3, nboxes=30, box=(2, 16), levels=2, dims=None):
    """Generate a scene on the device, build mips, rebuild the occupancy volume and
    levels. Returns (model, volume, host material array)."""
    nx, ny, nz = dims if dims else (N, N, N)
    model = api.Model(ctx, nx, ny, nz)
    model.generate(seed=seed, terrain_fill=fill, nboxes=nboxes, box_min=box[0], box_max=box[1])
    model.build_mips()
    vol = model.download(0)
    d = model.dims
    occ = api.Volume(ctx, d[0] // 2, d[1] // 2, d[2] // 2, nlevels=levels)
    occ.clear()
    occ.rebuild([api.make_instance(model.mip_ptr(0), d, IDENTITY, IDENTITY)])
    occ.build_levels()
    return model, occ, vol


def oracle_occ(O, vol):
    nz, ny, nx = vol.shape
    occ = np.zeros((nz // 2, ny // 2, nx // 2), np.uint8)
    O.rebuild(occ, [O.make_instance(vol, IDENTITY, IDENTITY)])
    return occ
