"""Pins for the oracle's rebuild (ShadowVoxSystem.cpp:7-201) and mips (VoxAsset.cpp:16-53)."""
import numpy as np
import pytest

import np_ref
from helpers import IDENTITY, assert_bitwise

f32 = np.float32


def test_set_volume_at_bit_formula(oracle):
    S = (3, 4, 5)
    for x in range(-1, 2 * S[0] + 1):
        for y in (0, 1, 2 * S[1] - 1, 2 * S[1]):
            for z in (-1, 0, 3, 2 * S[2] - 1):
                vol = np.zeros(S[::-1], np.uint8)
                oracle.set_volume_at(vol, x, y, z, 1)
                inside = 0 <= x < 2 * S[0] and 0 <= y < 2 * S[1] and 0 <= z < 2 * S[2]
                if not inside:
                    assert not vol.any()
                    continue
                assert vol.sum() == 1 << ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2))
                assert vol[z // 2, y // 2, x // 2] != 0
                oracle.set_volume_at(vol, x, y, z, 0)
                assert not vol.any()


def test_translate_order(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        m = rng.normal(size=16).astype(np.float32) * 10
        v = rng.normal(size=3).astype(np.float32)
        got = oracle.translate(m, v)
        assert_bitwise(got, np_ref.translate(m, v), "translate")
        ref = (m.reshape(4, 4).astype(np.float64).T @ np.r_[v.astype(np.float64), 1.0])
        assert np.allclose(got[12:], ref, rtol=1e-5, atol=1e-4)
    assert_bitwise(oracle.translate(IDENTITY, np.zeros(3, np.float32)), IDENTITY, "identity")


def test_identity_rebuild_sets_exactly_the_occupied_voxels(oracle):
    """With an identity matrix ivec3(x*0.1f*10f) == x for x < 4096, so the bit volume
    is the material volume thresholded at 16, packed 2x2x2."""
    rng = np.random.default_rng(4)
    m = (rng.integers(0, 60, (24, 20, 28))).astype(np.uint8)
    vol = np.zeros((12, 10, 14), np.uint8)
    regs = oracle.rebuild(vol, [oracle.make_instance(m, IDENTITY, IDENTITY)])
    occ = m >= 16
    packed = np.zeros_like(vol)
    for b in range(8):
        dz, dy, dx = (b >> 2) & 1, (b >> 1) & 1, b & 1
        packed |= (occ[dz::2, dy::2, dx::2].astype(np.uint8) << b)
    assert_bitwise(vol, packed, "packed")
    # AABB over fine coords [0, 27]x[0, 19]x[0, 23] -> texels /2, clamped
    assert regs.tolist() == [[0, 0, 0, 14, 10, 12]]
    x = np.arange(4096, dtype=np.float32)
    assert (np.trunc((x * f32(0.1)).astype(np.float32) * f32(10.0)) == x).all()


def _col(m):
    return np.ascontiguousarray(np.asarray(m, np.float64).T.reshape(16), dtype=np.float32)


def _random_affine(rng):
    a = rng.uniform(-np.pi, np.pi, 3)
    cx, sx, cy, sy, cz, sz = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1]), np.cos(a[2]), np.sin(a[2])
    R = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]]) @ np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]]) @ \
        np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(rng.uniform(0.6, 1.6, 3))
    M[:3, 3] = rng.uniform(-1.0, 7.0, 3)
    return _col(M)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rebuild_matches_numpy_restatement(oracle, seed):
    """Random affine instances with pivots, overlapping footprints, erase-before-write
    per instance in order, destroy operations and an invalid model, over two frames."""
    rng = np.random.default_rng(seed)
    S = (40, 32, 44)
    vol_o = np.zeros(S[::-1], np.uint8)
    vol_n = np.zeros(S[::-1], np.uint8)
    models = [((rng.random(tuple(int(4 * rng.integers(1, 8)) for _ in range(3))) < 0.5) *
               rng.integers(0, 64, 1)).astype(np.uint8) + (rng.integers(0, 40, 1)).astype(np.uint8)
              for _ in range(6)]
    insts = []
    for k, m in enumerate(models):
        insts.append(dict(model=m, prev=_random_affine(rng), cur=_random_affine(rng),
                          pivot=rng.uniform(-2, 2, 3).astype(np.float32), flags=2 if k == 4 else 1))
    insts.append(dict(model=None, prev=IDENTITY, cur=IDENTITY, pivot=np.zeros(3, np.float32), flags=1))
    for frame in range(2):
        oi = [oracle.make_instance(i["model"], i["prev"], i["cur"], i["pivot"], i["flags"]) for i in insts]
        ro = oracle.rebuild(vol_o, oi)
        rn = np_ref.rebuild(vol_n, insts)
        assert_bitwise(vol_o, vol_n, f"frame {frame}")
        assert [tuple(r) for r in ro.tolist()] == rn
        for i in insts:
            i["prev"], i["cur"] = i["cur"], _random_affine(rng)
    assert np.unpackbits(vol_o).sum() > 100


def test_rebuild_erase_before_write_ordering(oracle):
    """Two overlapping instances: B's erase of its previous footprint clears bits that
    A wrote in the same frame, because instances run in order (:119-191)."""
    m = np.full((4, 4, 4), 20, np.uint8)
    vol = np.zeros((8, 8, 8), np.uint8)
    T = np.eye(4)
    T[:3, 3] = 0.4          # 4 fine voxels
    a = oracle.make_instance(m, IDENTITY, IDENTITY)
    b = oracle.make_instance(m, IDENTITY, _col(T))      # previous footprint overlaps A
    oracle.rebuild(vol, [a, b])
    bits = np.zeros((16, 16, 16), bool)
    for z in range(16):
        for y in range(16):
            for x in range(16):
                bits[z, y, x] = (vol[z // 2, y // 2, x // 2] >> ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2))) & 1
    assert not bits[0:4, 0:4, 0:4].any()          # A's footprint erased by B's previous one
    assert bits[4:8, 4:8, 4:8].all()
    vol2 = np.zeros_like(vol)
    oracle.rebuild(vol2, [b, a])                  # reversed order: A rewrites its bits
    assert vol2.sum() > vol.sum()


def test_rebuild_region_quirks(oracle):
    """aabbmin starts at the texel extent - 1 (compared against fine coordinates) and
    the region is skipped when aabbmax stays (0,0,0) (:128-132,181)."""
    m = np.full((4, 4, 4), 20, np.uint8)
    vol = np.zeros((8, 8, 8), np.uint8)
    T = np.eye(4)
    T[:3, 3] = -1.0         # footprint at fine [-10, -7]: fully outside, aabbmax stays 0
    regs = oracle.rebuild(vol, [oracle.make_instance(m, _col(T), _col(T))])
    assert regs.shape[0] == 0 and not vol.any()
    T[:3, 3] = [2.0, 0.0, 0.0]   # fine x in [20, 23]: outside on +x; min stays extent-1 = 7
    regs = oracle.rebuild(vol, [oracle.make_instance(m, _col(T), _col(T))])
    x0 = min(7, 20) // 2
    # extent clamps to the volume: hi = min(23/2, 7) = 7 -> w = 7 - 3 + 1
    assert regs.tolist() == [[x0, 0, 0, 7 - x0 + 1, 2, 2]]


def test_mip_first_nonzero_rule(oracle):
    m = np.zeros((4, 4, 4), np.uint8)
    m[1, 1, 1] = 9          # child 7 of cell 0
    m[0, 1, 0] = 5          # child 2 of cell 0
    assert oracle.build_mip(m)[0, 0, 0] == 5
    m[0, 0, 1] = 3          # child 1 wins over child 2
    assert oracle.build_mip(m)[0, 0, 0] == 3
    m[0, 0, 0] = 200        # child 0
    assert oracle.build_mip(m)[0, 0, 0] == 200
    rng = np.random.default_rng(6)
    for _ in range(5):
        v = ((rng.random((16, 12, 20)) < 0.2) * rng.integers(1, 256, (16, 12, 20))).astype(np.uint8)
        assert_bitwise(oracle.build_mip(v), np_ref.build_mip(v), "mip")
        assert_bitwise(oracle.build_mips(v)[2], np_ref.build_mip(np_ref.build_mip(v)), "mip2")
