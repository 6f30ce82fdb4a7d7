"""Pins for the oracle's traversal (Light.frag:14-217): closed forms, exact-rational
brute force, invariants, special cases, and an independent numpy restatement."""
from fractions import Fraction

import numpy as np
import pytest

import np_ref
from helpers import assert_bitwise, edge_rays, random_rays

f32 = np.float32


def random_occ(rng, S, p=0.08):
    return ((rng.random((S, S, S)) < p) * rng.integers(1, 256, (S, S, S))).astype(np.uint8)


def set_fine(vol, x, y, z):
    vol[z // 2, y // 2, x // 2] |= np.uint8(1 << ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)))


# ------------------------------------------------------------------ getVolumeAt
def test_get_volume_at_bit_layout(oracle):
    vol = np.zeros((4, 4, 4), np.uint8)
    O = oracle.occ(vol)
    for (x, y, z) in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (5, 3, 6), (7, 7, 7)]:
        vol[:] = 0
        set_fine(vol, x, y, z)
        for (a, b, c) in [(x, y, z), (x ^ 1, y, z), (x, y ^ 1, z), (x, y, z ^ 1)]:
            want = (a, b, c) == (x, y, z)
            assert oracle.lib().vxo_get_volume_at(O, a, b, c, 0) == want
            assert oracle.lib().vxo_get_volume_at(O, a, b, c, 1) == 1   # odd mip: byte != 0
    # integer division truncates: fine x = -1 reads texel 0 with bit (-1 & 1) = 1
    vol[:] = 0
    set_fine(vol, 1, 0, 0)
    assert oracle.lib().vxo_get_volume_at(O, -1, 0, 0, 0) == 1
    # texels outside the volume read 0
    vol[:] = 0xFF
    assert oracle.lib().vxo_get_volume_at(O, 8, 0, 0, 1) == 0
    assert oracle.lib().vxo_get_volume_at(O, -2, 0, 0, 1) == 0


# ------------------------------------------------------------------ DDA closed forms
def test_dda_row_closed_form(oracle):
    """Nearly axis-aligned rays through an empty row: the first occupied fine voxel at
    x = k is hit after k - x0 steps, best_t is its exit time (k + 1 - o.x) and the
    normal is the exit face -sign(d)."""
    S = 32
    vol = np.zeros((S, S, S), np.uint8)
    set_fine(vol, 40, 21, 33)
    o = (10.5, 21.5, 33.5)
    for dx in (1.0, -1.0):
        d = (dx, 1e-6, -1e-6)
        org = o if dx > 0 else (60.5, 21.5, 33.5)
        r = oracle.dda(vol, org, d, 1000.0)
        assert r.hit == 1 and r.term == 1
        assert tuple(r.voxel) == (40, 21, 33)
        steps = abs(40 - int(org[0]))
        assert r.steps == steps and r.fetches == steps + 1
        exit_t = (41.0 - org[0]) if dx > 0 else (org[0] - 40.0)
        assert r.t == f32(exit_t)
        assert tuple(r.normal) == (-dx, -0.0, 0.0)


def test_dda_restart_counts_four_passes(oracle):
    """An empty region: each pass stops once totalt >= maxt (or after 256 steps) and
    the outer do/while repeats it four times (Light.frag:75,79)."""
    S = 64
    vol = np.zeros((S, S, S), np.uint8)
    r = oracle.dda(vol, (10.5, 50.5, 50.5), (1.0, 1e-7, 1e-7), 10.0)
    # exit times 0.5, 1.5, ...; the pass continues while totalt < 10 -> 11 steps
    assert (r.hit, r.term, r.steps, r.fetches) == (0, 0, 44, 44)
    r = oracle.dda(vol, (0.5, 50.5, 50.5), (1.0, 1e-7, 1e-7), 1e6)
    # fetches fine x = 0..129 (the bound 2*S+1 is inclusive), then x = 130 is out of
    # bounds: returns without restart
    assert (r.hit, r.term, r.steps, r.fetches) == (0, 2, 130, 130)
    # no ties (x and y would step together): 256 steps stay inside the volume
    r = oracle.dda(vol, (0.5, 0.5, 0.5), (0.48, 0.52, 0.7059745), 1e6)
    assert r.term == 0 and r.steps == 4 * 256 and r.fetches == 4 * 256


def test_dda_out_of_volume_origin_returns_immediately(oracle):
    vol = np.zeros((8, 8, 8), np.uint8)
    # the fine-voxel bound is [0, 2*S+1] = [0, 17]
    for o in [(-0.5, 2.0, 2.0), (2.0, 18.5, 2.0), (2.0, 2.0, 1e9)]:
        r = oracle.dda(vol, o, (0.3, 0.4, 0.5), 100.0)
        assert (r.hit, r.term, r.steps, r.fetches) == (0, 2, 0, 0)


def test_dda_origin_voxel_is_tested(oracle):
    vol = np.zeros((8, 8, 8), np.uint8)
    set_fine(vol, 3, 4, 5)
    r = oracle.dda(vol, (3.5, 4.25, 5.75), (0.3, 0.4, 0.5), 100.0)
    assert r.hit == 1 and r.steps == 0 and tuple(r.voxel) == (3, 4, 5)


def test_dda_zero_component_quirk(oracle):
    """A +0 direction component z gives sign 0, next bound cv.z + 0.5 and
    t_max.z = (cv.z + 0.5 - o.z) / +0.
    o.z frac 0.25: t_max.z = +inf, select.z = 0, best_t = inf*0 = NaN, so the pass
    ends after one step (totalt = NaN): 4 passes x 1 step.
    o.z frac 0.75: t_max.z = -inf is selected first, best_t = -inf, no move, then
    t_max.z = -inf + inf = NaN; the second step has best_t = NaN: 4 passes x 2 steps."""
    vol = np.zeros((16, 16, 16), np.uint8)
    r = oracle.dda(vol, (4.5, 4.25, 4.25), (0.6, 0.8, 0.0), 100.0)
    assert (r.hit, r.term, r.steps) == (0, 0, 4)
    r = oracle.dda(vol, (4.5, 4.25, 4.75), (0.6, 0.8, 0.0), 100.0)
    assert (r.hit, r.term, r.steps) == (0, 0, 8)


def test_dda_exact_diagonal_tie(oracle):
    """Equal t_max on all axes: select = (1,1,1), the step is diagonal and best_t is
    the sum of the tied values ((t+t)+t)."""
    vol = np.zeros((16, 16, 16), np.uint8)
    set_fine(vol, 9, 9, 9)
    d = np.full(3, f32(1.0) / np.sqrt(f32(3.0)), np.float32)
    r = oracle.dda(vol, (5.5, 5.5, 5.5), d, 100.0)
    assert r.hit == 1 and tuple(r.voxel) == (9, 9, 9) and r.steps == 4
    td = f32(1.0) / d[0]
    tm = f32(f32(6.0) - f32(5.5)) / d[0]
    for _ in range(4):
        tm = f32(tm + td)
    assert r.t == f32(f32(tm + tm) + tm)
    assert tuple(r.normal) == (-1.0, -1.0, -1.0)


def _exact_first_hit(vol, o, d, maxt):
    """Exact-rational Amanatides-Woo over fine voxels (origin voxel included), for a
    ray whose crossing times have no ties. Returns (voxel, steps) or None."""
    S = vol.shape[0]
    o = [Fraction(float(v)) for v in o]
    d = [Fraction(float(v)) for v in d]
    cv = [int(np.floor(float(v))) for v in o]
    step = [1 if v > 0 else -1 for v in d]
    tmax = [((cv[k] + (1 if step[k] > 0 else 0)) - o[k]) / d[k] for k in range(3)]
    tdel = [abs(1 / d[k]) for k in range(3)]
    total = Fraction(0)
    for n in range(256):
        if any(c < 0 or c > 2 * S + 1 for c in cv):
            return None
        x, y, z = cv
        if x < 2 * S and y < 2 * S and z < 2 * S:
            if (vol[z // 2, y // 2, x // 2] >> ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2))) & 1:
                return tuple(cv), n
        k = int(np.argmin([float(t) for t in tmax]))
        total = tmax[k]
        cv[k] += step[k]
        tmax[k] += tdel[k]
        if total >= maxt:
            return None
    return None


def test_dda_matches_exact_rational_traversal(oracle):
    """On rays whose exact crossing times are separated by more than 1e-3 (so fp32
    rounding cannot reorder them), the oracle visits the same voxels as exact
    arithmetic: same hit voxel and step count."""
    rng = np.random.default_rng(11)
    S = 24
    vol = random_occ(rng, S, 0.02)
    rays = random_rays(rng, 400, 2.0, 2 * S - 2.0, maxt=(40.0, 80.0))
    checked = 0
    for ry in rays:
        o, d, maxt = ry[:3], ry[3:6], float(ry[6])
        times = []
        for k in range(3):
            c = np.floor(o[k])
            first = ((c + (1 if d[k] > 0 else 0)) - o[k]) / d[k]
            times += [first + j / abs(d[k]) for j in range(120)]
        times = np.sort(np.array(times))
        times = times[times < 90]
        if np.diff(times).min() < 1e-3:
            continue
        exact = _exact_first_hit(vol, o, d, maxt)
        r = oracle.dda(vol, o, d, maxt)
        if exact is None:
            assert r.hit == 0
        else:
            assert r.hit == 1 and tuple(r.voxel) == exact[0] and r.steps == exact[1], (ry, exact, r.steps)
        checked += 1
    assert checked > 150


def test_dda_hit_invariants(oracle):
    rng = np.random.default_rng(12)
    S = 32
    vol = random_occ(rng, S, 0.05)
    rays = random_rays(rng, 3000, -4.0, 2 * S + 4.0)
    res = oracle.trace_rays(vol, rays, oracle.MARCH_DDA)
    for ry, r in zip(rays, res):
        if not r["hit"]:
            continue
        x, y, z = r["voxel"]
        assert (vol[z // 2, y // 2, x // 2] >> ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2))) & 1
        hp = (ry[:3] + ry[3:6] * r["t"]).astype(np.float32)
        assert_bitwise(r["hitpos"], hp, "hitpos")
        n = r["normal"]
        assert ((n == 0) | (n == -np.sign(ry[3:6]))).all() and np.abs(n).sum() >= 1
        assert r["fetches"] == r["steps"] + 1


@pytest.mark.parametrize("S,p", [(16, 0.05), (32, 0.1), (40, 0.02)])
def test_dda_matches_numpy_restatement(oracle, S, p):
    rng = np.random.default_rng(S)
    vol = random_occ(rng, S, p)
    rays = np.concatenate([random_rays(rng, 3000, -5.0, 2 * S + 5.0), edge_rays(S)])
    ref = np_ref.dda(vol, rays)
    got = oracle.trace_rays(vol, rays, oracle.MARCH_DDA)
    for f in ("hit", "term", "steps", "fetches"):
        assert_bitwise(got[f], ref[f], f)
    h = ref["hit"] == 1
    for f in ("t", "voxel", "hitpos", "normal"):
        assert_bitwise(got[f][h], ref[f][h], f)


# ------------------------------------------------------------------ sparse marchers
@pytest.mark.parametrize("super_,dist,want", [(False, 128.0, 31 + 112), (True, 128.0, 6 + 23),
                                              (False, 16.0, 31), (False, 20.0, 35), (False, 1000.0, 31 + 148),
                                              (True, 16.0, 6), (True, 32.0, 6 + 3), (False, 0.25, 31)])
def test_sparse_iteration_counts(oracle, super_, dist, want):
    vol = np.zeros((8, 8, 8), np.uint8)
    r = oracle.sparse(vol, (3.0, 3.0, 3.0), (0.3, 0.4, 0.5), dist, super_)
    assert r.steps == want and r.hit == 0 and r.t == f32(dist)


def test_sparse_sub_voxel_bit_uses_mod_half(oracle):
    """Phase 1 picks the bit from mod(pos, 0.5) > 0.25 (Light.frag:144-146), not from
    the fine-voxel parity used by SetVolumeAt."""
    vol = np.zeros((8, 8, 8), np.uint8)
    # pos (2.3, 4.1, 6.4): texel (1, 2, 3); mod bits x: 0.3>0.25 -> 1, y: 0.1 -> 0, z: 0.4 -> 4
    vol[3, 2, 1] = 1 << 5
    r = oracle.sparse(vol, (2.3, 4.1, 6.4), (0.0, 0.0, 1.0), 128.0)
    assert r.hit == 1 and r.t == f32(0.5) and r.steps == 1
    vol[3, 2, 1] = 1 << 0      # the parity bit of fine voxel (2, 4, 6)
    r = oracle.sparse(vol, (2.3, 4.1, 6.4), (1e-3, 0.0, 0.0), 15.0)
    assert r.hit == 0


def test_sparse_truncates_toward_zero(oracle):
    """ivec3(pos/2) truncates, so pos.x in (-2, 0) reads texel x = 0."""
    vol = np.zeros((4, 4, 4), np.uint8)
    vol[0, 0, 0] = 0xFF
    r = oracle.sparse(vol, (-1.6, 0.2, 0.2), (-1.0, 0.0, 0.0), 128.0)
    assert r.hit == 1 and r.t == f32(0.5)


def test_sparse_phase_two_is_byte_test(oracle):
    vol = np.zeros((64, 64, 64), np.uint8)
    vol[5, 5, 20] = 0x01     # texel (20, 5, 5): fine x 40..41
    r = oracle.sparse(vol, (0.25, 10.5, 10.5), (1.0, 0.0, 0.0), 128.0)
    # phase 1 ends at d = 16 (pos.x 15.75); phase 2 samples pos.x = 15.75 + k, d = 16 + k;
    # the byte of texel int(pos.x)/2 = 20 is first non-zero at int(pos.x) = 40 -> k = 25
    assert r.hit == 1 and r.t == f32(16.0 + 25.0) and r.steps == 31 + 26


@pytest.mark.parametrize("super_", [False, True])
def test_sparse_matches_numpy_restatement(oracle, super_):
    rng = np.random.default_rng(5 + super_)
    S = 24
    vol = random_occ(rng, S, 0.04)
    rays = np.concatenate([random_rays(rng, 3000, -10.0, 2 * S + 10.0, maxt=(1.0, 300.0)), edge_rays(S)])
    march = oracle.MARCH_SUPERSPARSE if super_ else oracle.MARCH_SPARSE
    got = oracle.trace_rays(vol, rays, march)
    d, it, hit = np_ref.sparse(vol, rays, 2.5 if super_ else 0.5)
    assert_bitwise(got["t"], d, "d")
    assert_bitwise(got["steps"], it, "iterations")
    assert_bitwise(got["hit"], hit, "hit")
