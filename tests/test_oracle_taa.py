"""Pins for the oracle's LightTAA (LightTAA.frag:37-141): the sampler, the sky
pass-through, and closed forms of the spiral / clip / variance / blend on
uniform fields, the motion term, the off-screen branch and the same-surface
test at a material edge."""
import numpy as np
import pytest

N = 48


def _pal():
    pal = np.zeros((256, 6), np.float32)
    pal[1] = [0.8, 0.6, 0.4, 0.5, 0.0, 0.0]
    pal[2] = [0.2, 0.3, 0.9, 0.9, 1.0, 0.0]
    return pal


def _frame(L=(0.5, 0.25, 0.125), last=(0.5, 0.25, 0.125), var=0.3, mat=1, depth=0.01, motion=(0.0, 0.0)):
    npx = N * N
    light = np.tile(np.float32(L), (npx, 1))
    lst = np.tile(np.float32(list(last) + [var]), (npx, 1))
    mo = np.tile(np.float32(motion), (npx, 1))
    dep = np.full(npx, depth, np.float32)
    nrm = np.tile(np.float32([0.0, 1.0, 0.0]), (npx, 1))
    m = np.full(npx, mat, np.uint8)
    return light, lst, mo, dep, nrm, m


def _inner(a, lo=16, hi=32):
    return a.reshape(N, N, -1)[lo:hi, lo:hi].reshape(-1, a.shape[-1])


def test_sampler_centres_midpoints_weights_and_border(oracle):
    rng = np.random.default_rng(0)
    img = rng.random((5, 7, 4)).astype(np.float32)
    h, w = 5, 7
    for y in range(h):
        for x in range(w):
            np.testing.assert_array_equal(oracle.taa_sample(img, (x + 0.5) / w, (y + 0.5) / h), img[y, x])
    got = oracle.taa_sample(img, 3.0 / w, 2.5 / h)                 # halfway between x = 2 and 3
    np.testing.assert_array_equal(got, img[2, 2] * np.float32(0.5) + img[2, 3] * np.float32(0.5))
    got = oracle.taa_sample(img, 2.8 / w, 2.5 / h)                 # fraction 0.3 -> 77/256
    f = np.float32(77 / 256)
    np.testing.assert_allclose(got, img[2, 2] * (1 - f) + img[2, 3] * f, rtol=1e-6)
    np.testing.assert_array_equal(oracle.taa_sample(img, -0.5 / w, 0.5 / h), [0, 0, 0, 1])   # border
    np.testing.assert_array_equal(oracle.taa_sample(img, (w + 0.5) / w, 2.5 / h), [0, 0, 0, 1])


def test_sky_pixels_pass_through(oracle):
    light, lst, mo, dep, nrm, m = _frame()
    dep[:] = 1.0
    m[:] = 0
    out = oracle.light_taa(N, N, 0, 1, light, lst, mo, dep, nrm, m, _pal())
    np.testing.assert_array_equal(out[:, :3], light)
    assert np.all(out[:, 3] == 1.0)                                # the sky's light alpha


@pytest.mark.parametrize("var", [0.0, 0.3, 0.9])
def test_uniform_field_closed_form(oracle, var):
    """Every neighbour counts with weight 1.212 - r/12: the spatial mean is
    the light itself, the clip is a no-op, the variance term is 0 and the
    running variance drops by diffSum / 13 * 0.08 with diffSum = 1."""
    L = np.float32([0.5, 0.25, 0.125])
    light, lst, mo, dep, nrm, m = _frame(L=L, var=var)
    out = oracle.light_taa(N, N, 3, 7, light, lst, mo, dep, nrm, m, _pal())
    inner = _inner(out)
    np.testing.assert_allclose(inner[:, :3], np.tile(L, (len(inner), 1)), rtol=2e-6)
    want_a = np.clip((var - 0.08 / 13.0) * 0.7, 0.0, 1.0)
    np.testing.assert_allclose(inner[:, 3], want_a, rtol=1e-5, atol=1e-7)


def test_last_light_is_clipped_to_the_neighbourhood(oracle):
    L = np.float32([0.5, 0.25, 0.125])
    light, lst, mo, dep, nrm, m = _frame(L=L, last=(3.0, -1.0, 0.125), var=0.3)
    out = oracle.light_taa(N, N, 0, 7, light, lst, mo, dep, nrm, m, _pal())
    np.testing.assert_allclose(_inner(out)[:, :3], np.tile(L, (16 * 16, 1)), rtol=2e-6)


def test_motion_raises_the_variance(oracle):
    mo_v = (0.01, -0.005)
    light, lst, mo, dep, nrm, m = _frame(var=0.3, motion=mo_v)
    out = oracle.light_taa(N, N, 0, 7, light, lst, mo, dep, nrm, m, _pal())
    want = np.clip((0.3 - 0.08 / 13.0 + np.hypot(*mo_v) * 20.0) * 0.7, 0.0, 1.0)
    np.testing.assert_allclose(_inner(out)[:, 3], want, rtol=1e-5)


def test_off_screen_history_uses_the_19x19_neighbourhood(oracle):
    L = np.float32([0.5, 0.25, 0.125])
    light, lst, mo, dep, nrm, m = _frame(L=L, motion=(2.0, 0.0))
    out = oracle.light_taa(N, N, 0, 7, light, lst, mo, dep, nrm, m, _pal())
    inner = _inner(out, 10, 38)
    np.testing.assert_allclose(inner[:, :3], np.tile(L, (len(inner), 1)), rtol=2e-6)
    assert np.all(inner[:, 3] == 1.0)


def test_other_surfaces_do_not_leak_across_a_material_edge(oracle):
    """Left half material 1 with light LA, right half material 2 with LB:
    pixels on the left keep LA even next to the edge (the colour term drops
    every sample touching the other surface); the right keeps LB."""
    LA, LB = np.float32([0.5, 0.25, 0.125]), np.float32([4.0, 0.0, 2.0])
    light, lst, mo, dep, nrm, m = _frame(L=LA, var=0.2)
    img = light.reshape(N, N, 3)
    img[:, N // 2:] = LB
    lst.reshape(N, N, 4)[:, N // 2:, :3] = LB
    m.reshape(N, N)[:, N // 2:] = 2
    out = oracle.light_taa(N, N, 0, 7, light, lst, mo, dep, nrm, m, _pal()).reshape(N, N, 4)
    rows = slice(14, 34)
    np.testing.assert_allclose(out[rows, N // 2 - 3:N // 2, :3], np.broadcast_to(LA, (20, 3, 3)), rtol=2e-6)
    np.testing.assert_allclose(out[rows, N // 2:N // 2 + 3, :3], np.broadcast_to(LB, (20, 3, 3)), rtol=2e-6)
    # the running variance there still follows the same-surface closed form
    np.testing.assert_allclose(out[rows, N // 2 - 3:N // 2, 3], (0.2 - 0.08 / 13.0) * 0.7, rtol=1e-5)


def test_sampler_far_and_non_finite_positions_read_the_border(oracle):
    img = np.ones((5, 7, 4), np.float32)
    for u, v in [(1e30, 0.5), (-1e30, 0.5), (np.inf, 0.5), (-np.inf, 0.5), (np.nan, 0.5), (0.5, np.nan),
                 (0.5, 3e9), (2.0, 0.5)]:
        np.testing.assert_array_equal(oracle.taa_sample(img, np.float32(u), np.float32(v)), [0, 0, 0, 1])
