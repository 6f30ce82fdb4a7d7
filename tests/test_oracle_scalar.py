"""Oracle scalar semantics: GLSL builtin conventions, the build's sin/cos and RNG."""
import math

import numpy as np


def test_f2i_truncates_saturates_and_maps_nan(oracle):
    assert oracle.f2i(2.9) == 2 and oracle.f2i(-2.9) == -2 and oracle.f2i(-0.5) == 0
    assert oracle.f2i(3e9) == 2**31 - 1 and oracle.f2i(-3e9) == -2**31
    assert oracle.f2i(float("nan")) == 0
    assert oracle.f2i(float("inf")) == 2**31 - 1 and oracle.f2i(float("-inf")) == -2**31


def test_sign_of_zero_and_nan(oracle):
    L = oracle.lib()
    assert L.vxo_sign(2.0) == 1.0 and L.vxo_sign(-1e-40) == -1.0
    for z in (0.0, -0.0, float("nan")):
        s = np.float32(L.vxo_sign(z))
        assert s.view(np.uint32) == 0, z    # +0, not -0 or NaN


def test_mod_is_x_minus_y_floor(oracle):
    L = oracle.lib()
    for x, y, want in [(2.3, 0.5, 0.3), (-0.3, 0.5, 0.2), (-2.0, 0.5, 0.0), (5.0, 2.0, 1.0), (-3.0, 2.0, 1.0)]:
        got = np.float32(L.vxo_mod(x, y))
        assert abs(got - want) < 1e-6, (x, y, got)
    rng = np.random.default_rng(0)
    xs = rng.uniform(-1000, 1000, 2000).astype(np.float32)
    m = np.array([L.vxo_mod(float(x), 0.5) for x in xs], np.float32)
    assert (m >= 0).all() and (m < 0.5).all()


def test_sincos_against_libm(oracle):
    xs = np.linspace(0.0, 6.3, 20001, dtype=np.float32)
    err_s = err_c = 0.0
    for x in xs[::7]:
        s, c = oracle.sincos(x)
        err_s = max(err_s, abs(float(s) - math.sin(float(x))))
        err_c = max(err_c, abs(float(c) - math.cos(float(x))))
    assert err_s < 3e-7 and err_c < 3e-7, (err_s, err_c)
    # quadrant boundaries and negative arguments
    for x in (0.0, math.pi / 2, math.pi, 3 * math.pi / 2, 2 * math.pi, -1.0, -4.0, 10.0, 100.0):
        s, c = oracle.sincos(x)
        assert abs(float(s) - math.sin(np.float32(x))) < 1e-6 and abs(float(c) - math.cos(np.float32(x))) < 1e-6


def _lowbias32(x):
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_hash_and_rand(oracle):
    for x in (0, 1, 2, 12345, 0xFFFFFFFF, 0x80000000, 0xDEADBEEF):
        assert oracle.hash32(x) == _lowbias32(x)
    k = oracle.pixkey(7, 0, 123)
    u = np.array([oracle.rand(k, i) for i in range(4000)])
    assert (u >= 0).all() and (u < 1).all()
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    # avalanche: flipping one input bit flips ~half the output bits
    flips = [bin(oracle.hash32(x) ^ oracle.hash32(x ^ (1 << b))).count("1") for x in range(50) for b in range(32)]
    assert 15.0 < np.mean(flips) < 17.0


def test_cosine_hemisphere_is_unit_and_cosine_weighted(oracle):
    import ctypes as C
    L = oracle.lib()
    rng = np.random.default_rng(1)
    zs = []
    for u1, u2 in rng.random((4000, 2)):
        out = (C.c_float * 3)()
        L.vxo_cosine_hemisphere(float(np.float32(u1)), float(np.float32(u2)), out)
        v = np.array(out, np.float64)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-6
        assert v[2] >= 0.0
        zs.append(v[2])
    assert abs(np.mean(zs) - 2.0 / 3.0) < 0.01     # E[cos(theta)] under p = cos/pi
