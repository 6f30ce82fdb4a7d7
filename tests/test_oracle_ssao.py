"""Pins for the oracle's screen-space occlusion (calculateOcclusion and
screenspaceOcclusion, LightAmbient.frag:59-79,89-109) against a float64
restatement from the reference's text and closed-form cases."""
import numpy as np
import pytest

NEAR, FAR = 0.1, 4096.0


def _frame(W=48, H=32, seed=5):
    """A view of a slanted depth ramp with a step, random world normals."""
    from voxelengine_amd import views
    vd = views.make_view((5.0, 8.0, 20.0), (6.0, 2.0, 0.0), W, H)
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    depth = (0.004 + 0.00002 * xx + 0.00003 * yy + np.where(xx > W // 2, 0.00004, 0.0)).astype(np.float32)
    depth[0, 0] = 1.0                                   # one sky pixel
    nrm = rng.normal(size=(H * W, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    mat = np.full(H * W, 3, np.uint8)
    mat[5] = 0                                          # one pixel without a hit
    return vd, depth.reshape(-1), nrm.astype(np.float32), mat


def _bilinear(img, W, H, u, v):
    """The build's sampler in float64: 8-bit weights, border 0."""
    def wt(x, size):
        x = min(max(x, -2.0), size + 1.0)
        fl = np.floor(x)
        f = np.floor((x - fl) * 256 + 0.5) / 256
        if f >= 1:
            fl, f = fl + 1, 0.0
        return int(fl), f
    x0, fx = wt(u * W - 0.5, W)
    y0, fy = wt(v * H - 0.5, H)

    def t(x, y):
        return img[y * W + x] if 0 <= x < W and 0 <= y < H else 0.0
    return ((t(x0, y0) * (1 - fx) + t(x0 + 1, y0) * fx) * (1 - fy) +
            (t(x0, y0 + 1) * (1 - fx) + t(x0 + 1, y0 + 1) * fx) * fy)


def _ssao64(oracle, vd, nrm, depth, seed, i, margin):
    """float64 calculateOcclusion of pixel i; None when a sample lies within
    `margin` of a branch (uv edge or depth window)."""
    W, H = vd["width"], vd["height"]
    px, py = i % W, i // W
    Vm = np.asarray(vd["view"], np.float64).reshape(4, 4).T
    P = np.asarray(vd["proj"], np.float64).reshape(4, 4).T
    IP = np.asarray(vd["inv_proj"], np.float64).reshape(4, 4).T
    N = (Vm @ np.r_[nrm[i], 0.0])[:3]
    t0 = np.array([0.0, -N[2], N[1]]) if abs(N[2]) > 0.5 else np.array([-N[1], N[0], 0.0])
    T = t0 / np.linalg.norm(t0)
    B = np.cross(N, T)
    B /= np.linalg.norm(B)
    u, v = (px + 0.5) / W, (py + 0.5) / H
    f4 = IP @ np.array([2 * u - 1, 1 - 2 * v, 1.0, 1.0])
    fv = f4[:3] / f4[3]
    pos = fv * (depth[i] * (1 + 1 / FAR) + NEAR / FAR)
    pk = oracle.pixkey(seed ^ 0x632be5ab, vd["frame"], i)
    occ = 0.0
    for s in range(4):
        a, b, dist = (float(oracle.rand(pk, 4 * s + k)) for k in range(3))
        rv = np.array([np.sqrt(a) * np.cos(6.283 * b), np.sqrt(a) * np.sin(6.283 * b), np.sqrt(max(0, 1 - a))])
        d = T * rv[0] + B * rv[1] + N * rv[2]
        mid = pos + d / np.linalg.norm(d) * 0.2 * dist
        mp = P @ np.r_[mid, 1.0]
        sd = (mp[3] - NEAR) / (FAR - NEAR)
        su, sv = (mp[0] / mp[3]) * 0.5 + 0.5, (-mp[1] / mp[3]) * 0.5 + 0.5
        if min(abs(su), abs(su - 1), abs(sv), abs(sv - 1)) < margin:
            return None
        if not (0 <= su <= 1 and 0 <= sv <= 1):
            continue
        mn = _bilinear(depth.astype(np.float64), W, H, su, sv)
        mx = mn + 0.2 / FAR
        if min(abs(sd - mn), abs(sd - mx)) < 1e-9:
            return None
        if mn < sd < mx:
            tt = min(max((sd - mx) / (mn - mx), 0.0), 1.0)
            occ += tt * tt * (3 - 2 * tt) * dist
    return min(max(1 - occ / 4 * 3, 0.0), 1.0)


def test_ssao_float64_restatement(oracle):
    vd, depth, nrm, mat = _frame()
    got = oracle.ssao(vd, nrm, mat, depth, seed=21)
    assert got[0] == 1.0 and got[5] == 1.0           # sky / no hit
    checked, occluded = 0, 0
    for i in range(6, depth.size, 3):
        want = _ssao64(oracle, vd, nrm, depth, 21, i, 1e-6)
        if want is None:
            continue
        assert abs(got[i] - want) < 2e-4, (i, got[i], want)
        checked += 1
        occluded += want < 1.0
    assert checked > 300 and occluded > 50


def test_ssao_sky_texture_occludes_nothing(oracle):
    """Depth 1 around every pixel (the sky behind): no sample finds a surface
    in front of it within the window, so the occlusion is 1 everywhere."""
    vd, depth, nrm, mat = _frame()
    far = np.ones_like(depth)
    far[::7] = 0.004                                  # the pixels' own depths only at a few places
    got = oracle.ssao(vd, nrm, mat, np.where(np.arange(depth.size) % 7 == 0, 0.004, 1.0).astype(np.float32))
    assert ((got == 1.0) | (np.arange(depth.size) % 7 == 0)).all()


def test_ssao_range_and_seed(oracle):
    vd, depth, nrm, mat = _frame()
    a = oracle.ssao(vd, nrm, mat, depth, seed=1)
    b = oracle.ssao(vd, nrm, mat, depth, seed=2)
    assert (a >= 0).all() and (a <= 1).all() and (a < 1).any()
    assert not np.array_equal(a, b)
