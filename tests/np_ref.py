"""Independent numpy float32 restatements of the traversal and rebuild.

A second, vectorised implementation written from the reference shaders
(Light.frag:14-217, ShadowVoxSystem.cpp:82-201, VoxAsset.cpp:26-47), used to
cross-check the C oracle. Every numpy float32 operation rounds once, so the
results must agree with the oracle bit for bit.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
HALF, ONE, ZERO = f32(0.5), f32(1.0), f32(0.0)


def f2i(x):
    """ivec3(): truncation toward zero, saturating, NaN -> 0."""
    x = np.asarray(x, np.float32)
    with np.errstate(invalid="ignore"):
        t = np.trunc(np.nan_to_num(x, nan=0.0, posinf=3e9, neginf=-3e9).astype(np.float64))
    return np.clip(t, -2147483648, 2147483647).astype(np.int64)


def gsign(x):
    return np.where(x > 0, ONE, np.where(x < 0, -ONE, ZERO)).astype(np.float32)


def gstep(e, x):
    return np.where(x < e, ZERO, ONE).astype(np.float32)


def texel(vol, t):
    """texelFetch with out-of-range reads returning 0; vol (sz,sy,sx), t (n,3) ints."""
    sz, sy, sx = vol.shape
    ok = (t[:, 0] >= 0) & (t[:, 1] >= 0) & (t[:, 2] >= 0) & (t[:, 0] < sx) & (t[:, 1] < sy) & (t[:, 2] < sz)
    out = np.zeros(t.shape[0], np.uint32)
    tt = t[ok]
    out[ok] = vol[tt[:, 2], tt[:, 1], tt[:, 0]]
    return out


def get_volume_at(vol, p, mip):
    """getVolumeAt (Light.frag:14-27) for integer fine positions p (n,3)."""
    bit = (p[:, 0] & 1) | ((p[:, 1] & 1) << 1) | ((p[:, 2] & 1) << 2)
    q = np.where(p >= 0, p // 2, -((-p) // 2))          # C-style truncating division
    v = texel(vol, q)
    if mip % 2 == 1:
        return v != 0
    return (v & (np.uint32(1) << bit.astype(np.uint32))) != 0


def dda(vol, rays):
    """raycastShadowVolume (Light.frag:29-81) at mip 0 for rays (n,7): o3, d3, maxt."""
    rays = np.asarray(rays, np.float32)
    n = rays.shape[0]
    o, d, maxt = rays[:, 0:3].copy(), rays[:, 3:6].copy(), rays[:, 6].copy()
    sz, sy, sx = vol.shape
    up = np.array([sx, sy, sz], np.int64) * 2 + 1
    with np.errstate(all="ignore"):
        s = gsign(d)
        td = (ONE / (d * s)).astype(np.float32)
        cv = f2i(np.floor(o))
        nb = (cv.astype(np.float32) + (s * HALF + HALF)).astype(np.float32)
        tm = ((nb - o) / d).astype(np.float32)
    active = np.ones(n, bool)
    nt = np.zeros(n, np.int64)
    term = np.zeros(n, np.int64)
    totalt = np.zeros(n, np.float32)
    res = {"hit": np.zeros(n, np.int32), "t": np.zeros(n, np.float32), "voxel": np.zeros((n, 3), np.int32),
           "hitpos": np.zeros((n, 3), np.float32), "normal": np.zeros((n, 3), np.float32)}
    for it in range(256):
        if not active.any():
            break
        with np.errstate(all="ignore"):
            sel = np.stack([gstep(tm[:, 0], tm[:, 2]) * gstep(tm[:, 0], tm[:, 1]),
                            gstep(tm[:, 1], tm[:, 0]) * gstep(tm[:, 1], tm[:, 2]),
                            gstep(tm[:, 2], tm[:, 1]) * gstep(tm[:, 2], tm[:, 0])], axis=1).astype(np.float32)
            oob = active & ((cv < 0) | (cv > up)).any(axis=1)
            term[oob] = 2
            active &= ~oob
            vox = np.zeros(n, bool)
            idx = np.nonzero(active)[0]
            vox[idx] = get_volume_at(vol, cv[idx], 0)
            best = ((tm[:, 0] * sel[:, 0] + tm[:, 1] * sel[:, 1]) + tm[:, 2] * sel[:, 2]).astype(np.float32)
            hit = active & vox
            if hit.any():
                h = np.nonzero(hit)[0]
                term[h] = 1
                res["hit"][h] = 1
                res["t"][h] = best[h]
                res["voxel"][h] = cv[h]
                res["normal"][h] = (-s[h]) * sel[h]
                res["hitpos"][h] = o[h] + d[h] * best[h, None]
                active &= ~hit
            a = np.nonzero(active)[0]
            cv[a] += f2i(sel[a] * s[a])
            tm[a] = (tm[a] + td[a] * sel[a]).astype(np.float32)
            totalt[a] = best[a]
            nt[a] += 1
            cont = (it + 1 < 256) & (totalt < maxt)
            active &= cont
    exhausted = term == 0
    fetches = np.where(exhausted, 4 * nt, nt + (term == 1))
    nt = np.where(exhausted, 4 * nt, nt)
    res.update(term=term.astype(np.int32), steps=nt.astype(np.int32), fetches=fetches.astype(np.int32))
    return res


def gmod(x, y):
    return (x - y * np.floor(x / y)).astype(np.float32)


def sparse(vol, rays, step_factor):
    """raycastShadowVolumeSparse / SuperSparse (Light.frag:131-217); returns (d, fetches, hit)."""
    rays = np.asarray(rays, np.float32)
    n = rays.shape[0]
    pos, d, dist = rays[:, 0:3].copy(), rays[:, 3:6], rays[:, 6]
    sf = f32(step_factor)
    sd = (d * sf).astype(np.float32)
    dd = np.full(n, sf, np.float32)
    out = dist.copy()
    it = np.zeros(n, np.int32)
    active = np.ones(n, bool)
    hitf = np.zeros(n, np.int32)
    with np.errstate(all="ignore"):
        while True:
            a = active & (dd < f32(16.0))
            if not a.any():
                break
            idx = np.nonzero(a)[0]
            it[idx] += 1
            vol_b = texel(vol, f2i(pos[idx] / f32(2.0)))
            m = gmod(pos[idx], HALF) > f32(0.25)
            bit = m[:, 0] * 1 + m[:, 1] * 2 + m[:, 2] * 4
            h = (vol_b & (np.uint32(1) << bit.astype(np.uint32))) != 0
            out[idx[h]] = dd[idx[h]]
            hitf[idx[h]] = 1
            active[idx[h]] = False
            g = idx[~h]
            pos[g] = (pos[g] + sd[g]).astype(np.float32)
            dd[g] = dd[g] + sf
        sf2 = f32(sf * f32(2.0))
        sd = (sd * f32(2.0)).astype(np.float32)
        lim = np.where(f32(164.0) < dist, f32(164.0), dist).astype(np.float32)
        while True:
            a = active & (dd < lim)
            if not a.any():
                break
            idx = np.nonzero(a)[0]
            it[idx] += 1
            h = get_volume_at(vol, f2i(pos[idx]), 1)
            out[idx[h]] = dd[idx[h]]
            hitf[idx[h]] = 1
            active[idx[h]] = False
            g = idx[~h]
            pos[g] = (pos[g] + sd[g]).astype(np.float32)
            dd[g] = dd[g] + sf2
    return out, it, hitf


def translate(m, v):
    """glm::translate (matrix_transform.inl:10-15) in float32, left to right."""
    m = np.asarray(m, np.float32).reshape(4, 4)    # [col, row]
    v = np.asarray(v, np.float32)
    r = m.copy()
    r[3] = ((m[0] * v[0] + m[1] * v[1]) + m[2] * v[2]) + m[3]
    return r.reshape(16)


def footprint(model, m, thr=16):
    """Fine coordinates ivec3(((o + x*dx) + y*dy) + z*dz)*10) of model voxels >= thr
    (ShadowVoxSystem.cpp:137-177); returns (k,3) int64 in the loop order."""
    m = np.asarray(m, np.float32).reshape(4, 4)
    o = m[3, :3]
    dx, dy, dz = m[0, :3] * f32(0.1), m[1, :3] * f32(0.1), m[2, :3] * f32(0.1)
    z, y, x = np.nonzero(model >= thr)
    xf, yf, zf = x.astype(np.float32)[:, None], y.astype(np.float32)[:, None], z.astype(np.float32)[:, None]
    wp = (((o + xf * dx) + yf * dy) + zf * dz).astype(np.float32)
    return f2i((wp * f32(10.0)).astype(np.float32))


def set_bits(vol, f, value):
    sz, sy, sx = vol.shape
    ok = (f[:, 0] >= 0) & (f[:, 1] >= 0) & (f[:, 2] >= 0) & (f[:, 0] < 2 * sx) & (f[:, 1] < 2 * sy) & \
         (f[:, 2] < 2 * sz)
    f = f[ok]
    bit = ((f[:, 0] & 1) | ((f[:, 1] & 1) << 1) | ((f[:, 2] & 1) << 2)).astype(np.uint8)
    t = f // 2
    flat = vol.reshape(-1)
    lin = t[:, 0] + t[:, 1] * sx + t[:, 2] * sx * sy
    masks = (np.uint8(1) << bit).astype(np.uint8)
    if value:
        np.bitwise_or.at(flat, lin, masks)
    else:
        np.bitwise_and.at(flat, lin, np.invert(masks))


def rebuild(vol, instances, thr=16):
    """instances: list of dict(model, prev, cur, pivot, flags). Sequential reference semantics."""
    sz, sy, sx = vol.shape
    regions = []
    for inst in instances:
        if inst["model"] is None:
            continue
        mn = np.array([sx - 1, sy - 1, sz - 1], np.int64)
        mx = np.zeros(3, np.int64)
        if inst["flags"] == 2:
            steps = [(inst["cur"], 0)]
        else:
            npv = -np.asarray(inst["pivot"], np.float32)
            steps = [(translate(inst["prev"], npv), 0), (translate(inst["cur"], npv), 1)]
        for m, val in steps:
            f = footprint(inst["model"], m, thr)
            if f.size:
                mn = np.minimum(mn, f.min(axis=0))
                mx = np.maximum(mx, f.max(axis=0))
            set_bits(vol, f, val)
        if (mx != 0).any():
            lo = np.maximum(np.where(mn >= 0, mn // 2, -((-mn) // 2)), 0)
            hi = np.minimum(np.where(mx >= 0, mx // 2, -((-mx) // 2)), np.array([sx - 1, sy - 1, sz - 1]))
            ext = (hi - lo + 1) & 0xFFFFFFFF
            regions.append((int(lo[0]), int(lo[1]), int(lo[2]), int(ext[0]), int(ext[1]), int(ext[2])))
    return regions


def build_mip(parent):
    """VoxAsset::Upload mip (VoxAsset.cpp:26-47): first non-zero of the children in
    x-fastest order (vi = 0..7; vi = 8 repeats child 0)."""
    pz, py, px = parent.shape
    c = parent.reshape(pz // 2, 2, py // 2, 2, px // 2, 2)
    out = np.zeros((pz // 2, py // 2, px // 2), np.uint8)
    for vi in range(7, -1, -1):
        ch = c[:, (vi >> 2) & 1, :, (vi >> 1) & 1, :, vi & 1]
        out = np.where(ch != 0, ch, out)
    return out
