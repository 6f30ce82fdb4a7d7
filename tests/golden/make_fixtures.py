"""Writes tests/golden/prefabs.npz from the reference's own asset data.

Reads (as data only) the ModernHouse and FarmHouse prefabs (.pf JSON,
PrefabAsset.cpp:30-142), their .v models (3 x u32 dims + X-fastest bytes,
VoxAsset.h:42-50) and resolves each VoxRenderer's Vox GUID as the FNV-1a 64
hash of "default/<prefab>/<file>" (std::hash of the asset path, Assets.h:12-13).
World matrices follow TransformSystem::RealculateMatrix
(TransformSystem.cpp:124-135): translate * rotZ * rotY * rotX * scale under
an identity parent, evaluated in float64 and stored as fp32 inputs.

The expected occupancy of the reference's 524x188x524 ShadowVox volume
(ShadowVoxSystem.cpp:56-58) after one rebuild frame, and after a second frame
that moves every instance, is computed by the oracle and cross-checked
against the independent numpy restatement before it is written.

Usage (in a container that has the VoxelEngine tree at /root/reference):
    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

REF = "/root/reference/Assets/Mods/default"
VOLUME = (524, 188, 524)


def fnv1a64(s: str) -> int:
    h = 14695981039346656037
    for b in s.encode():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def read_v(path):
    b = open(path, "rb").read()
    sx, sy, sz = struct.unpack("<3I", b[:12])
    return np.frombuffer(b[12:12 + sx * sy * sz], np.uint8).reshape(sz, sy, sx).copy()


def rot(axis, a):
    c, s = np.cos(a), np.sin(a)
    x, y, z = axis
    R = np.eye(4)
    R[:3, :3] = [[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                 [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                 [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]]
    return R


def world_matrix(tr):
    pos = [float(v) for v in tr["Position"].split()]
    r = [float(v) for v in tr["Rotation"].split()]
    sc = [float(v) for v in tr["Scale"].split()]
    T = np.eye(4)
    T[:3, 3] = pos
    M = T @ rot((0, 0, 1), r[2]) @ rot((0, 1, 0), r[1]) @ rot((1, 0, 0), r[0]) @ np.diag(sc + [1.0])
    return np.ascontiguousarray(M.T.reshape(16), dtype=np.float32)   # column-major


def load_prefab(name):
    files = {}
    d = os.path.join(REF, name)
    for f in os.listdir(d):
        if f.endswith(".v"):
            files[fnv1a64(f"default/{name}/{f}")] = f
    ents = json.load(open(os.path.join(REF, f"{name}.pf")))
    models, insts = [], []
    index = {}
    for e in ents:
        vr = e.get("VoxRenderer")
        if not vr:
            continue
        g = int(vr["Vox"], 16)
        f = files[g]
        if f not in index:
            index[f] = len(models)
            models.append(read_v(os.path.join(d, f)))
        piv = np.array([float(v) for v in vr["Pivot"].split()], np.float32)
        insts.append((index[f], world_matrix(e["Transform"]), piv))
    return models, insts


def main():
    from oracle import oracle as O
    import np_ref
    out = {}
    for name in ("ModernHouse", "FarmHouse"):
        models, insts = load_prefab(name)
        vol = np.zeros(VOLUME[::-1], np.uint8)
        vol_n = np.zeros(VOLUME[::-1], np.uint8)
        ident = np.eye(4, dtype=np.float32).reshape(16)
        frames = []
        moved = []
        for k, (mi, M, piv) in enumerate(insts):
            M2 = M.copy()
            M2[12:15] += np.array([0.37, 0.0, -0.52], np.float32) * np.float32(1 + k % 3)
            moved.append(M2)
        for frame in range(2):
            if frame == 0:
                pairs = [(ident, M) for (_, M, _) in insts]    # new entities: previous = identity
            else:
                pairs = [(M, M2) for (_, M, _), M2 in zip(insts, moved)]
            oi = [O.make_instance(models[mi], p, c, piv) for (mi, _, piv), (p, c) in zip(insts, pairs)]
            ni = [dict(model=models[mi], prev=p, cur=c, pivot=piv, flags=1) for (mi, _, piv), (p, c) in
                  zip(insts, pairs)]
            regs = O.rebuild(vol, oi)
            regs_n = np_ref.rebuild(vol_n, ni)
            assert np.array_equal(vol, vol_n), name
            assert [tuple(r) for r in regs.tolist()] == regs_n
            frames.append(dict(sha=hashlib.sha256(vol.tobytes()).hexdigest(),
                               popcount=int(np.unpackbits(vol).sum()), regions=np.asarray(regs, np.int64)))
            print(name, frame, frames[-1]["sha"][:16], frames[-1]["popcount"], len(regs))
        for i, m in enumerate(models):
            out[f"{name}_model{i}"] = m
        out[f"{name}_inst_model"] = np.array([mi for mi, _, _ in insts], np.int32)
        out[f"{name}_inst_matrix"] = np.stack([M for _, M, _ in insts])
        out[f"{name}_inst_moved"] = np.stack(moved)
        out[f"{name}_inst_pivot"] = np.stack([p for _, _, p in insts])
        for f, fr in enumerate(frames):
            out[f"{name}_frame{f}_sha256"] = np.frombuffer(bytes.fromhex(fr["sha"]), np.uint8)
            out[f"{name}_frame{f}_popcount"] = np.int64(fr["popcount"])
            out[f"{name}_frame{f}_regions"] = fr["regions"]
    out["volume_dims"] = np.array(VOLUME, np.int32)
    np.savez_compressed(os.path.join(HERE, "prefabs.npz"), **out)
    print("wrote", os.path.join(HERE, "prefabs.npz"), os.path.getsize(os.path.join(HERE, "prefabs.npz")), "bytes")


if __name__ == "__main__":
    main()
