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
    return models, insts, list(index)


def guid_map():
    """GUID -> mod-relative path of every file under the mod folder (Mod.h:86-94)."""
    out = {}
    mods = os.path.dirname(REF)
    for root, _, files in os.walk(REF):
        for f in files:
            rel = os.path.relpath(os.path.join(root, f), mods).replace(os.sep, "/")
            out[fnv1a64(rel)] = rel
    return out


def spawn(rel, gmap, ents, pools, parent_entity=-1):
    """PrefabAsset::Spawn (PrefabAsset.cpp:30-142) into `ents` (creation
    order); pools["vox"] / pools["light"] record the component creation order."""
    data = json.load(open(os.path.join(os.path.dirname(REF), rel)))
    idmap, root = {}, -1
    for E in data:
        parent = idmap[E["Parent"]] if E.get("Parent") is not None else -1
        if E.get("Instance") is None:
            e = len(ents)
            ents.append(dict(parent=-1))
        else:
            e = spawn(gmap[int(E["Instance"], 16)], gmap, ents, pools, parent)
        ents[e]["id"] = E["Id"]
        idmap[E["Id"]] = e
        if parent < 0:
            root = e
            ents[e]["parent"] = parent_entity
        else:
            ents[e]["parent"] = parent
        if "Transform" in E:
            ents[e]["tr"] = E["Transform"]
        if "VoxRenderer" in E:
            if "vr" not in ents[e]:
                pools["vox"].append(e)
            ents[e]["vr"] = E["VoxRenderer"]
        if "Light" in E:
            if "light" not in ents[e]:
                pools["light"].append(e)
            ents[e]["light"] = E["Light"]
    return root


def spawn_full(name, gmap):
    """Spawn with nested instances; world matrices in float64 (parent @ local)."""
    ents, pools = [], {"vox": [], "light": []}
    spawn(f"default/{name}.pf", gmap, ents, pools)
    for e in ents:
        if "tr" not in e:
            e["world64"] = np.eye(4)
            continue
        local = world_matrix(e["tr"]).astype(np.float64).reshape(4, 4).T
        par = ents[e["parent"]]["world64"] if e["parent"] >= 0 and "tr" in ents[e["parent"]] else np.eye(4)
        e["world64"] = par @ local
    return ents, pools


def main():
    from oracle import oracle as O
    import np_ref
    out = {}
    gmap = guid_map()
    # raw asset files the loader tests rebuild a mod folder from (data, not code)
    for name in ("ModernHouse", "FarmHouse", "player_ik"):
        out[f"pf_{name}"] = np.frombuffer(open(os.path.join(REF, f"{name}.pf"), "rb").read(), np.uint8)
    for d in ("ModernHouse", "FarmHouse", "Player"):
        out[f"palette_{d}"] = np.frombuffer(open(os.path.join(REF, d, f"{d}.p"), "rb").read(), np.uint8)
    pfiles = sorted(f for f in os.listdir(os.path.join(REF, "Player")) if f.endswith(".v"))
    for i, f in enumerate(pfiles):
        out[f"Player_model{i}"] = read_v(os.path.join(REF, "Player", f))
    out["Player_model_files"] = np.array(pfiles)
    # FarmHouse spawned as the reference's ECS orders it: VoxRenderer pool
    # walked newest first (EnTT 3.6 views), lights likewise, point then spot.
    ents, pools = spawn_full("FarmHouse", gmap)
    order = [e for e in reversed(pools["vox"]) if "tr" in ents[e] and int(ents[e]["vr"]["Vox"], 16) in gmap]
    out["FarmHouse_ref_order_ids"] = np.array([ents[e]["id"] for e in order], np.int32)
    lights = []
    for t in (0, 1):
        for e in reversed(pools["light"]):
            L = ents[e]["light"]
            if "tr" not in ents[e] or L["LightType"] != t:
                continue
            w = ents[e]["world64"].astype(np.float32)
            col = np.array([float(v) for v in L["Color"].split()], np.float32) * np.float32(L["Intensity"])
            lights.append([t, *w[:3, 3], L["Range"], *col, L["Attenuation"], *w[:3, 2], L["Angle"],
                           L["AngleAttenuation"]])
    out["FarmHouse_ref_lights"] = np.array(lights, np.float32)
    for name in ("ModernHouse", "FarmHouse"):
        models, insts, names = load_prefab(name)
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
        out[f"{name}_model_files"] = np.array(names)
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
    # FarmHouse's first frame in the reference order (expected volume by the oracle,
    # cross-checked against the numpy restatement)
    models, insts, _ = load_prefab("FarmHouse")
    ents, pools = spawn_full("FarmHouse", gmap)
    order = [e for e in reversed(pools["vox"]) if "tr" in ents[e] and int(ents[e]["vr"]["Vox"], 16) in gmap]
    by_id = {}
    for k, e in enumerate([x for x in ents if "vr" in x]):
        by_id[e["id"]] = k
    vol = np.zeros(VOLUME[::-1], np.uint8)
    vol_n = np.zeros(VOLUME[::-1], np.uint8)
    ident = np.eye(4, dtype=np.float32).reshape(16)
    oi, ni = [], []
    for e in order:
        mi, M, piv = insts[by_id[ents[e]["id"]]]
        oi.append(O.make_instance(models[mi], ident, M, piv))
        ni.append(dict(model=models[mi], prev=ident, cur=M, pivot=piv, flags=1))
    regs = O.rebuild(vol, oi)
    assert [tuple(r) for r in regs.tolist()] == np_ref.rebuild(vol_n, ni)
    assert np.array_equal(vol, vol_n)
    out["FarmHouse_ref_frame0_sha256"] = np.frombuffer(hashlib.sha256(vol.tobytes()).digest(), np.uint8)
    out["FarmHouse_ref_frame0_popcount"] = np.int64(np.unpackbits(vol).sum())
    out["FarmHouse_ref_frame0_regions"] = np.asarray(regs, np.int64)
    print("FarmHouse reference order", hashlib.sha256(vol.tobytes()).hexdigest()[:16], int(np.unpackbits(vol).sum()))
    out["volume_dims"] = np.array(VOLUME, np.int32)
    np.savez_compressed(os.path.join(HERE, "prefabs.npz"), **out)
    print("wrote", os.path.join(HERE, "prefabs.npz"), os.path.getsize(os.path.join(HERE, "prefabs.npz")), "bytes")


if __name__ == "__main__":
    main()
