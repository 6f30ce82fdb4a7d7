"""Rebuilds a reference-shaped mod folder (Mods/default/...) from the data in
tests/golden/prefabs.npz: the ModernHouse, FarmHouse and player_ik prefabs,
the .v models they reference and the three .p palettes."""
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "prefabs.npz")


def write_v(path, vox):
    sz, sy, sx = vox.shape
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(struct.pack("<3I", sx, sy, sz))
        f.write(np.ascontiguousarray(vox, np.uint8).tobytes())


def write_mod_folder(root):
    """Writes root/Mods and returns its path."""
    z = np.load(GOLDEN)
    mods = os.path.join(root, "Mods")
    d = os.path.join(mods, "default")
    os.makedirs(d, exist_ok=True)
    for name in ("ModernHouse", "FarmHouse", "player_ik"):
        with open(os.path.join(d, f"{name}.pf"), "wb") as f:
            f.write(z[f"pf_{name}"].tobytes())
    for folder in ("ModernHouse", "FarmHouse", "Player"):
        for i, fname in enumerate(z[f"{folder}_model_files"]):
            write_v(os.path.join(d, folder, str(fname)), z[f"{folder}_model{i}"])
        with open(os.path.join(d, folder, f"{folder}.p"), "wb") as f:
            f.write(z[f"palette_{folder}"].tobytes())
    return mods
