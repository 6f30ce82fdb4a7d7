"""Workload definitions: scenes, views, ray set-ups and palettes.

P  plain reference: 256^3 terrain (~30% fill), 640x360, 1 shadow + 4 AO rays capped at 16
   voxels, reference marchers (Sparse shadow, SuperSparse AO).
F  plain forward: 512^3 terrain + 2000 boxes (edges 2..32), 2 coarse levels, 1920x1080,
   1 shadow + 16 AO (length 32) + 1 GGX spec-occ ray per pixel, DDA marcher.
B  batched: 64 scenes, each axis a multiple of 8 in [64,512], fill 5..40%, one view each
   from 256x256 to 1920x1080, 1 shadow + 8 AO rays.
D  differentiable shading: 256^3 terrain, 1280x720, 255-entry fp32 palette, directional
   light, AO-weighted ambient, visibility fixed.
L  large: 1024^3 city blocks (20000 grounded boxes, edges 8..56: 5.2% fill), 3 coarse levels,
   four 3840x2160 views from inside the city (96% of the pixels see a block), 1 shadow + 32 AO +
   1 spec-occ ray per pixel, forward + backward.
All inputs are synthetic and seeded.
"""
from __future__ import annotations

import numpy as np

from . import views
from ._lib import MARCH_DDA, MARCH_SPARSE, MARCH_SUPERSPARSE

SEED = 20240607


def trace_params(n_shadow=1, n_ao=16, ao_maxt=32.0, n_spec=1, shadow_maxt=128.0, spec_maxt=256.0,
                 march="dda", seed=SEED, frame=0, jitter_shadow=1):
    if march == "dda":
        ms = ma = mp = MARCH_DDA
    elif march == "reference":   # LightAmbient.frag:121,167; LightReflection.frag:113
        ms, ma, mp = MARCH_SPARSE, MARCH_SUPERSPARSE, MARCH_SPARSE
    else:
        raise ValueError(march)
    return dict(shadow_march=ms, ao_march=ma, spec_march=mp, n_shadow=n_shadow, n_ao=n_ao, n_spec=n_spec,
                shadow_maxt=float(shadow_maxt), ao_maxt=float(ao_maxt), spec_maxt=float(spec_maxt),
                light_dir=views.sun_dir(), seed=seed, frame=frame, jitter_shadow=jitter_shadow)


CONFIGS = {
    "P": dict(dims=(256, 256, 256), fill=0.30, nboxes=0, box=(2, 32), grounded=False, levels=2,
              width=640, height=360, nviews=1,
              params=dict(n_shadow=1, n_ao=4, ao_maxt=16.0, n_spec=0, march="reference")),
    "F": dict(dims=(512, 512, 512), fill=0.30, nboxes=2000, box=(2, 32), grounded=False, levels=2,
              width=1920, height=1080, nviews=1,
              params=dict(n_shadow=1, n_ao=16, ao_maxt=32.0, n_spec=1, march="dda")),
    "D": dict(dims=(256, 256, 256), fill=0.30, nboxes=0, box=(2, 32), grounded=False, levels=2,
              width=1280, height=720, nviews=1,
              params=dict(n_shadow=1, n_ao=8, ao_maxt=32.0, n_spec=0, march="dda")),
    "L": dict(dims=(1024, 1024, 1024), fill=0.0, nboxes=20000, box=(8, 56), grounded=True, levels=3,
              width=3840, height=2160, nviews=4, camera="city",
              params=dict(n_shadow=1, n_ao=32, ao_maxt=32.0, n_spec=1, march="dda")),
}


def config_view(cfg: dict, width=None, height=None, variant=0, frame=0) -> dict:
    """The view of a config: inside the city for L, above a corner otherwise."""
    w, h = width or cfg["width"], height or cfg["height"]
    if cfg.get("camera") == "city":
        return views.city_view(cfg["dims"], w, h, frame=frame, variant=variant)
    return views.scene_view(cfg["dims"], w, h, frame=frame, variant=variant)


def rays_per_pixel(params: dict) -> int:
    return params["n_shadow"] + params["n_ao"] + params["n_spec"]


def make_palette(seed=SEED, n=256) -> np.ndarray:
    """256x6 fp32 {albedo rgb, roughness, metallic, emissive}; entry 0 unused (empty voxel).
    albedo U[0,1], roughness U[0.05,1], metallic Bernoulli(0.3), emissive 0 except 5%."""
    rng = np.random.default_rng(seed)
    pal = np.zeros((n, 6), np.float32)
    pal[1:, 0:3] = rng.random((n - 1, 3))
    pal[1:, 3] = 0.05 + 0.95 * rng.random(n - 1)
    pal[1:, 4] = (rng.random(n - 1) < 0.3).astype(np.float32)
    em = rng.random(n - 1) < 0.05
    pal[1:, 5] = np.where(em, rng.random(n - 1), 0.0)
    return pal


def perturbed_palette(pal: np.ndarray, seed=SEED + 1, sigma=0.05) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q = pal.copy()
    q[1:] = q[1:] + rng.normal(0.0, sigma, size=q[1:].shape).astype(np.float32)
    q[1:, 0:3] = np.clip(q[1:, 0:3], 0.0, 1.0)
    q[1:, 3] = np.clip(q[1:, 3], 0.05, 1.0)
    q[1:, 4] = np.clip(q[1:, 4], 0.0, 1.0)
    q[1:, 5] = np.clip(q[1:, 5], 0.0, None)
    return q.astype(np.float32)


def default_light() -> np.ndarray:
    """{dir3, rgb3}: SUN_DIR and SUN_COLOR = (0.9,0.9,0.8)*0.5 (LightAmbient.frag:15-16)."""
    d = views.sun_dir()
    c = np.array([0.9, 0.9, 0.8], np.float32) * np.float32(0.5)
    return np.concatenate([d, c]).astype(np.float32)


def batched_scenes(n=64, seed=SEED + 7):
    """Config B: n scenes of uneven size with one view each."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dims = tuple(int(8 * rng.integers(8, 65)) for _ in range(3))
        fill = float(rng.uniform(0.05, 0.40))
        w = int(rng.integers(256, 1921))
        h = int(rng.integers(256, 1081))
        out.append(dict(dims=dims, fill=fill, nboxes=int(rng.integers(0, 200)), box=(2, 16), seed=seed + 31 * i,
                        width=w, height=h, variant=i % 4))
    return out
