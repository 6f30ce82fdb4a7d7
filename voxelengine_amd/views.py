"""Camera and ViewData construction (host side).

Mirrors the reference's ViewData (Sources/Graphics/Renderer/View.h:16-31),
filled the way WorldRenderer::DrawWorld fills it (WorldRenderer.cpp:189-206):
view, inverse view, projection and inverse projection matrices (column-major
glm layout), resolution, camera position, jitter and frame. World units are
the reference's: one world unit is ten fine voxels.
"""
from __future__ import annotations

import numpy as np

NEAR = 0.1      # Common.frag:12
FAR = 4096.0    # Common.frag:13
FOV = 0.8       # EditorCamera.h:11


def _normalize(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """Right-handed look-at (glm::lookAtRH); returns a column-major 4x4 as m[col][row]."""
    eye = np.asarray(eye, np.float64)
    f = _normalize(np.asarray(target, np.float64) - eye)
    s = _normalize(np.cross(f, up))
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, 0], m[1, 0], m[2, 0] = s
    m[0, 1], m[1, 1], m[2, 1] = u
    m[0, 2], m[1, 2], m[2, 2] = -f
    m[3, 0] = -np.dot(s, eye)
    m[3, 1] = -np.dot(u, eye)
    m[3, 2] = np.dot(f, eye)
    return m  # indexed [col, row]


def perspective(fovy, aspect, near=NEAR, far=FAR):
    """glm::perspectiveRH_NO; column-major [col, row]."""
    t = np.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -1.0
    m[3, 2] = -(2.0 * far * near) / (far - near)
    return m


def _inv_colmajor(m):
    # m[col,row] -> row-major matrix is m.T
    return np.linalg.inv(m.T).T


def make_view(eye_world, target_world, width, height, fov=FOV, frame=0, jitter=(0.0, 0.0)):
    """Returns a dict with float32 column-major matrices (flattened, 16 floats each)."""
    V = look_at(eye_world, target_world)
    P = perspective(fov, width / height)
    f32 = lambda m: np.ascontiguousarray(np.asarray(m, np.float64).reshape(16), dtype=np.float32)
    return {
        "last_view": f32(V),
        "view": f32(V),
        "inv_view": f32(_inv_colmajor(V)),
        "proj": f32(P),
        "inv_proj": f32(_inv_colmajor(P)),
        "res": np.array([width, height], np.float32),
        "ires": np.array([1.0 / width, 1.0 / height], np.float32),
        "cam_pos": np.asarray(eye_world, np.float32),
        "jitter": np.asarray(jitter, np.float32),
        "frame": int(frame),
        "width": int(width),
        "height": int(height),
    }


def scene_view(dims, width, height, frame=0, variant=0):
    """Default camera for a scene of fine-voxel dims (nx, ny, nz): above one corner,
    looking at the middle of the terrain. variant rotates the corner."""
    nx, ny, nz = dims
    corners = [(-0.08, -0.08), (1.08, -0.08), (1.08, 1.08), (-0.08, 1.08)]
    cx, cz = corners[variant % 4]
    eye = np.array([cx * nx, 1.05 * ny, cz * nz]) / 10.0
    tgt = np.array([0.5 * nx, 0.15 * ny, 0.5 * nz]) / 10.0
    return make_view(eye, tgt, width, height, frame=frame)


def city_view(dims, width, height, frame=0, variant=0):
    """Camera inside a low city (config L): 0.293 ny above one corner of the
    ground, looking down the diagonal at the ground a quarter of the way to
    the centre, so that nearly every pixel sees a block. variant picks the
    corner."""
    nx, ny, nz = dims
    cx, cz = [(0.02, 0.02), (0.98, 0.02), (0.98, 0.98), (0.02, 0.98)][variant % 4]
    eye = np.array([cx * nx, 0.293 * ny, cz * nz]) / 10.0
    tgt = np.array([(cx + (0.5 - cx) * 0.5) * nx, 0.0, (cz + (0.5 - cz) * 0.5) * nz]) / 10.0
    return make_view(eye, tgt, width, height, frame=frame)


def sun_dir():
    """SUN_DIR = normalize(vec3(0.3,0.4,0.5)) (LightAmbient.frag:15), in fp32 as
    v / sqrt(dot(v, v))."""
    v = np.array([0.3, 0.4, 0.5], np.float32)
    d = np.float32(np.float32(v[0] * v[0]) + np.float32(v[1] * v[1])) + np.float32(v[2] * v[2])
    l = np.float32(np.sqrt(d))
    return (v / l).astype(np.float32)
