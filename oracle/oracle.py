"""ctypes bindings for the oracle (oracle/libvx_oracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement that checks the HIP product.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module. The product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvx_oracle.so")

MARCH_DDA, MARCH_SPARSE, MARCH_SUPERSPARSE = 0, 1, 2
INST_UPDATE, INST_DESTROY = 1, 2


class Occ(C.Structure):
    _fields_ = [("data", C.c_void_p), ("sx", C.c_int32), ("sy", C.c_int32), ("sz", C.c_int32)]


class Instance(C.Structure):
    _fields_ = [("model", C.c_void_p), ("sx", C.c_int32), ("sy", C.c_int32), ("sz", C.c_int32),
                ("prev", C.c_float * 16), ("cur", C.c_float * 16), ("pivot", C.c_float * 3),
                ("flags", C.c_int32)]


class RayResult(C.Structure):
    _fields_ = [("hit", C.c_int32), ("term", C.c_int32), ("steps", C.c_int32), ("fetches", C.c_int32),
                ("voxel", C.c_int32 * 3), ("t", C.c_float), ("hitpos", C.c_float * 3),
                ("normal", C.c_float * 3)]


RAY_RESULT_DTYPE = np.dtype([("hit", "<i4"), ("term", "<i4"), ("steps", "<i4"), ("fetches", "<i4"),
                             ("voxel", "<i4", 3), ("t", "<f4"), ("hitpos", "<f4", 3),
                             ("normal", "<f4", 3)])
assert RAY_RESULT_DTYPE.itemsize == C.sizeof(RayResult)


class View(C.Structure):
    _fields_ = [("last_view", C.c_float * 16), ("view", C.c_float * 16), ("inv_view", C.c_float * 16),
                ("proj", C.c_float * 16), ("inv_proj", C.c_float * 16), ("res", C.c_float * 2),
                ("ires", C.c_float * 2), ("cam_pos", C.c_float * 3), ("jitter", C.c_float * 2),
                ("frame", C.c_int32), ("width", C.c_int32), ("height", C.c_int32)]


class TraceParams(C.Structure):
    _fields_ = [("shadow_march", C.c_int32), ("ao_march", C.c_int32), ("spec_march", C.c_int32),
                ("n_shadow", C.c_int32), ("n_ao", C.c_int32), ("n_spec", C.c_int32),
                ("shadow_maxt", C.c_float), ("ao_maxt", C.c_float), ("spec_maxt", C.c_float),
                ("light_dir", C.c_float * 3), ("seed", C.c_uint32), ("frame", C.c_int32),
                ("jitter_shadow", C.c_int32), ("pad", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", C.c_float * 3), ("range", C.c_float), ("color", C.c_float * 3),
                ("attenuation", C.c_float), ("direction", C.c_float * 3), ("angle", C.c_float),
                ("angle_attenuation", C.c_float), ("pad", C.c_int32)]


class ReflectionParams(C.Structure):
    _fields_ = [("march", C.c_int32), ("maxt", C.c_float), ("seed", C.c_uint32), ("frame", C.c_int32)]


class LightParams(C.Structure):
    _fields_ = [("point_march", C.c_int32), ("spot_march", C.c_int32), ("seed", C.c_uint32), ("frame", C.c_int32)]


LIGHT_DTYPE = np.dtype([("type", np.int32), ("position", np.float32, 3), ("range", np.float32),
                        ("color", np.float32, 3), ("attenuation", np.float32), ("direction", np.float32, 3),
                        ("angle", np.float32), ("angle_attenuation", np.float32), ("pad", np.int32)])

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        f32p = C.POINTER(C.c_float)
        L.vxo_f2i.argtypes = [C.c_float]; L.vxo_f2i.restype = C.c_int32
        L.vxo_sign.argtypes = [C.c_float]; L.vxo_sign.restype = C.c_float
        L.vxo_mod.argtypes = [C.c_float, C.c_float]; L.vxo_mod.restype = C.c_float
        L.vxo_sincos.argtypes = [C.c_float, f32p, f32p]; L.vxo_sincos.restype = None
        L.vxo_hash.argtypes = [C.c_uint32]; L.vxo_hash.restype = C.c_uint32
        L.vxo_rand.argtypes = [C.c_uint32, C.c_uint32]; L.vxo_rand.restype = C.c_float
        L.vxo_pixkey.argtypes = [C.c_uint32, C.c_int32, C.c_uint32]; L.vxo_pixkey.restype = C.c_uint32
        L.vxo_cosine_hemisphere.argtypes = [C.c_float, C.c_float, f32p]; L.vxo_cosine_hemisphere.restype = None
        L.vxo_set_volume_at.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.vxo_translate.argtypes = [f32p, f32p, f32p]
        L.vxo_rebuild.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Instance), C.c_int32, C.c_int32, P]
        L.vxo_rebuild.restype = C.c_int32
        L.vxo_build_mip.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, P]
        L.vxo_get_volume_at.argtypes = [C.POINTER(Occ), C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.vxo_get_volume_at.restype = C.c_int32
        L.vxo_dda.argtypes = [C.POINTER(Occ), f32p, f32p, C.c_float, C.POINTER(RayResult)]
        L.vxo_dda.restype = C.c_int32
        L.vxo_sparse.argtypes = [C.POINTER(Occ), f32p, f32p, C.c_float, C.POINTER(RayResult)]
        L.vxo_sparse.restype = C.c_float
        L.vxo_supersparse.argtypes = [C.POINTER(Occ), f32p, f32p, C.c_float, C.POINTER(RayResult)]
        L.vxo_supersparse.restype = C.c_float
        L.vxo_trace_rays.argtypes = [C.POINTER(Occ), P, C.c_int32, C.c_int32, P, C.c_int32]
        L.vxo_clip_to_aabb.argtypes = [f32p, f32p, f32p, f32p]
        L.vxo_intersect_volume.argtypes = [C.POINTER(P), C.POINTER(C.c_int32), f32p, f32p, f32p, f32p, f32p,
                                           C.c_int32, f32p, f32p, C.POINTER(C.c_uint8)]
        L.vxo_intersect_volume.restype = C.c_int32
        L.vxo_primary.argtypes = [C.POINTER(P), C.POINTER(C.c_int32), C.POINTER(View), P, P, P, C.c_int32]
        L.vxo_primary_pixels.argtypes = [C.POINTER(P), C.POINTER(C.c_int32), C.POINTER(View), P, C.c_int32, P, P, P,
                                         C.c_int32]
        L.vxo_trace.argtypes = [C.POINTER(Occ), C.POINTER(View), C.POINTER(TraceParams), P, P, P, P,
                                P, P, P, P, P, C.c_int32, C.c_int32]
        L.vxo_pixel_rays.argtypes = [C.POINTER(View), C.POINTER(TraceParams), P, P, P, P, P, C.c_int32, P]
        L.vxo_sky_procedural.argtypes = [f32p, f32p]
        L.vxo_cube_sample.argtypes = [P, C.c_int32, f32p, f32p]
        L.vxo_reflection_rays.argtypes = [C.POINTER(View), C.POINTER(ReflectionParams), P, P, P, P, P, C.c_int32, P]
        L.vxo_reflection.argtypes = [C.POINTER(Occ), C.POINTER(View), C.POINTER(ReflectionParams), P, P, P, P, P, P,
                                     P, C.c_int32, P, P, C.c_int32]
        L.vxo_ssao.argtypes = [C.POINTER(View), P, P, P, C.c_uint32, P, C.c_int32]
        L.vxo_shade_f32.argtypes = [C.POINTER(View), P, P, P, P, P, P, P, P, P, C.c_int32]
        L.vxo_shade_f64.argtypes = [C.POINTER(View), P, P, P, P, P, P, P, P, P, C.c_int32]
        L.vxo_shade_backward_f64.argtypes = [C.POINTER(View), P, P, P, P, P, P, P, P, P, P, P]
        L.vxo_gen_terrain.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_float, C.c_int32]
        L.vxo_gen_boxes.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int32]
        L.vxo_log2.argtypes = [C.c_float]; L.vxo_log2.restype = C.c_float
        L.vxo_exp2.argtypes = [C.c_float]; L.vxo_exp2.restype = C.c_float
        L.vxo_pow.argtypes = [C.c_float, C.c_float]; L.vxo_pow.restype = C.c_float
        L.vxo_light_pass.argtypes = [C.POINTER(Occ), C.POINTER(View), P, C.c_int32, C.POINTER(LightParams),
                                     P, P, P, P, P, P, C.c_int32]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


# ---------------------------------------------------------------- scalars
def f2i(x):
    return lib().vxo_f2i(float(np.float32(x)))


def sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().vxo_sincos(float(np.float32(x)), C.byref(s), C.byref(c))
    return np.float32(s.value), np.float32(c.value)


def hash32(x):
    return lib().vxo_hash(int(x) & 0xFFFFFFFF)


def rand(pixkey, idx):
    return np.float32(lib().vxo_rand(pixkey & 0xFFFFFFFF, idx & 0xFFFFFFFF))


def pixkey(seed, frame, pixel):
    return lib().vxo_pixkey(seed & 0xFFFFFFFF, frame, pixel & 0xFFFFFFFF)


# ---------------------------------------------------------------- rebuild
def set_volume_at(vol: np.ndarray, x, y, z, value):
    sz, sy, sx = vol.shape
    lib().vxo_set_volume_at(_p(vol), sx, sy, sz, x, y, z, value)


def translate(m, v):
    m = np.ascontiguousarray(m, dtype=np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    v = np.asarray(v, np.float32)
    lib().vxo_translate(m.ctypes.data_as(C.POINTER(C.c_float)), v.ctypes.data_as(C.POINTER(C.c_float)),
                        out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def make_instance(model: np.ndarray, prev, cur, pivot=(0, 0, 0), flags=INST_UPDATE):
    """model: uint8 array shaped (sz, sy, sx) (X fastest). Matrices column-major 16 floats."""
    inst = Instance()
    inst.model = model.ctypes.data if model is not None else None
    if model is not None:
        inst.sz, inst.sy, inst.sx = model.shape
    inst.prev = (C.c_float * 16)(*np.asarray(prev, np.float32).reshape(16))
    inst.cur = (C.c_float * 16)(*np.asarray(cur, np.float32).reshape(16))
    inst.pivot = _f3(pivot)
    inst.flags = flags
    return inst


def rebuild(vol: np.ndarray, instances, threshold=16):
    """vol: uint8 (sz, sy, sx) texel volume, modified in place. Returns regions (n, 6) int64
    with extents as uint32 values."""
    sz, sy, sx = vol.shape
    n = len(instances)
    arr = (Instance * max(n, 1))(*instances)
    regions = np.zeros((max(n, 1), 6), np.int32)
    nr = lib().vxo_rebuild(_p(vol), sx, sy, sz, arr, n, threshold, _p(regions))
    r = regions[:nr].astype(np.int64)
    r[:, 3:] &= 0xFFFFFFFF
    return r


def build_mip(parent: np.ndarray):
    pz, py, px = parent.shape
    out = np.zeros((max(pz // 2, 1), max(py // 2, 1), max(px // 2, 1)), np.uint8)
    lib().vxo_build_mip(_p(parent), px, py, pz, _p(out))
    return out


def build_mips(model: np.ndarray):
    m1 = build_mip(model)
    m2 = build_mip(m1)
    return [np.ascontiguousarray(model), m1, m2]


# ---------------------------------------------------------------- traversal
def occ(vol: np.ndarray) -> Occ:
    sz, sy, sx = vol.shape
    o = Occ()
    o.data = vol.ctypes.data
    o.sx, o.sy, o.sz = sx, sy, sz
    return o


def dda(vol, o, d, maxt):
    r = RayResult()
    O = occ(vol)
    lib().vxo_dda(C.byref(O), _f3(o), _f3(d), float(np.float32(maxt)), C.byref(r))
    return r


def sparse(vol, o, d, dist, super_=False):
    r = RayResult()
    O = occ(vol)
    fn = lib().vxo_supersparse if super_ else lib().vxo_sparse
    fn(C.byref(O), _f3(o), _f3(d), float(np.float32(dist)), C.byref(r))
    return r


def trace_rays(vol, rays: np.ndarray, march, nthreads=0):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 7)
    out = np.zeros(rays.shape[0], RAY_RESULT_DTYPE)
    O = occ(vol)
    lib().vxo_trace_rays(C.byref(O), _p(rays), rays.shape[0], march, _p(out), nthreads)
    return out


def _mips_args(mips):
    ptrs = (C.c_void_p * 3)(*[m.ctypes.data for m in mips])
    d = mips[0].shape
    dims = (C.c_int32 * 3)(d[2], d[1], d[0])
    return ptrs, dims


def intersect_volume(mips, o, d, cam_local, uv, res, frame):
    ptrs, dims = _mips_args(mips)
    hit = (C.c_float * 3)(); nrm = (C.c_float * 3)(); mat = C.c_uint8()
    f2 = lambda v: (C.c_float * 2)(*[float(x) for x in v])
    h = lib().vxo_intersect_volume(ptrs, dims, _f3(o), _f3(d), _f3(cam_local), f2(uv), f2(res), frame,
                                   hit, nrm, C.byref(mat))
    return h, np.array(hit, np.float32), np.array(nrm, np.float32), mat.value


def view_struct(vd: dict) -> View:
    v = View()
    for k in ("last_view", "view", "inv_view", "proj", "inv_proj"):
        setattr(v, k, (C.c_float * 16)(*np.asarray(vd[k], np.float32).reshape(16)))
    v.res = (C.c_float * 2)(*vd["res"]); v.ires = (C.c_float * 2)(*vd["ires"])
    v.cam_pos = _f3(vd["cam_pos"]); v.jitter = (C.c_float * 2)(*vd["jitter"])
    v.frame = vd["frame"]; v.width = vd["width"]; v.height = vd["height"]
    return v


def primary(mips, vd: dict, nthreads=0, pixels=None):
    """G-buffer of the view: (pos, nrm, mat) for every pixel, or compact for the
    listed pixels."""
    ptrs, dims = _mips_args(mips)
    if pixels is not None:
        pixels = np.ascontiguousarray(pixels, np.int32)
        n = pixels.size
        pos = np.zeros((n, 3), np.float32); nrm = np.zeros((n, 3), np.float32); mat = np.zeros(n, np.uint8)
        v = view_struct(vd)
        lib().vxo_primary_pixels(ptrs, dims, C.byref(v), _p(pixels), n, _p(pos), _p(nrm), _p(mat), nthreads)
        return pos, nrm, mat
    W, H = vd["width"], vd["height"]
    pos = np.zeros((H * W, 3), np.float32); nrm = np.zeros((H * W, 3), np.float32)
    mat = np.zeros(H * W, np.uint8)
    v = view_struct(vd)
    lib().vxo_primary(ptrs, dims, C.byref(v), _p(pos), _p(nrm), _p(mat), nthreads)
    return pos, nrm, mat


class VoxCmd(C.Structure):
    _fields_ = [("mips", C.c_void_p * 3), ("dims", C.c_int32 * 3), ("world", C.c_float * 16),
                ("last_world", C.c_float * 16), ("rid", C.c_int32), ("palette", C.c_int32)]


def primary_instances(cmds, vd: dict, nthreads=0):
    """cmds: list of dict(mips=[3 uint8 arrays (sz,sy,sx)], world, last_world, rid, palette)."""
    arr = (VoxCmd * max(len(cmds), 1))()
    keep = []
    for k, c in enumerate(cmds):
        mips = [np.ascontiguousarray(m, np.uint8) for m in c["mips"]]
        keep.append(mips)
        arr[k].mips = (C.c_void_p * 3)(*[m.ctypes.data for m in mips])
        sz, sy, sx = mips[0].shape
        arr[k].dims = (C.c_int32 * 3)(sx, sy, sz)
        arr[k].world = (C.c_float * 16)(*np.asarray(c["world"], np.float32).reshape(16))
        lw = c.get("last_world", c["world"])
        arr[k].last_world = (C.c_float * 16)(*np.asarray(lw, np.float32).reshape(16))
        arr[k].rid, arr[k].palette = int(c.get("rid", 1)), int(c.get("palette", 0))
    W, H = vd["width"], vd["height"]
    pos = np.zeros((H * W, 3), np.float32); nrm = np.zeros((H * W, 3), np.float32)
    mat = np.zeros(H * W, np.uint8); motion = np.zeros((H * W, 2), np.float32)
    depth = np.zeros(H * W, np.float32); pal = np.zeros(H * W, np.uint16)
    v = view_struct(vd)
    lib().vxo_primary_instances(arr, len(cmds), C.byref(v), _p(pos), _p(nrm), _p(mat), _p(motion), _p(depth),
                                _p(pal), nthreads)
    return dict(pos=pos, nrm=nrm, mat=mat, motion=motion, depth=depth, palette=pal)


def light_taa(width, height, frame, seed, light, last, motion, depth, normal, mat, palette, nthreads=0):
    """LightTAA over numpy buffers; returns (npx, 4) float32."""
    npx = width * height
    f = lambda a, n: np.ascontiguousarray(np.asarray(a, np.float32).reshape(npx, n))
    light, last, depth, normal = f(light, 3), f(last, 4), f(depth, 1), f(normal, 3)
    motion = None if motion is None else f(motion, 2)
    mat = np.ascontiguousarray(np.asarray(mat, np.uint8).reshape(npx))
    palette = np.ascontiguousarray(np.asarray(palette, np.float32).reshape(256, 6))
    out = np.zeros((npx, 4), np.float32)
    lib().vxo_light_taa(C.c_int32(width), C.c_int32(height), C.c_int32(frame), C.c_uint32(seed & 0xFFFFFFFF),
                        _p(light), _p(last), _p(motion) if motion is not None else None, _p(depth), _p(normal),
                        _p(mat), _p(palette), _p(out), C.c_int32(nthreads))
    return out


def taa_sample(img, u, v):
    """The TAA sampler over an (h, w, n) float image at (u, v)."""
    img = np.ascontiguousarray(img, np.float32)
    h, w, n = img.shape
    out = (C.c_float * n)()
    lib().vxo_taa_sample(_p(img), C.c_int32(w), C.c_int32(h), C.c_int32(n), C.c_float(u), C.c_float(v), out)
    return np.array(out[:], np.float32)


def trace_params(p: dict) -> TraceParams:
    t = TraceParams()
    t.shadow_march = p["shadow_march"]; t.ao_march = p["ao_march"]; t.spec_march = p["spec_march"]
    t.n_shadow = p["n_shadow"]; t.n_ao = p["n_ao"]; t.n_spec = p["n_spec"]
    t.shadow_maxt = p["shadow_maxt"]; t.ao_maxt = p["ao_maxt"]; t.spec_maxt = p["spec_maxt"]
    t.light_dir = _f3(p["light_dir"]); t.seed = p["seed"] & 0xFFFFFFFF; t.frame = p["frame"]
    t.jitter_shadow = p["jitter_shadow"]
    return t


def trace(vol, vd: dict, params: dict, pos, nrm, mat, palette, pixels=None, nthreads=0):
    npx = vd["width"] * vd["height"]
    shadow = np.zeros(npx, np.uint8); ao = np.zeros(npx, np.float32); spec = np.zeros(npx, np.float32)
    fetches = np.zeros(npx, np.uint32)
    O = occ(vol); v = view_struct(vd); tp = trace_params(params)
    palette = np.ascontiguousarray(palette, np.float32)
    if pixels is not None:
        pixels = np.ascontiguousarray(pixels, np.int32)
        pp, npix = _p(pixels), pixels.size
    else:
        pp, npix = None, 0
    lib().vxo_trace(C.byref(O), C.byref(v), C.byref(tp), _p(pos), _p(nrm), _p(mat), _p(palette),
                    _p(shadow), _p(ao), _p(spec), _p(fetches), pp, npix, nthreads)
    return shadow, ao, spec, fetches


def pixel_rays(vd: dict, params: dict, pos, nrm, mat, palette, pixels):
    """The rays the trace marches for each listed pixel: (npixels, nrays, 7)
    float32 (origin3, dir3, maxt), zeros for sky pixels."""
    pixels = np.ascontiguousarray(pixels, np.int32)
    nr = params["n_shadow"] + params["n_ao"] + params["n_spec"]
    out = np.zeros((pixels.size, nr, 7), np.float32)
    v = view_struct(vd); tp = trace_params(params)
    lib().vxo_pixel_rays(C.byref(v), C.byref(tp), _p(np.ascontiguousarray(pos, np.float32)),
                         _p(np.ascontiguousarray(nrm, np.float32)), _p(np.ascontiguousarray(mat, np.uint8)),
                         _p(np.ascontiguousarray(palette, np.float32)), _p(pixels), pixels.size, _p(out))
    return out


def shade_f32(vd, pos, nrm, mat, shadow, ao, palette, light, ssao=None, nthreads=0):
    npx = vd["width"] * vd["height"]
    rgb = np.zeros((npx, 3), np.float32)
    v = view_struct(vd)
    light = np.ascontiguousarray(light, np.float32)
    palette = np.ascontiguousarray(palette, np.float32)
    lib().vxo_shade_f32(C.byref(v), _p(pos), _p(nrm), _p(mat), _p(shadow), _p(ao),
                        _p(ssao) if ssao is not None else None, _p(palette), _p(light), _p(rgb), nthreads)
    return rgb


def shade_f64(vd, pos, nrm, mat, shadow, ao, palette, light, ssao=None, nthreads=0):
    npx = vd["width"] * vd["height"]
    rgb = np.zeros((npx, 3), np.float64)
    v = view_struct(vd)
    light = np.ascontiguousarray(light, np.float64)
    palette = np.ascontiguousarray(palette, np.float64)
    lib().vxo_shade_f64(C.byref(v), _p(pos), _p(nrm), _p(mat), _p(shadow), _p(ao),
                        _p(ssao) if ssao is not None else None, _p(palette), _p(light), _p(rgb), nthreads)
    return rgb


def shade_backward_f64(vd, pos, nrm, mat, shadow, ao, palette, light, dl_drgb, ssao=None):
    v = view_struct(vd)
    light = np.ascontiguousarray(light, np.float64)
    palette = np.ascontiguousarray(palette, np.float64)
    g = np.ascontiguousarray(dl_drgb, np.float64)
    dpal = np.zeros((256, 6), np.float64); dlight = np.zeros(6, np.float64)
    lib().vxo_shade_backward_f64(C.byref(v), _p(pos), _p(nrm), _p(mat), _p(shadow), _p(ao),
                                 _p(ssao) if ssao is not None else None, _p(palette), _p(light), _p(g),
                                 _p(dpal), _p(dlight))
    return dpal, dlight


# ---------------------------------------------------------------- scenes
def gen_terrain(nx, ny, nz, seed, fill, nthreads=0):
    vol = np.zeros((nz, ny, nx), np.uint8)
    lib().vxo_gen_terrain(_p(vol), nx, ny, nz, seed & 0xFFFFFFFF, float(fill), nthreads)
    return vol


def gen_boxes(vol, seed, nboxes, emin, emax, grounded=0):
    nz, ny, nx = vol.shape
    lib().vxo_gen_boxes(_p(vol), nx, ny, nz, seed & 0xFFFFFFFF, nboxes, emin, emax, grounded)
    return vol


def light_params(point_march=1, spot_march=2, seed=0, frame=0) -> LightParams:
    return LightParams(point_march, spot_march, seed & 0xFFFFFFFF, frame)


def light_pass(vol, vd: dict, lights: np.ndarray, lp: LightParams, pos, nrm, mat, palette, rgb=None, nthreads=0):
    """lights: LIGHT_DTYPE array. Adds every light's direct term to rgb (npx x 3,
    zeros when None); returns (rgb, fetches)."""
    npx = vd["width"] * vd["height"]
    rgb = np.zeros((npx, 3), np.float32) if rgb is None else np.ascontiguousarray(rgb, np.float32).copy()
    fetches = np.zeros(npx, np.uint32)
    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    O = occ(vol); v = view_struct(vd)
    lib().vxo_light_pass(C.byref(O), C.byref(v), lights.ctypes.data, lights.size, C.byref(lp), _p(pos), _p(nrm),
                         _p(mat), _p(np.ascontiguousarray(palette, np.float32)), _p(rgb), _p(fetches), nthreads)
    return rgb, fetches


# ---------------------------------------------------------------- reflection pass
def reflection_params(march=1, maxt=256.0, seed=0, frame=0) -> ReflectionParams:
    return ReflectionParams(march, float(maxt), seed & 0xFFFFFFFF, frame)


def sky_procedural(e):
    out = (C.c_float * 3)()
    lib().vxo_sky_procedural(_f3(e), out)
    return np.array(out[:], np.float32)


def cube_sample(faces, d):
    """faces: (6, size, size, 4) float32."""
    faces = np.ascontiguousarray(faces, np.float32)
    out = (C.c_float * 4)()
    lib().vxo_cube_sample(_p(faces), faces.shape[1], _f3(d), out)
    return np.array(out[:], np.float32)


def reflection_rays(vd, rp: ReflectionParams, pos, nrm, mat, palette, pixels):
    """(npixels, 6): wcp (fine units) and direction of each listed pixel's reflection ray."""
    pixels = np.ascontiguousarray(pixels, np.int32)
    out = np.zeros((pixels.size, 6), np.float32)
    v = view_struct(vd)
    lib().vxo_reflection_rays(C.byref(v), C.byref(rp), _p(np.ascontiguousarray(pos, np.float32)),
                              _p(np.ascontiguousarray(nrm, np.float32)), _p(np.ascontiguousarray(mat, np.uint8)),
                              _p(np.ascontiguousarray(palette, np.float32)), _p(pixels), pixels.size, _p(out))
    return out


def reflection(vol, vd, rp: ReflectionParams, pos, nrm, mat, palette, depth, light, sky=None, nthreads=0):
    """LightReflection over numpy buffers: (out (npx, 4), t (npx,))."""
    npx = vd["width"] * vd["height"]
    out = np.zeros((npx, 4), np.float32)
    t = np.zeros(npx, np.float32)
    O = occ(vol); v = view_struct(vd)
    sky_a = None if sky is None else np.ascontiguousarray(sky, np.float32)
    lib().vxo_reflection(C.byref(O), C.byref(v), C.byref(rp), _p(np.ascontiguousarray(pos, np.float32)),
                         _p(np.ascontiguousarray(nrm, np.float32)), _p(np.ascontiguousarray(mat, np.uint8)),
                         _p(np.ascontiguousarray(palette, np.float32)),
                         _p(np.ascontiguousarray(depth, np.float32).reshape(npx)),
                         _p(np.ascontiguousarray(light, np.float32).reshape(npx, 4)),
                         _p(sky_a) if sky_a is not None else None, 0 if sky_a is None else sky_a.shape[1],
                         _p(out), _p(t), nthreads)
    return out, t


def ssao(vd, nrm, mat, depth, seed=0, nthreads=0):
    """calculateOcclusion per pixel (npx,), 1 where nothing was hit."""
    npx = vd["width"] * vd["height"]
    out = np.zeros(npx, np.float32)
    v = view_struct(vd)
    lib().vxo_ssao(C.byref(v), _p(np.ascontiguousarray(nrm, np.float32)), _p(np.ascontiguousarray(mat, np.uint8)),
                   _p(np.ascontiguousarray(depth, np.float32).reshape(npx)), seed & 0xFFFFFFFF, _p(out), nthreads)
    return out
