"""Python host API over the C ABI (device buffers are torch tensors on the HIP device).

Used by bench.py, the tests and smoke(). Every call lands in libvx_amd.so; a
missing library or device raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._lib import check


def _stream(stream=None):
    if stream is None:
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


def _ptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device buffers must be contiguous HIP tensors"
    return C.c_void_p(t.data_ptr())


def _hp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """vx_ctx: one HIP device."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(L.lib().vx_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def set_tuning(self, kernel=0, block=0, blocks_per_cu=0, refill=0, steps=0, tile=0, order=0, lights=0, chunks=0):
        t = L.TraceTuning(kernel, block, blocks_per_cu, refill, steps, tile, order, lights, chunks)
        check(L.lib().vx_ctx_set_trace_tuning(self.h, C.byref(t)))

    def reset_tuning(self):
        check(L.lib().vx_ctx_set_trace_tuning(self.h, None))

    def profile(self, enable: bool = True):
        """Start (or stop) timing the march kernel of every trace launch."""
        check(L.lib().vx_ctx_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        """(total march-kernel ms, launches) since profiling was enabled."""
        ms, n = C.c_double(), C.c_uint32()
        check(L.lib().vx_ctx_profile_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def tuning(self) -> dict:
        t = L.TraceTuning()
        check(L.lib().vx_ctx_get_trace_tuning(self.h, C.byref(t)))
        return dict(kernel=t.kernel, block=t.block, blocks_per_cu=t.blocks_per_cu, refill=t.refill, steps=t.steps,
                    tile=t.tile, order=t.order, lights=t.lights, chunks=t.chunks)

    def close(self):
        if self.h:
            L.lib().vx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """vx_model: VoxAsset material volume with 3 mips. Arrays are (sz, sy, sx), X fastest."""

    def __init__(self, ctx: Context, sx: int, sy: int, sz: int):
        h = C.c_void_p()
        check(L.lib().vx_model_create(ctx.h, sx, sy, sz, C.byref(h)))
        self.h, self.ctx = h, ctx
        d = (C.c_int32 * 3)()
        check(L.lib().vx_model_dims(h, d))
        self.dims = (d[0], d[1], d[2])

    def close(self):
        if self.h:
            L.lib().vx_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, vol: np.ndarray, stream=None):
        vol = np.ascontiguousarray(vol, np.uint8)
        check(L.lib().vx_model_upload(self.h, _hp(vol), vol.nbytes, _stream(stream)))

    def download(self, mip: int = 0, stream=None) -> np.ndarray:
        sx, sy, sz = (d >> mip for d in self.dims)
        out = np.empty((sz, sy, sx), np.uint8)
        check(L.lib().vx_model_download(self.h, mip, _hp(out), out.nbytes, _stream(stream)))
        return out

    def mip_ptr(self, mip: int = 0) -> int:
        p = C.c_void_p()
        check(L.lib().vx_model_mip_ptr(self.h, mip, C.byref(p)))
        return p.value

    def generate(self, seed, terrain_fill=0.3, nboxes=0, box_min=2, box_max=32, grounded=False, stream=None):
        d = L.SceneDesc(seed & 0xFFFFFFFF, float(terrain_fill), nboxes, box_min, box_max, int(bool(grounded)))
        check(L.lib().vx_model_generate(self.h, C.byref(d), _stream(stream)))

    def build_mips(self, stream=None):
        check(L.lib().vx_model_build_mips(self.h, _stream(stream)))


def region_tuple(r: L.Region):
    return (r.x, r.y, r.z, r.w, r.h, r.d)


def make_instance(model_ptr: int, dims, prev, cur, pivot=(0.0, 0.0, 0.0), flags=L.INST_UPDATE) -> L.Instance:
    inst = L.Instance()
    inst.model = model_ptr
    inst.sx, inst.sy, inst.sz = dims
    inst.prev = L.f16(*np.asarray(prev, np.float32).reshape(16))
    inst.cur = L.f16(*np.asarray(cur, np.float32).reshape(16))
    inst.pivot = (C.c_float * 3)(*[float(x) for x in pivot])
    inst.flags = flags
    return inst


IDENTITY = np.eye(4, dtype=np.float32).reshape(16)


class Volume:
    """vx_volume: ShadowVox occupancy volume (texel dims) with coarse levels."""

    def __init__(self, ctx: Context, sx: int, sy: int, sz: int, nlevels: int = 2):
        h = C.c_void_p()
        check(L.lib().vx_volume_create(ctx.h, sx, sy, sz, nlevels, C.byref(h)))
        self.h, self.ctx = h, ctx
        self.dims = (sx, sy, sz)
        self.nlevels = nlevels

    def close(self):
        if self.h:
            L.lib().vx_volume_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self, stream=None):
        check(L.lib().vx_volume_clear(self.h, _stream(stream)))

    def rebuild(self, instances, threshold=16, stream=None):
        n = len(instances)
        arr = (L.Instance * max(n, 1))(*instances)
        dirty = (L.Region * max(n, 1))()
        nd = C.c_uint32()
        check(L.lib().vx_volume_rebuild(self.h, arr, n, threshold, dirty, C.byref(nd), _stream(stream)))
        return [region_tuple(dirty[i]) for i in range(nd.value)]

    def build_levels(self, regions=None, stream=None):
        if regions:
            arr = (L.Region * len(regions))(*[L.Region(*r) for r in regions])
            check(L.lib().vx_volume_build_levels(self.h, arr, len(regions), _stream(stream)))
        else:
            check(L.lib().vx_volume_build_levels(self.h, None, 0, _stream(stream)))

    def upload_region(self, host: np.ndarray, regions, stream=None):
        host = np.ascontiguousarray(host, np.uint8)
        arr = (L.Region * len(regions))(*[L.Region(*r) for r in regions])
        check(L.lib().vx_volume_upload_region(self.h, _hp(host), arr, len(regions), _stream(stream)))

    def download(self, stream=None) -> np.ndarray:
        sx, sy, sz = self.dims
        out = np.empty((sz, sy, sx), np.uint8)
        check(L.lib().vx_volume_download(self.h, _hp(out), out.nbytes, _stream(stream)))
        return out

    def download_level(self, level: int, stream=None) -> np.ndarray:
        total = C.c_size_t()
        check(L.lib().vx_volume_download_level(self.h, level, None, 0, C.byref(total), _stream(stream)))
        out = np.empty(total.value, np.uint64)
        check(L.lib().vx_volume_download_level(self.h, level, _hp(out), out.size, None, _stream(stream)))
        return out


def view_struct(vd: dict) -> L.View:
    v = L.View()
    for k in ("last_view", "view", "inv_view", "proj", "inv_proj"):
        setattr(v, k, L.f16(*np.asarray(vd[k], np.float32).reshape(16)))
    v.res = (C.c_float * 2)(*[float(x) for x in vd["res"]])
    v.ires = (C.c_float * 2)(*[float(x) for x in vd["ires"]])
    v.cam_pos = (C.c_float * 3)(*[float(x) for x in vd["cam_pos"]])
    v.jitter = (C.c_float * 2)(*[float(x) for x in vd["jitter"]])
    v.frame, v.width, v.height = vd["frame"], vd["width"], vd["height"]
    return v


def trace_params(p: dict) -> L.TraceParams:
    t = L.TraceParams()
    t.shadow_march, t.ao_march, t.spec_march = p["shadow_march"], p["ao_march"], p["spec_march"]
    t.n_shadow, t.n_ao, t.n_spec = p["n_shadow"], p["n_ao"], p["n_spec"]
    t.shadow_maxt, t.ao_maxt, t.spec_maxt = p["shadow_maxt"], p["ao_maxt"], p["spec_maxt"]
    t.light_dir = (C.c_float * 3)(*[float(x) for x in p["light_dir"]])
    t.seed = p["seed"] & 0xFFFFFFFF
    t.frame = p["frame"]
    t.jitter_shadow = p["jitter_shadow"]
    return t


class GBuffer:
    def __init__(self, width, height, device="cuda"):
        n = width * height
        self.width, self.height = width, height
        self.pos = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.normal = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.material = torch.zeros(n, dtype=torch.uint8, device=device)

    def struct(self) -> L.GBuf:
        return L.GBuf(_ptr(self.pos), _ptr(self.normal), _ptr(self.material))


class GBufferExt:
    """Optional geometry-pass outputs: motion (2 per pixel), D24 depth read back
    as float, palette index."""

    def __init__(self, width, height, device="cuda"):
        n = width * height
        self.motion = torch.zeros((n, 2), dtype=torch.float32, device=device)
        self.depth = torch.zeros(n, dtype=torch.float32, device=device)
        self.palette = torch.zeros(n, dtype=torch.int16, device=device)

    def struct(self) -> L.GBufExt:
        return L.GBufExt(_ptr(self.motion), _ptr(self.depth), _ptr(self.palette))


def make_vox_cmd(model: "Model", world, last_world=None, volume_rid=1, palette=0) -> L.VoxCmd:
    c = L.VoxCmd()
    c.model = model.h
    c.world = L.f16(*np.asarray(world, np.float32).reshape(16))
    lw = world if last_world is None else last_world
    c.last_world = L.f16(*np.asarray(lw, np.float32).reshape(16))
    c.volume_rid, c.palette = int(volume_rid), int(palette)
    return c


def primary_instances(ctx: Context, cmds, vd: dict, gbuf: GBuffer, ext: GBufferExt = None, stream=None):
    """The geometry pass over voxel draws (list of L.VoxCmd) in draw order."""
    arr = (L.VoxCmd * max(len(cmds), 1))(*cmds)
    v = view_struct(vd); g = gbuf.struct()
    e = ext.struct() if ext is not None else None
    check(L.lib().vx_primary_instances(ctx.h, arr, len(cmds), C.byref(v), C.byref(g),
                                       C.byref(e) if e is not None else None, _stream(stream)))


def light_taa(ctx: Context, vd: dict, light, last, out, depth, normal, material, palette, motion=None, seed=0,
              stream=None):
    """LightTAA: out (npx x 4) from this frame's light (npx x 3), last frame's
    output (npx x 4), the G-buffer's depth / normal / material and motion."""
    io = L.TaaIO(_ptr(light), _ptr(last), _ptr(motion), _ptr(depth), _ptr(normal), _ptr(material), _ptr(palette),
                 _ptr(out))
    v = view_struct(vd)
    check(L.lib().vx_light_taa(ctx.h, C.byref(v), C.byref(io), seed & 0xFFFFFFFF, _stream(stream)))


class Visibility:
    def __init__(self, width, height, fetches=False, device="cuda"):
        n = width * height
        self.shadow = torch.zeros(n, dtype=torch.uint8, device=device)
        self.ao = torch.zeros(n, dtype=torch.float32, device=device)
        self.spec = torch.zeros(n, dtype=torch.float32, device=device)
        self.fetches = torch.zeros(n, dtype=torch.int32, device=device) if fetches else None

    def struct(self) -> L.TraceOut:
        return L.TraceOut(_ptr(self.shadow), _ptr(self.ao), _ptr(self.spec), _ptr(self.fetches))


def primary(ctx: Context, model: Model, vd: dict, gbuf: GBuffer, stream=None):
    v = view_struct(vd)
    g = gbuf.struct()
    check(L.lib().vx_primary(ctx.h, model.h, C.byref(v), C.byref(g), _stream(stream)))


def trace(ctx: Context, vol: Volume, vd: dict, gbuf: GBuffer, palette: torch.Tensor, params: dict,
          vis: Visibility, stream=None):
    v = view_struct(vd)
    g = gbuf.struct()
    tp = trace_params(params)
    o = vis.struct()
    check(L.lib().vx_trace(ctx.h, vol.h, C.byref(v), C.byref(g), _ptr(palette), C.byref(tp), C.byref(o),
                           _stream(stream)))


TRACE_HIT_DTYPE = np.dtype([("term", "<i4"), ("steps", "<i4"), ("fetches", "<i4"), ("voxel", "<i4", 3),
                            ("t", "<f4"), ("normal", "<f4", 3)])


def trace_hits_device(ctx: Context, vol: Volume, vd: dict, gbuf: GBuffer, palette: torch.Tensor, params: dict,
                      vis: Visibility, stream=None) -> torch.Tensor:
    """vx_trace_hits: the DDA trace plus every ray's hit record, as a device
    int32 tensor of (width * height * rays per pixel, 10) words in
    TRACE_HIT_DTYPE layout (term, steps, fetches, voxel xyz, t, normal xyz)."""
    nr = params["n_shadow"] + params["n_ao"] + params["n_spec"]
    npx = vd["width"] * vd["height"]
    hits = torch.empty((npx * nr * TRACE_HIT_DTYPE.itemsize // 8,), dtype=torch.int64, device="cuda")
    v = view_struct(vd)
    g = gbuf.struct()
    tp = trace_params(params)
    o = vis.struct()
    check(L.lib().vx_trace_hits(ctx.h, vol.h, C.byref(v), C.byref(g), _ptr(palette), C.byref(tp), C.byref(o),
                                _ptr(hits), _stream(stream)))
    return hits.view(torch.int32).view(npx * nr, TRACE_HIT_DTYPE.itemsize // 4)


def trace_hits(ctx: Context, vol: Volume, vd: dict, gbuf: GBuffer, palette: torch.Tensor, params: dict,
               vis: Visibility, stream=None) -> np.ndarray:
    """vx_trace_hits with the records on the host, shaped (width * height,
    rays per pixel)."""
    nr = params["n_shadow"] + params["n_ao"] + params["n_spec"]
    hits = trace_hits_device(ctx, vol, vd, gbuf, palette, params, vis, stream)
    return hits.cpu().numpy().view(TRACE_HIT_DTYPE).reshape(-1, nr)


RAY_RESULT_DTYPE = np.dtype([("hit", "<i4"), ("term", "<i4"), ("steps", "<i4"), ("fetches", "<i4"),
                             ("voxel", "<i4", 3), ("t", "<f4"), ("hitpos", "<f4", 3), ("normal", "<f4", 3)])


def trace_rays(ctx: Context, vol: Volume, rays: torch.Tensor, march: int, stream=None) -> np.ndarray:
    """rays: (n, 7) float32 device tensor (origin3, dir3, maxt). Returns host records."""
    n = rays.shape[0]
    out = torch.zeros((n, RAY_RESULT_DTYPE.itemsize // 4), dtype=torch.int32, device=rays.device)
    check(L.lib().vx_trace_rays(ctx.h, vol.h, _ptr(rays), n, march, _ptr(out), _stream(stream)))
    return out.cpu().numpy().view(RAY_RESULT_DTYPE).reshape(n)


class Batch:
    """vx_batch: many (volume, view, G-buffer, outputs) items in one graph-captured launch."""

    def __init__(self, ctx: Context, items, palette: torch.Tensor, params: dict):
        """items: list of (Volume, view dict, GBuffer, Visibility)."""
        arr = self._items(items)
        tp = trace_params(params)
        h = C.c_void_p()
        check(L.lib().vx_batch_create(ctx.h, arr, len(items), _ptr(palette), C.byref(tp), C.byref(h)))
        self.h, self.ctx, self.items, self.palette = h, ctx, items, palette

    @staticmethod
    def _items(items):
        arr = (L.BatchItem * len(items))()
        for k, (vol, vd, g, vis) in enumerate(items):
            arr[k].volume = vol.h
            arr[k].view = view_struct(vd)
            arr[k].gbuf = g.struct()
            arr[k].out = vis.struct()
        return arr

    def update(self, items):
        """vx_batch_update: new items (any count and view sizes), same palette
        and ray set-up; the launch is captured again."""
        arr = self._items(items)
        self.items = items
        check(L.lib().vx_batch_update(self.h, arr, len(items)))

    def info(self):
        """(item count, graph captures so far)."""
        n, caps = C.c_uint32(), C.c_uint32()
        check(L.lib().vx_batch_info(self.h, C.byref(n), C.byref(caps)))
        return n.value, caps.value

    def launch(self, stream=None):
        check(L.lib().vx_batch_launch(self.h, _stream(stream)))

    def launch_direct(self, stream=None):
        check(L.lib().vx_batch_launch_direct(self.h, _stream(stream)))

    def close(self):
        if self.h:
            L.lib().vx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _light(light):
    return (C.c_float * 6)(*[float(x) for x in np.asarray(light, np.float32).reshape(6)])


def shade_forward(ctx, vd, gbuf, vis, palette, light, rgb, ssao=None, stream=None):
    v = view_struct(vd); g = gbuf.struct(); o = vis.struct()
    check(L.lib().vx_shade_forward(ctx.h, C.byref(v), C.byref(g), C.byref(o), _ptr(ssao), _ptr(palette),
                                   _light(light), _ptr(rgb), _stream(stream)))


def shade_backward(ctx, vd, gbuf, vis, palette, light, dl_drgb, dpalette, dlight, ssao=None, stream=None):
    v = view_struct(vd); g = gbuf.struct(); o = vis.struct()
    check(L.lib().vx_shade_backward(ctx.h, C.byref(v), C.byref(g), C.byref(o), _ptr(ssao), _ptr(palette),
                                    _light(light), _ptr(dl_drgb), _ptr(dpalette), _ptr(dlight), _stream(stream)))


def shade_backward_batch(ctx, items, palette, light, dpalette, dlight, stream=None):
    """items: [(vd, gbuf, vis, dl_drgb, ssao or None)]; gradients of all views
    accumulated in one launch."""
    arr = (L.ShadeItem * len(items))()
    for k, (vd, gbuf, vis, dl, ssao) in enumerate(items):
        arr[k] = L.ShadeItem(view_struct(vd), gbuf.struct(), vis.struct(), _ptr(ssao), _ptr(dl))
    check(L.lib().vx_shade_backward_batch(ctx.h, arr, len(items), _ptr(palette), _light(light), _ptr(dpalette),
                                          _ptr(dlight), _stream(stream)))


LIGHT_DTYPE = np.dtype([("type", np.int32), ("position", np.float32, 3), ("range", np.float32),
                        ("color", np.float32, 3), ("attenuation", np.float32), ("direction", np.float32, 3),
                        ("angle", np.float32), ("angle_attenuation", np.float32), ("pad", np.int32)])


def lights_to_device(lights: np.ndarray) -> torch.Tensor:
    """LIGHT_DTYPE array -> device buffer of vx_light (64 B each)."""
    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    return torch.from_numpy(lights.view(np.uint8).copy()).cuda()


def light_pass(ctx, vol, vd, gbuf, palette, lights_dev, nlights, rgb, point_march=L.MARCH_SPARSE,
               spot_march=L.MARCH_SUPERSPARSE, seed=0, frame=0, fetches=None, stream=None):
    """Adds the point/spot lights' shadowed direct light to rgb (npx x 3 fp32)."""
    v = view_struct(vd); g = gbuf.struct()
    p = L.LightParams(point_march, spot_march, seed & 0xFFFFFFFF, frame)
    check(L.lib().vx_light_pass(ctx.h, vol.h, C.byref(v), C.byref(g), _ptr(palette),
                                _ptr(lights_dev) if nlights else None, nlights, C.byref(p), _ptr(rgb),
                                _ptr(fetches), _stream(stream)))


# ---------------------------------------------------------------- asset files (host)
def asset_guid(path: str) -> int:
    """AssetGUID of a mod-relative path ("default/FarmHouse/0.v")."""
    b = path.encode()
    return int(L.lib().vx_asset_guid(b, len(b)))


def parse_vox(data: bytes):
    """.v bytes -> ((sx, sy, sz), uint8 array (sz, sy, sx))."""
    buf = np.frombuffer(data, np.uint8)
    d = (C.c_int32 * 3)()
    off = C.c_size_t()
    check(L.lib().vx_vox_parse(_hp(buf) if buf.size else None, buf.size, d, C.byref(off)))
    dims = (d[0], d[1], d[2])
    vox = buf[off.value:off.value + dims[0] * dims[1] * dims[2]].reshape(dims[2], dims[1], dims[0])
    return dims, vox


def parse_palette(data: bytes) -> np.ndarray:
    """.p bytes -> 256 x 6 fp32 palette {albedo rgb, roughness, metallic, emissive}."""
    buf = np.frombuffer(data, np.uint8)
    out = np.empty((256, 6), np.float32)
    check(L.lib().vx_palette_parse(_hp(buf) if buf.size else None, buf.size,
                                   out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def load_palette(path: str) -> np.ndarray:
    out = np.empty((256, 6), np.float32)
    check(L.lib().vx_palette_load(str(path).encode(), out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def load_model(ctx: Context, path: str, stream=None) -> Model:
    """A .v file as a Model with its mips built."""
    h = C.c_void_p()
    check(L.lib().vx_model_load(ctx.h, str(path).encode(), C.byref(h), _stream(stream)))
    m = Model.__new__(Model)
    m.h, m.ctx = h, ctx
    d = (C.c_int32 * 3)()
    check(L.lib().vx_model_dims(h, d))
    m.dims = (d[0], d[1], d[2])
    return m


def _mat(a) -> np.ndarray:
    return np.ctypeslib.as_array(a).astype(np.float32).copy()


class Prefab:
    """A spawned .pf prefab: entities in creation order, the files they
    reference, the rebuild order and the light list of the reference's ECS."""

    def __init__(self, mods_dir: str, path: str):
        h = C.c_void_p()
        check(L.lib().vx_prefab_load(str(mods_dir).encode(), str(path).encode(), C.byref(h)))
        self.h = h
        ne, nf = C.c_uint32(), C.c_uint32()
        check(L.lib().vx_prefab_counts(h, C.byref(ne), C.byref(nf)))
        arr = (L.PrefabEntity * max(ne.value, 1))()
        check(L.lib().vx_prefab_entities(h, arr, ne.value))
        self.entities = list(arr)[:ne.value]
        buf = C.create_string_buffer(4096)
        self.files = []
        for i in range(nf.value):
            check(L.lib().vx_prefab_file(h, i, buf, len(buf)))
            self.files.append(buf.value.decode())

    def close(self):
        if self.h:
            L.lib().vx_prefab_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def world(self, i: int) -> np.ndarray:
        return _mat(self.entities[i].world)

    def rebuild_order(self) -> list:
        n = C.c_uint32()
        idx = (C.c_uint32 * max(len(self.entities), 1))()
        check(L.lib().vx_prefab_rebuild_order(self.h, idx, len(idx), C.byref(n)))
        return [int(idx[k]) for k in range(n.value)]

    def lights(self) -> np.ndarray:
        n = C.c_uint32()
        out = np.zeros(max(len(self.entities), 1), LIGHT_DTYPE)
        check(L.lib().vx_prefab_lights(self.h, out.ctypes.data_as(C.POINTER(L.Light)), out.size, C.byref(n)))
        return out[:n.value].copy()

    def load_models(self, ctx: Context, stream=None) -> dict:
        """Every referenced .v file as a Model, keyed by file index."""
        idx = sorted({e.model for e in self.entities if e.flags & L.ENT_VOX and e.model >= 0})
        return {i: load_model(ctx, self.files[i], stream) for i in idx}

    def vox_cmds(self, models: dict, prev=None) -> list:
        """The geometry pass's draws (L.VoxCmd) in the reference's order; models:
        file index -> Model; prev: per-entity previous world matrices or None."""
        nf = len(self.files)
        arr = (C.c_void_p * max(nf, 1))()
        for fi, m in models.items():
            arr[fi] = m.h.value if hasattr(m.h, "value") else m.h
        pv = None
        if prev is not None:
            pv = np.ascontiguousarray(np.asarray(prev, np.float32).reshape(len(self.entities), 16))
        n = C.c_uint32()
        out = (L.VoxCmd * max(len(self.entities), 1))()
        check(L.lib().vx_prefab_vox_cmds(self.h, arr, nf, pv.ctypes.data_as(C.POINTER(C.c_float)) if pv is not None
                                         else None, out, len(out), C.byref(n)))
        return list(out)[:n.value]

    def instances(self, models: dict, prev=None, order=None) -> list:
        """vx_instance list for Volume.rebuild in `order` (default: the
        reference's rebuild order); prev: per-entity previous world matrices
        (default identity, as for entities spawned this frame). Entities whose
        Vox did not resolve are skipped, as ShadowVoxSystem does."""
        out = []
        for i in (self.rebuild_order() if order is None else order):
            e = self.entities[i]
            if e.model < 0:
                continue
            m = models[e.model]
            p = IDENTITY if prev is None else prev[i]
            out.append(make_instance(m.mip_ptr(0), m.dims, p, self.world(i), tuple(e.pivot)))
        return out


def reflection(ctx: Context, vol: Volume, vd: dict, gbuf: GBuffer, palette: torch.Tensor, light: torch.Tensor,
               depth: torch.Tensor, out: torch.Tensor, t: torch.Tensor = None, sky: torch.Tensor = None,
               march: int = L.MARCH_SPARSE, maxt: float = 256.0, seed: int = 0, frame: int = 0, stream=None):
    """LightReflection: out (npx x 4) = (ambient * F * (1 - roughness), F) from
    the light buffer (npx x 4) and the depth attachment (npx); sky: a
    (6, size, size, 4) cube map or None for the procedural sky."""
    io = L.ReflectionIO(_ptr(light), _ptr(depth), _ptr(out), _ptr(t))
    sb = L.SkyBox(_ptr(sky), 0 if sky is None else int(sky.shape[1]), 0)
    rp = L.ReflectionParams(march, float(maxt), seed & 0xFFFFFFFF, frame)
    v = view_struct(vd)
    g = gbuf.struct()
    check(L.lib().vx_reflection(ctx.h, vol.h, C.byref(v), C.byref(g), _ptr(palette), C.byref(io), C.byref(sb),
                                C.byref(rp), _stream(stream)))


def gbuf_depth(ctx: Context, vd: dict, gbuf: GBuffer, depth: torch.Tensor, volume_rid: int = 0, stream=None):
    """The depth attachment (npx) of a G-buffer written by vx_primary."""
    v = view_struct(vd)
    g = gbuf.struct()
    check(L.lib().vx_gbuf_depth(ctx.h, C.byref(v), C.byref(g), volume_rid, _ptr(depth), _stream(stream)))


def ssao(ctx: Context, vd: dict, gbuf: GBuffer, depth: torch.Tensor, out: torch.Tensor, seed: int = 0, stream=None):
    """calculateOcclusion per pixel into out (npx), from the G-buffer normals
    and the depth attachment."""
    v = view_struct(vd)
    g = gbuf.struct()
    check(L.lib().vx_ssao(ctx.h, C.byref(v), C.byref(g), _ptr(depth), seed & 0xFFFFFFFF, _ptr(out), _stream(stream)))


FAIL_ALLOC, FAIL_LAUNCH = 0, 1


def inject_failure(kind: int, after: int):
    """vx_debug_inject_failure: make the device allocation (FAIL_ALLOC) or
    kernel launch (FAIL_LAUNCH) numbered `after` (0 = the next one) report
    failure once; -1 disarms. The environment variables VX_FAIL_ALLOC and
    VX_FAIL_LAUNCH arm the same counters when the library loads."""
    check(L.lib().vx_debug_inject_failure(int(kind), int(after)))
