"""ctypes bindings of the C ABI in include/vx/vx_api.h (libvx_amd.so).

The product path: every call goes to the HIP library. If the library is not
built, loading fails loudly; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# VX_AMD_LIB selects another build of the library (same ABI).
LIB_PATH = os.environ.get("VX_AMD_LIB") or os.path.join(HERE, "libvx_amd.so")

VX_OK = 0
VX_ERR_INVALID_ARGUMENT = -1
VX_ERR_OUT_OF_MEMORY = -2
VX_ERR_DEVICE = -3
VX_ERR_UNSUPPORTED = -4
VX_ERR_RANGE = -5
VX_ERR_STATE = -6

MARCH_DDA, MARCH_SPARSE, MARCH_SUPERSPARSE = 0, 1, 2
INST_UPDATE, INST_DESTROY = 1, 2
LIGHT_POINT, LIGHT_SPOT = 0, 1

P = C.c_void_p
f16 = C.c_float * 16


class Region(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32),
                ("w", C.c_uint32), ("h", C.c_uint32), ("d", C.c_uint32)]


class Instance(C.Structure):
    _fields_ = [("model", P), ("sx", C.c_int32), ("sy", C.c_int32), ("sz", C.c_int32),
                ("prev", f16), ("cur", f16), ("pivot", C.c_float * 3), ("flags", C.c_int32)]


class View(C.Structure):
    _fields_ = [("last_view", f16), ("view", f16), ("inv_view", f16), ("proj", f16), ("inv_proj", f16),
                ("res", C.c_float * 2), ("ires", C.c_float * 2), ("cam_pos", C.c_float * 3),
                ("jitter", C.c_float * 2), ("frame", C.c_int32), ("width", C.c_int32), ("height", C.c_int32)]


class TraceParams(C.Structure):
    _fields_ = [("shadow_march", C.c_int32), ("ao_march", C.c_int32), ("spec_march", C.c_int32),
                ("n_shadow", C.c_int32), ("n_ao", C.c_int32), ("n_spec", C.c_int32),
                ("shadow_maxt", C.c_float), ("ao_maxt", C.c_float), ("spec_maxt", C.c_float),
                ("light_dir", C.c_float * 3), ("seed", C.c_uint32), ("frame", C.c_int32),
                ("jitter_shadow", C.c_int32), ("pad", C.c_int32)]


class GBuf(C.Structure):
    _fields_ = [("pos", P), ("normal", P), ("material", P)]


class TraceOut(C.Structure):
    _fields_ = [("shadow", P), ("ao", P), ("spec", P), ("fetches", P)]


class RayResult(C.Structure):
    _fields_ = [("hit", C.c_int32), ("term", C.c_int32), ("steps", C.c_int32), ("fetches", C.c_int32),
                ("voxel", C.c_int32 * 3), ("t", C.c_float), ("hitpos", C.c_float * 3),
                ("normal", C.c_float * 3)]


class SceneDesc(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("terrain_fill", C.c_float), ("nboxes", C.c_int32),
                ("box_min", C.c_int32), ("box_max", C.c_int32), ("boxes_grounded", C.c_int32)]


class BatchItem(C.Structure):
    _fields_ = [("volume", P), ("view", View), ("gbuf", GBuf), ("out", TraceOut)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", C.c_float * 3), ("range", C.c_float), ("color", C.c_float * 3),
                ("attenuation", C.c_float), ("direction", C.c_float * 3), ("angle", C.c_float),
                ("angle_attenuation", C.c_float), ("pad", C.c_int32)]


class LightParams(C.Structure):
    _fields_ = [("point_march", C.c_int32), ("spot_march", C.c_int32), ("seed", C.c_uint32), ("frame", C.c_int32)]


class ShadeItem(C.Structure):
    _fields_ = [("view", View), ("gbuf", GBuf), ("vis", TraceOut), ("ssao", P), ("dl_drgb", P)]


class TraceTuning(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("refill", C.c_int32),
                ("steps", C.c_int32), ("tile", C.c_int32), ("order", C.c_int32), ("lights", C.c_int32),
                ("chunks", C.c_int32)]


class VoxCmd(C.Structure):
    _fields_ = [("model", P), ("world", C.c_float * 16), ("last_world", C.c_float * 16), ("volume_rid", C.c_int32),
                ("palette", C.c_int32)]


class GBufExt(C.Structure):
    _fields_ = [("motion", P), ("depth", P), ("palette", P)]


class TaaIO(C.Structure):
    _fields_ = [("light", P), ("last", P), ("motion", P), ("depth", P), ("normal", P), ("material", P),
                ("palette", P), ("out", P)]


class PrefabEntity(C.Structure):
    _fields_ = [("id", C.c_int32), ("parent", C.c_int32), ("flags", C.c_int32), ("model", C.c_int32),
                ("palette", C.c_int32), ("light_type", C.c_int32), ("vox_guid", C.c_uint64),
                ("palette_guid", C.c_uint64), ("position", C.c_float * 3), ("rotation", C.c_float * 3),
                ("scale", C.c_float * 3), ("pivot", C.c_float * 3), ("matrix", C.c_float * 16),
                ("world", C.c_float * 16), ("light_color", C.c_float * 3), ("light_intensity", C.c_float),
                ("light_attenuation", C.c_float), ("light_range", C.c_float), ("light_angle", C.c_float),
                ("light_angle_attenuation", C.c_float), ("pad", C.c_int32 * 2)]


ENT_TRANSFORM, ENT_VOX, ENT_LIGHT, ENT_INSTANCE = 1, 2, 4, 8

# Every symbol the header declares, with its signature.
class ReflectionParams(C.Structure):
    _fields_ = [("march", C.c_int32), ("maxt", C.c_float), ("seed", C.c_uint32), ("frame", C.c_int32)]


class SkyBox(C.Structure):
    _fields_ = [("faces", C.c_void_p), ("size", C.c_int32), ("pad", C.c_int32)]


class ReflectionIO(C.Structure):
    _fields_ = [("light", C.c_void_p), ("depth", C.c_void_p), ("out", C.c_void_p), ("t", C.c_void_p)]


SIGNATURES = {
    "vx_api_version": ([], C.c_int32),
    "vx_last_error": ([], C.c_char_p),
    "vx_ctx_create": ([C.c_int32, C.POINTER(P)], C.c_int32),
    "vx_ctx_destroy": ([P], C.c_int32),
    "vx_ctx_set_trace_tuning": ([P, C.POINTER(TraceTuning)], C.c_int32),
    "vx_ctx_get_trace_tuning": ([P, C.POINTER(TraceTuning)], C.c_int32),
    "vx_ctx_profile": ([P, C.c_int32], C.c_int32),
    "vx_ctx_profile_read": ([P, C.POINTER(C.c_double), C.POINTER(C.c_uint32)], C.c_int32),
    "vx_model_create": ([P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(P)], C.c_int32),
    "vx_model_destroy": ([P], C.c_int32),
    "vx_model_dims": ([P, C.POINTER(C.c_int32)], C.c_int32),
    "vx_model_mip_ptr": ([P, C.c_int32, C.POINTER(P)], C.c_int32),
    "vx_model_upload": ([P, P, C.c_size_t, P], C.c_int32),
    "vx_model_download": ([P, C.c_int32, P, C.c_size_t, P], C.c_int32),
    "vx_model_build_mips": ([P, P], C.c_int32),
    "vx_model_generate": ([P, C.POINTER(SceneDesc), P], C.c_int32),
    "vx_volume_create": ([P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(P)], C.c_int32),
    "vx_volume_destroy": ([P], C.c_int32),
    "vx_volume_dims": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int32),
    "vx_volume_clear": ([P, P], C.c_int32),
    "vx_volume_upload_region": ([P, P, C.POINTER(Region), C.c_uint32, P], C.c_int32),
    "vx_volume_download": ([P, P, C.c_size_t, P], C.c_int32),
    "vx_volume_rebuild": ([P, C.POINTER(Instance), C.c_uint32, C.c_int32, C.POINTER(Region),
                           C.POINTER(C.c_uint32), P], C.c_int32),
    "vx_volume_build_levels": ([P, C.POINTER(Region), C.c_uint32, P], C.c_int32),
    "vx_volume_download_level": ([P, C.c_int32, P, C.c_size_t, C.POINTER(C.c_size_t), P], C.c_int32),
    "vx_primary": ([P, P, C.POINTER(View), C.POINTER(GBuf), P], C.c_int32),
    "vx_trace": ([P, P, C.POINTER(View), C.POINTER(GBuf), P, C.POINTER(TraceParams), C.POINTER(TraceOut), P],
                 C.c_int32),
    "vx_debug_inject_failure": ([C.c_int32, C.c_int64], C.c_int32),
    "vx_trace_noalloc": ([P, P, C.POINTER(View), C.POINTER(GBuf), P, C.POINTER(TraceParams), C.POINTER(TraceOut),
                          P], C.c_int32),
    "vx_trace_hits": ([P, P, C.POINTER(View), C.POINTER(GBuf), P, C.POINTER(TraceParams), C.POINTER(TraceOut),
                       P, P], C.c_int32),
    "vx_trace_rays": ([P, P, P, C.c_uint32, C.c_int32, P, P], C.c_int32),
    "vx_batch_create": ([P, C.POINTER(BatchItem), C.c_uint32, P, C.POINTER(TraceParams), C.POINTER(P)], C.c_int32),
    "vx_batch_launch": ([P, P], C.c_int32),
    "vx_batch_update": ([P, C.POINTER(BatchItem), C.c_uint32], C.c_int32),
    "vx_batch_info": ([P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)], C.c_int32),
    "vx_batch_launch_direct": ([P, P], C.c_int32),
    "vx_batch_destroy": ([P], C.c_int32),
    "vx_shade_forward": ([P, C.POINTER(View), C.POINTER(GBuf), C.POINTER(TraceOut), P, P, C.POINTER(C.c_float),
                          P, P], C.c_int32),
    "vx_shade_backward": ([P, C.POINTER(View), C.POINTER(GBuf), C.POINTER(TraceOut), P, P,
                           C.POINTER(C.c_float), P, P, P, P], C.c_int32),
    "vx_light_pass": ([P, P, C.POINTER(View), C.POINTER(GBuf), P, P, C.c_uint32, C.POINTER(LightParams), P, P, P],
                      C.c_int32),
    "vx_shade_backward_batch": ([P, C.POINTER(ShadeItem), C.c_uint32, P, C.POINTER(C.c_float), P, P, P],
                                C.c_int32),
    "vx_asset_guid": ([C.c_char_p, C.c_size_t], C.c_uint64),
    "vx_vox_parse": ([P, C.c_size_t, C.POINTER(C.c_int32), C.POINTER(C.c_size_t)], C.c_int32),
    "vx_model_load": ([P, C.c_char_p, C.POINTER(P), P], C.c_int32),
    "vx_palette_parse": ([P, C.c_size_t, C.POINTER(C.c_float)], C.c_int32),
    "vx_palette_load": ([C.c_char_p, C.POINTER(C.c_float)], C.c_int32),
    "vx_prefab_load": ([C.c_char_p, C.c_char_p, C.POINTER(P)], C.c_int32),
    "vx_prefab_destroy": ([P], C.c_int32),
    "vx_prefab_counts": ([P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)], C.c_int32),
    "vx_prefab_entities": ([P, C.POINTER(PrefabEntity), C.c_uint32], C.c_int32),
    "vx_prefab_file": ([P, C.c_uint32, C.c_char_p, C.c_size_t], C.c_int32),
    "vx_prefab_rebuild_order": ([P, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint32)], C.c_int32),
    "vx_prefab_lights": ([P, C.POINTER(Light), C.c_uint32, C.POINTER(C.c_uint32)], C.c_int32),
    "vx_prefab_vox_cmds": ([P, C.POINTER(P), C.c_uint32, C.POINTER(C.c_float), C.POINTER(VoxCmd), C.c_uint32,
                            C.POINTER(C.c_uint32)], C.c_int32),
    "vx_light_taa": ([P, C.POINTER(View), C.POINTER(TaaIO), C.c_uint32, P], C.c_int32),
    "vx_reflection": ([P, P, C.POINTER(View), C.POINTER(GBuf), P, C.POINTER(ReflectionIO), C.POINTER(SkyBox),
                       C.POINTER(ReflectionParams), P], C.c_int32),
    "vx_gbuf_depth": ([P, C.POINTER(View), C.POINTER(GBuf), C.c_int32, P, P], C.c_int32),
    "vx_ssao": ([P, C.POINTER(View), C.POINTER(GBuf), P, C.c_uint32, P, P], C.c_int32),
    "vx_primary_instances": ([P, C.POINTER(VoxCmd), C.c_uint32, C.POINTER(View), C.POINTER(GBuf),
                              C.POINTER(GBufExt), P], C.c_int32),
}

_lib = None


class VXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vx error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status):
    if status != VX_OK:
        raise VXError(status, lib().vx_last_error().decode(errors="replace"))
    return status
