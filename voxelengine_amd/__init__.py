"""MI355X-native (gfx950) voxel-traversal lighting path of VoxelEngine.

The compute path is the HIP library libvx_amd.so (C ABI in include/vx/vx_api.h);
this package holds its ctypes bindings, a torch-tensor host API, the view and
workload definitions, and the build recipe.
"""
__all__ = ["api", "configs", "views", "build"]
