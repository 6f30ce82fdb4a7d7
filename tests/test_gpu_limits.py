"""Documented size limits of the C ABI, checked on a device: requests beyond
them return VX_ERR_RANGE before any allocation."""
import pytest

pytestmark = pytest.mark.gpu


def test_march_grid_beyond_32bit_word_indices_is_rejected(ctx):
    """The ray-parallel trace indexes the march grid (one u64 word per 4^3 fine
    voxels, a border word per side) with 32-bit word indices, so a volume whose
    grid reaches 2^32 words is refused at creation: with 4096 x 4096 texel
    slices, 2040 slices fit and 2041 do not (grid 2050 x 2050 x 1023 words)."""
    import torch
    from voxelengine_amd import _lib, api
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for dims in ((4096, 4096, 2041), (4096, 4096, 4096), (4096, 2048, 4096)):
        with pytest.raises(_lib.VXError) as e:
            api.Volume(ctx, *dims, nlevels=2)
        assert e.value.code == _lib.VX_ERR_RANGE, (dims, str(e.value))
        assert "32-bit" in str(e.value)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] >= free0 - (64 << 20)


def test_grid_word_count_formula():
    """The boundary above from the grid's own size rule: ((2 s + 2) >> 2) + 2
    words per axis."""
    g = lambda s: ((2 * s + 2) >> 2) + 2
    assert g(4096) == 2050
    assert g(4096) * g(4096) * g(2040) < 2 ** 32 <= g(4096) * g(4096) * g(2041)
    assert g(4096) * g(2048) * g(4096) >= 2 ** 32
