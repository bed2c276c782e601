"""CPU checks of the C ABI library and of the host-side block logic (no GPU calls)."""
import ctypes
import os

import numpy as np
import pytest

from mvr import decon, hip


def test_library_exports_every_declared_symbol():
    if not os.path.exists(hip.LIB_PATH):
        pytest.fail("libmvr_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(hip.LIB_PATH)
    names = hip.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_declares_reference_replacements():
    names = hip.header_functions()
    for n in ["mvr_get_num_devices", "mvr_convolution3d_fft_inplace", "mvr_block_run_iteration_seq",
              "mvr_block_run_iteration_mul", "mvr_decon_run", "mvr_init_psfs"]:
        assert n in names


def test_python_block_generator_matches_oracle(oracle):
    for img, ks, bs in [((1024, 1024, 1), (16, 32, 1), (384, 384, 1)), ((100, 60, 40), (11, 11, 11), (30, 30, 30)),
                        ((256, 256, 256), (81, 81, 81), (336, 336, 336)), ((70, 33, 20), (9, 5, 3), (20, 20, 20))]:
        a = decon.divide_into_blocks(img, ks, bs)
        b = oracle.divide_into_blocks(img, ks, bs)
        assert len(a) == len(b)
        for x, y in zip(a, b):
            assert x.offset == y["offset"] and x.effective_size == y["effectiveSize"]
            assert x.effective_offset == y["effectiveOffset"] and x.effective_local_offset == y["effectiveLocalOffset"]
    assert decon.divide_into_blocks((50, 50, 50), (81, 81, 81), (80, 80, 80)) is None


@pytest.mark.parametrize("min_req", [1, 8, 100])
def test_python_sorter_matches_oracle(oracle, min_req):
    img, ks, bs = (100, 60, 40), (11, 11, 11), (30, 30, 30)
    blocks = decon.divide_into_blocks(img, ks, bs)
    ours = decon.sort_blocks_by_smallest_footprint(blocks, img, min_req)
    theirs, _ = oracle.sort_blocks(img, ks, bs, min_req)
    assert [[blocks.index(b) for b in batch] for batch in ours] == theirs


def test_copy_paste_roundtrip():
    rng = np.random.default_rng(1)
    psi = rng.normal(size=(17, 19, 23)).astype(np.float32)
    blocks = decon.divide_into_blocks((23, 19, 17), (5, 5, 5), (12, 12, 12))
    out = np.zeros_like(psi)
    for b in blocks:
        decon.paste_block(out, b, decon.copy_block_mirror(psi, b))
    np.testing.assert_array_equal(out, psi)


def test_mirror_copy_law():
    psi = np.arange(5, dtype=np.float32).reshape(1, 1, 5)
    b = decon.Block((9, 1, 1), (-2, 0, 0), (5, 1, 1), (0, 0, 0), (2, 0, 0))
    np.testing.assert_array_equal(decon.copy_block_mirror(psi, b)[0, 0], [2, 1, 0, 1, 2, 3, 4, 3, 2])


def test_argument_errors_are_reported_before_any_device_work():
    # argument checks run on the host before the library touches a device
    import numpy as np
    for args in [((8, 8, 8), 0), ((8, 8, 8), 33), ((0, 8, 8), 1)]:
        with pytest.raises(hip.MvrError):
            hip.Decon(*args)
    h = ctypes.c_void_p()
    assert hip.lib().mvr_decon_create(ctypes.byref(h), 0, hip._dims((8, 8, 8)), 1, 7, 0, 0.0, 1e-4, 8) < 0
    assert "PSF type" in hip.lib().mvr_last_error().decode()
    with pytest.raises(hip.MvrError, match="views"):
        hip.UNet(9)
    img = np.ones((4, 4, 4), np.float32)
    eye = np.eye(3, 4)
    with pytest.raises(hip.MvrError, match="singular"):
        hip.transform_view(img, np.zeros((3, 4)), (0, 0, 0), (4, 4, 4))
    with pytest.raises(hip.MvrError, match="interpolation"):
        hip.transform_view(img, eye, (0, 0, 0), (4, 4, 4), interpolation=2)
    with pytest.raises(hip.MvrError, match="at least one view"):
        hip.fuse_group([], [], (0, 0, 0), (4, 4, 4))
    with pytest.raises(hip.MvrError, match="empty bounding box"):
        hip.fuse_group([img], [eye], (0, 0, 0), (0, 4, 4))


def test_noalloc_refuses_oversized_grids_before_device_work():
    # the scratch layout is checked on the host: a 200^3 volume with a 41^3 kernel needs a 240^3 grid
    fake = ctypes.c_void_p(256)     # never dereferenced: the call ends at the size check
    rc = hip.lib().mvr_fft_convolve_noalloc(0, fake, hip._dims((200, 200, 200)), 0, 0.0, fake, hip._dims((41, 41, 41)),
                                            fake, None)
    assert rc < 0
    msg = hip.lib().mvr_last_error().decode()
    assert "scratch" in msg and "240x240x240" in msg


def test_mirror_copy_fast_path_equals_index_form():
    # the reflect-pad form of Block.copyBlock equals mirror-single indexing of the whole volume, also where a
    # block reaches further than the volume is long (index form) and for one-voxel axes
    rng = np.random.default_rng(3)
    for shp in [(17, 19, 23), (1, 5, 9), (2, 2, 2)]:
        psi = rng.normal(size=shp).astype(np.float32)
        for off in [(-4, -4, -4), (0, 0, 0), (3, 2, 1), (-10, 5, -3), (10, 10, 10), (-30, 0, 0), (20, -2, 15)]:
            for bs in [(12, 12, 12), (31, 27, 25), (5, 5, 5), (40, 8, 9), (1, 1, 1)]:
                b = decon.Block(bs, off, bs, off, (0, 0, 0))
                idx = [decon._mirror_idx(np.arange(bs[d]) + off[d], shp[2 - d]) for d in range(3)]
                np.testing.assert_array_equal(decon.copy_block_mirror(psi, b), psi[np.ix_(idx[2], idx[1], idx[0])])
