"""Export of the deconvolved volume: ImageJ TIFF stacks and N5 datasets (round trips, format fields)."""
import json
import os
import struct

import numpy as np

from mvr import export


def test_tiff_stack_roundtrip_and_header(tmp_path):
    rng = np.random.default_rng(1)
    vol = rng.uniform(-5, 1e4, (7, 13, 17)).astype(np.float32)
    p = export.save_tiff_stack(vol, str(tmp_path / "psi.tif"), spacing_zyx=(2.5, 1.0, 1.0), unit="um")
    raw = open(p, "rb").read()
    assert raw[:4] == b"MM\x00\x2a"
    back, desc = export.read_tiff_stack(p)
    np.testing.assert_array_equal(back, vol)
    assert "images=7" in desc and "slices=7" in desc and "spacing=2.5" in desc and "unit=um" in desc
    # first IFD: 13 entries, width/length/bits/sample format as written
    (ifd,) = struct.unpack(">I", raw[4:8])
    (n,) = struct.unpack(">H", raw[ifd:ifd + 2])
    tags = {}
    for k in range(n):
        tag, typ, cnt = struct.unpack(">HHI", raw[ifd + 2 + 12 * k: ifd + 10 + 12 * k])
        tags[tag] = struct.unpack(">H", raw[ifd + 10 + 12 * k: ifd + 12 + 12 * k])[0] if typ == 3 else \
            struct.unpack(">I", raw[ifd + 10 + 12 * k: ifd + 14 + 12 * k])[0]
    assert tags[256] == 17 and tags[257] == 13 and tags[258] == 32 and tags[339] == 3 and tags[259] == 1


def test_tiff_single_plane(tmp_path):
    vol = np.arange(12, dtype=np.float32).reshape(3, 4)
    back, _ = export.read_tiff_stack(export.save_tiff_stack(vol, str(tmp_path / "a.tif")))
    np.testing.assert_array_equal(back[0], vol)


def test_n5_roundtrip_uneven_blocks(tmp_path):
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 1, (9, 20, 33)).astype(np.float32)
    for comp in ("raw", "gzip"):
        root = str(tmp_path / f"out_{comp}.n5")
        export.save_n5(vol, root, "deconvolved", block_size_xyz=(16, 8, 4), compression=comp)
        back, attrs = export.read_n5(root, "deconvolved")
        np.testing.assert_array_equal(back, vol)
        assert attrs["dimensions"] == [33, 20, 9] and attrs["blockSize"] == [16, 8, 4]
        assert attrs["compression"]["type"] == comp
        assert json.load(open(os.path.join(root, "attributes.json")))["n5"].startswith("4")
    # the last block along x holds 33 - 32 = 1 column: header {mode 0, 3 dims, 1 x 8 x 4}
    raw = open(os.path.join(str(tmp_path / "out_raw.n5"), "deconvolved", "2", "0", "0"), "rb").read()
    assert struct.unpack(">HHIII", raw[:16]) == (0, 3, 1, 8, 4)
    assert len(raw) == 16 + 1 * 8 * 4 * 4
    assert struct.unpack(">f", raw[16:20])[0] == vol[0, 0, 32]


def test_tiff_stack_readable_by_pillow(tmp_path):
    PIL = __import__("pytest").importorskip("PIL.Image")
    rng = np.random.default_rng(3)
    vol = rng.uniform(0, 100, (4, 6, 9)).astype(np.float32)
    p = export.save_tiff_stack(vol, str(tmp_path / "p.tif"))
    im = PIL.open(p)
    for z in range(4):
        im.seek(z)
        assert im.size == (9, 6)
        np.testing.assert_array_equal(np.asarray(im, np.float32), vol[z])


def test_psf_storage_n5_and_legacy_tif(tmp_path):
    rng = np.random.default_rng(3)
    psf = rng.uniform(0, 1, (31, 21, 25)).astype(np.float32)
    base = str(tmp_path)
    name = export.save_psf(base, 0, 3, psf)
    assert name == "psf_t0_v3"
    attrs = json.load(open(os.path.join(base, "psf.n5", name, "attributes.json")))
    assert attrs["blockSize"] == [128, 128, 128] and attrs["dimensions"] == [25, 21, 31]
    assert attrs["compression"] == {"type": "gzip", "useZlib": False, "level": 1}
    np.testing.assert_array_equal(export.load_psf(base, name), psf)
    # saving again replaces the dataset
    export.save_psf(base, 0, 3, psf[:5])
    np.testing.assert_array_equal(export.load_psf(base, name), psf[:5])
    # legacy PSF files: <base>/psf/<name>.tif
    os.makedirs(os.path.join(base, "psf"))
    export.save_tiff_stack(psf, os.path.join(base, "psf", "psf_t0_v1.tif"))
    np.testing.assert_array_equal(export.load_psf(base, "psf_t0_v1.tif"), psf)
