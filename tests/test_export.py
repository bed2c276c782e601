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


# ---------------------------------------------------------------- OME-ZARR
def test_ome_zarr_roundtrip_levels_and_metadata(tmp_path):
    rng = np.random.default_rng(5)
    vol = rng.uniform(0, 100, (20, 36, 50)).astype(np.float32)
    p = str(tmp_path / "psi.ome.zarr")
    for comp in ("raw", "gzip", "zlib"):
        export.save_ome_zarr(vol, p, voxel_size_zyx=(2.0, 0.5, 0.5), chunks_zyx=(8, 16, 16),
                             downsampling_zyx=((1, 1, 1), (2, 2, 2), (4, 4, 2)), compression=comp)
        back, attrs = export.read_ome_zarr(p, 0)
        np.testing.assert_array_equal(back, vol)
        # level 1: 2x2x2 means (even sizes: plain reshape means)
        l1, _ = export.read_ome_zarr(p, 1)
        want = vol.astype(np.float64).reshape(10, 2, 18, 2, 25, 2).mean((1, 3, 5))
        np.testing.assert_allclose(l1, want, rtol=1e-6)
        # level 2 from level 1 by (2, 2, 1): z 10 -> 5, y 18 -> 9, x 25 -> 25
        l2, _ = export.read_ome_zarr(p, 2)
        assert l2.shape == (5, 9, 25)
        np.testing.assert_allclose(l2, l1.astype(np.float64).reshape(5, 2, 9, 2, 25).mean((1, 3)), rtol=1e-6)
    ms = attrs["multiscales"][0]
    assert ms["version"] == "0.4"
    assert [a["name"] for a in ms["axes"]] == ["t", "c", "z", "y", "x"]
    assert [a["type"] for a in ms["axes"]] == ["time", "channel", "space", "space", "space"]
    assert [d["path"] for d in ms["datasets"]] == ["0", "1", "2"]
    sc = [d["coordinateTransformations"][0]["scale"] for d in ms["datasets"]]
    tr = [d["coordinateTransformations"][1]["translation"] for d in ms["datasets"]]
    assert sc[0] == [1, 1, 2.0, 0.5, 0.5] and sc[1] == [1, 1, 4.0, 1.0, 1.0] and sc[2] == [1, 1, 8.0, 2.0, 1.0]
    assert tr[0] == [0, 0, 0, 0, 0] and tr[1] == [0, 0, 1.0, 0.25, 0.25] and tr[2] == [0, 0, 3.0, 0.75, 0.25]
    meta = json.load(open(os.path.join(p, "0", ".zarray")))
    assert meta["shape"] == [1, 1, 20, 36, 50] and meta["chunks"] == [1, 1, 8, 16, 16]
    assert meta["dtype"] == "<f4" and meta["dimension_separator"] == "/" and meta["order"] == "C"
    assert json.load(open(os.path.join(p, ".zgroup"))) == {"zarr_format": 2}


def test_ome_zarr_chunk_layout_by_hand(tmp_path):
    # an edge chunk decoded by hand: key t/c/z/y/x chunk indices, whole chunk of C-order little-endian floats,
    # the part outside the volume filled with 0
    vol = np.arange(5 * 6 * 7, dtype=np.float32).reshape(5, 6, 7)
    p = str(tmp_path / "a.zarr")
    export.save_ome_zarr(vol, p, chunks_zyx=(4, 4, 4), downsampling_zyx=((1, 1, 1),), compression="raw")
    raw = open(os.path.join(p, "0", "0", "0", "1", "1", "1"), "rb").read()
    chunk = np.frombuffer(raw, "<f4").reshape(4, 4, 4)
    np.testing.assert_array_equal(chunk[:1, :2, :3], vol[4:, 4:, 4:])
    assert not chunk[1:].any() and not chunk[:, 2:].any() and not chunk[:, :, 3:].any()
    n = sum(len(f) for _, _, f in os.walk(os.path.join(p, "0"))) - 1   # chunk files (without .zarray)
    assert n == 2 * 2 * 2


def test_downsample_mean_odd_edges():
    v = np.arange(5, dtype=np.float32).reshape(1, 1, 5)
    np.testing.assert_allclose(export.downsample_mean(v, (1, 1, 2))[0, 0], [0.5, 2.5, 4.0])
