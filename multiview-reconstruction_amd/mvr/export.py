"""Writing the deconvolved volume: ImageJ TIFF stacks and N5 datasets (host I/O).

Mirrors the exporters the reference's deconvolution plug-in hands its result to
(src/main/java/net/preibisch/mvrecon/fiji/plugin/Image_Deconvolution.java:290-299 -> ImgExport):
  * save_tiff_stack: process/export/Save3dTIFF.java:40-110 (one uncompressed 32-bit float TIFF stack per
    result, ImageJ description with images/slices, unit and voxel spacing; file name <prefix><title>.tif);
  * save_n5: process/export/ExportN5Api.java (an N5 dataset: attributes.json with dimensions, blockSize,
    dataType and compression; one file per block at <dataset>/<i>/<j>/<k> holding a big-endian header
    {mode, number of dimensions, block size} and the big-endian block data, x fastest; raw or gzip).
  * save_ome_zarr: the OME-ZARR branch of ExportN5Api.exportImage (process/export/ExportN5Api.java:219-270,
    365-385, metadata OMEZarrAttibutes.createOMEZarrMetadata, fiji/spimdata/imgloaders/OMEZarrAttibutes.java):
    one zarr v2 group with a 5D array per resolution level (t, c, z, y, x; levels "0", "1", ...), chunks
    under "/"-separated keys, and OME-NGFF 0.4 "multiscales" metadata (axes time/channel/space, per level a
    scale of voxel size x downsampling factor and a half-pixel translation).
Volumes are numpy arrays indexed [z, y, x].  Readers for all formats are included for round trips.
PSFs are stored like fiji/spimdata/pointspreadfunctions/PointSpreadFunction.java:46-50, 123-177: one N5
dataset psf_t<timepoint>_v<setup> per view in <base>/psf.n5 (128^3 blocks, gzip level 1); legacy .tif
files in <base>/psf are read as well (save_psf / load_psf).
"""
from __future__ import annotations

import gzip
import json
import os
import struct

import numpy as np

# ---------------------------------------------------------------- TIFF (ImageJ stack)
_TAGS = {"width": 256, "length": 257, "bits": 258, "compression": 259, "photometric": 262, "description": 270,
         "strip_offsets": 273, "samples": 277, "rows_per_strip": 278, "strip_bytes": 279, "x_res": 282,
         "y_res": 283, "res_unit": 296, "sample_format": 339}


def imagej_description(nz, spacing_zyx=(1.0, 1.0, 1.0), unit="pixel", vmin=None, vmax=None):
    lines = ["ImageJ=1.54f", f"images={nz}", f"slices={nz}", f"unit={unit}", f"spacing={spacing_zyx[0]}",
             "loop=false"]
    if vmin is not None and vmax is not None:
        lines += [f"min={vmin}", f"max={vmax}"]
    return "\n".join(lines) + "\n\x00"


def save_tiff_stack(vol, path, spacing_zyx=(1.0, 1.0, 1.0), unit="pixel"):
    """Uncompressed big-endian 32-bit float TIFF stack, one IFD per z plane, ImageJ description in the first."""
    vol = np.ascontiguousarray(vol, np.float32)
    if vol.ndim == 2:
        vol = vol[None]
    nz, ny, nx = vol.shape
    desc = imagej_description(nz, spacing_zyx, unit, float(vol.min()), float(vol.max())).encode("latin-1")
    plane = nx * ny * 4
    data_off = 8
    desc_off = data_off + nz * plane
    res_off = desc_off + len(desc)
    ifd_off = res_off + 16
    ntags = 13
    ifd_size = 2 + ntags * 12 + 4
    with open(path, "wb") as f:
        f.write(b"MM" + struct.pack(">HI", 42, ifd_off))
        f.write(vol.astype(">f4").tobytes())
        f.write(desc)
        xr = 1.0 / spacing_zyx[2] if spacing_zyx[2] else 1.0
        yr = 1.0 / spacing_zyx[1] if spacing_zyx[1] else 1.0
        f.write(struct.pack(">IIII", int(round(xr * 1e6)), 1000000, int(round(yr * 1e6)), 1000000))
        for z in range(nz):
            entries = [
                (_TAGS["width"], 4, 1, nx), (_TAGS["length"], 4, 1, ny), (_TAGS["bits"], 3, 1, 32),
                (_TAGS["compression"], 3, 1, 1), (_TAGS["photometric"], 3, 1, 1),
                (_TAGS["description"], 2, len(desc), desc_off) if z == 0 else (_TAGS["description"], 2, 1, 0),
                (_TAGS["strip_offsets"], 4, 1, data_off + z * plane), (_TAGS["samples"], 3, 1, 1),
                (_TAGS["rows_per_strip"], 4, 1, ny), (_TAGS["strip_bytes"], 4, 1, plane),
                (_TAGS["x_res"], 5, 1, res_off), (_TAGS["y_res"], 5, 1, res_off + 8),
                (_TAGS["sample_format"], 3, 1, 3)]
            nxt = ifd_off + (z + 1) * ifd_size if z + 1 < nz else 0
            f.write(struct.pack(">H", ntags))
            for tag, typ, cnt, val in entries:
                if typ == 3:  # SHORT, left-justified in the value field
                    f.write(struct.pack(">HHIHH", tag, typ, cnt, val, 0))
                else:
                    f.write(struct.pack(">HHII", tag, typ, cnt, val))
            f.write(struct.pack(">I", nxt))
    return path


def read_tiff_stack(path):
    """Reads uncompressed 32-bit float TIFF stacks (either byte order, one strip or several per plane)."""
    with open(path, "rb") as f:
        buf = f.read()
    bo = ">" if buf[:2] == b"MM" else "<"
    (magic,) = struct.unpack(bo + "H", buf[2:4])
    assert magic == 42, "not a TIFF file"
    (off,) = struct.unpack(bo + "I", buf[4:8])
    planes = []
    description = None
    while off:
        (n,) = struct.unpack(bo + "H", buf[off:off + 2])
        tags = {}
        for k in range(n):
            e = buf[off + 2 + 12 * k: off + 14 + 12 * k]
            tag, typ, cnt = struct.unpack(bo + "HHI", e[:8])
            if typ == 3 and cnt == 1:
                (val,) = struct.unpack(bo + "H", e[8:10])
            elif typ == 3 and cnt == 2:
                val = list(struct.unpack(bo + "HH", e[8:12]))
            else:
                (val,) = struct.unpack(bo + "I", e[8:12])
            tags[tag] = (typ, cnt, val)
        w, h = tags[256][2], tags[257][2]
        assert tags[258][2] == 32 and tags.get(339, (0, 0, 1))[2] == 3 and tags.get(259, (0, 0, 1))[2] == 1
        if 270 in tags and description is None and tags[270][1] > 4:
            t = tags[270]
            description = buf[t[2]:t[2] + t[1]].rstrip(b"\x00").decode("latin-1")
        so, sb = tags[273], tags[279]
        if so[1] == 1:
            offs, cnts = [so[2]], [sb[2]]
        else:
            offs = list(struct.unpack(bo + "I" * so[1], buf[so[2]:so[2] + 4 * so[1]]))
            cnts = list(struct.unpack(bo + "I" * sb[1], buf[sb[2]:sb[2] + 4 * sb[1]]))
        raw = b"".join(buf[o:o + c] for o, c in zip(offs, cnts))
        planes.append(np.frombuffer(raw, dtype=bo + "f4").reshape(h, w))
        (off,) = struct.unpack(bo + "I", buf[off + 2 + 12 * n: off + 6 + 12 * n])
    return np.stack(planes).astype(np.float32), description


# ---------------------------------------------------------------- N5
def save_n5(vol, root, dataset, block_size_xyz=(64, 64, 64), compression="raw", attributes=None, level=-1):
    """Writes vol [z, y, x] as float32 N5 dataset `dataset` under `root` (gzip at `level`, -1 = default)."""
    vol = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = vol.shape
    os.makedirs(os.path.join(root, dataset), exist_ok=True)
    rattr = os.path.join(root, "attributes.json")
    if not os.path.exists(rattr):
        with open(rattr, "w") as f:
            json.dump({"n5": "4.0.0"}, f)
    comp = {"type": "raw"} if compression == "raw" else {"type": "gzip", "useZlib": False, "level": int(level)}
    attrs = {"dimensions": [nx, ny, nz], "blockSize": list(block_size_xyz), "dataType": "float32",
             "compression": comp}
    if attributes:
        attrs.update(attributes)
    with open(os.path.join(root, dataset, "attributes.json"), "w") as f:
        json.dump(attrs, f)
    bx, by, bz = block_size_xyz
    for k in range(0, nz, bz):
        for j in range(0, ny, by):
            for i in range(0, nx, bx):
                blk = vol[k:k + bz, j:j + by, i:i + bx]
                dz, dy, dx = blk.shape
                header = struct.pack(">HHIII", 0, 3, dx, dy, dz)
                data = np.ascontiguousarray(blk).astype(">f4").tobytes()
                if compression != "raw":
                    data = gzip.compress(data, compresslevel=9 if level < 0 else int(level))
                d = os.path.join(root, dataset, str(i // bx), str(j // by))
                os.makedirs(d, exist_ok=True)
                with open(os.path.join(d, str(k // bz)), "wb") as f:
                    f.write(header + data)
    return os.path.join(root, dataset)


def read_n5(root, dataset):
    with open(os.path.join(root, dataset, "attributes.json")) as f:
        attrs = json.load(f)
    nx, ny, nz = attrs["dimensions"]
    bx, by, bz = attrs["blockSize"]
    assert attrs["dataType"] == "float32"
    gz = attrs["compression"]["type"] == "gzip"
    out = np.zeros((nz, ny, nx), np.float32)
    for k in range(0, nz, bz):
        for j in range(0, ny, by):
            for i in range(0, nx, bx):
                p = os.path.join(root, dataset, str(i // bx), str(j // by), str(k // bz))
                if not os.path.exists(p):
                    continue
                with open(p, "rb") as f:
                    raw = f.read()
                mode, nd = struct.unpack(">HH", raw[:4])
                dims = struct.unpack(">" + "I" * nd, raw[4:4 + 4 * nd])
                body = raw[4 + 4 * nd:]
                if gz:
                    body = gzip.decompress(body)
                blk = np.frombuffer(body, ">f4").reshape(dims[2], dims[1], dims[0])
                out[k:k + dims[2], j:j + dims[1], i:i + dims[0]] = blk
    return out, attrs


# ---------------------------------------------------------------- PSFs of a dataset
PSF_N5 = "psf.n5"      # PointSpreadFunction.subPath
PSF_TIF_DIR = "psf"    # PointSpreadFunction.subDirFile (legacy .tif files)


def psf_file_name(timepoint, setup):
    """PointSpreadFunction.createPSFFilePath: psf_t<timepoint>_v<setup>."""
    return f"psf_t{int(timepoint)}_v{int(setup)}"


def save_psf(base_dir, timepoint, setup, psf):
    """PointSpreadFunction.save: the PSF as dataset psf_t<tp>_v<setup> of <base_dir>/psf.n5, replacing an
    existing one, 128^3 blocks, gzip level 1.  Returns the dataset name (the file the XML records)."""
    import shutil
    name = psf_file_name(timepoint, setup)
    root = os.path.join(base_dir, PSF_N5)
    if os.path.isdir(os.path.join(root, name)):
        shutil.rmtree(os.path.join(root, name))
    save_n5(psf, root, name, (128, 128, 128), "gzip", level=1)
    return name


def load_psf(base_dir, file):
    """PointSpreadFunction.load: a legacy .tif from <base_dir>/psf, otherwise dataset `file` of <base_dir>/psf.n5."""
    if file.endswith(".tif"):
        return read_tiff_stack(os.path.join(base_dir, PSF_TIF_DIR, file))[0]
    return read_n5(os.path.join(base_dir, PSF_N5), file)[0]


# ---------------------------------------------------------------- OME-ZARR (zarr v2 + OME-NGFF 0.4 multiscales)
def downsample_mean(vol, factors_zyx):
    """Block mean by integer factors per axis; an edge block shorter than its factor averages what it holds.
    (The reference's pyramid uses imglib2-algorithm's Downsample, version managed by pom-scijava 43.0.0, for the
    same 2^k averaging; its rounding at odd edges is not pinned here.)"""
    out = np.asarray(vol, np.float64)
    for axis, f in enumerate(factors_zyx):
        f = int(f)
        if f == 1:
            continue
        n = out.shape[axis]
        m = -(-n // f)
        pad = m * f - n
        if pad:
            widths = [(0, 0)] * out.ndim
            widths[axis] = (0, pad)
            summed = np.pad(out, widths).reshape(out.shape[:axis] + (m, f) + out.shape[axis + 1:]).sum(axis + 1)
            counts = np.full(m, float(f))
            counts[-1] = f - pad
            shape = [1] * out.ndim
            shape[axis] = m
            out = summed / counts.reshape(shape)
        else:
            out = out.reshape(out.shape[:axis] + (m, f) + out.shape[axis + 1:]).mean(axis + 1)
    return out.astype(np.float32)


def ome_zarr_multiscales(name, voxel_size_zyx, unit, downsampling_zyx, levels_path=None):
    """The "multiscales" attribute (OME-NGFF 0.4) of a 5D t, c, z, y, x image: per level a scale of voxel size x
    downsampling factor and a translation of (factor - 1) / 2 voxels (the mipmap transform of a half-pixel
    downsampling), 1 and 0 for t and c."""
    axes = [{"name": "t", "type": "time", "unit": "millisecond"}, {"name": "c", "type": "channel"}] + \
           [{"name": a, "type": "space", "unit": unit} for a in ("z", "y", "x")]
    datasets = []
    for s, f in enumerate(downsampling_zyx):
        scale = [1.0, 1.0] + [float(v) * float(k) for v, k in zip(voxel_size_zyx, f)]
        trans = [0.0, 0.0] + [float(v) * (float(k) - 1.0) / 2.0 for v, k in zip(voxel_size_zyx, f)]
        datasets.append({"path": (levels_path or str)(s),
                         "coordinateTransformations": [{"type": "scale", "scale": scale},
                                                       {"type": "translation", "translation": trans}]})
    return [{"version": "0.4", "name": name, "axes": axes, "datasets": datasets,
             "coordinateTransformations": [{"type": "scale", "scale": [1.0] * 5}]}]


def _zarr_compressor(compression, level):
    if compression == "raw":
        return None
    if compression == "gzip":
        return {"id": "gzip", "level": 6 if level < 0 else int(level)}
    if compression == "zlib":
        return {"id": "zlib", "level": 6 if level < 0 else int(level)}
    raise ValueError(f"unknown compression {compression}")


def _zarr_encode(data, comp):
    import zlib
    if comp is None:
        return data
    if comp["id"] == "gzip":
        return gzip.compress(data, compresslevel=comp["level"])
    return zlib.compress(data, comp["level"])


def _zarr_decode(raw, comp):
    import zlib
    if comp is None:
        return raw
    if comp["id"] == "gzip":
        return gzip.decompress(raw)
    if comp["id"] == "zlib":
        return zlib.decompress(raw)
    raise ValueError(f"unsupported compressor {comp}")


def save_ome_zarr(vol, path, voxel_size_zyx=(1.0, 1.0, 1.0), unit="micrometer", name="deconvolved",
                  chunks_zyx=(64, 128, 128), downsampling_zyx=((1, 1, 1), (2, 2, 2), (4, 4, 4)), compression="gzip",
                  level=-1):
    """vol [z, y, x] as an OME-ZARR image at `path`: a zarr v2 group whose arrays "0", "1", ... hold the resolution
    levels as float32 5D arrays (1, 1, z, y, x), chunked (1, 1, chunks_zyx) with "/" separated chunk keys.
    downsampling_zyx: the absolute factor of every level (the first must be 1, each a multiple of the previous).
    The reference's default block size is 128 x 128 x 64 (x, y, z; ExportN5Api.java:110-112)."""
    vol = np.ascontiguousarray(vol, np.float32)
    if tuple(downsampling_zyx[0]) != (1, 1, 1):
        raise ValueError("level 0 is the full resolution")
    comp = _zarr_compressor(compression, level)
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, ".zgroup"), "w") as f:
        json.dump({"zarr_format": 2}, f)
    with open(os.path.join(path, ".zattrs"), "w") as f:
        json.dump({"multiscales": ome_zarr_multiscales(name, voxel_size_zyx, unit, downsampling_zyx)}, f, indent=1)
    cur, prev = vol, (1, 1, 1)
    for s, fac in enumerate(downsampling_zyx):
        rel = tuple(int(a) // int(b) for a, b in zip(fac, prev))
        if any(int(a) % int(b) for a, b in zip(fac, prev)):
            raise ValueError("each downsampling factor must be a multiple of the previous level's")
        cur = downsample_mean(cur, rel) if s else cur
        prev = tuple(fac)
        _write_zarr_array(cur, os.path.join(path, str(s)), chunks_zyx, comp)
    return path


def _write_zarr_array(vol, adir, chunks_zyx, comp):
    nz, ny, nx = vol.shape
    cz, cy, cx = chunks_zyx
    os.makedirs(adir, exist_ok=True)
    meta = {"zarr_format": 2, "shape": [1, 1, nz, ny, nx], "chunks": [1, 1, cz, cy, cx], "dtype": "<f4",
            "compressor": comp, "fill_value": 0.0, "order": "C", "filters": None, "dimension_separator": "/"}
    with open(os.path.join(adir, ".zarray"), "w") as f:
        json.dump(meta, f, indent=1)
    for k in range(0, nz, cz):
        for j in range(0, ny, cy):
            for i in range(0, nx, cx):
                chunk = np.zeros((cz, cy, cx), np.float32)      # edge chunks are stored whole, padded with 0
                blk = vol[k:k + cz, j:j + cy, i:i + cx]
                chunk[:blk.shape[0], :blk.shape[1], :blk.shape[2]] = blk
                d = os.path.join(adir, "0", "0", str(k // cz), str(j // cy))
                os.makedirs(d, exist_ok=True)
                with open(os.path.join(d, str(i // cx)), "wb") as f:
                    f.write(_zarr_encode(chunk.astype("<f4").tobytes(), comp))


def read_ome_zarr(path, level=0):
    """(vol [z, y, x] of resolution level `level`, the group's attributes)."""
    with open(os.path.join(path, ".zattrs")) as f:
        attrs = json.load(f)
    ds = attrs["multiscales"][0]["datasets"][level]["path"]
    adir = os.path.join(path, ds)
    with open(os.path.join(adir, ".zarray")) as f:
        meta = json.load(f)
    assert meta["zarr_format"] == 2 and meta["dtype"] in ("<f4", "|f4") and meta["order"] == "C"
    sep = meta.get("dimension_separator", ".")
    t, c, nz, ny, nx = meta["shape"]
    _, _, cz, cy, cx = meta["chunks"]
    out = np.full((nz, ny, nx), meta["fill_value"] or 0.0, np.float32)
    for k in range(0, nz, cz):
        for j in range(0, ny, cy):
            for i in range(0, nx, cx):
                p = os.path.join(adir, sep.join(["0", "0", str(k // cz), str(j // cy), str(i // cx)]))
                if not os.path.exists(p):
                    continue
                with open(p, "rb") as f:
                    chunk = np.frombuffer(_zarr_decode(f.read(), meta["compressor"]), "<f4").reshape(cz, cy, cx)
                blk = out[k:k + cz, j:j + cy, i:i + cx]
                blk[...] = chunk[:blk.shape[0], :blk.shape[1], :blk.shape[2]]
    return out, attrs
