"""Seeded synthetic multiview inputs: blob phantom, rotated Gaussian PSFs,
blurred + Poisson-noised views and blending weights.

The generator is shared by the tests, the benchmark and the GPU smoke run so
that every path consumes identical FP32 arrays.  Arrays are indexed [z, y, x]
(x fastest, the flat order of the reference's ImgLib2 images).

Sources of the formulas:
* blending weights: BlendingRealRandomAccess.computeWeight and its 1001-entry
  cosine look-up table (process/fusion/transformed/weights/BlendingRealRandomAccess.java:46-130),
  applied along z only (border 0, blending range in voxels);
* weight normalisation: NormalizingRandomAccess.get / hardWeights with
  osemspeedup 1 (process/deconvolution/normalization/NormalizingRandomAccess.java:75-214);
* observed-image clamp to minValueImg = 1 for covered voxels
  (process/fusion/transformed/AbstractTransformedImgRandomAccess.java:83,
  MultiViewDeconvolution.java:48-50).

Random numbers come from numpy's PCG64 generator with fixed seeds
(1000+c phantom, 2000+100c+v Poisson noise of view v).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    shape_zyx: tuple
    n_views: int
    psf_size: int
    sigma_xy: float
    sigma_z: float
    angles_deg: list
    n_blobs: int
    r_min: float
    r_max: float
    peak: float | None
    weights: str          # "uniform" or "zramp"
    blend_range: int
    psf_type: str
    lam: float
    iterations: int
    seed_index: int
    extra: dict = field(default_factory=dict)


# The two single-GPU workloads named by the benchmark specification.
C1 = Workload("c1_128cube_3views_eb", (128, 128, 128), 3, 31, 1.5, 4.5, [0, 120, 240], 200, 2, 6, None,
              "uniform", 0, "EFFICIENT_BAYESIAN", 0.0, 10, 1)
C2 = Workload("c2_256cube_6views_opt2", (256, 256, 256), 6, 41, 1.8, 5.0, [0, 60, 120, 180, 240, 300], 1500, 2, 6,
              1000.0, "zramp", 20, "OPTIMIZATION_II", 0.0006, 20, 2)
# Views of uneven extents (x, y, z) placed centred in a common 400 x 352 x 224 frame; each view covers
# only its box (image 0 and weight 0 outside), weights blend along z at the box faces.  The reference
# runs it in 128^3 effective blocks (block size 128 + 2(K - 1) = 216).
C3 = Workload("c3_400x352x224_4views_eb", (224, 352, 400), 4, 45, 1.8, 5.0, [0, 90, 180, 270], 3000, 2, 6, None,
              "zramp_box", 20, "EFFICIENT_BAYESIAN", 0.0006, 30, 3,
              extra={"extents_xyz": [(384, 320, 200), (352, 352, 180), (400, 300, 220), (360, 330, 190)],
                     "block_effective": 128})


# Large combined run: 8 views of a 1024 x 1024 x 512 frame, 61^3 PSFs, Efficient-Bayesian, 40 iterations;
# the reference runs it in 512^3 effective blocks (block size 512 + 2(K - 1) = 632).
C5 = Workload("c5_1024x1024x512_8views_eb", (512, 1024, 1024), 8, 61, 2.0, 6.0, [0, 45, 90, 135, 180, 225, 270, 315],
              50000, 1, 8, 2000.0, "zramp", 20, "EFFICIENT_BAYESIAN", 0.0006, 40, 5,
              extra={"block_effective": 512})
# Training volumes of the 8-view learned initial estimate: C5's phantom density, PSFs, noise and weights
# on 256^3 (50000 blobs per 1024^2 x 512 voxels -> 1562 per 256^3).
C5_TRAIN = Workload("c5_train_256cube_8views_eb", (256, 256, 256), 8, 61, 2.0, 6.0,
                    [0, 45, 90, 135, 180, 225, 270, 315], 1562, 1, 8, 2000.0, "zramp", 20, "EFFICIENT_BAYESIAN",
                    0.0006, 40, 6)


def small_workload(n=24, views=3, k=7, psf_type="EFFICIENT_BAYESIAN", lam=0.0, iterations=3, weights="uniform",
                   seed_index=7, n_blobs=12, peak=None, shape=None):
    shape = shape or (n, n, n)
    ang = [i * 360.0 / views for i in range(views)]
    return Workload(f"small_{shape}_{views}v_k{k}", shape, views, k, 1.0, 2.0, ang, n_blobs, 1.0, 2.5, peak,
                    weights, max(2, min(shape) // 6), psf_type, lam, iterations, seed_index)


def blob_phantom(shape_zyx, n_blobs, r_min, r_max, seed, background=10.0):
    """Background + Gaussian blobs (sigma = radius), truncated at 4 sigma, summed in FP64."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape_zyx
    vol = np.full(shape_zyx, background, np.float64)
    cz = rng.uniform(0, nz, n_blobs)
    cy = rng.uniform(0, ny, n_blobs)
    cx = rng.uniform(0, nx, n_blobs)
    r = rng.uniform(r_min, r_max, n_blobs)
    amp = rng.uniform(50.0, 500.0, n_blobs)
    for i in range(n_blobs):
        ext = 4.0 * r[i]
        lo = [max(0, int(math.floor(c - ext))) for c in (cz[i], cy[i], cx[i])]
        hi = [min(n, int(math.ceil(c + ext)) + 1) for c, n in zip((cz[i], cy[i], cx[i]), shape_zyx)]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        z = np.arange(lo[0], hi[0]) - cz[i]
        y = np.arange(lo[1], hi[1]) - cy[i]
        x = np.arange(lo[2], hi[2]) - cx[i]
        d2 = z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2
        g = amp[i] * np.exp(-d2 / (2.0 * r[i] ** 2))
        g[d2 > ext * ext] = 0.0
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += g
    return vol


def gaussian_psf(size, sigma_xy, sigma_z, angle_deg):
    """Anisotropic Gaussian on a size^3 grid rotated about y; FP64, normalised to sum 1, cast to FP32."""
    c = size // 2
    t = math.radians(angle_deg)
    ry = np.array([[math.cos(t), 0, math.sin(t)], [0, 1, 0], [-math.sin(t), 0, math.cos(t)]])
    cov = ry @ np.diag([sigma_xy ** 2, sigma_xy ** 2, sigma_z ** 2]) @ ry.T   # (x, y, z) order
    icov = np.linalg.inv(cov)
    o = np.arange(size) - c
    z, y, x = np.meshgrid(o, o, o, indexing="ij")
    p = np.stack([x, y, z], -1).astype(np.float64)
    q = np.einsum("...i,ij,...j->...", p, icov, p)
    k = np.exp(-0.5 * q)
    k /= k.sum()
    return k.astype(np.float32)


def convolve_mirror_f64(vol, psf):
    """FP64 linear convolution of the mirror-single extended volume with psf (odd size)."""
    from scipy import fft as sfft
    c = [s // 2 for s in psf.shape]
    pad = np.pad(vol, [(ci, s - 1 - ci) for ci, s in zip(c, psf.shape)], mode="reflect")
    # circular convolution of length >= the padded extent only wraps into the first K-1
    # outputs, which the 'valid' crop discards
    L = [sfft.next_fast_len(n, real=True) for n in pad.shape]
    f = sfft.rfftn(pad, L, workers=-1) * sfft.rfftn(psf.astype(np.float64), L, workers=-1)
    full = sfft.irfftn(f, L, workers=-1)
    k = psf.shape
    return full[k[0] - 1:k[0] - 1 + vol.shape[0], k[1] - 1:k[1] - 1 + vol.shape[1], k[2] - 1:k[2] - 1 + vol.shape[2]]


_LUT = None


def _blend_lut():
    """BlendingRealRandomAccess static look-up table (:51-57), accumulated exactly as the Java loop does."""
    global _LUT
    if _LUT is None:
        lut = np.zeros(1001, np.float64)
        d = 0.0
        while d <= 1.0001:
            lut[int(d * 1000.0 + 0.5)] = (math.cos((1 - d) * math.pi) + 1) / 2
            d = d + 0.001
        _LUT = lut
    return _LUT


def blend_weight_1d(n, blending, border=0.0):
    """computeWeight restricted to one axis (:95-127), float arithmetic as in Java."""
    lut = _blend_lut()
    out = np.zeros(n, np.float32)
    dim_minus1 = np.float32(n - 1)
    for i in range(n):
        l = np.float32(i)
        tmp = min(np.float32(l - np.float32(border)), np.float32(dim_minus1 - l - np.float32(border)))
        if tmp <= 0:
            out[i] = 0.0
            continue
        rel = np.float32(tmp / np.float32(blending))
        w = np.float32(1.0)
        if rel < 1:
            w = np.float32(w * lut[int(float(rel) * 1000.0 + 0.5)])
        out[i] = w
    return out


def normalize_weights(ws):
    """NormalizingRandomAccess.get with hard weights and osemspeedup 1 (:75-110, :205-214)."""
    ws = np.asarray(ws, np.float32)
    clipped = np.minimum(1.0, ws.astype(np.float64))
    sumw = clipped.sum(0)
    out = np.empty_like(ws)
    for v in range(ws.shape[0]):
        mine = clipped[v].astype(np.float32)
        hard = np.divide(mine.astype(np.float64), sumw, out=np.zeros_like(sumw), where=sumw > 1).astype(np.float32)
        val = np.where(sumw > 1, hard, mine).astype(np.float64)
        out[v] = np.minimum(1.0, val * 1.0).astype(np.float32)
    return out


def view_boxes(w: Workload):
    """Per-view covered box ((z0, z1), (y0, y1), (x0, x1)), centred in the frame, or None."""
    ext = w.extra.get("extents_xyz") if w.extra else None
    if not ext:
        return None
    out = []
    for ex, ey, ez in ext[:w.n_views]:
        out.append(tuple(((n - e) // 2, (n - e) // 2 + e) for n, e in zip(w.shape_zyx, (ez, ey, ex))))
    return out


def make_inputs(w: Workload):
    """Returns dict(phantom f64, imgs [V,z,y,x] f32, weights f32, psfs [V,k,k,k] f32)."""
    V = w.n_views
    ph = blob_phantom(w.shape_zyx, w.n_blobs, w.r_min, w.r_max, 1000 + w.seed_index)
    psfs = np.stack([gaussian_psf(w.psf_size, w.sigma_xy, w.sigma_z, a) for a in w.angles_deg[:V]])
    imgs = np.empty((V,) + tuple(w.shape_zyx), np.float32)
    boxes = view_boxes(w)
    for v in range(V):
        b = convolve_mirror_f64(ph, psfs[v])
        if w.peak is not None:
            b = b * (w.peak / b.max())
        rng = np.random.default_rng(2000 + 100 * w.seed_index + v)
        noisy = rng.poisson(np.maximum(b, 0.0)).astype(np.float64)
        imgs[v] = np.maximum(1.0, noisy).astype(np.float32)
        if boxes is not None:  # uncovered voxels are 0 (outsideValueImg)
            (z0, z1), (y0, y1), (x0, x1) = boxes[v]
            cover = np.zeros(w.shape_zyx, bool)
            cover[z0:z1, y0:y1, x0:x1] = True
            imgs[v][~cover] = 0.0
    if w.weights == "uniform":
        weights = np.ones_like(imgs)
    elif w.weights == "zramp_box":
        raw = np.zeros(imgs.shape, np.float32)
        for v in range(V):
            (z0, z1), (y0, y1), (x0, x1) = boxes[v]
            wz = blend_weight_1d(z1 - z0, w.blend_range)
            raw[v, z0:z1, y0:y1, x0:x1] = wz[:, None, None]
        weights = np.ascontiguousarray(normalize_weights(raw))
    else:
        wz = blend_weight_1d(w.shape_zyx[0], w.blend_range)
        raw = np.broadcast_to(wz[None, :, None, None], imgs.shape).astype(np.float32)
        weights = normalize_weights(raw)
    return dict(phantom=ph, imgs=imgs, weights=np.ascontiguousarray(weights), psfs=psfs)


# ---------------------------------------------------------------- device-side generator (large workloads)
def _blob_params(shape_zyx, n_blobs, r_min, r_max, seed):
    """The blob centres, radii and amplitudes blob_phantom draws (same generator, same order)."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape_zyx
    cz = rng.uniform(0, nz, n_blobs)
    cy = rng.uniform(0, ny, n_blobs)
    cx = rng.uniform(0, nx, n_blobs)
    r = rng.uniform(r_min, r_max, n_blobs)
    amp = rng.uniform(50.0, 500.0, n_blobs)
    return cz, cy, cx, r, amp


def blob_phantom_torch(shape_zyx, n_blobs, r_min, r_max, seed, background=10.0, device="cuda",
                       chunk_voxels=1 << 26):
    """blob_phantom as an FP64 torch tensor on `device`: the same blobs (each voxel within 4 sigma of a centre
    gets amp exp(-d^2 / 2 sigma^2)), splatted in groups of equal box size; sums differ from the host loop
    only in FP64 summation order."""
    import torch
    cz, cy, cx, r, amp = _blob_params(shape_zyx, n_blobs, r_min, r_max, seed)
    nz, ny, nx = shape_zyx
    vol = torch.full(tuple(shape_zyx), float(background), dtype=torch.float64, device=device)
    flat = vol.view(-1)
    ext = 4.0 * r
    half = np.ceil(ext).astype(np.int64) + 1          # offsets -half..half around floor(centre) cover +-ext
    dims = torch.tensor([nz, ny, nx], device=device)
    for h in np.unique(half):
        sel = np.nonzero(half == h)[0]
        side = 2 * int(h) + 1
        per = max(1, chunk_voxels // side ** 3)
        o = torch.arange(-int(h), int(h) + 1, device=device, dtype=torch.int64)
        for a in range(0, len(sel), per):
            ids = sel[a:a + per]
            c = torch.as_tensor(np.stack([cz[ids], cy[ids], cx[ids]], 1), device=device)       # [b, 3]
            p = torch.floor(c).to(torch.int64)[:, :, None] + o[None, None, :]                   # [b, 3, side]
            dd = (p.to(torch.float64) - c[:, :, None]) ** 2
            inside = (p >= 0) & (p < dims[None, :, None])
            d2 = dd[:, 0, :, None, None] + dd[:, 1, None, :, None] + dd[:, 2, None, None, :]
            rr = torch.as_tensor(r[ids], device=device)[:, None, None, None]
            e2 = torch.as_tensor(ext[ids] ** 2, device=device)[:, None, None, None]
            g = torch.as_tensor(amp[ids], device=device)[:, None, None, None] * torch.exp(-d2 / (2.0 * rr * rr))
            mask = (d2 <= e2) & inside[:, 0, :, None, None] & inside[:, 1, None, :, None] & inside[:, 2, None, None, :]
            idx = (p[:, 0, :, None, None] * ny + p[:, 1, None, :, None]) * nx + p[:, 2, None, None, :]
            flat.index_put_((idx[mask],), g[mask], accumulate=True)
    return vol


def blur_mirror_torch(vol, psf, lengths=None):
    """convolve_mirror_f64 on a torch FP64 tensor (any device): mirror-single padding, circular convolution of
    the given lengths (>= the padded extent per axis), 'valid' crop."""
    import torch
    import torch.nn.functional as F
    k = psf.shape
    c = [s // 2 for s in k]
    pad = F.pad(vol[None, None], (c[2], k[2] - 1 - c[2], c[1], k[1] - 1 - c[1], c[0], k[0] - 1 - c[0]),
                mode="reflect")[0, 0]
    L = lengths or tuple(pad.shape)
    f = torch.fft.rfftn(pad, s=L)
    del pad
    kt = torch.as_tensor(np.asarray(psf, np.float64), device=vol.device)
    f *= torch.fft.rfftn(kt, s=L)
    full = torch.fft.irfftn(f, s=L)
    del f
    return full[k[0] - 1:k[0] - 1 + vol.shape[0], k[1] - 1:k[1] - 1 + vol.shape[1],
                k[2] - 1:k[2] - 1 + vol.shape[2]].contiguous()


def _smooth_length(n):
    """Smallest 2^a 3^b >= n (FFT lengths with stock kernels in every FFT library)."""
    best = 1 << (int(n) - 1).bit_length()
    p3 = 1
    while p3 < 2 * n:
        m = p3
        while m < n:
            m *= 2
        best = min(best, m)
        p3 *= 3
    return best


def make_inputs_device(w: Workload, device="cuda", log=None):
    """make_inputs for volumes too large for the host generator, on `device` (torch): the same phantom, PSFs,
    FP64 mirror blur, peak scaling and clamp; the Poisson noise comes from torch's generator seeded
    2000 + 100 c + v.  Returns dict(imgs=[V float32 tensors], weights=[V tensors, one shared z profile when
    the views share it], psfs=[V, k, k, k] numpy)."""
    import torch
    V = w.n_views
    psfs = np.stack([gaussian_psf(w.psf_size, w.sigma_xy, w.sigma_z, a) for a in w.angles_deg[:V]])
    ph = blob_phantom_torch(w.shape_zyx, w.n_blobs, w.r_min, w.r_max, 1000 + w.seed_index, device=device)
    if log:
        log(f"phantom {tuple(w.shape_zyx)} with {w.n_blobs} blobs on {device}")
    L = tuple(_smooth_length(n + w.psf_size - 1) for n in w.shape_zyx)
    imgs = []
    for v in range(V):
        b = blur_mirror_torch(ph, psfs[v], L)
        if w.peak is not None:
            b *= w.peak / float(b.max())
        g = torch.Generator(device=device)
        g.manual_seed(2000 + 100 * w.seed_index + v)
        noisy = torch.poisson(b.clamp_min_(0.0), generator=g)
        del b
        imgs.append(noisy.clamp_min_(1.0).to(torch.float32))
        del noisy
        if log:
            log(f"view {v + 1}/{V} blurred and noised")
    del ph
    if w.weights == "uniform":
        wt = torch.ones(tuple(w.shape_zyx), dtype=torch.float32, device=device)
        weights = [wt] * V
    elif w.weights == "zramp":
        wz = blend_weight_1d(w.shape_zyx[0], w.blend_range)
        prof = normalize_weights(np.broadcast_to(wz[None, :, None, None], (V, len(wz), 1, 1)).astype(np.float32))
        assert all(np.array_equal(prof[0], prof[v]) for v in range(V))
        wt = torch.as_tensor(prof[0], device=device).expand(tuple(w.shape_zyx)).contiguous()
        weights = [wt] * V
    else:
        raise ValueError(f"device generator: weights '{w.weights}' not supported")
    return dict(imgs=imgs, weights=weights, psfs=psfs)
