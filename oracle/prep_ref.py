"""FP64 numpy restatement of the deconvolution's input preparation (test infrastructure only).

Only tests may import this module; it is the checker for the HIP kernels in csrc/prep.hip.
Reference (src/main/java/net/preibisch/mvrecon/process/):
  transform_view      fusion/transformed/TransformView.java:59-73, TransformedInputRandomAccess.java:61-84,
                      AbstractTransformedImgRandomAccess.getInsideValue, AbstractTransformedIntervalRandomAccess
                      .intersectsLinearInterpolation (strictly inside), trilinear / nearest interpolation
                      (ImgLib2 NLinearInterpolator / NearestNeighborInterpolator, third-party: imglib2 8.0.0);
  transform_blending  fusion/transformed/TransformWeight.java:82-134, weights/BlendingRealRandomAccess.java:46-127
                      (float arithmetic, the 1001-entry cosine table);
  normalize_weights   deconvolution/normalization/NormalizingRandomAccess.java:76-108, 183-214;
  transform_psf       psf/PSFExtraction.java:182-193, 367-470;
  fuse_group          deconvolution/util/ProcessInputImages.java:279-399 (one group): views intensity-adjusted
                      (fusion/intensityadjust/IntensityAdjuster.java:43-54) and resampled, AVG fusion
                      (fusion/transformed/FusedRandomAccess.java:67-88) and summed decon weights
                      (weightcombination/CombineWeightsSumRandomAccess.java:40-50).
Pins (tests/test_prep.py): identity and integer-translation transforms reproduce the input; trilinear
interpolation reproduces affine functions exactly; the blending weight with the identity transform is the
product of the per-axis weights; the normalisation's closed forms and special cases; the PSF transform
with the identity keeps the normalised PSF, a 90 degree rotation permutes its axes exactly; group fusion of
one view is the transformed view, a duplicated view keeps the average and doubles the decon weight, views with
fusion weight 0 are skipped while constant-1 weights average in the outside zeros, and the intensity adjustment
with the min clamp follows its closed form.
"""
from __future__ import annotations

import math

import numpy as np


def inverse_affine(a):
    a = np.asarray(a, np.float64).reshape(3, 4)
    A = np.linalg.inv(a[:, :3])
    return np.concatenate([A, -(A @ a[:, 3])[:, None]], axis=1)


def _input_coords(affine, bb_min_xyz, bb_shape_zyx):
    inv = inverse_affine(affine)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in bb_shape_zyx], indexing="ij")
    s = np.stack([x + bb_min_xyz[0], y + bb_min_xyz[1], z + bb_min_xyz[2]], -1)
    t = s @ inv[:, :3].T + inv[:, 3]
    return t[..., 0], t[..., 1], t[..., 2]


def transform_view(img, affine, bb_min_xyz, bb_shape_zyx, min_value=1.0, has_min=True, outside=0.0, interpolation=1):
    img = np.asarray(img, np.float32).astype(np.float64)
    nz, ny, nx = img.shape
    tx, ty, tz = _input_coords(affine, bb_min_xyz, bb_shape_zyx)
    inside = (tx > 0) & (ty > 0) & (tz > 0) & (tx < nx - 1) & (ty < ny - 1) & (tz < nz - 1)
    out = np.full(tuple(bb_shape_zyx), np.float64(outside))
    X, Y, Z = tx[inside], ty[inside], tz[inside]
    if interpolation == 1:
        x0, y0, z0 = np.floor(X).astype(int), np.floor(Y).astype(int), np.floor(Z).astype(int)
        fx, fy, fz = X - x0, Y - y0, Z - z0
        val = np.zeros_like(X)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                    val += w * img[z0 + dz, y0 + dy, x0 + dx]
    else:
        val = img[np.floor(Z + 0.5).astype(int), np.floor(Y + 0.5).astype(int), np.floor(X + 0.5).astype(int)]
    val = val.astype(np.float32)
    if has_min:
        val = np.maximum(np.float32(min_value), val)
    out[inside] = val
    return out.astype(np.float32)


def blend_lut():
    lut = np.zeros(1001)
    d = 0.0
    while d <= 1.0001:
        lut[int(d * 1000.0 + 0.5)] = (math.cos((1 - d) * math.pi) + 1) / 2
        d = d + 0.001
    return lut


def blending_weight(loc_xyz, dims_xyz, border_xyz, blending_xyz):
    """BlendingRealRandomAccess.computeWeight for float locations (arrays), interval [0, dims - 1]."""
    lut = blend_lut()
    w = np.ones(np.shape(loc_xyz[0]), np.float32)
    zero = np.zeros(np.shape(loc_xyz[0]), bool)
    dist = []
    for d in range(3):
        l = np.asarray(loc_xyz[d], np.float32)
        dm = np.float32(dims_xyz[d] - 1)
        b = np.float32(border_xyz[d])
        t = np.minimum((l - b).astype(np.float32), ((dm - l).astype(np.float32) - b).astype(np.float32))
        dist.append(t)
        zero |= t <= 0
    for d in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):   # blending range 0: rel = inf, weight kept
            rel = (dist[d] / np.float32(blending_xyz[d])).astype(np.float32)
        idx = np.clip((np.where((rel < 1) & np.isfinite(rel), rel, 0).astype(np.float64) * 1000.0 + 0.5).astype(np.int64), 0, 1000)
        w = np.where(rel < 1, (w.astype(np.float64) * lut[idx]).astype(np.float32), w)
    return np.where(zero, np.float32(0), w).astype(np.float32)


def transform_blending(img_shape_zyx, border_xyz, blending_xyz, affine, bb_min_xyz, bb_shape_zyx):
    tx, ty, tz = _input_coords(affine, bb_min_xyz, bb_shape_zyx)
    dims = (img_shape_zyx[2], img_shape_zyx[1], img_shape_zyx[0])
    return blending_weight((tx.astype(np.float32), ty.astype(np.float32), tz.astype(np.float32)), dims, border_xyz,
                           blending_xyz)


def fuse_group(images, affines, bb_min_xyz, bb_shape_zyx, intensity_adjust=None, fusion_blending=None,
               decon_blending=None):
    """(fused image, summed decon weights) of one group; blending args per view (border_xyz, range_xyz)."""
    sumI = np.zeros(tuple(bb_shape_zyx))
    sumW = np.zeros(tuple(bb_shape_zyx))
    sumD = np.zeros(tuple(bb_shape_zyx))
    for j, (img, a) in enumerate(zip(images, affines)):
        img = np.asarray(img, np.float32)
        if intensity_adjust is not None:
            m00, m01 = intensity_adjust[j]
            img = (img.astype(np.float64) * m00 + m01).astype(np.float32)
        v = transform_view(img, a, bb_min_xyz, bb_shape_zyx, 1.0, True, 0.0, 1).astype(np.float64)
        if decon_blending is None:
            sumD += 1.0
        else:
            b, r = decon_blending[j]
            sumD += transform_blending(img.shape, b, r, a, bb_min_xyz, bb_shape_zyx).astype(np.float64)
        if fusion_blending is None:
            wf = np.ones(tuple(bb_shape_zyx))
        else:
            b, r = fusion_blending[j]
            wf = transform_blending(img.shape, b, r, a, bb_min_xyz, bb_shape_zyx).astype(np.float64)
        use = wf != 0
        sumI[use] += v[use] * wf[use]
        sumW[use] += wf[use]
    fused = np.where(sumW > 0, sumI / np.where(sumW > 0, sumW, 1.0), 0.0)
    return fused.astype(np.float32), sumD.astype(np.float32)


def normalize_weights(weights, osem_speedup=1.0, smooth=False, max_diff_range=0.1, scaling_range=0.05):
    w = np.asarray(weights, np.float32)
    clipped = np.minimum(1.0, w.astype(np.float64))
    sumw = clipped.sum(0)
    out = np.empty_like(w)
    for v in range(w.shape[0]):
        my = clipped[v].astype(np.float32)
        if smooth:
            safe = np.where(sumw > 0, sumw, 1.0)
            ideal = (my.astype(np.float64) / safe).astype(np.float32)
            diff = (my - ideal).astype(np.float32)
            yv = np.maximum(np.float32(0), ((np.float32(max_diff_range) - np.abs(diff)).astype(np.float32) *
                                            np.float32(np.float32(1.0) / np.float32(max_diff_range))).astype(np.float32))
            scale = ((yv * my).astype(np.float32) * np.float32(scaling_range)).astype(np.float32)
            val = np.where(sumw <= 0, 0.0, (np.minimum(my, ideal) - scale).astype(np.float32).astype(np.float64))
        else:
            hard = (my.astype(np.float64) / np.where(sumw > 1, sumw, 1.0)).astype(np.float32).astype(np.float64)
            val = np.where(sumw > 1, hard, my.astype(np.float64))
        out[v] = np.minimum(1.0, val * osem_speedup).astype(np.float32)
    return out


def transform_psf(psf, affine):
    """PSFExtraction.getTransformedNormalizedPSF (psf/PSFExtraction.java:182-193, 367-470): min-max normalise,
    bounds of the transformed interval [0, dim - 1] -> odd size, centre dim / 2 kept at the new centre,
    linear interpolation of the zero-extended PSF.  FP64."""
    psf = np.asarray(psf, np.float32).astype(np.float64)
    psf = ((psf - psf.min()) / (psf.max() - psf.min())).astype(np.float32).astype(np.float64)
    a = np.asarray(affine, np.float64).reshape(3, 4)
    nz, ny, nx = psf.shape
    corners = np.array([[x, y, z] for z in (0, nz - 1) for y in (0, ny - 1) for x in (0, nx - 1)], np.float64)
    t = corners @ a[:, :3].T + a[:, 3]
    size = t.max(0) - t.min(0)
    n = [int(s) + 1 for s in size]
    n = [k + 1 if k % 2 == 0 else k for k in n]
    ctr = a[:, :3] @ np.array([nx // 2, ny // 2, nz // 2], np.float64) + a[:, 3]
    off = [ctr[d] - (n[d] // 2) for d in range(3)]
    inv = inverse_affine(affine)
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    s = np.stack([x + off[0], y + off[1], z + off[2]], -1).astype(np.float64)
    q = s @ inv[:, :3].T + inv[:, 3]
    X, Y, Z = q[..., 0], q[..., 1], q[..., 2]
    x0, y0, z0 = np.floor(X).astype(int), np.floor(Y).astype(int), np.floor(Z).astype(int)
    fx, fy, fz = X - x0, Y - y0, Z - z0
    out = np.zeros(X.shape)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                xi, yi, zi = x0 + dx, y0 + dy, z0 + dz
                ok = (xi >= 0) & (xi < nx) & (yi >= 0) & (yi < ny) & (zi >= 0) & (zi < nz)
                w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                out[ok] += w[ok] * psf[zi[ok], yi[ok], xi[ok]]
    return out.astype(np.float32)
