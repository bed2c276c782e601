"""Input and PSF preparation for the deconvolution (ProcessInputImages / PSFPreparation), on the GPU.

Mirrors src/main/java/net/preibisch/mvrecon/process/deconvolution/util/ProcessInputImages.java:279-399.
Groups of one view (``prepare_views``): every view is resampled into the deconvolution bounding box
(``hip.transform_view``, minValueImg 1 inside, 0 outside), its blending weights are transformed
(``hip.transform_blending``) and all weights are normalised across views (``hip.normalize_weights``,
NormalizingRandomAccess).  Groups of several views (``prepare_groups``): each group is fused in one GPU pass
(``hip.fuse_group``: intensity adjustment, resampling, AVG fusion with the fusion blending, summed decon
blending weights), then the group weights are normalised across groups.  PSFs follow deconvolution/util/PSFPreparation.java:41-89: each is normalised and
resampled by its view's affine (``hip.transform_psf``), a group's PSFs are averaged over their smallest
common size (PSFCombination.computeAverageImage, psf/PSFCombination.java:74-135) and all are brought to one
size (PSFCombination.makeSameSize, :182-221).  The resampling runs in the HIP library; the array
bookkeeping on tiny PSFs is host code.
"""
from __future__ import annotations

import numpy as np

from . import hip

MIN_VALUE_IMG = 1.0      # MultiViewDeconvolution.minValueImg
OUTSIDE_VALUE_IMG = 0.0  # MultiViewDeconvolution.outsideValueImg


def prepare_views(images, affines, bb_min_xyz, bb_shape_zyx, blending_range_xyz=(0.0, 0.0, 0.0),
                  blending_border_xyz=(0.0, 0.0, 0.0), osem_speedup=1.0, smooth_blending=False,
                  max_diff_range=0.1, scaling_range=0.05, dev=0):
    """Views resampled into the box and their normalised weights: (imgs [V, z, y, x], weights [V, z, y, x]).

    Without blending ranges the decon weight of a view is 1 on the box (ConstantRandomAccessible); with them,
    the transformed blending function of the view's image interval.  Blending values are taken as given
    (after FusionTools.adjustBlending)."""
    V = len(images)
    imgs = np.empty((V,) + tuple(bb_shape_zyx), np.float32)
    ws = np.empty_like(imgs)
    use_blend = any(b > 0 for b in blending_range_xyz)
    for v in range(V):
        imgs[v] = hip.transform_view(images[v], affines[v], bb_min_xyz, bb_shape_zyx, min_value=MIN_VALUE_IMG,
                                     has_min=True, outside=OUTSIDE_VALUE_IMG, interpolation=1, dev=dev)
        if use_blend:
            ws[v] = hip.transform_blending(np.shape(images[v]), blending_border_xyz, blending_range_xyz, affines[v],
                                           bb_min_xyz, bb_shape_zyx, dev=dev)
        else:
            ws[v] = 1.0
    ws = hip.normalize_weights(ws, osem_speedup, smooth_blending, max_diff_range, scaling_range, dev=dev)
    return imgs, ws


def prepare_groups(groups, bb_min_xyz, bb_shape_zyx, blending_range_fusion=None, blending_border_fusion=None,
                   blending_range_decon=None, blending_border_decon=None, osem_speedup=1.0, smooth_blending=False,
                   max_diff_range=0.1, scaling_range=0.05, dev=0):
    """ProcessInputImages.fuseGroups + weight normalisation: (imgs [G, z, y, x], weights [G, z, y, x]).

    groups: list of groups, each a list of views {"img": array z, y, x, "affine": 3 x 4 input -> world,
    optional "intensity": (m00, m01)}.  Blending ranges / borders (x, y, z, after FusionTools.adjustBlending)
    apply to every view; None keeps the constant weight 1 of the reference's no-blending branch."""
    G = len(groups)
    imgs = np.empty((G,) + tuple(bb_shape_zyx), np.float32)
    ws = np.empty_like(imgs)
    for g, views in enumerate(groups):
        n = len(views)
        adj = None
        if any("intensity" in v for v in views):
            adj = [v.get("intensity", (1.0, 0.0)) for v in views]
        fb = None if blending_range_fusion is None else [list(blending_border_fusion) + list(blending_range_fusion)] * n
        db = None if blending_range_decon is None else [list(blending_border_decon) + list(blending_range_decon)] * n
        imgs[g], ws[g] = hip.fuse_group([v["img"] for v in views], [v["affine"] for v in views], bb_min_xyz,
                                        bb_shape_zyx, adj, fb, db, dev=dev)
    ws = hip.normalize_weights(ws, osem_speedup, smooth_blending, max_diff_range, scaling_range, dev=dev)
    return imgs, ws


def average_psf(psfs):
    """PSFCombination.computeAverageImage with the minimal common size (max = false): PSFs centred on
    dim / 2 summed into the smallest size, then divided by their count."""
    psfs = [np.asarray(p, np.float32) for p in psfs]
    size = tuple(min(p.shape[d] for p in psfs) for d in range(3))
    acc = np.zeros(size, np.float32)
    c = [s // 2 for s in size]
    for p in psfs:
        pc = [s // 2 for s in p.shape]
        lo = [c[d] - pc[d] for d in range(3)]  # position of the PSF's origin in the average
        src = tuple(slice(max(0, -lo[d]), min(p.shape[d], size[d] - lo[d])) for d in range(3))
        dst = tuple(slice(max(0, lo[d]), max(0, lo[d]) + (src[d].stop - src[d].start)) for d in range(3))
        acc[dst] += p[src]
    return (acc.astype(np.float64) / len(psfs)).astype(np.float32)


def make_same_size(psf, size_zyx):
    """PSFCombination.makeSameSize: centred crop or pad, padding with the PSF's minimum."""
    psf = np.asarray(psf, np.float32)
    out = np.full(tuple(size_zyx), psf.min(), np.float32)
    off = [psf.shape[d] // 2 - size_zyx[d] // 2 for d in range(3)]  # out index + off = psf index
    src = tuple(slice(max(0, off[d]), min(psf.shape[d], size_zyx[d] + off[d])) for d in range(3))
    dst = tuple(slice(src[d].start - off[d], src[d].stop - off[d]) for d in range(3))
    out[dst] = psf[src]
    return out


def prepare_psfs(raw_psfs, affines, same_size_for_all=True, dev=0):
    """One PSF per view (groups of one view): normalised, transformed; optionally all padded to one size."""
    return prepare_group_psfs([[p] for p in raw_psfs], [[a] for a in affines], same_size_for_all, dev)


def prepare_group_psfs(group_psfs, group_affines, same_size_for_all=True, dev=0):
    """PSFPreparation.loadGroupTransformPSFs: per group, every view's PSF normalised and transformed by its
    registration, averaged over the group's smallest common size; optionally all groups padded to one size."""
    out = [average_psf([hip.transform_psf(p, a, dev=dev) for p, a in zip(ps, afs)])
           for ps, afs in zip(group_psfs, group_affines)]
    if same_size_for_all:
        size = tuple(max(p.shape[d] for p in out) for d in range(3))
        out = [make_same_size(p, size) for p in out]
    return out
