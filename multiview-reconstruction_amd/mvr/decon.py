"""Host-side mirror of the reference's deconvolution surface over the HIP library.

The Java side of the reference keeps its drivers (MultiViewDeconvolutionSeq/Mul,
DeconView, Block, BlockSorter) and plugs a compute thread in through
ComputeBlockThreadFactory.  This module mirrors that surface in Python, with the
same names and argument meaning, so the block-level drop-in can be exercised
end to end: every block iteration runs in ``libmvr_hip.so`` through
``mvr_block_run_iteration_{seq,mul}``.  The resident single-GPU path
(``mvr.hip.Decon``) runs the whole driver inside the library instead.

Arrays are numpy [z, y, x] float32; sizes in the Block records are (x, y, z),
the order of the reference's long[] arrays.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import hip

MIN_VALUE = np.float32(1e-4)          # MultiViewDeconvolution.minValue (D/MultiViewDeconvolution.java:50)


@dataclass
class Block:
    """C/Block.java:91-133 (sizes and offsets in x, y, z order)."""
    block_size: tuple
    offset: tuple
    effective_size: tuple
    effective_offset: tuple
    effective_local_offset: tuple


def divide_into_blocks(img_size, kernel_size, block_size):
    """BlockGeneratorFixedSizePrecise.divideIntoBlocks (C/BlockGeneratorFixedSizePrecise.java:59-131).
    Returns None when a block is not larger than the kernel."""
    n = len(img_size)
    eff = [block_size[d] - kernel_size[d] + 1 for d in range(n)]
    if min(eff) <= 0:
        return None
    loc = [kernel_size[d] // 2 for d in range(n)]
    nb = [img_size[d] // eff[d] + (1 if img_size[d] % eff[d] else 0) for d in range(n)]
    blocks = []
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                cb = (bx, by, bz)
                eo = tuple(cb[d] * eff[d] for d in range(n))
                off = tuple(eo[d] - kernel_size[d] // 2 for d in range(n))
                es = tuple(min(eff[d], img_size[d] - eo[d]) for d in range(n))
                blocks.append(Block(tuple(block_size), off, es, eo, tuple(loc)))
    return blocks


def sort_blocks_by_smallest_footprint(blocks, img_size, min_required_blocks=1):
    """BlockSorter.sortBlocksBySmallestFootprint (C/BlockSorter.java:55-143): layers of
    non-interfering blocks along the dimension with the smallest qualifying layer size.
    Ties resolve to the last dimension (HashMap overwrite order)."""
    n = len(img_size)
    e = blocks[0].effective_size
    nb = [img_size[d] // e[d] + (1 if img_size[d] % e[d] else 0) for d in range(n)]
    size_to_dim, sizes = {}, []
    for d in range(n):
        s = int(np.prod([nb[k] for k in range(n) if k != d]))
        sizes.append(s)
        size_to_dim[s] = d
    sizes.sort()
    min_dim = -1
    for i in range(n):
        if min_dim != -1:
            break
        if sizes[i] >= min_required_blocks or i == n - 1:
            min_dim = size_to_dim[sizes[i]]
    min_off = blocks[0].offset[min_dim]
    batches, total = [], 0
    for i in range(nb[min_dim]):
        off = min_off + i * e[min_dim]
        layer = [b for b in blocks if b.offset[min_dim] == off]
        total += len(layer)
        batches.append(layer)
    if total != len(blocks):
        batches = [list(blocks)]
    return batches


def _mirror_idx(p, n):
    if n == 1:
        return np.zeros_like(p)
    per = 2 * n - 2
    p = np.mod(p, per)
    return np.where(p >= n, per - p, p)


def copy_block_mirror(psi, b: Block):
    """Block.copyBlock from Views.extendMirrorSingle(psi) (C/Block.java:158-197, 277-315).  Blocks that reach at
    most one period beyond the volume are one reflect-pad (mirror-single) and a slice; others index per axis."""
    lo = [max(0, -b.offset[d]) for d in range(3)]
    hi = [max(0, b.offset[d] + b.block_size[d] - psi.shape[2 - d]) for d in range(3)]
    src = [slice(max(0, b.offset[d]), min(psi.shape[2 - d], b.offset[d] + b.block_size[d])) for d in range(3)]
    size = [s_.stop - s_.start for s_ in src]
    # a padded side starts (ends) at the volume's first (last) voxel, so reflect-padding the slice mirrors the
    # volume as long as the slice is longer than the padding on each side
    if all(lo[d] < size[d] and hi[d] < size[d] for d in range(3)):
        core = psi[src[2], src[1], src[0]]
        if not any(lo) and not any(hi):
            return np.ascontiguousarray(core)
        return np.ascontiguousarray(np.pad(core, [(lo[2], hi[2]), (lo[1], hi[1]), (lo[0], hi[0])], mode="reflect"))
    idx = [_mirror_idx(np.arange(b.block_size[d]) + b.offset[d], psi.shape[2 - d]) for d in range(3)]
    return np.ascontiguousarray(psi[np.ix_(idx[2], idx[1], idx[0])])


def copy_block_zero(img, b: Block):
    """Views.zeroMin(Views.interval(Views.extendZero(img), block))."""
    out = np.zeros(b.block_size[::-1], np.float32)
    src = [slice(max(0, b.offset[d]), min(img.shape[2 - d], b.offset[d] + b.block_size[d])) for d in range(3)]
    dst = [slice(s.start - b.offset[d], s.stop - b.offset[d]) for d, s in enumerate(src)]
    if all(s.stop > s.start for s in src):
        out[dst[2], dst[1], dst[0]] = img[src[2], src[1], src[0]]
    return out


def paste_block(psi, b: Block, blk):
    """Block.pasteBlock (C/Block.java:357-403): the effective region only."""
    g = [slice(b.effective_offset[d], b.effective_offset[d] + b.effective_size[d]) for d in range(3)]
    l = [slice(b.effective_local_offset[d], b.effective_local_offset[d] + b.effective_size[d]) for d in range(3)]
    psi[g[2], g[1], g[0]] = blk[l[2], l[1], l[0]]


class DeconView:
    """D/DeconView.java:118-184: image, weight, raw kernel and the block lists."""

    def __init__(self, image, weight, kernel, block_size_xyz=None, min_required_blocks=1,
                 filter_blocks_for_content=True):
        self.image = np.ascontiguousarray(image, np.float32)
        self.weight = np.ascontiguousarray(weight, np.float32)
        self.kernel = np.ascontiguousarray(kernel, np.float32)
        img_size = self.image.shape[::-1]
        ksize = [2 * k - 1 for k in self.kernel.shape[::-1]]
        if block_size_xyz is None:
            block_size_xyz = tuple(img_size[d] + ksize[d] - 1 for d in range(3))
        self.block_size = tuple(block_size_xyz)
        blocks = divide_into_blocks(img_size, ksize, self.block_size)
        if blocks is None:
            self.num_blocks, self.batches = -1, None
            return
        self.num_blocks = len(blocks)
        self.batches = sort_blocks_by_smallest_footprint(blocks, img_size, min_required_blocks)
        if filter_blocks_for_content:
            # DeconView.filterBlocksForContent (:204-274)
            self.batches = [[b for b in batch if copy_block_zero(self.weight, b).any()] for batch in self.batches]
            self.batches = [batch for batch in self.batches if batch]
        self.kernel1 = None
        self.kernel2 = None


class DeconViews:
    """D/DeconViews.java:44-71 + DeconViewPSF.init for every view, in list order."""

    def __init__(self, views, psf_type="EFFICIENT_BAYESIAN", nT=8, dev=0):
        shapes = {v.image.shape for v in views}
        if len(shapes) != 1:
            raise ValueError("all views must have the same dimensions")
        self.views = views
        self.psf_type = psf_type
        k1, k2 = hip.init_psfs(np.stack([v.kernel for v in views]), psf_type, nT, dev)
        for i, v in enumerate(views):
            v.kernel1, v.kernel2 = k1[i], k2[i]


class ComputeBlockThreadHIPFactory:
    """ComputeBlockSeqThreadCUDAFactory's role for the HIP library: one block thread per device."""

    def __init__(self, block_size_xyz, lam=0.0, min_value=MIN_VALUE, devices=(0,)):
        self.block_size = tuple(block_size_xyz)
        self.lam = lam
        self.min_value = min_value
        self.devices = list(devices)

    def num_parallel_blocks(self):
        return len(self.devices)

    def create(self, i):
        return hip.BlockThread(self.block_size, self.min_value, self.lam, i, self.devices[i])


def psi_init_avg_precise(views):
    """PsiInitAvgPrecise (D/init/PsiInitAvgPrecise.java:52-112) in FP64 on the host."""
    imgs = np.stack([v.image for v in views])
    pos = imgs > 0
    cnt = pos.sum(0)
    s = np.where(pos, imgs.astype(np.float64), 0.0).sum(0)
    cov = cnt > 0
    avg = float((s[cov] / cnt[cov]).sum() / max(1, cov.sum()))
    if np.isnan(avg):
        avg = 1.0
    mx = np.where(pos, imgs, 0).reshape(len(views), -1).max(1).astype(np.float32)
    return avg, mx


class MultiViewDeconvolution:
    """MultiViewDeconvolution{Seq,Mul} (D/MultiViewDeconvolution.java:90-212,
    D/MultiViewDeconvolutionSeq.java:58-180, D/MultiViewDeconvolutionMul.java:116-245)
    driving HIP block threads."""

    def __init__(self, views: DeconViews, num_iterations, factory: ComputeBlockThreadHIPFactory, scheme="SEQ",
                 psi0=None, resident_views=True):
        """resident_views: each thread holds every view's image and weights on its device (uploaded once) and cuts
        the blocks there (mvr_block_run_iteration_*_at); otherwise the host cuts and sends every block, as the
        reference's drivers hand Views.extendZero intervals to the thread."""
        self.views = views
        self.num_iterations = num_iterations
        self.scheme = scheme
        self.resident_views = resident_views
        self.threads = [factory.create(i) for i in range(factory.num_parallel_blocks())]
        shape = views.views[0].image.shape
        self.avg, self.max = psi_init_avg_precise(views.views)
        self.psi = np.full(shape, np.float32(self.avg), np.float32) if psi0 is None else \
            np.ascontiguousarray(psi0, np.float32).copy()
        for t in self.threads:
            for i, v in enumerate(views.views):
                t.set_view_kernels(i, v.kernel1, v.kernel2)
                if resident_views:
                    t.set_view_data(i, v.image, v.weight)
        self.it = 0
        self.stats = []

    def run_iterations(self):
        while self.it < self.num_iterations:
            self.run_next_iteration()

    def _block_loop(self, batches, total_blocks, compute):
        sum_change, max_change = 0.0, -1.0
        previous = []
        t = self.threads[0]
        for batch in batches:
            current = []
            for b in batch:
                blk = copy_block_mirror(self.psi, b)
                out, st = compute(t, b, blk)
                sum_change += st[0]
                max_change = max(max_change, st[1])
                if total_blocks == 1:
                    paste_block(self.psi, b, out)
                else:
                    current.append((b, out))
            for b, out in previous:          # lagged write-back (:150-157)
                paste_block(self.psi, b, out)
            previous = current
        for b, out in previous:
            paste_block(self.psi, b, out)
        return sum_change, max_change

    def run_next_iteration(self):
        self.it += 1
        vs = self.views.views
        if self.scheme == "SEQ":
            for i, v in enumerate(vs):
                if self.resident_views:
                    def compute(t, b, blk, i=i):
                        return t.run_iteration_at(i, blk, b.offset, self.max[i])
                else:
                    def compute(t, b, blk, i=i, v=v):
                        return t.run_iteration(i, blk, copy_block_zero(v.image, b), copy_block_zero(v.weight, b),
                                               self.max[i])
                self.stats.append(self._block_loop(v.batches, v.num_blocks, compute))
        else:
            if self.resident_views:
                def compute(t, b, blk):
                    return t.run_iteration_mul_at(blk, b.offset, self.max)
            else:
                def compute(t, b, blk):
                    return t.run_iteration_mul(blk, [copy_block_zero(v.image, b) for v in vs],
                                               [copy_block_zero(v.weight, b) for v in vs], self.max)
            self.stats.append(self._block_loop(vs[0].batches, vs[0].num_blocks, compute))

    def close(self):
        for t in self.threads:
            t.close()
