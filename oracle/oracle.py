"""ctypes front-end of the CPU oracle (``mvr_oracle.cpp``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The HIP product never calls it.

The oracle restates the reference's CPU path (Java, not executable here:
no JDK/Maven; see the header of ``mvr_oracle.cpp`` for what is restated from
third-party definitions).  How each function is pinned (``tests/test_oracle_*.py``):

=============================  ==============================================================
oracle function                pinned by
=============================  ==============================================================
divide_into_portions           known-answer table from the Java formula (FusionTools:1287-1329)
sum_img_quirk / norm_to_sum1   closed form S/(S+S0) with numpy sums; invariance checks
mirror                         numpy flip (odd sizes); middle-pair law (even sizes)
fft1d / rfft3d                 numpy.fft (float64)
fft_convolve (mirror/zero/1)   brute-force direct convolution (tiny) and scipy.ndimage.convolve
init_psfs (4 PSF types)        delta-PSF known answers, scipy.signal.convolve(mode='same'),
                               symmetry invariants
next_value / next_value_mul    closed forms + exhaustive special cases (<=0, NaN, lambda=0, w=0/1)
divide_into_blocks/sort_blocks known-answer tables incl. the reference Block.main inputs
psi_init_avg_precise           closed form with numpy
psi_init_avg_approx            closed form (slice mean / max), invariance to voxels off the slice
psi_init_blurred_fused         constant images, delta response = separable kernel product, kernel
                               sum 1, borders against numpy reflect padding + direct convolution
deconvolve (Seq/Mul)           block-independence invariant, numpy/scipy global-form iteration,
                               FP32-vs-FP64 agreement, flux preservation
=============================  ==============================================================
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmvr_oracle.so")

PSF_TYPES = {"OPTIMIZATION_II": 0, "OPTIMIZATION_I": 1, "EFFICIENT_BAYESIAN": 2, "INDEPENDENT": 3}
SCHEMES = {"SEQ": 0, "MUL": 1}

_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -ffp-contract=off -fopenmp)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = ctypes.c_int
        L.oracle_fast_length.argtypes = [ctypes.c_int]
        L.oracle_divide_into_portions.argtypes = [ctypes.c_longlong, ctypes.c_int, _i64p, _i64p, ctypes.c_int]
        L.oracle_sum_img_quirk_f32.argtypes = [_f32p, ctypes.c_longlong, ctypes.c_int]
        L.oracle_sum_img_quirk_f32.restype = ctypes.c_double
        L.oracle_norm_to_sum1_f32.argtypes = [_f32p, ctypes.c_longlong, ctypes.c_int]
        L.oracle_norm_to_sum1_f64.argtypes = [_f64p, ctypes.c_longlong, ctypes.c_int]
        L.oracle_mirror_f32.argtypes = [_f32p, _i32p]
        L.oracle_fft1d_f64.argtypes = [_f64p, ctypes.c_int, ctypes.c_int]
        L.oracle_rfft3d_f64.argtypes = [_f64p, _i32p, _f64p]
        L.oracle_fft_convolve_f32.argtypes = [_f32p, _i32p, ctypes.c_int, ctypes.c_float, _f32p, _i32p, _f32p]
        L.oracle_fft_convolve_f64.argtypes = [_f64p, _i32p, ctypes.c_int, ctypes.c_double, _f64p, _i32p, _f64p]
        L.oracle_next_value_f32.argtypes = [ctypes.c_float] * 3 + [ctypes.c_double] + [ctypes.c_float] * 2
        L.oracle_next_value_f32.restype = ctypes.c_float
        L.oracle_next_value_f64.argtypes = [ctypes.c_double] * 6
        L.oracle_next_value_f64.restype = ctypes.c_double
        L.oracle_next_value_mul_f32.argtypes = [ctypes.c_float, _f32p, _f32p, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_float, ctypes.c_float]
        L.oracle_next_value_mul_f32.restype = ctypes.c_float
        L.oracle_compute_quotient_f32.argtypes = [_f32p, _f32p, ctypes.c_longlong]
        L.oracle_compute_final_values_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_longlong, ctypes.c_double,
                                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, _f64p]
        L.oracle_compute_final_values_mul_f32.argtypes = [_f32p, ctypes.c_int, _f32p, _f32p, ctypes.c_longlong,
                                                          ctypes.c_double, ctypes.c_float, ctypes.c_float,
                                                          ctypes.c_int, _f64p]
        L.oracle_divide_into_blocks.argtypes = [_i64p, _i64p, _i64p, _i64p, ctypes.c_int]
        L.oracle_sort_blocks.argtypes = [_i64p, _i64p, _i64p, ctypes.c_int, _i32p, _i32p, _i32p]
        L.oracle_init_psfs_f32.argtypes = [_f32p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.oracle_init_psfs_f64.argtypes = [_f64p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        L.oracle_psi_init_avg_precise.argtypes = [_f32p, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, _f32p]
        L.oracle_psi_init_avg_precise.restype = ctypes.c_double
        L.oracle_block_seq_f32.argtypes = [_f32p, _i32p, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_int, _f64p]
        L.oracle_block_mul_f32.argtypes = [_f32p, _i32p, ctypes.c_int, _f32p, _f32p, _f32p, _f32p, _i32p,
                                           ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_int, _f64p]
        for suf, fp in (("f32", _f32p), ("f64", _f64p)):
            f = getattr(L, "oracle_deconvolve_" + suf)
            f.argtypes = [_i32p, _f32p, _f32p, _f32p, _f32p, _f32p, fp, _f32p, _f64p, fp, _f64p, _f64p]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def _dims_xyz(a: np.ndarray) -> np.ndarray:
    """numpy arrays are indexed [z, y, x] (x fastest, ImgLib2 flat order)."""
    return np.array([a.shape[2], a.shape[1], a.shape[0]], dtype=np.int32)


def fast_length(n: int) -> int:
    return lib().oracle_fast_length(int(n))


def divide_into_portions(n: int, nT: int):
    cap = max(1, nT) + int(n // (64 ** 3)) + 2
    s = np.zeros(cap, np.int64)
    l = np.zeros(cap, np.int64)
    k = lib().oracle_divide_into_portions(int(n), int(nT), s, l, cap)
    return list(zip(s[:k].tolist(), l[:k].tolist()))


def sum_img_quirk(a: np.ndarray, nT: int) -> float:
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    return lib().oracle_sum_img_quirk_f32(a, a.size, nT)


def norm_to_sum1(a: np.ndarray, nT: int) -> np.ndarray:
    if a.dtype == np.float64:
        out = np.ascontiguousarray(a, dtype=np.float64).copy()
        lib().oracle_norm_to_sum1_f64(out.ravel(), out.size, nT)
    else:
        out = np.ascontiguousarray(a, dtype=np.float32).copy()
        lib().oracle_norm_to_sum1_f32(out.ravel(), out.size, nT)
    return out


def mirror(a: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.float32).copy()
    lib().oracle_mirror_f32(out.ravel(), _dims_xyz(out))
    return out


def fft1d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).ravel(), dtype=np.float64)
    lib().oracle_fft1d_f64(buf, x.size, 1 if inverse else -1)
    return buf[0::2] + 1j * buf[1::2]


def rfft3d(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    pz, py, px = x.shape
    out = np.zeros(pz * py * (px // 2 + 1) * 2, np.float64)
    lib().oracle_rfft3d_f64(x.ravel(), np.array([px, py, pz], np.int32), out)
    return (out[0::2] + 1j * out[1::2]).reshape(pz, py, px // 2 + 1)


EXT_MIRROR, EXT_VALUE = 0, 1


def fft_convolve(img: np.ndarray, kernel: np.ndarray, ext: str = "mirror", value: float = 0.0,
                 dtype=np.float32) -> np.ndarray:
    """util.FFTConvolution: out[x] = sum_j K[j] ext(img)[x + c - j]."""
    img = np.ascontiguousarray(img, dtype=dtype)
    kernel = np.ascontiguousarray(kernel, dtype=dtype)
    out = np.zeros_like(img)
    e = EXT_MIRROR if ext == "mirror" else EXT_VALUE
    f = lib().oracle_fft_convolve_f32 if dtype == np.float32 else lib().oracle_fft_convolve_f64
    f(img.ravel(), _dims_xyz(img), e, value, kernel.ravel(), _dims_xyz(kernel), out.ravel())
    return out


def next_value(last, integral, weight, lam, min_i, max_i, dtype=np.float32):
    if dtype == np.float32:
        return np.float32(lib().oracle_next_value_f32(last, integral, weight, lam, min_i, max_i))
    return lib().oracle_next_value_f64(last, integral, weight, lam, min_i, max_i)


def next_value_mul(last, integrals, weights, lam, min_i, max_i):
    i = np.ascontiguousarray(integrals, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    return np.float32(lib().oracle_next_value_mul_f32(last, i, w, i.size, lam, min_i, max_i))


def compute_quotient(img, blurred):
    """DeconvolutionMethods.computeQuotient over a flat array: img > 0 ? img / blurred : 1."""
    img = np.ascontiguousarray(img, np.float32).ravel()
    out = np.ascontiguousarray(blurred, np.float32).ravel().copy()
    lib().oracle_compute_quotient_f32(img, out, out.size)
    return out


def compute_final_values(psi, integral, weight, lam, min_i, max_i, nT=8):
    """DeconvolutionMethods.computeFinalValues over flat arrays: (new psi, (sumChange, maxChange))."""
    out = np.ascontiguousarray(psi, np.float32).ravel().copy()
    st = np.zeros(2, np.float64)
    lib().oracle_compute_final_values_f32(out, np.ascontiguousarray(integral, np.float32).ravel(),
                                          np.ascontiguousarray(weight, np.float32).ravel(), out.size, lam, min_i,
                                          max_i, nT, st)
    return out, (float(st[0]), float(st[1]))


def compute_final_values_mul(psi, integrals, weights, lam, min_i, max_i, nT=8):
    """DeconvolutionMethods.computeFinalValuesMul: integrals / weights are [V, n]."""
    out = np.ascontiguousarray(psi, np.float32).ravel().copy()
    i = np.ascontiguousarray(integrals, np.float32).reshape(-1, out.size)
    w = np.ascontiguousarray(weights, np.float32).reshape(-1, out.size)
    st = np.zeros(2, np.float64)
    lib().oracle_compute_final_values_mul_f32(out, i.shape[0], i.ravel(), w.ravel(), out.size, lam, min_i, max_i, nT,
                                              st)
    return out, (float(st[0]), float(st[1]))


def divide_into_blocks(img_xyz, kernel_size_xyz, block_size_xyz):
    """Returns a list of dicts (x,y,z order) or None when the block is too small."""
    cap = 1
    for n, k, b in zip(img_xyz, kernel_size_xyz, block_size_xyz):
        e = b - k + 1
        cap *= max(1, -(-n // e) if e > 0 else 1)
    out = np.zeros(15 * cap, np.int64)
    k = lib().oracle_divide_into_blocks(np.array(img_xyz, np.int64), np.array(kernel_size_xyz, np.int64),
                                        np.array(block_size_xyz, np.int64), out, cap)
    if k < 0:
        return None
    res = []
    for i in range(k):
        r = out[15 * i:15 * i + 15]
        res.append(dict(blockSize=tuple(r[0:3]), offset=tuple(r[3:6]), effectiveSize=tuple(r[6:9]),
                        effectiveOffset=tuple(r[9:12]), effectiveLocalOffset=tuple(r[12:15])))
    return res


def sort_blocks(img_xyz, kernel_size_xyz, block_size_xyz, min_required=1):
    blocks = divide_into_blocks(img_xyz, kernel_size_xyz, block_size_xyz)
    n = len(blocks)
    batch_of = np.zeros(n, np.int32)
    order = np.zeros(n, np.int32)
    dim = np.zeros(1, np.int32)
    nb = lib().oracle_sort_blocks(np.array(img_xyz, np.int64), np.array(kernel_size_xyz, np.int64),
                                  np.array(block_size_xyz, np.int64), int(min_required), batch_of, order, dim)
    batches = [[] for _ in range(nb)]
    for i in order.tolist():
        batches[batch_of[i]].append(i)
    return batches, int(dim[0])


def init_psfs(kernels: np.ndarray, psf_type: str, nT: int, dtype=np.float32):
    """DeconViewPSF.init for all views (list order).  kernels: [V, kz, ky, kx]."""
    k = np.ascontiguousarray(kernels, dtype=dtype)
    V = k.shape[0]
    kd = np.array([k.shape[3], k.shape[2], k.shape[1]], np.int32)
    k1 = np.zeros_like(k)
    k2 = np.zeros_like(k)
    f = lib().oracle_init_psfs_f32 if dtype == np.float32 else lib().oracle_init_psfs_f64
    f(k.ravel(), V, kd, PSF_TYPES[psf_type], nT, k1.ravel(), k2.ravel())
    return k1, k2


def psi_init_avg_precise(imgs: np.ndarray, nT: int):
    imgs = np.ascontiguousarray(imgs, np.float32)
    V = imgs.shape[0]
    mx = np.zeros(V, np.float32)
    avg = lib().oracle_psi_init_avg_precise(imgs.ravel(), V, imgs[0].size, nT, mx)
    return avg, mx


def psi_init_avg_approx(imgs):
    """PsiInitAvgApprox (src/main/java/net/preibisch/mvrecon/process/deconvolution/init/PsiInitAvgApprox.java:47-99,
    PsiInitAvgApproxThread.java:57-85): per view, min/max/sum over the hyperslice x = nx/2, visited once per
    dimension (3 passes, the count tripled); avg = mean over views of sum/count; max[v] = (float) slice max.
    Returns (avg, max[V]).  FP64 numpy."""
    imgs = np.asarray(imgs, np.float32)
    V = imgs.shape[0]
    means = np.zeros(V)
    mx = np.zeros(V, np.float32)
    for v in range(V):
        sl = imgs[v][:, :, imgs.shape[3] // 2].astype(np.float64)
        total, count = 0.0, 0
        for _ in range(3):
            total += float(sl.sum())
            count += sl.size
        means[v] = total / count
        mx[v] = np.float32(sl.max())
    avg = float(means.sum() / V)
    return (1.0 if np.isnan(avg) else avg), mx


def gauss3_half_kernel(sigma):
    """Half kernel of ImgLib2's Gauss3 (imglib2-algorithm, version managed by pom-scijava 43.0.0; the
    published Gauss3.halfkernelsizes / halfkernel): H = max(2, (int)(3 sigma + 0.5) + 1) taps,
    k[0] = 1, k[x] = exp(-x^2 / (2 sigma^2)), normalised so that the symmetric kernel sums to 1."""
    H = max(2, int(3 * sigma + 0.5) + 1)
    k = np.ones(H)
    for x in range(1, H):
        k[x] = np.exp(-(x * x) / (2 * sigma * sigma))
    total = 2 * (0.5 + k[1:].sum())
    return k / total


def psi_init_blurred_fused(imgs, weights, sigma=5.0):
    """PsiInitBlurredFused (src/main/java/net/preibisch/mvrecon/process/deconvolution/init/PsiInitBlurredFused.java:62-127,
    deconvolution/util/FusedNonZeroRandomAccess.java:57-96): psi = sum(I w) / sum(w) over views with I > 0
    (0 without weight), stored as float, then Gauss3 with mirror-single borders along x, y, z.  FP64.
    Returns psi (float64)."""
    from scipy import ndimage
    imgs = np.asarray(imgs, np.float32).astype(np.float64)
    weights = np.asarray(weights, np.float32).astype(np.float64)
    pos = imgs > 0
    sI = np.where(pos, imgs * weights, 0.0).sum(0)
    sW = np.where(pos, weights, 0.0).sum(0)
    fused = np.where(sW > 0, sI / np.where(sW > 0, sW, 1.0), 0.0).astype(np.float32).astype(np.float64)
    half = gauss3_half_kernel(sigma)
    full = np.concatenate([half[:0:-1], half])
    out = fused
    for axis in (2, 1, 0):
        out = ndimage.correlate1d(out, full, axis=axis, mode="mirror")
    return out


def block_seq(psi_blk, img_blk, w_blk, k1, k2, lam, min_value, max_i, nT):
    """One ComputeBlockSeqThreadCPU.runIteration on a block (kernels already final)."""
    psi = np.ascontiguousarray(psi_blk, np.float32).copy()
    stats = np.zeros(2, np.float64)
    lib().oracle_block_seq_f32(psi.ravel(), _dims_xyz(psi), np.ascontiguousarray(img_blk, np.float32).ravel(),
                               np.ascontiguousarray(w_blk, np.float32).ravel(),
                               np.ascontiguousarray(k1, np.float32).ravel(),
                               np.ascontiguousarray(k2, np.float32).ravel(), _dims_xyz(k1), lam, min_value, max_i,
                               nT, stats)
    return psi, stats


def block_mul(psi_blk, img_blks, w_blks, k1s, k2s, lam, min_value, max_is, nT):
    """One ComputeBlockMulThreadCPU.runIteration on a block (kernels already final)."""
    psi = np.ascontiguousarray(psi_blk, np.float32).copy()
    imgs = np.ascontiguousarray(img_blks, np.float32)
    ws = np.ascontiguousarray(w_blks, np.float32)
    k1s = np.ascontiguousarray(k1s, np.float32)
    k2s = np.ascontiguousarray(k2s, np.float32)
    stats = np.zeros(2, np.float64)
    lib().oracle_block_mul_f32(psi.ravel(), _dims_xyz(psi), imgs.shape[0], imgs.ravel(), ws.ravel(), k1s.ravel(),
                               k2s.ravel(), _dims_xyz(k1s[0]), lam, min_value,
                               np.ascontiguousarray(max_is, np.float32), nT, stats)
    return psi, stats


def deconvolve(imgs, weights, kernels, *, psf_type="EFFICIENT_BAYESIAN", scheme="SEQ", block_size=None,
               lam=0.0, min_value=1e-4, iterations=10, nT=8, psi0=None, filter_blocks=True, dtype=np.float32,
               sample_views=0):
    """Run the restated reference driver (MultiViewDeconvolution{Seq,Mul}).

    imgs, weights: [V, z, y, x] float32.  kernels: [V, kz, ky, kx] float32 (raw PSFs,
    normalised inside as DeconViewPSF.init does).  block_size (x,y,z) defaults to
    one block covering the volume: N + 2(K-1).
    sample_views > 0 (Seq only) stops after that many view passes, for timing samples.
    Returns dict(psi, max, stats[it, v, 2] or [it, 2], k2, avg, loop_seconds).
    """
    imgs = np.ascontiguousarray(imgs, np.float32)
    weights = np.ascontiguousarray(weights, np.float32)
    kernels = np.ascontiguousarray(kernels, np.float32)
    V, nz, ny, nx = imgs.shape
    kz, ky, kx = kernels.shape[1:]
    if block_size is None:
        block_size = (nx + 2 * (kx - 1), ny + 2 * (ky - 1), nz + 2 * (kz - 1))
    ip = np.array([nx, ny, nz, kx, ky, kz, *block_size, V, PSF_TYPES[psf_type], SCHEMES[scheme], iterations, nT,
                   0 if psi0 is None else 1, 1 if filter_blocks else 0, sample_views], np.int32)
    fp = np.array([lam, min_value], np.float32)
    psi_in = np.ascontiguousarray(psi0 if psi0 is not None else np.zeros(1), np.float32)
    psi = np.zeros((nz, ny, nx), dtype)
    mx = np.zeros(V, np.float32)
    nst = iterations * (V if scheme == "SEQ" else 1) * 2
    stats = np.zeros(nst, np.float64)
    k2 = np.zeros((V, kz, ky, kx), dtype)
    avg = np.zeros(1, np.float64)
    secs = np.zeros(1, np.float64)
    f = lib().oracle_deconvolve_f32 if dtype == np.float32 else lib().oracle_deconvolve_f64
    rc = f(ip, fp, imgs.ravel(), weights.ravel(), kernels.ravel(), psi_in.ravel(), psi.ravel(), mx, stats, k2.ravel(),
           avg, secs)
    if rc != 0:
        raise ValueError("oracle deconvolve failed (block smaller than kernel?)")
    shape = (iterations, V, 2) if scheme == "SEQ" else (iterations, 2)
    return dict(psi=psi, max=mx, stats=stats.reshape(shape), k2=k2, avg=float(avg[0]), loop_seconds=float(secs[0]))
