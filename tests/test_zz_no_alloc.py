"""mvr_fft_convolve_noalloc (the entry point that allocates nothing) against mvr_fft_convolve, the entry point it
is a variant of, through the C ABI: consecutive calls that share the per-device scratch array (the second one
smaller, then one on a non-default stream), the largest FFT grid that fits the array, and the refusals (grid too
large, host pointers).  Outputs start as NaN so that every voxel must be written.
"""
import numpy as np
import pytest

from mvr import hip

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _run(img, k, ext, val, stream=None):
    import torch
    ti = torch.from_numpy(img).cuda()
    tk = torch.from_numpy(k).cuda()
    to = torch.full_like(ti, float("nan"))
    torch.cuda.synchronize()
    hip.fft_convolve_noalloc(ti.data_ptr(), img.shape, tk.data_ptr(), k.shape, to.data_ptr(), ext, val,
                             stream=stream)
    torch.cuda.synchronize()
    return to.cpu().numpy()


def test_noalloc_matches_fft_convolve(oracle):
    import torch
    rng = np.random.default_rng(77)
    cases = [((40, 44, 36), (9, 7, 11), "mirror", 0.0), ((20, 18, 24), (5, 5, 3), "value", 1.0),
             ((33, 12, 50), (7, 3, 9), "mirror", 0.0)]
    for i, (shape, kshape, ext, val) in enumerate(cases):
        img = rng.uniform(0.5, 2.0, shape).astype(np.float32)
        k = rng.uniform(0.0, 1.0, kshape).astype(np.float32)
        want = hip.fft_convolve(img, k, ext, val)
        if i == 2:  # on a stream of torch's (default priority)
            s = torch.cuda.Stream()
            got = _run(img, k, ext, val, stream=s.cuda_stream)
        else:
            got = _run(img, k, ext, val)
        assert np.isfinite(got).all()
        e = rel(got, want)
        e64 = rel(got, oracle.fft_convolve(img.astype(np.float64), k.astype(np.float64), ext, val, dtype=np.float64))
        print(f"{shape} * {kshape} {ext}: rel-L2 vs mvr_fft_convolve {e:.2e}, vs FP64 oracle {e64:.2e}")
        assert e < 1e-6
        assert e64 < 3e-6


def test_noalloc_largest_grid_and_refusals():
    rng = np.random.default_rng(78)
    # 140^3 with a 41^3 kernel runs on a 180^3 grid: about 50 MB of the 64 MiB array
    img = rng.uniform(0.5, 2.0, (140, 140, 140)).astype(np.float32)
    k = rng.uniform(0.0, 1.0, (41, 41, 41)).astype(np.float32)
    got = _run(img, k, "mirror", 0.0)
    want = hip.fft_convolve(img, k, "mirror", 0.0)
    assert np.isfinite(got).all()
    assert rel(got, want) < 1e-6
    # a 240^3 grid does not fit: refused before any device work
    big = np.ones((200, 200, 200), np.float32)
    with pytest.raises(hip.MvrError, match="scratch"):
        _run(big, k, "mirror", 0.0)
    # host pointers are refused (the kernels would fault on them)
    small = np.ones((8, 8, 8), np.float32)
    ks = np.ones((3, 3, 3), np.float32)
    out = np.zeros_like(small)
    with pytest.raises(hip.MvrError, match="device memory"):
        hip.fft_convolve_noalloc(small.ctypes.data, small.shape, ks.ctypes.data, ks.shape, out.ctypes.data)
    # the array keeps working after the refusals
    got = _run(img[:30, :30, :30].copy(), k[:5, :5, :5].copy(), "value", 1.0)
    want = hip.fft_convolve(img[:30, :30, :30], k[:5, :5, :5], "value", 1.0)
    assert rel(got, want) < 1e-6


@pytest.mark.gpu
def test_noalloc_repeat_after_other_sizes(oracle):
    # the scratch array keeps the last call's contents; a call after calls of other sizes still matches
    rng = np.random.default_rng(79)
    for shape, kshape in [((64, 64, 64), (31, 31, 31)), ((9, 10, 11), (3, 3, 3)), ((64, 64, 64), (31, 31, 31))]:
        img = rng.uniform(0.5, 2.0, shape).astype(np.float32)
        k = rng.uniform(0.0, 1.0, kshape).astype(np.float32)
        got = _run(img, k, "mirror", 0.0)
        assert rel(got, hip.fft_convolve(img, k, "mirror", 0.0)) < 1e-6
