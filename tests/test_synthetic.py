"""The device-side input generator (torch) against the host generator, on the CPU."""
import numpy as np
import pytest

from mvr import synthetic as syn

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("shape,n,rmin,rmax", [((40, 36, 44), 60, 1, 8), ((24, 30, 20), 25, 2, 6)])
def test_blob_phantom_torch_matches_host(shape, n, rmin, rmax):
    a = syn.blob_phantom(shape, n, rmin, rmax, 1234)
    b = syn.blob_phantom_torch(shape, n, rmin, rmax, 1234, device="cpu", chunk_voxels=20000).numpy()
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_blur_mirror_torch_matches_host():
    a = syn.blob_phantom((24, 30, 20), 25, 2, 6, 7)
    k = syn.gaussian_psf(9, 1.5, 3.0, 45)
    want = syn.convolve_mirror_f64(a, k)
    got = syn.blur_mirror_torch(torch.from_numpy(a), k, tuple(syn._smooth_length(s + 8) for s in a.shape)).numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def _three_smooth(c):
    for p in (2, 3):
        while c % p == 0:
            c //= p
    return c == 1


def test_smooth_length():
    # the smallest 2^a 3^b >= n
    for n in range(1, 3000):
        m = syn._smooth_length(n)
        assert m >= n and _three_smooth(m)
        assert not any(_three_smooth(c) for c in range(n, m))


def test_device_generator_small_workload():
    w = syn.small_workload(n=24, views=3, k=7, weights="zramp", peak=500.0)
    host = syn.make_inputs(w)
    dev = syn.make_inputs_device(w, "cpu")
    for v in range(3):
        img = dev["imgs"][v].numpy()
        assert img.dtype == np.float32 and img.shape == (24, 24, 24) and img.min() >= 1.0
        # same phantom, blur and peak: the noisy images agree in mean to Poisson accuracy
        assert abs(img.mean() / host["imgs"][v].mean() - 1.0) < 0.02
        np.testing.assert_array_equal(dev["weights"][v].numpy(), host["weights"][v])
    np.testing.assert_array_equal(dev["psfs"], host["psfs"])
