"""Pins of the CPU oracle (oracle/mvr_oracle.cpp) against closed forms, brute force,
numpy/scipy and known-answer tables.  CPU only.

The reference (Java) cannot run in this container, and its own test tree holds
no deconvolution tests or fixtures, so every pin here is independent of the
oracle's own code: a different algorithm (direct convolution, numpy FFT), a
closed form, or an integer table worked out from the Java formulas.
"""
import math

import numpy as np
import pytest
from scipy import ndimage, signal

RNG = np.random.default_rng(12345)


# ----------------------------------------------------------------- portions
@pytest.mark.parametrize("n,nT,expect", [
    (5, 8, [(i, 1) for i in range(5)]),
    (8, 8, [(i, 1) for i in range(8)]),
    (100, 8, [(0, 12), (12, 12), (24, 12), (36, 12), (48, 12), (60, 12), (72, 12), (84, 16)]),
    (100, 4, [(0, 25), (25, 25), (50, 25), (75, 25)]),
    (0, 8, []),
])
def test_divide_into_portions_table(oracle, n, nT, expect):
    assert oracle.divide_into_portions(n, nT) == expect


def test_divide_into_portions_large(oracle):
    n = 64 ** 3 * 20 + 7
    p = oracle.divide_into_portions(n, 8)
    assert len(p) == 20
    assert p[-1] == (19 * (n // 20), n // 20 + n % 20)


# ----------------------------------------------------------------- normToSum1 quirk
@pytest.mark.parametrize("shape,nT", [((31, 31, 31), 4), ((31, 31, 31), 8), ((41, 41, 41), 8), ((3, 1, 1), 8)])
def test_norm_to_sum1_closed_form(oracle, shape, nT):
    a = RNG.uniform(0.0, 1.0, shape).astype(np.float32)
    flat = a.ravel().astype(np.float64)
    portions = oracle.divide_into_portions(flat.size, nT)
    s0 = flat[portions[0][0]:portions[0][0] + portions[0][1]].sum()
    S = flat.sum()
    assert oracle.sum_img_quirk(a, nT) == pytest.approx(S + s0, rel=1e-13)
    n = oracle.norm_to_sum1(a, nT)
    assert n.dtype == np.float32
    np.testing.assert_allclose(n.astype(np.float64), (flat / (S + s0)).astype(np.float32).reshape(shape), rtol=0,
                               atol=0)
    assert n.astype(np.float64).sum() == pytest.approx(S / (S + s0), rel=1e-6)


def test_norm_quirk_magnitude_c1_psf(oracle):
    """The double-counted first portion of a 31^3 Gaussian is small but far above 1e-4."""
    from mvr import synthetic as syn
    k = syn.gaussian_psf(31, 1.5, 4.5, 0.0)
    n4 = oracle.norm_to_sum1(k, 4).astype(np.float64).sum()
    n8 = oracle.norm_to_sum1(k, 8).astype(np.float64).sum()
    assert 0.9 < n4 < 0.9999 and n4 < n8 < 1.0


# ----------------------------------------------------------------- Mirror
@pytest.mark.parametrize("shape", [(5, 7, 3), (1, 9, 11), (31, 31, 31)])
def test_mirror_odd_is_flip(oracle, shape):
    a = RNG.normal(size=shape).astype(np.float32)
    np.testing.assert_array_equal(oracle.mirror(a), a[::-1, ::-1, ::-1])


def test_mirror_even_middle_pair(oracle):
    a = RNG.normal(size=(1, 1, 6)).astype(np.float32)
    m = oracle.mirror(a)[0, 0]
    x = a[0, 0]
    # sequential Mirror.mirror: pairs (0,5), (1,4) swapped; (2,3) swapped twice
    np.testing.assert_array_equal(m, np.array([x[5], x[4], x[2], x[3], x[1], x[0]], np.float32))


# ----------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", list(range(1, 33)) + [49, 77, 121, 143, 150, 160, 300, 378])
def test_fft1d_vs_numpy(oracle, n):
    x = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    np.testing.assert_allclose(oracle.fft1d(x), np.fft.fft(x), rtol=0, atol=1e-11 * max(1, n))
    np.testing.assert_allclose(oracle.fft1d(x, inverse=True), np.fft.ifft(x) * n, rtol=0, atol=1e-11 * max(1, n))


@pytest.mark.parametrize("shape", [(1, 1, 2), (5, 6, 8), (7, 12, 30), (4, 9, 14)])
def test_rfft3d_vs_numpy(oracle, shape):
    x = RNG.normal(size=shape)
    np.testing.assert_allclose(oracle.rfft3d(x), np.fft.rfftn(x), rtol=0, atol=1e-10)


def test_fast_length(oracle):
    assert oracle.fast_length(296) == 300
    assert oracle.fast_length(376) == 378
    assert oracle.fast_length(17) == 18
    assert oracle.fast_length(2) == 2


# ----------------------------------------------------------------- FFTConvolution
def _ext_index(p, n, mode):
    if mode == "mirror":
        if n == 1:
            return 0
        per = 2 * n - 2
        p %= per
        return per - p if p >= n else p
    return p if 0 <= p < n else None


def brute_conv(img, k, mode, value=0.0):
    """out[x] = sum_j K[j] ext(img)[x + c - j], c = K // 2 per axis (A.3)."""
    nz, ny, nx = img.shape
    kz, ky, kx = k.shape
    cz, cy, cx = kz // 2, ky // 2, kx // 2
    out = np.zeros(img.shape, np.float64)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                s = 0.0
                for jz in range(kz):
                    for jy in range(ky):
                        for jx in range(kx):
                            iz = _ext_index(z + cz - jz, nz, mode)
                            iy = _ext_index(y + cy - jy, ny, mode)
                            ix = _ext_index(x + cx - jx, nx, mode)
                            v = value if None in (iz, iy, ix) else img[iz, iy, ix]
                            s += k[jz, jy, jx] * v
                out[z, y, x] = s
    return out


@pytest.mark.parametrize("mode,value", [("mirror", 0.0), ("value", 0.0), ("value", 1.0)])
@pytest.mark.parametrize("kshape", [(3, 5, 3), (4, 2, 3), (1, 1, 1), (5, 5, 5)])
def test_fft_convolve_brute_force(oracle, mode, value, kshape):
    img = RNG.uniform(0.5, 2.0, (6, 7, 8))
    k = RNG.uniform(0.0, 1.0, kshape)
    got = oracle.fft_convolve(img, k, ext=mode, value=value, dtype=np.float64)
    np.testing.assert_allclose(got, brute_conv(img, k, mode, value), rtol=1e-12, atol=1e-12)


def test_fft_convolve_vs_scipy(oracle):
    img = RNG.uniform(0.5, 2.0, (20, 22, 18))
    k = RNG.uniform(0.0, 1.0, (7, 9, 5))
    np.testing.assert_allclose(oracle.fft_convolve(img, k, "mirror", dtype=np.float64),
                               ndimage.convolve(img, k, mode="mirror"), rtol=1e-11)
    np.testing.assert_allclose(oracle.fft_convolve(img, k, "value", 1.0, dtype=np.float64),
                               ndimage.convolve(img, k, mode="constant", cval=1.0), rtol=1e-11)
    f32 = oracle.fft_convolve(img, k, "mirror", dtype=np.float32).astype(np.float64)
    ref = ndimage.convolve(img, k, mode="mirror")
    assert np.linalg.norm(f32 - ref) / np.linalg.norm(ref) < 2e-7


# ----------------------------------------------------------------- PSF types (DeconViewPSF.init)
def _norm(a):
    return a / a.sum()


def _same(a, b):
    return signal.convolve(a, b, mode="same", method="direct")


def _flip(a):
    return a[::-1, ::-1, ::-1]


@pytest.mark.parametrize("psf_type", ["INDEPENDENT", "OPTIMIZATION_II", "OPTIMIZATION_I", "EFFICIENT_BAYESIAN"])
def test_init_psfs_vs_scipy(oracle, psf_type):
    V = 3
    ks = RNG.uniform(0.0, 1.0, (V, 5, 7, 5))   # asymmetric: a dropped flip changes the result
    k1, k2 = oracle.init_psfs(ks, psf_type, 8, dtype=np.float64)
    for v in range(V):
        np.testing.assert_allclose(_norm(k1[v]), _norm(ks[v]), rtol=1e-12)
        kv = ks[v]
        if psf_type == "INDEPENDENT":
            want = _flip(kv)
        elif psf_type == "OPTIMIZATION_II":
            want = _flip(kv ** V)
        elif psf_type == "OPTIMIZATION_I":
            t = kv.copy()
            for w in range(V):
                if w != v:
                    t = t * _same(kv, _flip(ks[w]))
            want = _flip(t)
        else:
            t = _flip(kv)
            for w in range(V):
                if w != v:
                    t = t * _same(_same(_flip(kv), ks[w]), _flip(ks[w]))
            want = t
        np.testing.assert_allclose(_norm(k2[v]), _norm(want), rtol=1e-9, atol=1e-15)


def test_init_psfs_delta_known_answer(oracle):
    """Delta PSFs: every compound kernel is the delta again (scaled by the normalisation)."""
    k = np.zeros((2, 5, 5, 5), np.float32)
    k[:, 2, 2, 2] = 1.0
    for t in ["INDEPENDENT", "OPTIMIZATION_II", "OPTIMIZATION_I", "EFFICIENT_BAYESIAN"]:
        k1, k2 = oracle.init_psfs(k, t, 8)
        for v in range(2):
            assert k2[v][2, 2, 2] > 0 and np.count_nonzero(k2[v]) == 1


# ----------------------------------------------------------------- update function (DeconvolutionMethods)
def f32(x):
    return np.float32(x)


def test_next_value_special_cases(oracle):
    nv = oracle.next_value
    m = f32(1e-4)
    assert nv(2.0, 1.5, 1.0, 0.0, m, 100.0) == f32(3.0)
    assert nv(2.0, 1.5, 0.0, 0.0, m, 100.0) == f32(2.0)          # weight 0: unchanged
    assert nv(2.0, 1.5, 0.5, 0.0, m, 100.0) == f32(2.5)
    # the blend last + (next - last) * w is evaluated in float, as in Java
    step_to_min = f32(f32(2.0) + f32(m - f32(2.0)) * f32(1.0))
    assert step_to_min != m and abs(float(step_to_min) - 1e-4) < 1e-7
    assert nv(2.0, 0.0, 1.0, 0.0, m, 100.0) == step_to_min       # value <= 0 -> minValue
    assert nv(2.0, -3.0, 1.0, 0.0, m, 100.0) == step_to_min
    assert nv(2.0, float("nan"), 1.0, 0.0, m, 100.0) == step_to_min   # NaN -> value > 0 false
    assert nv(1e-6, 1.0, 1.0, 0.0, m, 100.0) == f32(f32(1e-6) + f32(m - f32(1e-6)))   # clamp to minValue
    # Tikhonov closed form (sqrt(1+2 lambda v) - 1) / lambda with v normalised by maxIntensity
    lam, mx = 0.5, 2.0
    want = f32(f32((math.sqrt(1 + 2 * lam * float(f32(8.0) / f32(mx))) - 1) / lam) * f32(mx))
    assert nv(4.0, 2.0, 1.0, lam, m, mx) == want
    assert float(want) == pytest.approx((math.sqrt(5.0) - 1) / 0.5 * 2.0, rel=1e-7)
    # maxIntensity 0 -> inf * 0 = NaN -> minValue
    assert nv(4.0, 2.0, 1.0, lam, m, 0.0) == f32(f32(4.0) + f32(m - f32(4.0)))


def test_next_value_mul_geometric_mean(oracle):
    m = f32(1e-4)
    assert oracle.next_value_mul(1.0, [2.0, 8.0], [0.3, 0.9], 0.0, m, 10.0) == f32(4.0)   # sqrt(16), min(1, 1.2)
    assert oracle.next_value_mul(1.0, [2.0, 8.0], [0.25, 0.25], 0.0, m, 10.0) == f32(1.0 + 3.0 * 0.5)
    assert oracle.next_value_mul(1.0, [2.0, 0.0], [1.0, 1.0], 0.0, m, 10.0) == f32(f32(1.0) + f32(m - f32(1.0)))


def test_next_value_f64_tracks_f32(oracle):
    for _ in range(200):
        last, integ, w = RNG.uniform(0.01, 50), RNG.uniform(0.1, 3), RNG.uniform(0, 1)
        a = oracle.next_value(last, integ, w, 6e-4, 1e-4, 300.0)
        b = oracle.next_value(last, integ, w, 6e-4, 1e-4, 300.0, dtype=np.float64)
        assert float(a) == pytest.approx(b, rel=1e-6)


def next_value_numpy(last, integral, weight, lam, min_i, max_i):
    """computeNextValue (DeconvolutionMethods.java:320-358) vectorised in numpy: every float32 and float64
    operation rounds on its own, as Java's do."""
    last, integral, weight = (np.asarray(a, np.float32) for a in (last, integral, weight))
    m, mx = np.float32(min_i), np.float32(max_i)
    with np.errstate(all="ignore"):
        value = last * integral
        if lam > 0:
            vn = (value / mx).astype(np.float64)
            adj = ((np.sqrt(1.0 + (2.0 * lam) * vn) - 1.0) / lam).astype(np.float32) * mx
        else:
            adj = value
        adj = np.where(value > 0, adj, m).astype(np.float32)
        nxt = np.where(np.isnan(adj), m, np.maximum(m, adj)).astype(np.float32)
        return last + (nxt - last) * weight


def update_inputs(n, seed):
    """psi, integral, weight with the special values of the update: integral <= 0, NaN, inf and tiny; weights 0
    and 1; psi at minValue."""
    rng = np.random.default_rng(seed)
    last = rng.uniform(1e-4, 800.0, n).astype(np.float32)
    last[::97] = np.float32(1e-4)
    integral = rng.uniform(0.2, 3.0, n).astype(np.float32)
    integral[::13] = 0.0
    integral[::17] = -1.5
    integral[::19] = np.nan
    integral[::23] = np.inf
    integral[::29] = np.float32(1e-30)
    weight = rng.uniform(0.0, 1.0, n).astype(np.float32)
    weight[::7] = 0.0
    weight[::11] = 1.0
    return last, integral, weight


@pytest.mark.parametrize("lam,max_i", [(0.0, 300.0), (6e-4, 300.0), (0.006, 1234.5), (6e-4, 0.0)])
def test_final_values_array_matches_numpy_restatement(oracle, lam, max_i):
    last, integral, weight = update_inputs(5003, 3)
    got, (s, mx) = oracle.compute_final_values(last, integral, weight, lam, 1e-4, max_i, nT=8)
    want = next_value_numpy(last, integral, weight, lam, 1e-4, max_i)
    np.testing.assert_array_equal(got, want)
    # an infinite integral with weight 0 gives NaN (inf * 0); Java's Math.max keeps a NaN maxChange
    ch = (want - last).astype(np.float64)
    assert np.isnan(ch).any() == np.isnan(mx) == np.isnan(s)
    assert np.isnan(ch).any() or max_i == 0.0   # with maxIntensity 0 every positive value becomes minValue
    fin = np.isfinite(integral)
    got, (s, mx) = oracle.compute_final_values(last[fin], integral[fin], weight[fin], lam, 1e-4, max_i, nT=8)
    ch = (got - last[fin]).astype(np.float64)
    assert mx == ch.max()
    assert s == pytest.approx(ch.sum(), rel=1e-12, abs=1e-9)
    # scalar entry point agrees with the array one
    for i in range(0, fin.sum(), 251):
        assert oracle.next_value(last[fin][i], integral[fin][i], weight[fin][i], lam, 1e-4, max_i) == got[i]


def test_final_values_mul_array_matches_numpy_restatement(oracle):
    rng = np.random.default_rng(4)
    n, V = 3001, 3
    last = rng.uniform(1e-4, 500.0, n).astype(np.float32)
    integ = rng.uniform(0.3, 3.0, (V, n)).astype(np.float32)
    integ[1, ::31] = 0.0
    w = rng.uniform(0.0, 0.6, (V, n)).astype(np.float32)
    got, _ = oracle.compute_final_values_mul(last, integ, w, 6e-4, 1e-4, 250.0)
    prod = np.ones(n)
    sw = np.zeros(n)
    for v in range(V):                       # Java multiplies and adds view by view, in double
        prod = prod * integ[v].astype(np.float64)
        sw = sw + w[v].astype(np.float64)
    prod = np.power(prod, 1.0 / V)
    sw = np.minimum(1.0, sw)
    with np.errstate(all="ignore"):
        value = last * prod.astype(np.float32)
        vn = (value / np.float32(250.0)).astype(np.float64)
        adj = ((np.sqrt(1.0 + (2.0 * 6e-4) * vn) - 1.0) / 6e-4).astype(np.float32) * np.float32(250.0)
        adj = np.where(value > 0, adj, np.float32(1e-4)).astype(np.float32)
        nxt = np.maximum(np.float32(1e-4), adj)
        want = last + (nxt - last) * sw.astype(np.float32)
    np.testing.assert_array_equal(got, want)


def test_quotient_array(oracle):
    rng = np.random.default_rng(5)
    img = rng.uniform(-2.0, 50.0, 4000).astype(np.float32)
    img[::9] = 0.0
    blurred = rng.uniform(0.1, 40.0, 4000).astype(np.float32)
    got = oracle.compute_quotient(img, blurred)
    np.testing.assert_array_equal(got, np.where(img > 0, img / blurred, np.float32(1.0)).astype(np.float32))


# ----------------------------------------------------------------- blocks (BlockGeneratorFixedSizePrecise, BlockSorter)
def test_blocks_reference_main_inputs(oracle):
    """Block.main (C/Block.java:421-436): 1024^2 image, kernel 16x32, block 384^2 (z kept trivial)."""
    b = oracle.divide_into_blocks((1024, 1024, 1), (16, 32, 1), (384, 384, 1))
    assert len(b) == 9
    ex, ey = 384 - 16 + 1, 384 - 32 + 1
    for i, blk in enumerate(b):
        bx, by = i % 3, i // 3                      # x fastest enumeration
        assert blk["effectiveOffset"] == (bx * ex, by * ey, 0)
        assert blk["offset"] == (bx * ex - 8, by * ey - 16, 0)
        assert blk["effectiveSize"] == (min(ex, 1024 - bx * ex), min(ey, 1024 - by * ey), 1)
        assert blk["effectiveLocalOffset"] == (8, 16, 0)
        assert blk["blockSize"] == (384, 384, 1)
    assert b[2]["effectiveSize"][0] == 286 and b[6]["effectiveSize"][1] == 318


def test_blocks_single_block_c2(oracle):
    b = oracle.divide_into_blocks((256, 256, 256), (81, 81, 81), (336, 336, 336))
    assert len(b) == 1 and b[0]["offset"] == (-40, -40, -40) and b[0]["effectiveSize"] == (256, 256, 256)
    assert oracle.divide_into_blocks((256, 256, 256), (81, 81, 81), (80, 80, 80)) is None


@pytest.mark.parametrize("min_req,dim,layers,per_layer", [(1, 0, 5, 6), (8, 1, 3, 10), (100, 2, 2, 15)])
def test_sort_blocks_known(oracle, min_req, dim, layers, per_layer):
    batches, d = oracle.sort_blocks((100, 60, 40), (11, 11, 11), (30, 30, 30), min_req)
    assert d == dim and len(batches) == layers and all(len(b) == per_layer for b in batches)
    blocks = oracle.divide_into_blocks((100, 60, 40), (11, 11, 11), (30, 30, 30))
    for i, batch in enumerate(batches):
        assert {blocks[j]["effectiveOffset"][dim] for j in batch} == {i * 20}


def test_sort_blocks_tie_takes_last_dimension(oracle):
    batches, d = oracle.sort_blocks((40, 40, 40), (11, 11, 11), (30, 30, 30), 1)
    assert d == 2 and len(batches) == 2


# ----------------------------------------------------------------- PsiInitAvgPrecise
def test_psi_init_closed_form(oracle):
    imgs = RNG.uniform(-1.0, 5.0, (3, 9, 8, 7)).astype(np.float32)
    imgs[:, 0, 0, :] = 0.0
    avg, mx = oracle.psi_init_avg_precise(imgs, 8)
    pos = imgs > 0
    cnt = pos.sum(0)
    s = np.where(pos, imgs.astype(np.float64), 0.0).sum(0)
    cov = cnt > 0
    want = (s[cov] / cnt[cov]).sum() / cov.sum()
    assert avg == pytest.approx(want, rel=1e-13)
    np.testing.assert_array_equal(mx, np.where(pos, imgs, 0).reshape(3, -1).max(1))


# ----------------------------------------------------------------- full iteration
def _small_inputs(n=(16, 18, 20), V=3, k=5, seed=3):
    rng = np.random.default_rng(seed)
    imgs = rng.uniform(5.0, 50.0, (V,) + n).astype(np.float32)
    imgs[0, :2] = 0.0                                  # uncovered voxels -> ratio 1
    w = rng.uniform(0.0, 1.0, (V,) + n).astype(np.float32)
    ks = rng.uniform(0.0, 1.0, (V, k, k + 2, k)).astype(np.float32)
    return imgs, w, ks


def global_form_seq(imgs, w, k1, k2, psi, lam, min_value, mx, iterations):
    """A.9 global form of the Seq scheme with direct spatial convolutions (scipy.ndimage)."""
    psi = psi.astype(np.float64).copy()
    for _ in range(iterations):
        for v in range(imgs.shape[0]):
            blurred = ndimage.convolve(psi, k1[v], mode="mirror")
            img = imgs[v].astype(np.float64)
            ratio = np.where(img > 0, img / np.where(img > 0, blurred, 1.0), 1.0)
            integral = ndimage.convolve(ratio, k2[v], mode="constant", cval=1.0)
            value = psi * integral
            if lam > 0:
                adj = (np.sqrt(1 + 2 * lam * np.maximum(value, 0) / mx[v]) - 1) / lam * mx[v]
            else:
                adj = value
            adj = np.where(value > 0, adj, min_value)
            nxt = np.maximum(min_value, adj)
            psi = psi + (nxt - psi) * w[v]
    return psi


@pytest.mark.parametrize("psf_type,lam", [("EFFICIENT_BAYESIAN", 0.0), ("OPTIMIZATION_II", 6e-4),
                                          ("OPTIMIZATION_I", 0.006), ("INDEPENDENT", 0.0)])
def test_deconvolve_f64_vs_global_form(oracle, psf_type, lam):
    imgs, w, ks = _small_inputs()
    r = oracle.deconvolve(imgs, w, ks, psf_type=psf_type, lam=lam, iterations=2, dtype=np.float64)
    k1, k2 = oracle.init_psfs(ks, psf_type, 8, dtype=np.float64)
    avg, mx = oracle.psi_init_avg_precise(imgs, 8)
    psi0 = np.full(imgs.shape[1:], avg)
    lam_d = float(np.float32(lam))                    # lambda is a Java float widened to double
    want = global_form_seq(imgs, w, k1, k2, psi0, lam_d, float(np.float32(1e-4)), mx.astype(np.float64), 2)
    assert np.linalg.norm(r["psi"] - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
def test_block_independence(oracle, scheme):
    """1 block == 8 blocks when the effective size E >= K-1 (lagged write-back is exact)."""
    imgs, w, ks = _small_inputs(n=(16, 16, 16), k=5)
    ks = ks[:, :, 1:-1, :].copy()                    # 5^3 kernels -> kernelSize 9, margin 4
    one = oracle.deconvolve(imgs, w, ks, scheme=scheme, iterations=2, dtype=np.float64)
    many = oracle.deconvolve(imgs, w, ks, scheme=scheme, iterations=2, block_size=(16, 16, 16), dtype=np.float64)
    assert np.linalg.norm(one["psi"] - many["psi"]) / np.linalg.norm(one["psi"]) < 1e-11


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
def test_f32_tracks_f64(oracle, scheme):
    imgs, w, ks = _small_inputs()
    a = oracle.deconvolve(imgs, w, ks, scheme=scheme, lam=6e-4, iterations=4)
    b = oracle.deconvolve(imgs, w, ks, scheme=scheme, lam=6e-4, iterations=4, dtype=np.float64)
    assert np.linalg.norm(a["psi"] - b["psi"]) / np.linalg.norm(b["psi"]) < 1e-5


def test_mul_global_form(oracle):
    """Mul scheme: geometric mean of the per-view integrals, min(1, sum w), mean of the maxima."""
    imgs, w, ks = _small_inputs(V=2)
    lam = 6e-4
    r = oracle.deconvolve(imgs, w, ks, scheme="MUL", psf_type="OPTIMIZATION_I", lam=lam, iterations=1,
                          dtype=np.float64)
    k1, k2 = oracle.init_psfs(ks, "OPTIMIZATION_I", 8, dtype=np.float64)
    avg, mx = oracle.psi_init_avg_precise(imgs, 8)
    psi = np.full(imgs.shape[1:], avg)
    integ = []
    for v in range(2):
        blurred = ndimage.convolve(psi, k1[v], mode="mirror")
        img = imgs[v].astype(np.float64)
        ratio = np.where(img > 0, img / np.where(img > 0, blurred, 1.0), 1.0)
        integ.append(ndimage.convolve(ratio, k2[v], mode="constant", cval=1.0))
    lam = float(np.float32(lam))
    prod = np.sqrt(integ[0] * integ[1])
    sw = np.minimum(1.0, w[0].astype(np.float64) + w[1])
    m = float(np.float32((float(mx[0]) + float(mx[1])) / 2))
    value = psi * prod
    adj = (np.sqrt(1 + 2 * lam * value / m) - 1) / lam * m
    mv = float(np.float32(1e-4))
    nxt = np.maximum(mv, np.where(value > 0, adj, mv))
    want = psi + (nxt - psi) * sw
    assert np.linalg.norm(r["psi"] - want) / np.linalg.norm(want) < 1e-12


def test_constant_image_is_fixed_point(oracle):
    """Constant data c, psi0 = avg = c: K1*psi = s*c, ratio = 1/s, K2*ratio = s2/s, so with K2 = flip(K1)
    (s2 = s, whatever the normalisation quirk made s) the interior stays exactly c after an iteration."""
    c = 37.0
    img = np.full((1, 14, 15, 16), c, np.float32)
    k = np.random.default_rng(9).uniform(0.1, 1.0, (1, 3, 5, 3)).astype(np.float32)
    r = oracle.deconvolve(img, np.ones_like(img), k, psf_type="INDEPENDENT", iterations=1, dtype=np.float64)
    inner = r["psi"][2:-2, 3:-3, 2:-2]
    np.testing.assert_allclose(inner, c, rtol=1e-12)
    assert not np.allclose(r["psi"][0], c, rtol=1e-6)    # ratio is extended with 1, not c/s: the faces move


def test_stats_are_signed_net_change(oracle):
    imgs, w, ks = _small_inputs()
    r = oracle.deconvolve(imgs, w, ks, iterations=1, dtype=np.float64)
    r0 = oracle.deconvolve(imgs, w, ks, iterations=0, dtype=np.float64)
    # one block, Seq: sum over the block includes its mirrored margins; the sign is kept (no abs)
    assert r["stats"].shape == (1, 3, 2)
    assert np.all(r["stats"][0, :, 1] >= -1)
    assert r0["psi"].std() == 0.0


# ---------------------------------------------------------------- PsiInitAvgApprox / PsiInitBlurredFused
def test_avg_approx_closed_form(oracle):
    rng = np.random.default_rng(31)
    imgs = rng.uniform(0, 50, (2, 5, 6, 7)).astype(np.float32)
    imgs[0, :, :, 3] = np.arange(30, dtype=np.float32).reshape(5, 6)  # x = 7 // 2 = 3
    imgs[1, :, :, 3] = 4.0
    avg, mx = oracle.psi_init_avg_approx(imgs)
    assert avg == pytest.approx((14.5 + 4.0) / 2, rel=1e-14)
    assert mx[0] == 29.0 and mx[1] == 4.0
    # voxels off the middle slice do not matter
    imgs2 = imgs.copy()
    imgs2[:, :, :, :3] = 1e6
    assert oracle.psi_init_avg_approx(imgs2)[0] == avg


def test_gauss3_half_kernel():
    import oracle as o
    for sigma in (0.5, 1.0, 2.3, 5.0):
        k = o.gauss3_half_kernel(sigma)
        assert len(k) == max(2, int(3 * sigma + 0.5) + 1)
        assert k[0] + 2 * k[1:].sum() == pytest.approx(1.0, abs=1e-14)
        assert np.all(np.diff(k) < 0)
    k = o.gauss3_half_kernel(5.0)
    assert len(k) == 16 and k[3] / k[0] == pytest.approx(np.exp(-9 / 50), rel=1e-14)


def test_blurred_fused_constant_and_delta(oracle):
    # constant positive views: the fusion and its blur are that constant
    imgs = np.full((2, 9, 10, 11), 7.0, np.float32)
    ws = np.full_like(imgs, 0.5)
    assert np.abs(oracle.psi_init_blurred_fused(imgs, ws, 2.0) - 7.0).max() < 1e-12
    # one bright voxel far from the borders: the response is the separable product of the kernel
    imgs = np.zeros((1, 21, 21, 21), np.float32)
    imgs[0, 10, 10, 10] = 1.0
    ws = np.ones_like(imgs)
    out = oracle.psi_init_blurred_fused(imgs, ws, 1.5)
    k = oracle.gauss3_half_kernel(1.5)
    for dz, dy, dx in ((0, 0, 0), (1, 0, 0), (2, 3, 1), (0, 4, 4)):
        assert out[10 + dz, 10 + dy, 10 + dx] == pytest.approx(k[dz] * k[dy] * k[dx], rel=1e-12)
    # positive-only fusion: a zero view does not pull the weighted mean down
    imgs = np.stack([np.full((4, 4, 4), 3.0, np.float32), np.zeros((4, 4, 4), np.float32)])
    assert np.abs(oracle.psi_init_blurred_fused(imgs, np.ones_like(imgs), 1.0) - 3.0).max() < 1e-12


def test_blurred_fused_borders_against_reflect_padding(oracle):
    rng = np.random.default_rng(5)
    imgs = rng.uniform(0.5, 9.0, (2, 6, 7, 8)).astype(np.float32)
    ws = rng.uniform(0.1, 1.0, imgs.shape).astype(np.float32)
    sigma = 1.2
    got = oracle.psi_init_blurred_fused(imgs, ws, sigma)
    fused = ((imgs.astype(np.float64) * ws).sum(0) / ws.astype(np.float64).sum(0)).astype(np.float32).astype(np.float64)
    k = oracle.gauss3_half_kernel(sigma)
    full = np.concatenate([k[:0:-1], k])
    H = len(k) - 1
    out = fused
    for axis in (2, 1, 0):
        pad = [(0, 0)] * 3
        pad[axis] = (H, H)
        p = np.pad(out, pad, mode="reflect")  # mirror-single: d c b | a b c d
        out = np.apply_along_axis(lambda r: np.convolve(r, full, mode="valid"), axis, p)
    assert np.abs(got - out).max() < 1e-12
