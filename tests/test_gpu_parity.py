"""Parity of the HIP library (called through its C ABI) with the CPU oracle.

Tolerances: the parity bar of the path is FP32 results within rel-L2 1e-4 of
the oracle's FP32 restatement after 10 iterations (SPEC).  The deconvolution
cases reach 1e-7 to 1e-6, so they assert BAR = 1e-5: a regression by an order of
magnitude fails long before the spec bar is crossed.  Single FFT convolutions
within 3e-6 and one block iteration within 1e-5; every deconvolution is also
checked against the FP64 oracle.
"""
import numpy as np
import pytest

from mvr import decon, hip
from mvr import synthetic as syn

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)
SPEC = 1e-4   # parity bar of the path (rel-L2 after 10 iterations)
BAR = 1e-5    # what the deconvolution cases below reach, with an order of magnitude to spare


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_device_query():
    assert hip.num_devices() >= 1
    name = hip.device_name(0)
    print("device", name)
    assert hip.lib().mvr_get_mem_device(0) > 0


@pytest.mark.parametrize("shape,kshape,ext,value", [
    ((20, 22, 18), (7, 9, 5), "mirror", 0.0),
    ((16, 16, 16), (5, 5, 5), "value", 1.0),
    ((12, 10, 30), (3, 4, 6), "value", 0.0),
    ((33, 7, 40), (9, 3, 11), "mirror", 0.0),
    ((1, 8, 8), (1, 3, 3), "mirror", 0.0),
    ((264, 6, 444), (7, 3, 7), "mirror", 0.0),   # FFT lengths 270 (z) and 450 (x: 225-point halves)
    ((10, 264, 84), (5, 7, 7), "value", 1.0),    # 270 along y
    ((4, 5, 3), (13, 11, 9), "mirror", 0.0),     # kernel wider than the volume: the mirror wraps repeatedly
    ((6, 2, 7), (5, 9, 5), "mirror", 0.0),
])
def test_fft_convolve_matches_oracle(oracle, shape, kshape, ext, value):
    img = RNG.uniform(0.5, 2.0, shape).astype(np.float32)
    k = RNG.uniform(0.0, 1.0, kshape).astype(np.float32)
    got = hip.fft_convolve(img, k, ext, value)
    want = oracle.fft_convolve(img.astype(np.float64), k.astype(np.float64), ext, value, dtype=np.float64)
    assert rel(got, want) < 3e-6


def test_every_compiled_fft_length(oracle):
    # each compiled length L as the FFT length of one axis at a time (extent L - 2, 3-voxel kernel, the other
    # axes short): the register FFTs of every length, x halves included, against the FP64 oracle
    lengths = hip.fft_lengths()
    assert lengths == sorted(lengths) and lengths[-1] >= 1152
    rng = np.random.default_rng(5)
    worst = 0.0
    for L in lengths:
        for axis in range(3):
            shape = [3, 4, 5]            # z, y, x
            shape[2 - axis] = L - 2
            img = rng.uniform(0.5, 2.0, shape).astype(np.float32)
            k = rng.uniform(0.0, 1.0, (3, 3, 3)).astype(np.float32)
            got = hip.fft_convolve(img, k, "mirror", 0.0)
            want = oracle.fft_convolve(img.astype(np.float64), k.astype(np.float64), "mirror", dtype=np.float64)
            e = rel(got, want)
            worst = max(worst, e)
            assert e < 3e-6, (L, axis, e)
    print(f"{len(lengths)} lengths x 3 axes, worst rel-L2 {worst:.2e}")


@pytest.mark.parametrize("shape", [(574, 6, 1150), (6, 1150, 1278)])
def test_resident_decon_long_axes(oracle, shape):
    # the x-ratio / x-update passes and the strided passes at the long lengths of a 1024^2 x 512 frame with
    # 61^3 PSFs (FFT grid 1152 x 1152 x 576), here with short PSFs and thin volumes
    rng = np.random.default_rng(sum(shape))
    V = 2
    imgs = rng.uniform(1.0, 40.0, (V,) + shape).astype(np.float32)
    ws = rng.uniform(0.0, 1.0, (V,) + shape).astype(np.float32)
    ks = rng.uniform(0.01, 1.0, (V, 3, 3, 3)).astype(np.float32)
    ref = oracle.deconvolve(imgs, ws, ks, psf_type="EFFICIENT_BAYESIAN", lam=6e-4, iterations=3)
    out = []
    for widths in ((8, 8), (16, 16)):   # 8- and 16-column tiles of the 1152-long passes
        d = hip.Decon(shape, V, "EFFICIENT_BAYESIAN", "SEQ", 6e-4)
        for v in range(V):
            d.set_view(v, imgs[v], ws[v], ks[v])
        d.init()
        d.set_tile_widths(*widths)
        d.run(3)
        P = d.geometry()["P"]
        out.append(d.psi())
        e = rel(out[-1], ref["psi"])
        d.close()
        print(f"{shape} tiles {widths}: FFT grid {P}, rel-L2 vs oracle {e:.2e}")
        assert max(P) >= 1152
        assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    np.testing.assert_array_equal(out[0], out[1])   # the tiling does not change the arithmetic


@pytest.mark.parametrize("px", [720, 800, 864, 900, 960, 1080])
def test_resident_decon_long_x_rows(oracle, px):
    # x lengths whose quotient pass forms img / blurred in the registers between its two transforms (rows of
    # px / 2 complex values split with more than 16 per thread); zero image pixels take the quotient's 1 branch
    rng = np.random.default_rng(px)
    shape = (5, 6, px - 2)
    V = 2
    imgs = rng.uniform(1.0, 40.0, (V,) + shape).astype(np.float32)
    imgs[:, :, :, ::7] = 0.0
    ws = rng.uniform(0.0, 1.0, (V,) + shape).astype(np.float32)
    ks = rng.uniform(0.01, 1.0, (V, 3, 3, 3)).astype(np.float32)
    ref = oracle.deconvolve(imgs, ws, ks, psf_type="EFFICIENT_BAYESIAN", lam=6e-4, iterations=3)
    d = hip.Decon(shape, V, "EFFICIENT_BAYESIAN", "SEQ", 6e-4)
    for v in range(V):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    d.run(3)
    P = d.geometry()["P"]
    e = rel(d.psi(), ref["psi"])
    d.close()
    print(f"{shape}: FFT grid {P}, rel-L2 vs oracle {e:.2e}")
    assert P[0] == px
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"


def test_inplace_convolution_is_mirror_convolution(oracle):
    img = RNG.uniform(0.5, 2.0, (24, 20, 16)).astype(np.float32)
    k = RNG.uniform(0.0, 1.0, (5, 7, 3)).astype(np.float32)
    got = hip.convolution3d_fft_inplace(img, k)
    want = oracle.fft_convolve(img.astype(np.float64), k.astype(np.float64), "mirror", dtype=np.float64)
    assert rel(got, want) < 3e-6


@pytest.mark.parametrize("psf_type", ["INDEPENDENT", "OPTIMIZATION_II", "OPTIMIZATION_I", "EFFICIENT_BAYESIAN"])
def test_init_psfs_matches_oracle(oracle, psf_type):
    ks = np.stack([syn.gaussian_psf(9, 1.2, 2.5, a) for a in (0, 100, 230)])
    k1, k2 = hip.init_psfs(ks, psf_type, 8)
    o1, o2 = oracle.init_psfs(ks, psf_type, 8)
    np.testing.assert_array_equal(k1, o1)          # normToSum1 is host arithmetic: bit-exact
    assert rel(k2, o2) < 2e-6


def _block_case(V=3, B=(24, 26, 22), K=(5, 7, 5), seed=5):
    rng = np.random.default_rng(seed)
    Bz = B[::-1]
    psi = rng.uniform(5, 30, Bz).astype(np.float32)
    imgs = rng.uniform(1, 60, (V,) + Bz).astype(np.float32)
    imgs[0, :, :3] = 0.0
    ws = rng.uniform(0, 1, (V,) + Bz).astype(np.float32)
    ks = np.stack([syn.gaussian_psf(K[0], 1.0, 2.0, a) for a in np.linspace(0, 300, V)])
    ks = ks[:, :, 1:-1, :] if K[1] < K[0] else np.pad(ks, ((0, 0), (0, 0), (1, 1), (0, 0)))
    return psi, imgs, ws, ks


@pytest.mark.parametrize("lam", [0.0, 6e-4])
def test_block_seq_matches_oracle(oracle, lam):
    psi, imgs, ws, ks = _block_case()
    k1, k2 = oracle.init_psfs(ks, "EFFICIENT_BAYESIAN", 8)
    B = psi.shape[::-1]
    t = hip.BlockThread(B, 1e-4, lam)
    for v in range(3):
        t.set_view_kernels(v, k1[v], k2[v])
    cur_g, cur_o = psi.copy(), psi.copy()
    for v in range(3):
        cur_g, st_g = t.run_iteration(v, cur_g, imgs[v], ws[v], 55.0)
        cur_o, st_o = oracle.block_seq(cur_o, imgs[v], ws[v], k1[v], k2[v], lam, 1e-4, 55.0, 8)
        assert rel(cur_g, cur_o) < 1e-5
        assert st_g[0] == pytest.approx(st_o[0], rel=1e-3, abs=1e-3)
        assert st_g[1] == pytest.approx(st_o[1], rel=1e-4)
    t.close()


def test_block_mul_matches_oracle(oracle):
    psi, imgs, ws, ks = _block_case()
    k1, k2 = oracle.init_psfs(ks, "OPTIMIZATION_I", 8)
    B = psi.shape[::-1]
    t = hip.BlockThread(B, 1e-4, 6e-4)
    for v in range(3):
        t.set_view_kernels(v, k1[v], k2[v])
    mx = np.array([60.0, 58.0, 59.5], np.float32)
    g, st_g = t.run_iteration_mul(psi, imgs, ws, mx)
    o, st_o = oracle.block_mul(psi, imgs, ws, k1, k2, 6e-4, 1e-4, mx, 8)
    assert rel(g, o) < 1e-5
    assert st_g[1] == pytest.approx(st_o[1], rel=1e-4)
    t.close()


def test_block_iteration_at_equals_host_cut_blocks():
    # the device-side cut of the resident view data (zero outside the volume) gives bit-identical results to the
    # host's copy_block_zero, for blocks that reach past every face of the volume and for an interior one
    rng = np.random.default_rng(9)
    V, N = 3, (26, 30, 34)                     # z, y, x
    imgs = rng.uniform(1, 60, (V,) + N).astype(np.float32)
    ws = rng.uniform(0, 1, (V,) + N).astype(np.float32)
    ks = np.stack([syn.gaussian_psf(5, 1.0, 2.0, a) for a in (0, 100, 230)])
    k1, k2 = hip.init_psfs(ks, "EFFICIENT_BAYESIAN", 8)
    B = (20, 18, 16)                            # x, y, z
    t = hip.BlockThread(B, 1e-4, 6e-4)
    for v in range(V):
        t.set_view_kernels(v, k1[v], k2[v])
        t.set_view_data(v, imgs[v], ws[v])
    for off in [(-4, -4, -4), (20, 16, 14), (5, 6, 7), (-30, 2, 3), (40, 40, 40)]:
        b = decon.Block(B, off, B, off, (0, 0, 0))
        psi = rng.uniform(5, 30, B[::-1]).astype(np.float32)
        for v in range(V):
            a, sa = t.run_iteration(v, psi, decon.copy_block_zero(imgs[v], b), decon.copy_block_zero(ws[v], b), 55.0)
            c, sc = t.run_iteration_at(v, psi, off, 55.0)
            np.testing.assert_array_equal(a, c)
            assert sa == sc
        mx = np.array([60.0, 58.0, 59.5], np.float32)
        a, sa = t.run_iteration_mul(psi, [decon.copy_block_zero(imgs[v], b) for v in range(V)],
                                    [decon.copy_block_zero(ws[v], b) for v in range(V)], mx)
        c, sc = t.run_iteration_mul_at(psi, off, mx)
        np.testing.assert_array_equal(a, c)
        assert sa == sc
    t.close()


def _expected_blocks_with_content(ws, K, B):
    """Per view, the blocks of BlockGeneratorFixedSizePrecise whose interval (effective region at i E, margin
    K - 1, size B) holds a non-zero weight, counted with numpy on the zero-extended weights."""
    V, nz, ny, nx = ws.shape
    E = [b - 2 * k + 2 for b, k in zip(B, K)]
    nb = [-(-n // e) for n, e in zip((nx, ny, nz), E)]
    counts = []
    for v in range(V):
        c = 0
        for k in range(nb[2]):
            for j in range(nb[1]):
                for i in range(nb[0]):
                    lo = [max(0, q * e - (kk - 1)) for q, e, kk in zip((i, j, k), E, K)]
                    hi = [min(n, q * e - (kk - 1) + b) for q, e, kk, b, n in zip((i, j, k), E, K, B, (nx, ny, nz))]
                    c += bool(ws[v, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]].any())
        counts.append(c)
    return nb[0] * nb[1] * nb[2], counts


@pytest.mark.parametrize("psf_type,lam", [("EFFICIENT_BAYESIAN", 6e-4), ("OPTIMIZATION_II", 0.0)])
def test_batched_blocks_match_blocked_oracle(oracle, psf_type, lam):
    # the reference's block structure in one launch per pass: 22^3 blocks of a 60 x 26 x 30 volume with 5^3 PSFs
    # (effective 14^3, uneven tail blocks); view 1 has no weight beyond x = 35, so DeconView's content filter
    # drops its last two block columns
    rng = np.random.default_rng(31)
    V, shape, K = 3, (30, 26, 60), (5, 5, 5)
    imgs = rng.uniform(1.0, 40.0, (V,) + shape).astype(np.float32)
    ws = rng.uniform(0.0, 1.0, (V,) + shape).astype(np.float32)
    ws[1, :, :, 36:] = 0.0
    ks = np.stack([syn.gaussian_psf(5, 1.0, 2.0, a) for a in (0, 120, 240)])
    B = (22, 22, 22)
    ref = oracle.deconvolve(imgs, ws, ks, psf_type=psf_type, lam=lam, iterations=4, block_size=B)
    d = hip.Decon(shape, V, psf_type, "SEQ", lam)
    for v in range(V):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.set_block_size(B)
    d.init()
    n, per = d.blocks()
    n_want, per_want = _expected_blocks_with_content(ws, K, B)
    print(f"{n} blocks, per view {per} (numpy {per_want}), FFT grid {d.geometry()['P']}")
    assert n == n_want and per == per_want and per[1] < per[0]
    d.run(4)
    e = rel(d.psi(), ref["psi"])
    st = d.stats()
    d.close()
    print(f"{psf_type}: batched blocks vs blocked oracle rel-L2 {e:.2e}")
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    # statistics count every voxel once (effective regions); the reference's blocked drivers also count the
    # discarded block margins, so only their form is checked here (one {sum, max} per view update)
    assert st.shape == (4, V, 2) and np.isfinite(st).all()


def test_batched_blocks_equal_frame_where_no_block_is_dropped():
    # with content in every block, the batched blocks and the one-frame schedule compute the same deconvolution
    inp = _small_problem(n=(30, 26, 28), K=5)
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    out = []
    for B in ((0, 0, 0), (22, 22, 22), (24, 20, 30)):
        d = hip.Decon(imgs.shape[1:], 3, "EFFICIENT_BAYESIAN", "SEQ", 6e-4)
        for v in range(3):
            d.set_view(v, imgs[v], ws[v], ks[v])
        d.set_block_size(B)
        d.init()
        d.run(3)
        out.append((d.psi(), d.stats()))
        d.close()
    for o, st in out[1:]:
        e = rel(o, out[0][0])
        assert e < 1e-6, e
        # each voxel updated once: the same statistics as the frame (sumChange to 1e-6 of the flux: it cancels)
        np.testing.assert_allclose(st[..., 1], out[0][1][..., 1], rtol=1e-4)
        assert np.abs(st[..., 0] - out[0][1][..., 0]).max() <= 1e-6 * float(out[0][0].sum(dtype=np.float64))


def _small_problem(n=(28, 30, 26), V=3, K=7, seed=11, weights="uniform"):
    w = syn.small_workload(views=V, k=K, seed_index=seed, n_blobs=25, weights=weights, shape=n)
    return syn.make_inputs(w)


@pytest.mark.parametrize("psf_type,scheme,lam", [
    ("EFFICIENT_BAYESIAN", "SEQ", 0.0),
    ("OPTIMIZATION_II", "SEQ", 6e-4),
    ("OPTIMIZATION_I", "SEQ", 0.006),
    ("INDEPENDENT", "SEQ", 0.0),
    ("EFFICIENT_BAYESIAN", "MUL", 6e-4),
    ("OPTIMIZATION_II", "MUL", 0.0),
])
def test_resident_decon_matches_oracle(oracle, psf_type, scheme, lam):
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    it = 10
    ref = oracle.deconvolve(imgs, ws, ks, psf_type=psf_type, scheme=scheme, lam=lam, iterations=it)
    ref64 = oracle.deconvolve(imgs, ws, ks, psf_type=psf_type, scheme=scheme, lam=lam, iterations=it,
                              dtype=np.float64)
    d = hip.Decon(imgs.shape[1:], imgs.shape[0], psf_type, scheme, lam)
    for v in range(imgs.shape[0]):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    mx, avg = d.max_avg()
    np.testing.assert_array_equal(mx, ref["max"])
    assert avg == pytest.approx(ref["avg"], rel=1e-12)
    d.run(it)
    psi = d.psi()
    e32, e64 = rel(psi, ref["psi"]), rel(psi, ref64["psi"])
    print(f"{psf_type} {scheme} lam={lam}: rel-L2 vs oracle FP32 {e32:.2e}, vs FP64 {e64:.2e}, "
          f"oracle FP32 vs FP64 {rel(ref['psi'], ref64['psi']):.2e}")
    assert e32 < BAR, f"rel-L2 {e32:.2e} (bar {BAR}, spec {SPEC})"
    assert e64 < BAR, f"rel-L2 {e64:.2e} (bar {BAR}, spec {SPEC})"
    assert d.stats().shape == ((it, 3, 2) if scheme == "SEQ" else (it, 1, 2))
    d.close()


@pytest.mark.parametrize("shape,kshape,psf_type,scheme", [
    ((20, 22, 18), (6, 8, 4), "EFFICIENT_BAYESIAN", "SEQ"),   # even PSF sizes: the double-swap Mirror
    ((17, 15, 19), (5, 7, 6), "OPTIMIZATION_I", "MUL"),
    ((1, 40, 36), (1, 7, 9), "OPTIMIZATION_II", "SEQ"),       # a single z plane (2D data)
])
def test_resident_decon_edge_shapes(oracle, shape, kshape, psf_type, scheme):
    rng = np.random.default_rng(sum(shape) + sum(kshape))
    V = 3
    imgs = rng.uniform(1.0, 40.0, (V,) + shape).astype(np.float32)
    ws = rng.uniform(0.0, 1.0, (V,) + shape).astype(np.float32)
    ks = rng.uniform(0.01, 1.0, (V,) + kshape).astype(np.float32)
    ref = oracle.deconvolve(imgs, ws, ks, psf_type=psf_type, scheme=scheme, lam=6e-4, iterations=5)
    d = hip.Decon(shape, V, psf_type, scheme, 6e-4)
    for v in range(V):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    d.run(5)
    e = rel(d.psi(), ref["psi"])
    print(f"{shape} K {kshape} {psf_type} {scheme}: rel-L2 vs oracle {e:.2e}, fft grid {d.geometry()['P']}")
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    d.close()


@pytest.mark.parametrize("shape,kshape,scheme", [((1, 1, 1), (1, 1, 1), "SEQ"), ((1, 1, 5), (1, 1, 3), "MUL"),
                                                 ((2, 2, 2), (3, 3, 3), "SEQ"), ((3, 1, 7), (3, 1, 5), "SEQ")])
def test_resident_decon_degenerate_sizes(oracle, shape, kshape, scheme):
    # a single voxel, a single row, volumes smaller than their PSF (mirror extension wraps several periods)
    rng = np.random.default_rng(sum(shape) * 7 + sum(kshape))
    V = 2
    imgs = rng.uniform(1.0, 40.0, (V,) + shape).astype(np.float32)
    ws = rng.uniform(0.2, 1.0, (V,) + shape).astype(np.float32)
    ks = rng.uniform(0.01, 1.0, (V,) + kshape).astype(np.float32)
    ref = oracle.deconvolve(imgs, ws, ks, psf_type="EFFICIENT_BAYESIAN", scheme=scheme, lam=6e-4, iterations=4)
    d = hip.Decon(shape, V, "EFFICIENT_BAYESIAN", scheme, 6e-4)
    for v in range(V):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    d.run(4)
    e = rel(d.psi(), ref["psi"])
    print(f"{shape} K {kshape} {scheme}: rel-L2 vs oracle {e:.2e}")
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    d.close()


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
def test_graph_replay_matches_eager(scheme):
    # one captured iteration replayed: the same kernels in the same order, so psi and stats are identical;
    # the replays are interleaved with eager host work and synchronisations
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    out = []
    for graph in (False, True):
        d = hip.Decon(imgs.shape[1:], 3, "EFFICIENT_BAYESIAN", scheme, 6e-4)
        for v in range(3):
            d.set_view(v, imgs[v], ws[v], ks[v])
        d.init()
        d.use_graph(graph)
        d.run(2)
        mid = d.psi()
        d.run(3)
        out.append((mid, d.psi(), d.stats()))
        d.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
def test_pass_timing_counts_and_leaves_results_unchanged(scheme):
    # the timed strided launches carry start/stop events of their own kernel; the timed run must give the same
    # psi and stats as an untimed one, and every pass must be counted once per launch with a positive time
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    V, its = 3, 2
    out = []
    for timed in (False, True):
        d = hip.Decon(imgs.shape[1:], V, "EFFICIENT_BAYESIAN", scheme, 6e-4)
        for v in range(V):
            d.set_view(v, imgs[v], ws[v], ks[v])
        d.init()
        d.profile(None if timed else [])
        d.run(its)
        out.append((d.psi(), d.stats(), d.pass_times()))
        d.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    t = out[1][2]
    if scheme == "SEQ":  # per view update: x_fwd, z_fwd, y_conv, z_inv, x_ratio, z_fwd, y_conv, z_inv, x_update;
        # the statistics of an iteration's view updates are reduced in one launch
        want = {"x_fwd": 1, "z_fwd": 2, "y_conv": 2, "z_inv": 2, "x_ratio": 1, "x_update": 1}
        want = {k: n * V * its for k, n in want.items()}
        want["reduce"] = its
    else:  # per iteration: shared x_fwd, z_fwd; per view y_conv, z_inv, x_ratio, z_fwd, y_conv, z_inv, x_store
        want = {"x_fwd": its, "z_fwd": its * (1 + V), "y_conv": 2 * V * its, "z_inv": 2 * V * its,
                "x_ratio": V * its, "x_store": V * its, "update_mul": its, "reduce": its}
    for name, n in want.items():
        ms, count = t[name]
        assert count == n, (name, count, n)
        assert 0.0 < ms < 1e4, (name, ms)
    assert all(t[name][1] == 0 for name in hip.PASSES if name not in want)
    assert all(c == 0 for _, c in out[0][2].values())


def test_sampled_pass_timing_counts_the_first_of_every_group():
    # profile(passes, every=k) times the launches 0, k, 2k, ... of each enabled pass: ceil(n / k) of n launches;
    # results are the same as untimed
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    V, its = 3, 2
    d = hip.Decon(imgs.shape[1:], V, "EFFICIENT_BAYESIAN", "SEQ", 6e-4)
    for v in range(V):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    d.run(its)
    ref = d.psi()
    for every in (1, 5, 7, 100):
        d.reset_psi()
        d.profile(["y_conv", "x_ratio"], every=every)
        d.run(its)
        t = d.pass_times()
        n_y, n_r = 2 * V * its, V * its
        assert t["y_conv"][1] == -(-n_y // every) and t["x_ratio"][1] == -(-n_r // every), (every, t)
        assert t["y_conv"][0] > 0.0 and t["z_fwd"][1] == 0
        np.testing.assert_array_equal(d.psi(), ref)
    with pytest.raises(hip.MvrError):
        d.profile(["y_conv"], every=0)
    d.profile([])
    d.close()


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
def test_autotuned_lengths_keep_parity(oracle, scheme):
    # the Mul tuner times one view (V swapped to 1 during the trials), then rebuilds T2 and every view's spectra
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    ref = oracle.deconvolve(imgs, ws, ks, psf_type="OPTIMIZATION_II", scheme=scheme, lam=6e-4, iterations=5)
    d = hip.Decon(imgs.shape[1:], 3, "OPTIMIZATION_II", scheme, 6e-4)
    for v in range(3):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    P, ms = d.autotune(3)
    need = [n + k - 1 for n, k in zip(imgs.shape[1:][::-1], ks.shape[1:][::-1])]
    tw = d.geometry()["tile_widths_yz"]
    print(f"{scheme}: autotuned FFT lengths {P} (need >= {need}), tile widths y/z {tw}, {ms:.3f} ms per update")
    assert all(w in (0, 8, 16) for w in tw)
    assert all(p >= q for p, q in zip(P, need))
    assert d.geometry()["P"] == P
    d.run(5)
    e = rel(d.psi(), ref["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    d.init()  # a later init keeps the tuned lengths
    assert d.geometry()["P"] == P
    d.run(5)
    e = rel(d.psi(), ref["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} after a later init"
    d.close()


@pytest.mark.parametrize("scheme,fail_at", [("SEQ", 2), ("MUL", 1), ("MUL", 4)])
def test_failed_autotune_keeps_previous_schedule(oracle, monkeypatch, scheme, fail_at):
    # a trial that fails (injected from the host) must leave a complete plan: every view's spectra, T2 for all
    # views (Mul), the previous FFT lengths, and psi at its initial estimate
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    ref = oracle.deconvolve(imgs, ws, ks, psf_type="EFFICIENT_BAYESIAN", scheme=scheme, lam=6e-4, iterations=3)
    d = hip.Decon(imgs.shape[1:], 3, "EFFICIENT_BAYESIAN", scheme, 6e-4)
    for v in range(3):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    geo = d.geometry()
    monkeypatch.setenv("MVR_TEST_FAIL_AUTOTUNE_TRIAL", str(fail_at))
    with pytest.raises(hip.MvrError, match="injected"):
        d.autotune(3)
    monkeypatch.delenv("MVR_TEST_FAIL_AUTOTUNE_TRIAL")
    assert d.geometry() == geo
    d.run(3)
    e = rel(d.psi(), ref["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} after a failed autotune"
    d.close()


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
def test_tile_widths_give_identical_results(scheme):
    # 8- and 16-column tiles of the strided passes run the same per-column arithmetic
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    out = []
    for wy, wz in ((0, 0), (8, 8), (16, 16), (8, 16)):
        d = hip.Decon(imgs.shape[1:], 3, "OPTIMIZATION_I", scheme, 6e-4)
        for v in range(3):
            d.set_view(v, imgs[v], ws[v], ks[v])
        d.init()
        d.set_tile_widths(wy, wz)
        assert d.geometry()["tile_widths_yz"] == (wy, wz)
        d.run(3)
        out.append(d.psi())
        d.close()
    for o in out[1:]:
        np.testing.assert_array_equal(o, out[0])


def test_psi_init_avg_approx(oracle):
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    avg, mx = oracle.psi_init_avg_approx(imgs)
    d = hip.Decon(imgs.shape[1:], 3)
    for v in range(3):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init(mode="AVG_APPROX")
    gmx, gavg = d.max_avg()
    np.testing.assert_array_equal(gmx, mx)
    assert gavg == pytest.approx(avg, rel=1e-12)
    assert np.all(d.psi() == np.float32(avg))
    ref = oracle.deconvolve(imgs, ws, ks, iterations=3, psi0=d.psi())
    d.run(3)
    # the oracle driver takes max[] from AvgPrecise; with lambda 0 max[] does not enter the update
    e = rel(d.psi(), ref["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    d.close()


def test_psi_init_blurred_fused(oracle):
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    ref = oracle.psi_init_blurred_fused(imgs, ws, 5.0)
    avg, mx = oracle.psi_init_avg_precise(imgs, 8)
    d = hip.Decon(imgs.shape[1:], 3, "OPTIMIZATION_II", "SEQ", 6e-4)
    for v in range(3):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init(mode="BLURRED_FUSED")
    psi = d.psi()
    e = rel(psi, ref)
    print(f"blurred fused init: rel-L2 vs FP64 oracle {e:.2e}")
    assert e < 1e-6
    gmx, gavg = d.max_avg()
    np.testing.assert_array_equal(gmx, mx)
    assert gavg == pytest.approx(avg, rel=1e-12)
    out = oracle.deconvolve(imgs, ws, ks, psf_type="OPTIMIZATION_II", lam=6e-4, iterations=10, psi0=psi)
    d.run(10)
    e = rel(d.psi(), out["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    d.close()


def test_resident_psi_init_from_array(oracle):
    inp = _small_problem()
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    psi0 = RNG.uniform(10, 40, imgs.shape[1:]).astype(np.float32)
    ref = oracle.deconvolve(imgs, ws, ks, iterations=3, psi0=psi0)
    d = hip.Decon(imgs.shape[1:], 3)
    for v in range(3):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init(psi0)
    d.run(3)
    e = rel(d.psi(), ref["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    d.reset_psi()
    np.testing.assert_array_equal(d.psi(), psi0)
    d.close()


@pytest.mark.parametrize("scheme,mode", [("SEQ", "AVG_PRECISE"), ("MUL", "AVG_APPROX")])
def test_resident_decon_with_callers_kernels_and_init(scheme, mode):
    # the Java driver's order: DeconViews has run DeconViewPSF.init (final kernels handed over) and the driver's
    # PsiInit has produced psi and max[] (handed over as an array and the maxima): the same run, bit for bit
    inp = _small_problem(weights="zramp")
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    a = hip.Decon(imgs.shape[1:], 3, "OPTIMIZATION_I", scheme, 6e-4)
    for v in range(3):
        a.set_view(v, imgs[v], ws[v], ks[v])
    a.init(mode=mode)
    psi0, (mx, _) = a.psi(), a.max_avg()
    kern = [a.kernels(v, ks.shape[1:]) for v in range(3)]
    a.run(4)
    b = hip.Decon(imgs.shape[1:], 3, "OPTIMIZATION_I", scheme, 6e-4)
    for v in range(3):
        b.set_view(v, imgs[v], ws[v], kern[v][0])
        b.set_view_kernels(v, *kern[v])
    b.init(psi0)
    b.set_max(mx)
    np.testing.assert_array_equal(b.max_avg()[0], mx)
    b.run(4)
    np.testing.assert_array_equal(a.psi(), b.psi())
    np.testing.assert_array_equal(a.stats(), b.stats())
    a.close()
    b.close()


@pytest.mark.parametrize("scheme", ["SEQ", "MUL"])
@pytest.mark.parametrize("resident_views", [True, False])
def test_blocked_driver_matches_oracle(oracle, scheme, resident_views):
    """The Java-side driver (mirrored in mvr.decon) with several blocks per view, every
    block iteration in the HIP library, against the oracle with the same block size; the view data either
    resident on the device (blocks cut there) or cut on the host and sent per block."""
    inp = _small_problem(n=(30, 26, 28), K=5)
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    bs = (22, 22, 22)
    ref = oracle.deconvolve(imgs, ws, ks, scheme=scheme, iterations=4, block_size=bs, lam=6e-4)
    views = [decon.DeconView(imgs[v], ws[v], ks[v], bs) for v in range(3)]
    assert views[0].num_blocks > 1
    dv = decon.DeconViews(views, "EFFICIENT_BAYESIAN")
    m = decon.MultiViewDeconvolution(dv, 4, decon.ComputeBlockThreadHIPFactory(bs, 6e-4), scheme,
                                     resident_views=resident_views)
    m.run_iterations()
    e = rel(m.psi, ref["psi"])
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR}, spec {SPEC})"
    m.close()


def test_errors_are_reported():
    with pytest.raises(hip.MvrError):
        hip.Decon((8, 8, 8), 0)
    d = hip.Decon((8, 8, 8), 1)
    with pytest.raises(hip.MvrError):
        d.run(1)                                         # init first
    with pytest.raises(hip.MvrError):
        hip.BlockThread((8, 8, 8)).run_iteration(0, np.zeros((8, 8, 8)), np.zeros((8, 8, 8)),
                                                 np.zeros((8, 8, 8)), 1.0)
    d.close()
