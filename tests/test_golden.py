"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces the committed small fixture.
GPU: the HIP library reproduces the small fixture and the full-size workloads
(128^3 x 3 views, 10 iterations; 256^3 x 6 views, 20 iterations; the uneven-view
frame against the blocked oracle) within rel-L2 1e-5 of the oracle's FP32 result
(also bounded against its FP64 result).  The parity bar of the path is 1e-4; the
runs reach 5e-7 to 1.1e-6, so 1e-5 catches a regression long before that bar.
"""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def load_small():
    z = np.load(os.path.join(HERE, "small.npz"), allow_pickle=False)
    variants = [s.split("|") for s in z["variants"].tolist()]
    return z, [(a, b, float(c), int(d)) for a, b, c, d in variants]


def test_oracle_reproduces_small_fixture(oracle):
    z, variants = load_small()
    for i, (pt, sc, lam, nT) in enumerate(variants):
        r = oracle.deconvolve(z["imgs"], z["weights"], z["psfs"], psf_type=pt, scheme=sc, lam=lam, iterations=3,
                              nT=nT)
        assert rel(r["psi"], z[f"psi_{i}"]) < 1e-6, (pt, sc)
        np.testing.assert_array_equal(r["max"], z[f"max_{i}"])
        assert rel(r["k2"], z[f"k2_{i}"]) < 1e-6


def test_small_fixture_fp32_vs_fp64_oracle():
    z, _ = load_small()
    for i in (0, 5):
        assert rel(z[f"psi_{i}"], z[f"psi64_{i}"]) < 1e-5


@pytest.mark.gpu
def test_hip_reproduces_small_fixture():
    from mvr import hip
    z, variants = load_small()
    imgs, ws, ks = z["imgs"], z["weights"], z["psfs"]
    for i, (pt, sc, lam, nT) in enumerate(variants):
        d = hip.Decon(imgs.shape[1:], imgs.shape[0], pt, sc, lam, nT=nT)
        for v in range(imgs.shape[0]):
            d.set_view(v, imgs[v], ws[v], ks[v])
        d.init()
        d.run(3)
        psi = d.psi()
        _, k2 = d.kernels(0, ks.shape[1:])
        d.close()
        e = rel(psi, z[f"psi_{i}"])
        print(f"{pt} {sc} lam={lam} nT={nT}: rel-L2 {e:.2e}")
        assert e < 1e-5, (pt, sc, lam, nT)
        assert rel(k2, z[f"k2_{i}"][0]) < 1e-5
        if f"psi64_{i}" in z:
            assert rel(psi, z[f"psi64_{i}"]) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_hip_full_size_workload(name):
    path = os.path.join(HERE, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{name}.npz not generated")
    from mvr import hip
    from mvr import synthetic as syn
    z = np.load(path, allow_pickle=False)
    w = {"c1": syn.C1, "c2": syn.C2, "c3": syn.C3}[name]
    inp = syn.make_inputs(w)
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    cs = np.array([imgs.astype(np.float64).sum(), ws.astype(np.float64).sum(), ks.astype(np.float64).sum()])
    print(f"{name}: input checksums {cs} (fixture {z['input_checksums']})")
    np.testing.assert_allclose(cs, z["input_checksums"], rtol=1e-9)
    d = hip.Decon(w.shape_zyx, w.n_views, w.psf_type, "SEQ", w.lam)
    for v in range(w.n_views):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.init()
    mx, avg = d.max_avg()
    np.testing.assert_array_equal(mx, z["max_f32"])
    d.run(int(z["iterations"][0]))
    psi = d.psi().astype(np.float64)
    d.close()
    s = int(z["step"][0])
    sub = psi[::s, ::s, ::s]
    e32 = rel(sub, z["sub_f32"])
    if "sub_f64" in z.files:
        e64 = rel(sub, z["sub_f64"])
        print(f"{name}: subsampled rel-L2 vs oracle FP32 {e32:.2e}, vs FP64 {e64:.2e}; "
              f"oracle FP32 vs FP64 {rel(z['sub_f32'], z['sub_f64']):.2e}")
        assert e64 < 1e-5, f"rel-L2 vs FP64 {e64:.2e} (bar 1e-5, spec 1e-4)"
    else:
        # the oracle ran the reference's blocked driver (block_size); the GPU runs one resident domain
        print(f"{name}: subsampled rel-L2 vs oracle FP32 (blocks {z['block_size']}) {e32:.2e}")
    assert e32 < 1e-5, f"rel-L2 vs FP32 {e32:.2e} (bar 1e-5, spec 1e-4)"
    tag = "f64" if "sum_f64" in z.files else "f32"
    assert psi.sum() == pytest.approx(float(z[f"sum_{tag}"][0]), rel=1e-5)
    assert np.linalg.norm(psi) == pytest.approx(float(z[f"norm_{tag}"][0]), rel=1e-5)


@pytest.mark.gpu
def test_hip_batched_blocks_full_size_c3():
    """The uneven-view workload (400 x 352 x 224, 4 views, 45^3 PSFs) in the reference's own 216^3 blocks (128^3
    effective, uneven tail blocks, views' empty blocks dropped), all blocks of a view in one launch per pass,
    against the oracle's blocked run of the same block size (tests/golden/c3.npz, 10 iterations)."""
    from mvr import hip
    from mvr import synthetic as syn
    z = np.load(os.path.join(HERE, "c3.npz"), allow_pickle=False)
    w = syn.C3
    inp = syn.make_inputs(w)
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    d = hip.Decon(w.shape_zyx, w.n_views, w.psf_type, "SEQ", w.lam)
    for v in range(w.n_views):
        d.set_view(v, imgs[v], ws[v], ks[v])
    d.set_block_size(tuple(int(b) for b in z["block_size"]))
    d.init()
    n, per = d.blocks()
    d.run(int(z["iterations"][0]))
    psi = d.psi().astype(np.float64)
    d.close()
    s = int(z["step"][0])
    e32 = rel(psi[::s, ::s, ::s], z["sub_f32"])
    print(f"c3 batched blocks {tuple(z['block_size'])}: {n} blocks, per view {per}; rel-L2 vs blocked oracle {e32:.2e}")
    assert e32 < 1e-5, f"rel-L2 {e32:.2e} (bar 1e-5, spec 1e-4)"
    assert psi.sum() == pytest.approx(float(z["sum_f32"][0]), rel=1e-5)
