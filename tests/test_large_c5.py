"""The large combined workload at full size on the GPU: 1024 x 1024 x 512 voxels, 8 views, 61^3 PSFs,
Efficient-Bayesian, lambda 0.0006.

The reference runs it in blocks of 632^3 (512^3 effective, margin K - 1 = 60,
BlockGeneratorFixedSizePrecise.java:59-131).  The library runs one resident frame; on the effective region
of a block, one Seq view update of the frame equals the block iteration of ComputeBlockSeqThreadCPU
(ComputeBlockSeqThreadCPU.java:79-169), because the margin covers both convolutions.  The test compares the
frame after view 0's update with the oracle's block iteration of the first block (effective region at the
origin) on the same inputs, then runs two whole iterations and checks properties of the result.

Inputs are generated on the GPU (mvr.synthetic.make_inputs_device; torch only as plumbing).  The oracle
block takes about 10 s on 16 host threads.
"""
import numpy as np
import pytest

from mvr import hip
from mvr import synthetic as syn

pytestmark = pytest.mark.gpu

BAR = 1e-5  # the path's parity bar is 1e-4 after 10 iterations; one view update reaches about 1e-7


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c5_full_size_first_block_and_two_iterations(oracle):
    import torch
    w = syn.C5
    V = w.n_views
    print(f"C5 {w.shape_zyx} x {V} views, generating inputs on the GPU", flush=True)
    inp = syn.make_inputs_device(w, "cuda")
    ks = inp["psfs"]
    K = ks.shape[1:]
    E = w.extra["block_effective"]
    m = [k - 1 for k in K]                                # block margin K - 1 per axis
    B = [E + 2 * mk for mk in m]                          # 632^3
    hi = [min(n, E + mk) for n, mk in zip(w.shape_zyx, m)]
    img0 = inp["imgs"][0][:hi[0], :hi[1], :hi[2]].cpu().numpy()
    w0 = inp["weights"][0][:hi[0], :hi[1], :hi[2]].cpu().numpy()
    d = hip.Decon(w.shape_zyx, V, w.psf_type, "SEQ", w.lam)
    for v in range(V):
        d.set_view_device(v, inp["imgs"][v].data_ptr(), inp["weights"][v].data_ptr(), ks[v])
    del inp
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    d.init()
    mx, avg = d.max_avg()
    k1, k2 = d.kernels(0, K)
    print(f"init done: FFT grid {d.geometry()['P']}, avg {avg:.4f}, max[0] {mx[0]}", flush=True)
    d.run_view(0)
    psi = d.psi()
    st = d.stats()
    assert st.shape == (1, 2) and np.isfinite(st).all()
    got = psi[:E, :E, :E].copy()
    del psi

    # the oracle's block iteration on the first block of view 0: psi is the AvgPrecise constant (its mirror copy
    # is that constant), image and weight zero outside the volume
    img_b = np.zeros(B, np.float32)
    w_b = np.zeros(B, np.float32)
    img_b[m[0]:m[0] + hi[0], m[1]:m[1] + hi[1], m[2]:m[2] + hi[2]] = img0
    w_b[m[0]:m[0] + hi[0], m[1]:m[1] + hi[1], m[2]:m[2] + hi[2]] = w0
    del img0, w0
    psi_b = np.full(B, np.float32(avg), np.float32)
    print("oracle block iteration (632^3)", flush=True)
    out, ost = oracle.block_seq(psi_b, img_b, w_b, k1, k2, w.lam, 1e-4, float(mx[0]), 8)
    del img_b, w_b, psi_b
    want = out[m[0]:m[0] + E, m[1]:m[1] + E, m[2]:m[2] + E]
    e = rel(got, want)
    print(f"first 512^3 block after view 0: rel-L2 vs oracle {e:.2e}, max abs {np.abs(got - want).max():.3e}")
    assert e < BAR, f"rel-L2 {e:.2e} (bar {BAR})"
    del out, got, want

    # two whole iterations from the initial estimate: finite, bounded below by minValue (up to the rounding of
    # the weighted blend), one {sum, max} change per view update
    d.reset_psi()
    d.run(2)
    psi = d.psi()
    st = d.stats()
    d.close()
    assert st.shape == (2, V, 2)
    assert np.isfinite(st).all()
    assert np.isfinite(psi).all()
    assert psi.min() >= 1e-4 - 1e-6
    s0 = float(psi.astype(np.float64).sum())
    print(f"2 iterations: psi min {psi.min():.3e} max {psi.max():.1f} sum {s0:.6e}; "
          f"max change of the last update {st[-1, -1, 1]:.3e}")
