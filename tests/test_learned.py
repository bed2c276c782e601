"""What the trained networks gain over the plain initial estimate (held-out seed of the 256^3 / 6-view
Optimization-II workload, and of the 8-view Efficient-Bayesian training workload of the 1024^2 x 512 run,
committed weights).

The distance to plain RL run to 200 iterations (rel-L2, FP64 on the GPU) of the learned start must be
smaller than that of plain RL after the workload's 20 iterations, and the inference must take less time
than those 20 iterations.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload", ["c2", "c5"])
def test_learned_init_beats_twenty_plain_iterations(workload):
    from mvr import hip
    from mvr import learned as L
    w, weights = L.TRAINING[workload]
    if not os.path.exists(weights):
        pytest.skip("no trained weights file")
    inp = L.generate_volume(w, L.HELD_OUT_SEED)
    d = hip.Decon(w.shape_zyx, w.n_views, w.psf_type, "SEQ", w.lam)
    for v in range(w.n_views):
        d.set_view(v, inp["imgs"][v], inp["weights"][v], inp["psfs"][v])
    d.init()
    d.run(200)
    d.set_reference(d.psi())
    d.reset_psi()
    start_plain = d.rel_l2()
    t0 = time.perf_counter()
    d.run(20)
    t_plain20 = time.perf_counter() - t0
    rel_plain20 = d.rel_l2()
    net = L.load(weights)
    d.init_learned(net)  # first call allocates the activations
    t0 = time.perf_counter()
    d.init_learned(net)
    t_inf = time.perf_counter() - t0
    rel_learned = d.rel_l2()
    psi = d.psi()
    d.run(20)
    rel_learned20 = d.rel_l2()
    print(f"{workload}: rel-L2 to RL-200: plain start {start_plain:.4f}, plain 20 it {rel_plain20:.4f} ({t_plain20 * 1e3:.1f} ms); "
          f"learned start {rel_learned:.4f} ({t_inf * 1e3:.1f} ms), learned + 20 it {rel_learned20:.4f}")
    assert np.isfinite(psi).all() and psi.min() >= 1e-4
    assert rel_learned < rel_plain20
    assert rel_learned20 < rel_plain20
    assert t_inf < t_plain20
    d.close()
    net.close()
