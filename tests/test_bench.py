"""Host logic of bench.py on the CPU: the one-GPU guard, the byte accounting, and the time-to-convergence study
(converged reference, trajectories, tolerance table) driven by a stand-in for the deconvolution handle whose
estimate follows a known closed form, so that every iteration count the study reports can be computed by hand.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_pass_bytes_follow_the_rows_each_pass_moves():
    n, k = (256, 256, 256), (41, 41, 41)
    E, N = 296 ** 3, 256 ** 3
    rows = 296 * 256 * 256
    ycols = 296 * 256 * 296
    # x_fwd writes psi's own rows only; the update pass moves whole volumes
    assert bench.pass_bytes("x_fwd", n, k) == 4 * N + 4 * rows
    assert bench.pass_bytes("x_update", n, k) == 4 * rows + 12 * N
    # the forward passes move only the rows and y planes the quotient region or psi's own rows select; the y
    # convolution reads the whole kernel spectrum
    assert bench.pass_bytes("z_fwd", n, k) == 4 * rows + 4 * ycols
    assert bench.pass_bytes("y_conv", n, k) == 4 * E + 8 * ycols
    # pruned passes never exceed their unpruned counts, and a view update totals the documented 1.88 GB
    assert bench.pass_bytes("y_conv", n, k) < 12 * E and bench.pass_bytes("z_inv", n, k) < 8 * E
    seq = ["x_fwd", "z_fwd", "y_conv", "z_inv", "x_ratio", "z_fwd", "y_conv", "z_inv", "x_update"]
    assert abs(sum(bench.pass_bytes(p, n, k) for p in seq) - 1.88e9) < 0.005e9
    # an axis with no padding (K = 1) leaves nothing to prune: the inverse passes move whole spectra
    assert bench.pass_bytes("z_inv", n, (1, 1, 1)) == 8 * N


class FakeDecon:
    """Scalar stand-in: after k iterations from start s the estimate is 1 + (s - 1) q^k (q < 1), i.e. every
    start converges geometrically to 1 along its own path.  rel_l2 is |estimate - reference| / |reference|."""

    Q = 0.999

    def __init__(self):
        self.starts = {"AVG_PRECISE": 0.1, "BLURRED_FUSED": 0.2}
        self.start = self.starts["AVG_PRECISE"]
        self.k = 0
        self.ref = None

    def _value(self):
        return 1.0 + (self.start - 1.0) * self.Q ** self.k

    def init(self, mode="AVG_PRECISE"):
        self.start, self.k = self.starts[mode], 0

    def init_learned(self, net):
        self.start, self.k = 0.95, 0

    def reset_psi(self):
        self.start, self.k = self.starts["AVG_PRECISE"], 0

    def run(self, n):
        self.k += n

    def psi(self):
        return np.array([self._value()])

    def set_reference(self, ref):
        self.ref = float(np.asarray(ref)[0])

    def rel_l2(self):
        return abs(self._value() - self.ref) / abs(self.ref)


def first_k(start, ref, tol):
    # smallest k >= 1 with |1 + (start - 1) q^k - ref| / ref <= tol, for ref near 1 (by scanning)
    q = FakeDecon.Q
    for k in range(1, 200000):
        if abs(1 + (start - 1) * q ** k - ref) / ref <= tol:
            return k
    return None


def test_converged_reference_stops_at_the_change_criterion():
    d = FakeDecon()
    ref, its, ch = bench.converged_reference(d)
    q = FakeDecon.Q
    # the change over 100 iterations at iteration k (k - 100 -> k), relative to the value at k - 100 (the
    # reference the handle measures against)
    def change(k):
        a, b = 1 + (0.1 - 1) * q ** (k - 100), 1 + (0.1 - 1) * q ** k
        return abs(b - a) / a
    assert its % bench.CONVERGED_BLOCK == 0
    assert ch <= bench.CONVERGED_CHANGE < change(its - bench.CONVERGED_BLOCK)
    assert ch == pytest.approx(change(its), rel=1e-9)
    assert ref[0] == pytest.approx(1 + (0.1 - 1) * q ** its)


def test_time_to_converge_table(monkeypatch):
    from mvr import learned as L
    from mvr import synthetic as syn

    class Net:
        def close(self):
            pass

    monkeypatch.setattr(L, "load", lambda path, dev=0: Net())
    d = FakeDecon()
    res = bench.time_to_converge(d, syn.C2, 1e-2, 0, "c2")
    ref_its = res["reference_iterations"]
    ref = 1 + (0.1 - 1) * FakeDecon.Q ** ref_its
    assert res["reference_last_change"] <= bench.CONVERGED_CHANGE
    by = {r["tol"]: r for r in res["by_tolerance"]}
    assert sorted(by) == sorted(set((1e-2,) + bench.TOLERANCES))
    for tol, row in by.items():
        assert row["plain_iterations"] == first_k(0.1, ref, tol)
        assert row["blurred_fused_iterations"] == first_k(0.2, ref, tol)
        want = first_k(0.95, ref, tol) if abs(0.95 - ref) / ref > tol else 0
        assert row["learned_iterations"] == want
        if row["plain_seconds"] and row["learned_seconds"]:
            assert row["speedup"] == pytest.approx(row["plain_seconds"] / row["learned_seconds"])
    # the headline entries are the 1e-2 row; the deep-target network is the one used against a converged reference
    assert res["plain"]["iterations"] == by[1e-2]["plain_iterations"]
    assert res["speedup"] == by[1e-2]["speedup"]
    assert res["weights"].endswith("unet_c2_6views_rl4000.npz")
    assert res["learned"]["weights_meta"]["target_iterations"] == 4000


def test_time_to_converge_fixed_reference_caps_at_its_length(monkeypatch):
    from mvr import learned as L
    from mvr import synthetic as syn

    class Net:
        def close(self):
            pass

    monkeypatch.setattr(L, "load", lambda path, dev=0: Net())
    d = FakeDecon()
    res = bench.time_to_converge(d, syn.C2, 1e-3, 50, "c2")
    # plain RL meets a fixed-length reference at that length; nothing runs longer than it
    by = {r["tol"]: r for r in res["by_tolerance"]}
    assert by[1e-3]["plain_iterations"] == 50
    assert res["reference_iterations"] == 50 and res["reference_last_change"] is None
    assert res["weights"].endswith("unet_c2_6views.npz")
    assert by[1e-3]["blurred_fused_iterations"] is None or by[1e-3]["blurred_fused_iterations"] <= 50
