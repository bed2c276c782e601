"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

Inputs come from the seeded generator in mvr.synthetic; every fixture stores
checksums of the inputs it was made from so a test can tell an input
difference from a result difference.

  small.npz       : one 20x22x18 three-view input set (stored in full) and the
                    oracle's FP32 psi after 3 iterations for 8 variants
                    (4 PSF types, Seq/Mul, lambda 0 / 6e-4, nT 4 / 8), FP64 psi for two of them.
  <workload>.npz  : the full-size workloads (128^3 x 3 views EB 10 it;
                    256^3 x 6 views Opt-II 20 it; 400x352x224 x 4 uneven views EB
                    10 it in 128^3 effective blocks): oracle FP32 (and FP64 for
                    c1/c2) psi subsampled every 8th (16th) voxel per axis,
                    full-volume sums and norms, per-view maxima and the average.

Run:  python tests/golden/make_golden.py [small] [c1] [c2] [c3]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "multiview-reconstruction_amd"), os.path.join(REPO, "oracle")]

import oracle  # noqa: E402
from mvr import synthetic as syn  # noqa: E402

SMALL_VARIANTS = [
    ("EFFICIENT_BAYESIAN", "SEQ", 0.0, 8),
    ("EFFICIENT_BAYESIAN", "SEQ", 6e-4, 4),
    ("OPTIMIZATION_I", "SEQ", 6e-4, 8),
    ("OPTIMIZATION_II", "SEQ", 0.0, 8),
    ("INDEPENDENT", "SEQ", 6e-4, 8),
    ("EFFICIENT_BAYESIAN", "MUL", 6e-4, 8),
    ("OPTIMIZATION_II", "MUL", 0.0, 4),
    ("INDEPENDENT", "MUL", 0.0, 8),
]


def checksums(imgs, ws, ks):
    return np.array([imgs.astype(np.float64).sum(), ws.astype(np.float64).sum(), ks.astype(np.float64).sum()])


def make_small():
    w = syn.small_workload(views=3, k=7, seed_index=21, n_blobs=20, weights="zramp", shape=(20, 22, 18))
    inp = syn.make_inputs(w)
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    out = dict(imgs=imgs, weights=ws, psfs=ks, variants=np.array([f"{a}|{b}|{c}|{d}" for a, b, c, d in SMALL_VARIANTS]))
    for i, (pt, sc, lam, nT) in enumerate(SMALL_VARIANTS):
        r = oracle.deconvolve(imgs, ws, ks, psf_type=pt, scheme=sc, lam=lam, iterations=3, nT=nT)
        out[f"psi_{i}"] = r["psi"]
        if i in (0, 5):
            r64 = oracle.deconvolve(imgs, ws, ks, psf_type=pt, scheme=sc, lam=lam, iterations=3, nT=nT,
                                    dtype=np.float64)
            out[f"psi64_{i}"] = r64["psi"]
        out[f"k2_{i}"] = r["k2"]
        out[f"stats_{i}"] = r["stats"]
        out[f"max_{i}"] = r["max"]
        out[f"avg_{i}"] = np.array([r["avg"]])
    np.savez_compressed(os.path.join(HERE, "small.npz"), **out)


def make_full(name, precisions=("f32", "f64")):
    w = {"c1": syn.C1, "c2": syn.C2, "c3": syn.C3}[name]
    step = {"c1": 8, "c2": 16, "c3": 8}[name]
    inp = syn.make_inputs(w)
    imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
    # c3: 10 iterations (the parity bar's count) keep the CPU run near 20 minutes
    iterations = 10 if name == "c3" else w.iterations
    out = dict(input_checksums=checksums(imgs, ws, ks), step=np.array([step]), iterations=np.array([iterations]))
    # c3 runs at the reference's tiling: 128^3 effective blocks, block size 128 + 2(K - 1)
    block = None
    if w.extra and "block_effective" in w.extra:
        block = tuple(w.extra["block_effective"] + 2 * (w.psf_size - 1) for _ in range(3))
        out["block_size"] = np.array(block)
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        if tag not in precisions:
            continue
        t = time.time()
        r = oracle.deconvolve(imgs, ws, ks, psf_type=w.psf_type, lam=w.lam, iterations=iterations, dtype=dt,
                              block_size=block)
        psi = r["psi"].astype(np.float64)
        out[f"sub_{tag}"] = psi[::step, ::step, ::step].copy()
        out[f"sum_{tag}"] = np.array([psi.sum()])
        out[f"norm_{tag}"] = np.array([np.linalg.norm(psi)])
        out[f"max_{tag}"] = r["max"]
        out[f"avg_{tag}"] = np.array([r["avg"]])
        out[f"seconds_{tag}"] = np.array([time.time() - t])
        print(name, tag, f"{time.time() - t:.1f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


if __name__ == "__main__":
    oracle.build()
    oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1)
    todo = sys.argv[1:] or ["small", "c1", "c2"]
    if "small" in todo:
        make_small()
    for n in ("c1", "c2"):
        if n in todo:
            make_full(n)
    if "c3" in todo:
        make_full("c3", ("f32",))
