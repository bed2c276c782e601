#!/usr/bin/env python3
"""Benchmark of the multiview Richardson-Lucy deconvolution path on one MI355X.

Workload (default): 256^3 float32 volume, 6 views, 41^3 rotated Gaussian PSFs,
Optimization-II compound kernels, Tikhonov lambda = 0.0006, 20 iterations, one
block (blob phantom of 1500 blobs, Poisson noise at peak 1000, cosine z-ramp
weights).  A step is one full deconvolution of that workload: psi reset to
its initial estimate, then 20 Seq iterations over all 6 views, with the views,
weights, PSF spectra and psi resident in HBM.

Printed: one JSON line with value = voxel-iterations/s (psi voxels x
iterations / second) and, for the dominant kernel, its achieved HBM bandwidth
from algorithmic bytes (HIP start / stop events of one in seven of its launches
in the timed region, recorded by the kernel launch itself, so that they hold the
kernel's own execution), plus a CPU baseline: the oracle restatement of the reference's CPU
path timed on a bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "multiview-reconstruction_amd")]

METRIC = "RL voxel-iters/s; wall time to rel-L2≤1e-2, learned-init vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0

# Algorithmic (unique) HBM bytes of one launch of each pass.  The complex half spectrum counts 4 B per voxel of
# the FFT extent E (N + K - 1 per axis; padding to the FFT length is never credited), the real volumes (psi, the
# view image, the weights) 4 B per volume voxel (N per axis), each read or written once.  On one frame a pass
# moves only the rows its neighbours use: a convolution's inverse passes produce the rows the next x pass reads
# (the output region's y and z; the quotient region is the output region here); the first convolution
# transforms only psi's own rows (y, z in N), its mirror copies being read from their sources; the quotient pass
# writes only the quotient rows, and the rows and y planes of ones around them are put in by the passes that
# read them (their spectra are known):
#   x_fwd   : psi 4 N + write of psi's rows 4 Ex Ny Nz
#   z_fwd   : read of psi's (or the quotient's) rows 4 Ex Ny Nz + write of their y planes, 4 Ex Ny Ez
#   y_conv  : read of those y, 4 Ex Ny Ez + kernel spectrum 4 E + write of the output region's y, 4 Ex Ny Ez
#   z_inv   : read of the output region's y, 4 Ex Ny Ez + write of its y and z, 4 Ex Ny Nz
#   x_ratio : spectrum rows of the output region 4 Ex Ny Nz + view image 4 N + write of those rows 4 Ex Ny Nz
#   x_update: spectrum rows of the output region 4 Ex Ny Nz + psi read 4 N + weights 4 N + psi write 4 N
#   x_store : spectrum rows of the output region 4 Ex Ny Nz + store 4 N
# The whole view update is also priced on the survey's budget of 60 B per extent voxel (five passes with the xy
# transforms fused), RUN_BYTES_PER_EXTENT_VOXEL.
RUN_BYTES_PER_EXTENT_VOXEL = 60
EVENT_STRIDE = 7


def pass_bytes(name, n_zyx, k_zyx):
    nz, ny, nx = n_zyx
    ez, ey, ex = (n + k - 1 for n, k in zip(n_zyx, k_zyx))
    ext = ex * ey * ez
    vol = nz * ny * nx
    rows = ex * ny * nz     # spectrum rows of the output region
    ycols = ex * ny * ez    # the output region's y on every z plane
    return {"x_fwd": 4 * vol + 4 * rows, "z_fwd": 4 * rows + 4 * ycols, "y_conv": 4 * ext + 8 * ycols,
            "z_inv": 4 * ycols + 4 * rows, "x_ratio": 8 * rows + 4 * vol,
            "x_update": 4 * rows + 12 * vol, "x_store": 4 * rows + 4 * vol}[name]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c5"])
    ap.add_argument("--autotune", type=int, default=0, help="autotune FFT lengths over this many candidates per axis")
    ap.add_argument("--graph", action="store_true", help="replay a graph-captured iteration outside profiled steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-views", type=int, default=0,
                    help="view updates the CPU baseline times (0: two iterations of a one-block workload, "
                         "two view updates of a blocked one, about 10 s of CPU work either way)")
    ap.add_argument("--iterations", type=int, default=0,
                    help="override the workload's iteration count (short runs under a profiler)")
    ap.add_argument("--no-learned", action="store_true", help="skip the time-to-convergence comparison")
    ap.add_argument("--weights", default=None, help="weights file of the learned start (default: the workload's)")
    ap.add_argument("--tol", type=float, default=1e-2)
    ap.add_argument("--reference-iterations", type=int, default=0,
                    help="reference of the time-to-convergence comparison: plain RL run this many iterations; 0 "
                         "(default): converged plain RL, iterated until the rel-L2 change over 100 iterations is "
                         "at most 1e-3 (the 256^3 workload); the 1024^2 x 512 workload defaults to 200")
    ap.add_argument("--block-path", type=int, default=0,
                    help="also time this many iterations through the per-block drop-in ABI (host buffers)")
    ap.add_argument("--block-size", type=int, nargs=3, default=None, metavar=("BX", "BY", "BZ"),
                    help="run in the reference's block structure (DeconView block size per axis, all blocks of a "
                         "view in one launch per pass) instead of one frame; e.g. 216 216 216 for 128^3 effective "
                         "blocks with 45^3 PSFs")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this size (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding this path over several "
              "GPUs is not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from mvr import hip
    from mvr import synthetic as syn

    w = {"c1": syn.C1, "c2": syn.C2, "c3": syn.C3, "c5": syn.C5}[args.workload]
    if args.iterations > 0:
        import dataclasses
        w = dataclasses.replace(w, iterations=args.iterations)
    if args.calibrate_copy_mib > 0:
        # plain streaming kernels of known size: a fill (writes N bytes) and an elementwise
        # multiply (reads N, writes N), N = calibrate_copy_mib MiB
        import torch
        src = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    V = w.n_views
    nz, ny, nx = w.shape_zyx
    nvox = nx * ny * nz
    t0 = time.time()
    device_inputs = nvox * V > (1 << 30)   # the host generator would take minutes: generate on the GPU
    crop0 = None
    if device_inputs:
        import torch
        inp = syn.make_inputs_device(w, "cuda", log)
        ks = inp["psfs"]
        d = hip.Decon(w.shape_zyx, V, w.psf_type, "SEQ", w.lam)
        for v in range(V):
            d.set_view_device(v, inp["imgs"][v].data_ptr(), inp["weights"][v].data_ptr(), ks[v])
        if not args.no_cpu_baseline:   # the CPU sample block's inputs (view 0)
            crop0 = cpu_block_inputs(inp, w, ks.shape[1:])
        del inp
        torch.cuda.empty_cache()
        imgs = ws = None
    else:
        inp = syn.make_inputs(w)
        imgs, ws, ks = inp["imgs"], inp["weights"], inp["psfs"]
        d = hip.Decon(w.shape_zyx, V, w.psf_type, "SEQ", w.lam)
        for v in range(V):
            d.set_view(v, imgs[v], ws[v], ks[v])
    log(f"inputs generated and uploaded in {time.time() - t0:.1f}s")
    if args.block_size:
        d.set_block_size(args.block_size)
    t0 = time.time()
    d.init()
    if args.block_size:
        log(f"blocks of {tuple(args.block_size)}: {d.blocks()[0]} in the volume, per view {d.blocks()[1]}")
    log(f"psf init + spectra + psi init in {time.time() - t0:.2f}s, fft grid {d.geometry()}")
    if args.autotune:
        P, ms = d.autotune(args.autotune)
        log(f"autotuned FFT lengths {P}, tile widths y/z {d.geometry()['tile_widths_yz']}: {ms:.3f} ms per view update")
    d.use_graph(args.graph)

    # warmup; the first warmup step times every pass to find the dominant kernel
    d.profile(None)
    for _ in range(max(1, args.warmup)):
        d.reset_psi()
        d.run(w.iterations)
    times = d.pass_times()
    steps_w = max(1, args.warmup)
    per_pass = {k: v[0] / steps_w for k, v in times.items() if v[1]}
    dominant = max(per_pass, key=per_pass.get)
    log("per-pass ms per step: " + ", ".join(f"{k} {v:.2f}" for k, v in per_pass.items()))

    # timed region: only the dominant kernel carries HIP events, on one launch in EVENT_STRIDE (odd, so that
    # both of its positions in a view update are sampled): a timed launch's events delay the stream by about
    # 5 us, which timing every launch would add to the measured run
    d.profile([dominant], every=EVENT_STRIDE)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d.reset_psi()
        d.run(w.iterations)
    elapsed = time.perf_counter() - t0
    ms_dom, n_dom = d.pass_times()[dominant]
    d.profile([])
    ms_per_step = elapsed * 1e3 / args.steps
    value = nvox * w.iterations * args.steps / elapsed

    geo = d.geometry()
    K = ks.shape[1:]
    ext = (nz + K[0] - 1) * (ny + K[1] - 1) * (nx + K[2] - 1)
    alg_bytes = pass_bytes(dominant, w.shape_zyx, K)
    avg_s = ms_dom / 1e3 / max(1, n_dom)
    achieved = alg_bytes / avg_s / 1e9
    # the whole run: the survey's five-pass budget and this design's own nine passes, both on unique bytes
    run_bytes = RUN_BYTES_PER_EXTENT_VOXEL * ext * V * w.iterations
    seq_passes = ("x_fwd", "z_fwd", "y_conv", "z_inv", "x_ratio", "z_fwd", "y_conv", "z_inv", "x_update")
    own_bytes = sum(pass_bytes(p, w.shape_zyx, K) for p in seq_passes) * V * w.iterations
    per_pass_frac = {k: pass_bytes(k, w.shape_zyx, K) * (2 if k in ("z_fwd", "y_conv", "z_inv") else 1) * V *
                     w.iterations / (v / 1e3) / 1e9 / HBM_PEAK_GBS
                     for k, v in per_pass.items() if k != "reduce" and v > 0}
    traffic = None
    tfile = os.path.join(REPO, "profiles", f"{dominant}_{w.name}_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")

    d_blocks = d.blocks()
    psi = d.psi()
    finite = bool(np.isfinite(psi).all())
    parity = golden_parity(d, psi, w, args.workload)

    learned = None
    if not args.no_learned and args.workload in ("c2", "c5"):
        ref_its = args.reference_iterations or (200 if args.workload == "c5" else 0)
        learned = time_to_converge(d, w, args.tol, ref_its, args.workload, args.weights)
    block_path = None
    if args.block_path > 0 and imgs is not None:
        block_path = time_block_path(imgs, ws, ks, w, args.block_path)
    gpu_block = None
    if crop0 is not None:
        crop0["k1"], crop0["k2"] = d.kernels(0, ks.shape[1:])
        crop0["psi"], crop0["max"] = d.max_avg()[1], d.max_avg()[0][0]
        # the GPU side of the block parity: view 0's update of the resident frame from PsiInitAvgPrecise (init
        # again: the learned comparison leaves the network's estimate as the initial psi)
        E = w.extra["block_effective"]
        d.init()
        d.run_view(0)
        gpu_block = d.psi()[:E, :E, :E].copy()
    d.close()

    cpu = None
    if not args.no_cpu_baseline and crop0 is not None:
        cpu, oracle_block = cpu_baseline_block(crop0, w, V)
        m = crop0["margin"]
        E = w.extra["block_effective"]
        want = oracle_block[m[0]:m[0] + E, m[1]:m[1] + E, m[2]:m[2] + E]
        parity = {"reference": "oracle restatement of ComputeBlockSeqThreadCPU on the first "
                               f"{'x'.join(map(str, crop0['block']))} block of view 0 (effective {E}^3 at the origin), "
                               "against the resident frame after view 0's update (one view update from PsiInitAvgPrecise)",
                  "rel_l2_vs_oracle_f32": rel_l2(gpu_block, want),
                  "max_abs_diff": float(np.abs(gpu_block.astype(np.float64) - want).max()),
                  "sum_rel": abs(float(gpu_block.sum(dtype=np.float64)) - float(want.sum(dtype=np.float64))) /
                  abs(float(want.sum(dtype=np.float64)))}
        parity["ok"] = parity["rel_l2_vs_oracle_f32"] <= PARITY_BAR
    elif not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        threads = oracle.set_threads(threads)
        block = None
        if w.extra and "block_effective" in w.extra:
            block = tuple(w.extra["block_effective"] + 2 * (k - 1) for k in K[::-1])
        sv = args.cpu_sample_views or (2 if block else 2 * V)
        sv = max(1, min(2 * V, sv))
        r = oracle.deconvolve(imgs, ws, ks, psf_type=w.psf_type, lam=w.lam, iterations=(sv + V - 1) // V,
                              sample_views=sv, block_size=block)
        cpu_val = nvox * (sv / V) / r["loop_seconds"]
        blocks = (f"{w.extra['block_effective']}^3 effective blocks of {block[0]}^3" if block else
                  f"one {nx + 2 * (K[2] - 1)}x{ny + 2 * (K[1] - 1)}x{nz + 2 * (K[0] - 1)} block")
        cpu = {"value": cpu_val, "unit": "voxel-iters/s", "cores": threads, "kind": "port",
               "sample": f"{sv} view updates ({V} per iteration) from iteration 1 of the same workload, reference block "
                         f"structure ({blocks}, padded FFT), oracle restatement "
                         f"in FP32 on {threads} OpenMP threads; loop time {r['loop_seconds']:.2f}s"}

    out = {
        "metric": METRIC, "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "volume_zyx": list(w.shape_zyx), "views": V, "psf": list(K),
                   "psf_type": w.psf_type, "lambda": w.lam, "iterations": w.iterations, "scheme": "SEQ",
                   "fft_grid_xyz": list(geo["P"]), "tile_widths_yz": list(geo["tile_widths_yz"]),
                   "parallelism": "single GPU", "psi_finite": finite,
                   "blocks": ({"block_size_xyz": list(args.block_size), "blocks": d_blocks[0],
                               "blocks_with_content_per_view": d_blocks[1]} if args.block_size else "one frame")},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_us": avg_s * 1e6, "launches": n_dom,
                     "run_frac": run_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "run_bytes": run_bytes, "run_bytes_rule": f"{RUN_BYTES_PER_EXTENT_VOXEL} B per extent voxel per "
                                                           "view update (five passes, xy transforms fused)",
                     "run_frac_nine_passes": own_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "per_pass_ms_per_step": per_pass, "per_pass_frac": per_pass_frac},
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    if learned:
        out["learned_init"] = learned
    if block_path:
        out["block_path"] = block_path
    print(json.dumps(out))
    if parity is not None and not parity.get("ok", True):
        log(f"parity check failed: {parity}")
        sys.exit(3)


PARITY_BAR = 1e-4   # rel-L2 of the path's FP32 result against the oracle's


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def golden_parity(d, psi, w, workload):
    """Error of the benchmarked result against the oracle fixtures of the same workload (tests/golden/<c>.npz:
    the oracle's FP32 and FP64 results on every s-th voxel per axis, their sums and norms).  The one-block
    workloads compare the psi of the last timed step; the uneven-view frame (fixture: the oracle's blocked
    run over 10 iterations) is rerun for the fixture's iteration count.  None where no fixture exists."""
    import numpy as np
    path = os.path.join(REPO, "tests", "golden", f"{workload}.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    its = int(z["iterations"][0])
    if its != w.iterations:
        d.reset_psi()
        d.run(its)
        psi = d.psi()
    s = int(z["step"][0])
    sub = psi[::s, ::s, ::s]
    tag = "f64" if "sum_f64" in z.files else "f32"
    p64 = psi.astype(np.float64)
    out = {"reference": f"oracle restatement of the reference CPU path (tests/golden/{workload}.npz, every {s}th "
                        f"voxel per axis{', blocks ' + 'x'.join(map(str, z['block_size'])) if 'block_size' in z.files else ''})",
           "iterations": its,
           "rel_l2_vs_oracle_f32": rel_l2(sub, z["sub_f32"]),
           "rel_l2_vs_oracle_f64": rel_l2(sub, z["sub_f64"]) if "sub_f64" in z.files else None,
           "sum_rel": abs(float(p64.sum()) - float(z[f"sum_{tag}"][0])) / abs(float(z[f"sum_{tag}"][0])),
           "norm_rel": abs(float(np.linalg.norm(p64)) - float(z[f"norm_{tag}"][0])) / float(z[f"norm_{tag}"][0])}
    worst = max(v for v in (out["rel_l2_vs_oracle_f32"], out["rel_l2_vs_oracle_f64"]) if v is not None)
    out["ok"] = worst <= PARITY_BAR
    if its != w.iterations:
        d.reset_psi()
    return out


def time_block_path(imgs, ws, ks, w, iterations):
    """The same workload through the per-block drop-in (one mvr_block_run_iteration_seq* call per block and view),
    driven like MultiViewDeconvolutionSeq with the reference's default block (volume + kernel size - 1 per axis):
    PCIe copies and the host driver's block copies are inside the time.  Two variants: the view data resident
    on the device (uploaded once, blocks cut there; only psi crosses the bus per call) and every image and weight
    block cut on the host and sent per call, as the reference's drivers hand them over."""
    from mvr import decon as MD
    V = w.n_views
    nvox = imgs[0].size
    out = {"iterations": iterations, "unit": "voxel-iters/s",
           "note": "per-block drop-in ABI with host buffers; host-device copies and the host driver's block "
                   "copies are inside the time"}
    for resident in (True, False):
        views = [MD.DeconView(imgs[v], ws[v], ks[v]) for v in range(V)]
        dv = MD.DeconViews(views, w.psf_type)
        m = MD.MultiViewDeconvolution(dv, iterations + 1, MD.ComputeBlockThreadHIPFactory(views[0].block_size, w.lam),
                                      "SEQ", resident_views=resident)
        m.run_next_iteration()  # warm-up: device buffers, kernel spectra
        t0 = time.perf_counter()
        for _ in range(iterations):
            m.run_next_iteration()
        el = time.perf_counter() - t0
        m.close()
        key = "resident_views" if resident else "host_blocks"
        log(f"per-block drop-in path ({key}): {el * 1e3 / iterations:.1f} ms per iteration")
        out[key] = {"value": nvox * iterations / el, "ms_per_iteration": el * 1e3 / iterations}
        out["block_xyz"] = list(views[0].block_size)
        out["blocks_per_view"] = views[0].num_blocks
    out["value"] = out["resident_views"]["value"]
    return out


def cpu_block_inputs(inp, w, K):
    """Inputs of the reference's first block of view 0 (effective block at the origin, margin K - 1, the image
    and weight zero outside the volume): host copies of the covered part."""
    be = w.extra["block_effective"]
    m = [k - 1 for k in K]
    hi = [min(n, be + mk) for n, mk in zip(w.shape_zyx, m)]
    img = inp["imgs"][0][:hi[0], :hi[1], :hi[2]].cpu().numpy()
    wt = inp["weights"][0][:hi[0], :hi[1], :hi[2]].cpu().numpy()
    return {"img": img, "w": wt, "margin": m, "block": [be + 2 * mk for mk in m]}


def cpu_baseline_block(c, w, V):
    """The oracle's ComputeBlockSeqThreadCPU.runIteration on the first block of view 0 in iteration 1 (psi is
    the AvgPrecise constant there, so its mirror copy is that constant), timed on this host."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = oracle.set_threads(threads)
    B, m = c["block"], c["margin"]
    img = np.zeros(B, np.float32)
    wt = np.zeros(B, np.float32)
    sz = c["img"].shape
    img[m[0]:m[0] + sz[0], m[1]:m[1] + sz[1], m[2]:m[2] + sz[2]] = c["img"]
    wt[m[0]:m[0] + sz[0], m[1]:m[1] + sz[1], m[2]:m[2] + sz[2]] = c["w"]
    psi = np.full(B, np.float32(c["psi"]), np.float32)
    t0 = time.perf_counter()
    out, _ = oracle.block_seq(psi, img, wt, c["k1"], c["k2"], w.lam, 1e-4, float(c["max"]), 8)
    secs = time.perf_counter() - t0
    eff = w.extra["block_effective"] ** 3
    return {"value": eff / V / secs, "unit": "voxel-iters/s", "cores": threads, "kind": "port",
            "sample": f"one {B[2]}x{B[1]}x{B[0]} block ({w.extra['block_effective']}^3 effective) of view 0 in "
                      f"iteration 1 of the same workload (1/{V} iteration of {eff} voxels), reference block "
                      f"structure, oracle restatement in FP32 on {threads} OpenMP threads; {secs:.2f}s"}, out


def unet_flops(shape_zyx, n_views):
    """Multiply-adds x 2 of one U-Net inference: the 3^3 layers at their level's voxel count, the first layer
    counted on its real input channels (the views and their weighted mean; the kernel pads them to 16, and
    the padding is not credited), plus the 1x1 head."""
    from mvr import hip
    nz, ny, nx = ((s + 3) // 4 * 4 for s in shape_zyx)
    v0 = nz * ny * nx
    level = (0, 0, 1, 1, 2, 2, 1, 1, 0, 0)
    cin = (n_views + 1,) + tuple(hip.UNet.LAYERS_CIN[1:])
    f = 0
    for l in range(10):
        f += 2 * 27 * cin[l] * hip.UNet.LAYERS_COUT[l] * v0 / 8 ** level[l]
    return f + 2 * 16 * v0


def _trajectory(d, start_seconds, max_iterations, stop_tol=0.0):
    """(seconds since the start, rel-L2) after each of up to max_iterations iterations, stopping once rel-L2 <=
    stop_tol; the distance is reduced on the GPU after every iteration, its cost is inside the times."""
    out = []
    t0 = time.perf_counter() - start_seconds
    for _ in range(max_iterations):
        d.run(1)
        out.append((time.perf_counter() - t0, d.rel_l2()))
        if out[-1][1] <= stop_tol:
            break
    return out


def _first_below(traj, start, tol):
    if start[1] <= tol:
        return 0, start[0]
    for i, (t, r) in enumerate(traj):
        if r <= tol:
            return i + 1, t
    return None, None


TOLERANCES = (1e-2, 0.02, 0.05, 0.1, 0.15, 0.2, 0.3)
CONVERGED_CHANGE = 1e-3     # rel-L2 change over CONVERGED_BLOCK iterations that ends the reference run
CONVERGED_BLOCK = 100
CONVERGED_CAP = 20000


def converged_reference(d):
    """Plain RL from the handle's initial estimate until the rel-L2 change over CONVERGED_BLOCK iterations is at
    most CONVERGED_CHANGE (at most CONVERGED_CAP iterations): (psi, iterations, last change)."""
    d.reset_psi()
    prev = d.psi()
    it, ch = 0, float("inf")
    while it < CONVERGED_CAP:
        d.set_reference(prev)
        d.run(CONVERGED_BLOCK)
        it += CONVERGED_BLOCK
        ch = d.rel_l2()
        prev = d.psi()
        if ch <= CONVERGED_CHANGE:
            break
    return prev, it, ch


def time_to_converge(d, w, tol, ref_iterations, workload="c2", weights=None):
    """Wall time to rel-L2 <= tol to a reference: plain PsiInitAvgPrecise start vs PsiInitBlurredFused vs the
    learned start (U-Net inference included).  The reference is converged plain RL (ref_iterations 0: until the
    rel-L2 change over 100 iterations is <= 1e-3) or plain RL run to ref_iterations.  The common set-up (PSFs,
    spectra, AvgPrecise pass) is outside the timed regions.  Also reported for the looser tolerances of
    TOLERANCES.  Every trajectory runs until it reaches tol or a cap: the converged reference's iteration count
    + 3000, or a fixed-length reference's own iteration count (plain RL meets that reference there)."""
    from mvr import learned as L
    if weights is None:
        # against a converged reference, the network trained on deep targets when there is one
        deep = L.DEEP_WEIGHTS.get(workload) if ref_iterations == 0 else None
        weights = deep if deep and os.path.exists(deep) else L.TRAINING[workload][1]
    t0 = time.perf_counter()
    if ref_iterations > 0:
        d.reset_psi()
        d.run(ref_iterations)
        ref, ref_its, ref_change = d.psi(), ref_iterations, None
        desc = f"plain RL, {ref_iterations} iterations from PsiInitAvgPrecise"
    else:
        ref, ref_its, ref_change = converged_reference(d)
        desc = (f"converged plain RL from PsiInitAvgPrecise: {ref_its} iterations, until the rel-L2 change over "
                f"{CONVERGED_BLOCK} iterations was <= {CONVERGED_CHANGE} (last {ref_change:.2e})")
    d.set_reference(ref)
    log(f"reference: {desc} ({time.perf_counter() - t0:.1f}s)")
    cap = ref_its if ref_iterations > 0 else ref_its + 3000
    d.reset_psi()
    plain0 = (0.0, d.rel_l2())
    plain = _trajectory(d, 0.0, cap, tol)
    log(f"plain trajectory done: {len(plain)} iterations")
    res = {"tol": tol, "reference": desc, "reference_iterations": ref_its, "reference_last_change": ref_change,
           "plain": {"rel_l2_start": plain0[1], "rel_l2_after_20": plain[min(19, len(plain) - 1)][1]}}
    # PsiInitBlurredFused, the reference GUI's default initialiser: its fusion + blur time is the init
    # difference to AvgPrecise (both inits also rebuild PSFs and spectra)
    t0 = time.perf_counter()
    d.init()
    t_avg = time.perf_counter() - t0
    t0 = time.perf_counter()
    d.init(mode="BLURRED_FUSED")
    t_blur = max(0.0, time.perf_counter() - t0 - t_avg)
    blurred0 = (t_blur, d.rel_l2())
    blurred = _trajectory(d, t_blur, cap, tol)
    log(f"BlurredFused trajectory done: {len(blurred)} iterations")
    res["blurred_fused"] = {"init_seconds_over_avg_precise": t_blur, "rel_l2_start": blurred0[1],
                            "rel_l2_after_20": blurred[min(19, len(blurred) - 1)][1]}
    if not os.path.exists(weights):
        res["learned"] = None
        res["note"] = "no trained weights file"
        return res
    net = L.load(weights)
    d.init_learned(net)  # warm-up: activation buffers
    t_inf = []
    for _ in range(3):
        t0 = time.perf_counter()
        d.init_learned(net)
        t_inf.append(time.perf_counter() - t0)
    net.close()
    inf = min(t_inf)
    learned0 = (inf, d.rel_l2())
    learned = _trajectory(d, inf, cap, tol)
    log(f"learned trajectory done: {len(learned)} iterations")
    best = min(range(len(learned)), key=lambda i: learned[i][1])
    fl = unet_flops(w.shape_zyx, w.n_views)
    res["learned"] = {"inference_seconds": inf, "rel_l2_start": learned0[1], "rel_l2_min": learned[best][1],
                      "rel_l2_min_at_iteration": best + 1, "inference_tflops": fl / inf / 1e12,
                      "inference_mfma_frac_bf16_dense": fl / inf / 1e12 / 2500.0,
                      "weights_meta": {k: v for k, v in L.meta(weights).items()
                                       if k in ("target_iterations", "seeds", "patches", "epochs")}}
    res["weights"] = os.path.relpath(weights, REPO)
    table = []
    for t in sorted(set((tol,) + TOLERANCES)):
        ip, tp = _first_below(plain, plain0, t)
        il, tl = _first_below(learned, learned0, t)
        ib, tb = _first_below(blurred, blurred0, t)
        table.append({"tol": t, "plain_iterations": ip, "plain_seconds": tp, "blurred_fused_iterations": ib,
                      "blurred_fused_seconds": tb, "learned_iterations": il, "learned_seconds": tl,
                      "speedup": (tp / tl) if (tp and tl) else None,
                      "speedup_vs_blurred_fused": (tb / tl) if (tb and tl) else None})
    res["by_tolerance"] = table
    head = next(r for r in table if r["tol"] == tol)
    res["plain"].update(iterations=head["plain_iterations"], seconds=head["plain_seconds"])
    res["learned"].update(iterations=head["learned_iterations"], seconds=head["learned_seconds"])
    res["blurred_fused"].update(iterations=head["blurred_fused_iterations"], seconds=head["blurred_fused_seconds"])
    res["speedup"] = head["speedup"]
    res["speedup_vs_blurred_fused"] = head["speedup_vs_blurred_fused"]
    return res


if __name__ == "__main__":
    main()
