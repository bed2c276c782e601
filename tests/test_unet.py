"""Learned initial estimate: the FP64 network restatement (oracle/unet_ref.py) pinned on the CPU, and
the HIP U-Net (BF16 MFMA, FP32 accumulation) checked against it through the C ABI.

Tolerances (GPU vs oracle on the same parameters and inputs):
  - forward, oracle rounding its stored tensors to bf16 like the kernels: rel-L2 <= 2e-3;
  - forward vs the plain FP64 network: rel-L2 <= 2e-2 (bf16 activations);
  - gradients per layer vs the bf16-rounding oracle: rel-L2 <= 2e-2 and cosine >= 0.999;
  - gradients vs the plain FP64 network: rel-L2 <= 1e-1 (reported per layer).
"""
import numpy as np
import pytest

import unet_ref as U

RNG = np.random.default_rng(77)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- oracle pins (CPU)
def test_conv3_matches_brute_force():
    x = RNG.normal(size=(1, 4, 5, 6, 3))
    w = RNG.normal(size=(27, 2, 3))
    b = RNG.normal(size=2)
    out = U.conv3(x, w, b)
    ref = np.zeros((1, 4, 5, 6, 2))
    for z in range(4):
        for y in range(5):
            for xx in range(6):
                for t, (dz, dy, dx) in enumerate(U.TAPS):
                    zz, yy, x2 = z + dz, y + dy, xx + dx
                    if 0 <= zz < 4 and 0 <= yy < 5 and 0 <= x2 < 6:
                        ref[0, z, y, xx] += w[t] @ x[0, zz, yy, x2]
                ref[0, z, y, xx] += b
    assert np.abs(out - ref).max() < 1e-12


def test_conv3_center_tap_is_pointwise_matmul():
    x = RNG.normal(size=(2, 4, 4, 4, 5))
    w = np.zeros((27, 3, 5))
    w[13] = RNG.normal(size=(3, 5))
    assert np.abs(U.conv3(x, w) - x @ w[13].T).max() < 1e-12


def test_pool_up_adjoint_and_inverse():
    x = RNG.normal(size=(1, 4, 6, 8, 3))
    g = RNG.normal(size=(1, 2, 3, 4, 3))
    # <pool x, g> = <x, pool^T g>;  <up g, x> = <g, up^T x>;  pool(up g) = g
    assert abs(np.sum(U.pool(x) * g) - np.sum(x * U.pool_backward(g))) < 1e-10
    assert abs(np.sum(U.up(g) * x) - np.sum(g * U.up_backward(x))) < 1e-10
    assert np.abs(U.pool(U.up(g)) - g).max() < 1e-12


def test_conv3_backward_adjoint():
    x = RNG.normal(size=(1, 4, 4, 4, 3))
    w = RNG.normal(size=(27, 2, 3))
    g = RNG.normal(size=(1, 4, 4, 4, 2))
    gx, gw, gb = U.conv3_backward(x, w, g)
    # <conv(x), g> is bilinear: its x-gradient and w-gradient reproduce it
    val = np.sum(U.conv3(x, w) * g)
    assert abs(np.sum(gx * x) - val) < 1e-9
    assert abs(np.sum(gw * w) - val) < 1e-9
    assert np.abs(gb - g.reshape(-1, 2).sum(0)).max() < 1e-12


def _small_problem(V=2, S=8, seed=3):
    rng = np.random.default_rng(seed)
    imgs = [rng.uniform(0.5, 2.0, (S, S, S)).astype(np.float32) for _ in range(V)]
    ws = [rng.uniform(0.2, 1.0, (S, S, S)).astype(np.float32) for _ in range(V)]
    scale = U.weighted_mean(imgs, ws)
    x0, avg = U.pack_input(imgs, ws, (0, 0, 0), (S, S, S), 1.0 / scale)
    target = avg * (1.0 + 0.3 * rng.normal(size=avg.shape))
    return x0[None], avg[None], target[None]


def test_oracle_gradient_matches_finite_differences():
    V = 2
    x0, avg, target = _small_problem(V)
    p = U.random_params(V, seed=5).astype(np.float64)
    loss, g = U.loss_and_grads(p, x0, avg, target)
    # central differences on a sample of parameters of every layer (and the head)
    rng = np.random.default_rng(9)
    offs = np.cumsum([0] + [27 * U.COUT[l] * U.CIN[l] + U.COUT[l] for l in range(10)])
    idx = []
    for l in range(10):
        nw = 27 * U.COUT[l] * U.CIN[l]
        w = p[offs[l]:offs[l] + nw].reshape(27, U.COUT[l], U.CIN[l])
        live = np.flatnonzero(w.ravel() != 0)  # unused input channels of layer 0 have no gradient path
        idx += list(offs[l] + rng.choice(live, 3, replace=False))
        idx.append(offs[l] + nw + rng.integers(U.COUT[l]))  # a bias
    idx += [offs[10] + 3, offs[10] + 16]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (U.loss_and_grads(pp, x0, avg, target)[0] - U.loss_and_grads(pm, x0, avg, target)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_oracle_unused_input_channels_get_no_gradient():
    V = 2
    x0, avg, target = _small_problem(V)
    p = U.random_params(V, seed=6).astype(np.float64)
    _, g = U.loss_and_grads(p, x0, avg, target)
    g0 = g[:27 * 16 * 16].reshape(27, 16, 16)
    assert np.all(g0[:, :, V + 1:] == 0)
    assert np.abs(g0[:, :, :V + 1]).max() > 0


def test_oracle_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-8, 65504.0, 1.0 + 2 ** -8], np.float32)
    r = U.bf16(x)
    # 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[5] == 1.0
    assert r[2] == 1.0078125
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


# ---------------------------------------------------------------- HIP kernels (GPU)
def _gpu_net(V, seed):
    from mvr import hip
    net = hip.UNet(V, seed=seed)
    assert net.param_count() == U.param_count()
    return net


@pytest.mark.gpu
def test_unet_param_roundtrip():
    V = 3
    net = _gpu_net(V, 1)
    p = net.params()
    layers, _ = U.unpack(p)
    assert np.all(layers[0][0][:, :, V + 1:] == 0)
    q = U.random_params(V, seed=2)
    net.set_params(q)
    assert np.array_equal(net.params(), q)
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("shape,V", [((16, 16, 16), 2), ((20, 24, 12), 6), ((18, 17, 15), 3), ((8, 12, 72), 2),
                                     ((12, 8, 40), 3), ((16, 12, 20), 8)])
def test_unet_forward_matches_oracle(shape, V):
    rng = np.random.default_rng(sum(shape) + V)
    imgs = [rng.uniform(0.0, 3.0, shape).astype(np.float32) for _ in range(V)]
    ws = [rng.uniform(0.0, 1.0, shape).astype(np.float32) for _ in range(V)]
    p = U.random_params(V, seed=11)
    net = _gpu_net(V, 0)
    net.set_params(p)
    psi = net.infer(imgs, ws, min_value=1e-4)
    net.close()
    ref_b = U.infer(p, imgs, ws, 1e-4, bf16_mode=True)
    ref_d = U.infer(p, imgs, ws, 1e-4, bf16_mode=False)
    eb, ed = rel(psi, ref_b), rel(psi, ref_d)
    print(f"forward {shape} V={V}: rel-L2 vs bf16 oracle {eb:.2e}, vs FP64 oracle {ed:.2e}")
    assert eb <= 2e-3
    assert ed <= 2e-2


def _train_problem(V, S, nvol, seed):
    rng = np.random.default_rng(seed)
    vols = []
    for _ in range(nvol):
        imgs = [rng.uniform(0.0, 3.0, (S, S, S)).astype(np.float32) for _ in range(V)]
        ws = [rng.uniform(0.1, 1.0, (S, S, S)).astype(np.float32) for _ in range(V)]
        tgt = (np.mean([a * b for a, b in zip(imgs, ws)], axis=0) *
               rng.uniform(0.5, 1.5, (S, S, S))).astype(np.float32)
        vols.append((imgs, ws, tgt))
    return vols


def _oracle_batch(vols, vidx, origins, P):
    xs, avgs, tg = [], [], []
    for vi, o in zip(vidx, origins):
        imgs, ws, tgt = vols[vi]
        s = np.float32(U.weighted_mean(imgs, ws))
        inv = np.float32(1.0) / s
        x0, avg = U.pack_input(imgs, ws, o, (P, P, P), inv)
        xs.append(x0)
        avgs.append(avg)
        tg.append((tgt[o[0]:o[0] + P, o[1]:o[1] + P, o[2]:o[2] + P] * inv).astype(np.float64))
    return np.stack(xs), np.stack(avgs), np.stack(tg)


@pytest.mark.gpu
def test_unet_gradients_match_oracle():
    V, S, P = 3, 24, 16
    vols = _train_problem(V, S, 2, 21)
    net = _gpu_net(V, 0)
    p = U.random_params(V, seed=12)
    net.set_params(p)
    for imgs, ws, tgt in vols:
        net.add_volume(imgs, ws, tgt)
    vidx = [0, 1]
    origins = [(0, 4, 8), (8, 8, 0)]
    loss = net.train_step(vidx, origins, P, lr=0.0, update=False)
    g = net.grads()
    x0, avg, tg = _oracle_batch(vols, vidx, origins, P)
    lb, gb = U.loss_and_grads(p, x0, avg, tg, bf16_mode=True)
    ld, gd = U.loss_and_grads(p, x0, avg, tg, bf16_mode=False)
    print(f"loss gpu {loss:.6e} oracle-bf16 {lb:.6e} oracle-fp64 {ld:.6e}")
    assert abs(loss - lb) <= 2e-3 * lb
    offs = np.cumsum([0] + [27 * U.COUT[l] * U.CIN[l] + U.COUT[l] for l in range(10)] + [17])
    names = [f"c{l}" for l in range(10)] + ["head"]
    for k in range(11):
        a, b, d = g[offs[k]:offs[k + 1]], gb[offs[k]:offs[k + 1]], gd[offs[k]:offs[k + 1]]
        cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)))
        print(f"{names[k]}: rel-L2 vs bf16 oracle {rel(a, b):.2e} (cos {cos:.6f}), vs FP64 oracle {rel(a, d):.2e}")
        assert rel(a, b) <= 2e-2 and cos >= 0.999, names[k]
        assert rel(a, d) <= 1e-1, names[k]
    net.close()


@pytest.mark.gpu
def test_unet_adam_step():
    V, S, P = 2, 16, 8
    vols = _train_problem(V, S, 1, 31)
    net = _gpu_net(V, 0)
    p = U.random_params(V, seed=13)
    net.set_params(p)
    net.add_volume(*vols[0])
    lr = 1e-3
    net.train_step([0], [(4, 0, 8)], P, lr=lr, update=True)
    g = net.grads().astype(np.float64)
    q = net.params().astype(np.float64)
    # first Adam step from zero moments: w - lr * g / (|g| + eps) (bias corrections cancel)
    expect = p.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    assert np.abs(q - expect).max() <= 1e-6
    net.close()


@pytest.mark.gpu
def test_unet_training_reduces_loss():
    V, S, P = 2, 32, 16
    vols = _train_problem(V, S, 2, 41)
    net = _gpu_net(V, 7)
    for v in vols:
        net.add_volume(*v)
    rng = np.random.default_rng(5)
    losses = []
    for step in range(60):
        vidx = rng.integers(0, 2, 2)
        origins = rng.integers(0, S - P + 1, (2, 3))
        losses.append(net.train_step(vidx, origins, P, lr=2e-3))
    print("loss first/last", np.mean(losses[:5]), np.mean(losses[-5:]))
    assert np.mean(losses[-5:]) < 0.8 * np.mean(losses[:5])
    net.close()


@pytest.mark.gpu
def test_decon_init_learned_uses_network_estimate():
    from mvr import hip
    from mvr import synthetic as syn
    w = syn.small_workload(n=24, views=2, k=7)
    inp = syn.make_inputs(w)
    net = _gpu_net(2, 3)
    psi_net = net.infer(inp["imgs"], inp["weights"], min_value=1e-4)
    d = hip.Decon(w.shape_zyx, 2, "EFFICIENT_BAYESIAN", "SEQ", 0.0)
    for v in range(2):
        d.set_view(v, inp["imgs"][v], inp["weights"][v], inp["psfs"][v])
    d.init_learned(net)
    assert np.array_equal(d.psi(), psi_net)
    d.run(2)
    d.reset_psi()
    assert np.array_equal(d.psi(), psi_net)
    # the distance check: psi to itself is 0, to a scaled copy is the scale error
    d.set_reference(psi_net)
    assert d.rel_l2() == 0.0
    d.set_reference(psi_net * 2.0)
    assert abs(d.rel_l2() - 0.5) < 1e-12
    d.close()
    net.close()


def test_receptive_field_within_inference_halo():
    # With every ReLU active (large biases) the network is affine in its input, so the planes an input
    # plane reaches are its structural receptive field; the inference slabs' halo must exceed it for the
    # slab-wise result to equal the single-window one.  Four plane alignments cover the pooling grid.
    from mvr import hip
    V = 2
    layers, head = U.unpack(U.random_params(V, seed=5))
    p = U.pack([(w, b + 50.0) for w, b in layers], head).astype(np.float64)
    rng = np.random.default_rng(0)
    D, H, W = 112, 8, 8
    x0 = np.zeros((1, D, H, W, 16))
    x0[..., :V + 1] = rng.uniform(0, 1, (1, D, H, W, V + 1))
    avg = np.zeros((1, D, H, W))
    base, _ = U.forward(p, x0, avg)
    radius = 0
    for k in range(52, 56):
        x = x0.copy()
        x[0, k, :, :, :V + 1] += 1.0
        out, _ = U.forward(p, x, avg)
        reach = np.nonzero(np.abs(out - base).max(axis=(0, 2, 3)) > 1e-9 * np.abs(base).max())[0]
        radius = max(radius, k - reach.min(), reach.max() - k)
    assert radius == 23   # 4 planes at full resolution, 8 at half, 8 at quarter, pooling grid offset
    halo = hip.UNet.inference_halo()
    assert halo > radius and halo % 4 == 0


@pytest.mark.gpu
@pytest.mark.parametrize("shape,slab", [((150, 20, 28), 8), ((150, 20, 28), 36), ((512, 128, 64), 36),
                                        ((516, 132, 68), 36)])
def test_unet_slab_inference_equals_single_window(shape, slab):
    # z slabs with their halo reproduce the single-window estimate exactly, also for a depth that is not a
    # multiple of 4 (the last slab carries the padded planes).  At 512 x 128 x 64 the single window's
    # full-resolution convolutions have 16384 plane tiles of 16 x 16: the 16 -> 16 layers run on the LDS-DMA
    # kernel (8 planes per work-group), the others 4 planes per work-group, the slabs' one plane per work-group
    # on the register-staged kernel.  516 x 132 x 68 adds partial tiles in x and y and a last z group of 4 planes
    from mvr import hip
    V = 3
    rng = np.random.default_rng(slab + shape[0])
    imgs = [rng.uniform(0.0, 3.0, shape).astype(np.float32) for _ in range(V)]
    ws = [rng.uniform(0.0, 1.0, shape).astype(np.float32) for _ in range(V)]
    net = _gpu_net(V, 0)
    net.set_params(U.random_params(V, seed=13))
    whole = net.infer(imgs, ws)
    halo = hip.UNet.inference_halo()
    net.set_inference_window((slab + 2 * halo) * shape[1] * shape[2])
    parts = net.infer(imgs, ws)
    net.set_inference_window(0)
    again = net.infer(imgs, ws)
    net.close()
    print(f"slab {slab} + 2 x {halo} planes: max |diff| {np.abs(parts - whole).max():.3e}")
    np.testing.assert_array_equal(parts, whole)
    np.testing.assert_array_equal(again, whole)


def test_weights_file_read_by_the_library(tmp_path):
    # mvr_unet_read_weights (host code, no GPU) against numpy's reader on the committed files, and its errors
    from mvr import hip
    from mvr import learned as L
    for _, path in L.TRAINING.values():
        with np.load(path, allow_pickle=False) as f:
            want, views = f["params"], int(f["views"])
        got, v = hip.read_weights(path)
        np.testing.assert_array_equal(got, want)
        assert v == views and got.size == U.param_count()
    p = U.random_params(3, seed=1)
    np.savez(tmp_path / "w.npz", params=p, views=np.int64(3))
    got, v = hip.read_weights(tmp_path / "w.npz")
    np.testing.assert_array_equal(got, p)
    assert v == 3
    np.savez_compressed(tmp_path / "c.npz", params=p, views=np.int32(3))
    with pytest.raises(hip.MvrError, match="compressed"):
        hip.read_weights(tmp_path / "c.npz")
    np.savez(tmp_path / "d.npz", params=p.astype(np.float64), views=np.int32(3))
    with pytest.raises(hip.MvrError, match="float32"):
        hip.read_weights(tmp_path / "d.npz")
    raw = open(tmp_path / "w.npz", "rb").read()
    open(tmp_path / "t.npz", "wb").write(raw[:1000])
    with pytest.raises(hip.MvrError):
        hip.read_weights(tmp_path / "t.npz")
    with pytest.raises(hip.MvrError, match="cannot open"):
        hip.read_weights(tmp_path / "missing.npz")


@pytest.mark.gpu
def test_unet_load_from_weights_file():
    from mvr import hip
    from mvr import learned as L
    path = L.TRAINING["c5"][1]
    net = hip.UNet.load(path)
    with np.load(path, allow_pickle=False) as f:
        np.testing.assert_array_equal(net.params(), f["params"])
        assert net.V == int(f["views"]) == 8
    net.close()
