"""FP64 numpy restatement of the learned initial estimate's network (test infrastructure only).

This module is the checker for the HIP U-Net kernels: only tests, the smoke check and the bench's
CPU leg may import it.  The network has no counterpart in the reference; it replaces the estimate of
PsiInitAvgPrecise (src/main/java/net/preibisch/mvrecon/process/deconvolution/init/PsiInitAvgPrecise.java:52-112)
and plugs in where PsiInitFromRAI does (deconvolution/init/PsiInitFromRAI.java:65-82).

Network (tensors [n][z][y][x][c], channels padded to multiples of 16):
  x0   = [w_v img_v / s for v < V] + [sum_v w_v img_v / (s sum_v w_v)] + zeros   (16 channels)
  e1   = relu(c1(relu(c0(x0))))                 16
  e2   = relu(c3(relu(c2(pool(e1)))))           32
  b3   = relu(c5(relu(c4(pool(e2)))))           64
  d2   = relu(c7(relu(c6([up(b3), e2]))))       32   (concat: up(b3) in channels 0..63)
  d1   = relu(c9(relu(c8([up(d2), e1]))))       16
  pred = weighted-mean channel (FP32) + d1 . wo + bo
s = sum w_v img_v / sum w_v over the volume (the weighted mean intensity); psi = max(minValue, s pred).
3^3 convolutions with zero padding, tap t = 9 (dz+1) + 3 (dy+1) + (dx+1), out[v] = sum_t in[v + off_t] w[t];
pool = 2x2x2 mean, up = nearest x2.  Training loss: mean of (pred - target / s)^2.

With ``bf16=True`` every tensor the kernels store in bfloat16 is rounded (round to nearest even) at
the same point: packed input, conv outputs, pooled tensors, activation gradients; weights of the 3^3
layers are used in bfloat16.  ``bf16=False`` is the plain FP64 network (weights still as given).

Pins (tests/test_unet.py): the backward pass against central finite differences of the forward pass
in FP64; the 3^3 convolution against a brute-force loop; pool/up adjointness.
"""
from __future__ import annotations

import numpy as np

CIN = (16, 16, 16, 32, 32, 64, 96, 32, 48, 16)
COUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 16)
TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def param_count():
    return sum(27 * COUT[l] * CIN[l] + COUT[l] for l in range(10)) + 17


def unpack(flat):
    """flat FP32 vector -> list of (w[27][cout][cin], b[cout]) for 10 layers, head (wo[16], bo)."""
    flat = np.asarray(flat, np.float64)
    layers, off = [], 0
    for l in range(10):
        n = 27 * COUT[l] * CIN[l]
        w = flat[off:off + n].reshape(27, COUT[l], CIN[l])
        off += n
        b = flat[off:off + COUT[l]]
        off += COUT[l]
        layers.append((w, b))
    wo = flat[off:off + 16]
    bo = flat[off + 16]
    assert off + 17 == flat.size
    return layers, (wo, bo)


def pack(layers_grads, head_grads):
    parts = []
    for w, b in layers_grads:
        parts += [np.ravel(w), np.ravel(b)]
    parts += [np.ravel(head_grads[0]), np.atleast_1d(head_grads[1])]
    return np.concatenate(parts)


def random_params(V, seed=0, scale=1.0):
    """He-normal weights (zero for unused input channels of layer 0), small random biases and head."""
    rng = np.random.default_rng(seed)
    layers = []
    for l in range(10):
        fan = 27 * (V + 1 if l == 0 else CIN[l])
        w = rng.normal(0, np.sqrt(2.0 / fan) * scale, (27, COUT[l], CIN[l]))
        if l == 0:
            w[:, :, V + 1:] = 0
        b = rng.normal(0, 0.05, COUT[l])
        layers.append((w, b))
    head = (rng.normal(0, 0.1, 16), 0.01)
    return pack(layers, head).astype(np.float32)


# ---------------------------------------------------------------- building blocks
def conv3(x, w, b=None):
    """x [N,D,H,W,Ci], w [27,Co,Ci] -> [N,D,H,W,Co], zero padding."""
    N, D, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((N, D, H, W, w.shape[1]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :] @ w[t].T
    if b is not None:
        out += b
    return out


def conv3_backward(x, w, gout):
    """gradients of conv3 w.r.t. (x, w, b) for output gradient gout."""
    N, D, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    g2 = gout.reshape(-1, gout.shape[-1])
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = (slice(None), slice(1 + dz, 1 + dz + D), slice(1 + dy, 1 + dy + H), slice(1 + dx, 1 + dx + W))
        gw[t] = g2.T @ xp[sl].reshape(-1, Ci)
        gxp[sl] += gout @ w[t]
    return gxp[:, 1:-1, 1:-1, 1:-1, :], gw, g2.sum(axis=0)


def pool(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_backward(g):
    N, D, H, W, C = g.shape
    return g.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def pack_input(imgs, weights, origin_zyx, size_zyx, inv_scale):
    """Network input of one sample: region origin..origin+size of the volumes (indices past the last
    voxel repeat it).  Channels v < V: w_v img_v * inv_scale; channel V: their weighted mean
    sum_v w_v img_v inv_scale / sum_v w_v (0 where no view has weight).  Float32 arithmetic like the
    kernel; returns (x [D,H,W,16], mean [D,H,W]) as float64 arrays."""
    V = len(imgs)
    nz, ny, nx = imgs[0].shape
    z = np.minimum(origin_zyx[0] + np.arange(size_zyx[0]), nz - 1)
    y = np.minimum(origin_zyx[1] + np.arange(size_zyx[1]), ny - 1)
    x = np.minimum(origin_zyx[2] + np.arange(size_zyx[2]), nx - 1)
    ix = np.ix_(z, y, x)
    out = np.zeros(tuple(size_zyx) + (16,), np.float32)
    s = np.zeros(tuple(size_zyx), np.float32)
    sw = np.zeros(tuple(size_zyx), np.float32)
    for v in range(V):
        wv = np.asarray(weights[v], np.float32)[ix]
        val = (np.asarray(imgs[v], np.float32)[ix] * wv) * np.float32(inv_scale)
        out[..., v] = val
        s += val
        sw += wv
    mean = np.where(sw > 0, s / np.where(sw > 0, sw, np.float32(1)), np.float32(0)).astype(np.float32)
    out[..., V] = mean
    return out.astype(np.float64), mean.astype(np.float64)


def weighted_mean(imgs, weights):
    """sum over voxels and views of w_v img_v divided by the sum of w_v (1 without weight)."""
    s = sw = 0.0
    for a, b in zip(imgs, weights):
        b32 = np.asarray(b, np.float32)
        s += float((np.asarray(a, np.float32) * b32).astype(np.float64).sum())
        sw += float(b32.astype(np.float64).sum())
    return max(1e-6, s / sw) if sw > 0 else 1.0


# ---------------------------------------------------------------- network
def forward(params, x0, avg, bf16_mode=False):
    """x0 [N,D,H,W,16], avg [N,D,H,W] -> (pred [N,D,H,W], cache)."""
    layers, (wo, bo) = unpack(params)
    R = bf16 if bf16_mode else (lambda a: a)
    Wl = [(R(w), b) for w, b in layers] if bf16_mode else layers

    def cr(l, x):
        return R(np.maximum(conv3(x, Wl[l][0], Wl[l][1]), 0.0))

    c = {}
    c["x0"] = R(x0)
    c["a1"] = cr(0, c["x0"])
    c["e1"] = cr(1, c["a1"])
    c["p1"] = R(pool(c["e1"]))
    c["a2"] = cr(2, c["p1"])
    c["e2"] = cr(3, c["a2"])
    c["p2"] = R(pool(c["e2"]))
    c["a3"] = cr(4, c["p2"])
    c["b3"] = cr(5, c["a3"])
    c["cat2"] = np.concatenate([up(c["b3"]), c["e2"]], axis=-1)
    c["a4"] = cr(6, c["cat2"])
    c["d2"] = cr(7, c["a4"])
    c["cat1"] = np.concatenate([up(c["d2"]), c["e1"]], axis=-1)
    c["a5"] = cr(8, c["cat1"])
    c["d1"] = cr(9, c["a5"])
    pred = avg + c["d1"] @ wo + bo
    c["W"] = Wl
    c["head"] = (wo, bo)
    return pred, c


def loss_and_grads(params, x0, avg, target, bf16_mode=False):
    """Mean squared error to target and its gradient as a flat vector (layout of the parameters)."""
    pred, c = forward(params, x0, avg, bf16_mode)
    R = bf16 if bf16_mode else (lambda a: a)
    Wl = c["W"]
    wo, bo = c["head"]
    n = pred.size
    e = pred - target
    loss = float(np.mean(e * e))
    dr = 2.0 * e / n
    g_head = ((dr[..., None] * c["d1"]).reshape(-1, 16).sum(axis=0), float(dr.sum()))
    gl = [None] * 10

    def back(l, x, post, g):
        g = g * (post > 0)
        gx, gw, gb = conv3_backward(x, Wl[l][0], g)
        gl[l] = (gw, gb)
        return R(gx)

    g = R(dr[..., None] * wo)
    g = back(9, c["a5"], c["d1"], g)
    g = back(8, c["cat1"], c["a5"], g)
    g_up_d2, g_e1 = g[..., :32], g[..., 32:]
    g = R(up_backward(g_up_d2))
    g = back(7, c["a4"], c["d2"], g)
    g = back(6, c["cat2"], c["a4"], g)
    g_up_b3, g_e2 = g[..., :64], g[..., 64:]
    g = R(up_backward(g_up_b3))
    g = back(5, c["a3"], c["b3"], g)
    g = back(4, c["p2"], c["a3"], g)
    g_e2 = R(g_e2 + R(pool_backward(g)))
    g = back(3, c["a2"], c["e2"], g_e2)
    g = back(2, c["p1"], c["a2"], g)
    g_e1 = R(g_e1 + R(pool_backward(g)))
    g = back(1, c["a1"], c["e1"], g_e1)
    back(0, c["x0"], c["a1"], g)
    return loss, pack(gl, g_head)


def infer(params, imgs, weights, min_value=1e-4, bf16_mode=False):
    """psi estimate for whole volumes (z, y, x); extents padded up to multiples of 4 like the kernels."""
    shape = imgs[0].shape
    P = tuple((s + 3) // 4 * 4 for s in shape)
    scale = weighted_mean(imgs, weights)
    x0, avg = pack_input(imgs, weights, (0, 0, 0), P, np.float32(1.0 / np.float32(scale)))
    pred, _ = forward(params, x0[None], avg[None], bf16_mode)
    psi = np.maximum(min_value, np.float32(scale) * pred[0])
    return psi[:shape[0], :shape[1], :shape[2]]
