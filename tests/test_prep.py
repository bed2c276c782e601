"""Input preparation: the numpy restatement (oracle/prep_ref.py) pinned on the CPU, and the HIP kernels
(affine resampling, blending weights, weight normalisation) checked against it through the C ABI.

Tolerances: resampled values rel-L2 <= 1e-6 (FP64 vs FP32 interpolation; the rounding order of ImgLib2's
interpolator is third-party and not pinned), identical inside/outside masks; blending weights and
normalised weights within 1e-6 absolute.
"""
import math

import numpy as np
import pytest

import prep_ref as P

RNG = np.random.default_rng(404)


def rot(axis, deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = np.eye(3)
    i, j = [(1, 2), (0, 2), (0, 1)][axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def affine(A, t):
    return np.concatenate([np.asarray(A, np.float64), np.asarray(t, np.float64)[:, None]], axis=1)


IDENT = affine(np.eye(3), [0, 0, 0])


# ---------------------------------------------------------------- pins (CPU)
def test_identity_transform_keeps_the_strict_interior():
    img = RNG.uniform(2, 9, (6, 7, 8)).astype(np.float32)
    out = P.transform_view(img, IDENT, (0, 0, 0), img.shape, min_value=1.0, outside=0.0)
    np.testing.assert_array_equal(out[1:-1, 1:-1, 1:-1], img[1:-1, 1:-1, 1:-1])
    # voxels on the input's faces are not strictly inside: outside value
    assert np.all(out[0] == 0) and np.all(out[:, :, -1] == 0)


def test_integer_translation_shifts_the_image():
    img = RNG.uniform(2, 9, (9, 10, 11)).astype(np.float32)
    a = affine(np.eye(3), [2, -1, 3])  # input pixel (x, y, z) -> world (x + 2, y - 1, z + 3)
    out = P.transform_view(img, a, (0, 0, 0), (9, 10, 11), min_value=0.0, has_min=False, outside=-1.0)
    for z, y, x in [(4, 4, 5), (5, 2, 3), (7, 7, 9)]:
        assert out[z, y, x] == img[z - 3, y + 1, x - 2]
    near = P.transform_view(img, a, (0, 0, 0), (9, 10, 11), has_min=False, outside=-1.0, interpolation=0)
    np.testing.assert_array_equal(near[out != -1.0], out[out != -1.0])


def test_trilinear_reproduces_affine_functions():
    nz, ny, nx = 12, 13, 14
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = lambda X, Y, Z: 3.0 + 0.5 * X - 0.25 * Y + 0.125 * Z  # exactly representable on the grid
    img = f(x, y, z).astype(np.float32)
    a = affine(rot(2, 17) @ rot(0, 9) @ np.diag([1.1, 0.9, 1.0]), [1.5, -0.7, 0.3])
    out = P.transform_view(img, a, (0, 0, 0), (nz, ny, nx), has_min=False, outside=np.nan)
    tx, ty, tz = P._input_coords(a, (0, 0, 0), (nz, ny, nx))
    ok = ~np.isnan(out)
    assert ok.sum() > 500
    np.testing.assert_allclose(out[ok], f(tx, ty, tz)[ok], rtol=1e-6)


def test_nearest_rounds_half_up():
    img = np.arange(4 * 4 * 4, dtype=np.float32).reshape(4, 4, 4)
    a = affine(np.eye(3), [-0.5, 0, 0])  # world x = input x - 0.5: world 1 -> input 1.5 -> pixel 2
    out = P.transform_view(img, a, (1, 1, 1), (1, 1, 1), has_min=False, interpolation=0)
    assert out[0, 0, 0] == img[1, 1, 2]


def test_blending_identity_is_product_of_axis_weights():
    from mvr import synthetic as syn
    shape = (30, 45, 50)
    w = P.transform_blending(shape, (0, 0, 0), (20, 10, 10), IDENT, (0, 0, 0), shape)
    wz = syn.blend_weight_1d(shape[0], 10)
    wy = syn.blend_weight_1d(shape[1], 10)
    wx = syn.blend_weight_1d(shape[2], 20)
    ref = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
    np.testing.assert_allclose(w, ref, rtol=1e-6, atol=0)
    assert w[15, 22, 25] == 1.0 and np.all(w[0] == 0)


def test_normalisation_closed_forms():
    w = np.array([[0.3, 0.8, 1.5, 0.6, 0.55], [0.4, 0.9, 0.2, 0.6, 0.55]], np.float32)[:, None, None, :]
    hard = P.normalize_weights(w)
    np.testing.assert_allclose(hard[:, 0, 0, 0], [0.3, 0.4], rtol=1e-7)          # sum <= 1: unchanged
    np.testing.assert_allclose(hard[:, 0, 0, 1], [0.8 / 1.7, 0.9 / 1.7], rtol=1e-6)
    np.testing.assert_allclose(hard[:, 0, 0, 2], [1 / 1.2, 0.2 / 1.2], rtol=1e-6)  # 1.5 clipped to 1
    osem = P.normalize_weights(w, osem_speedup=2.0)
    np.testing.assert_allclose(osem[:, 0, 0, 0], [0.6, 0.8], rtol=1e-6)
    assert np.all(osem <= 1.0)
    sm = P.normalize_weights(w, smooth=True)
    assert sm[0, 0, 0, 3] == pytest.approx(0.5, abs=1e-7)        # diff 0.1 = maxDiffRange: no scaling
    assert sm[0, 0, 0, 4] == pytest.approx(0.5 - 0.5 * 0.55 * 0.05, abs=1e-6)


# ---------------------------------------------------------------- HIP kernels (GPU)
def test_fuse_group_single_view_is_the_transformed_view():
    img = RNG.uniform(2, 60, (12, 14, 16)).astype(np.float32)
    a = affine(rot(2, 25), [1.5, -2.0, 0.5])
    bb_min, bb_shape = (-3, -3, -2), (16, 20, 22)
    fused, w = P.fuse_group([img], [a], bb_min, bb_shape)
    np.testing.assert_array_equal(fused, P.transform_view(img, a, bb_min, bb_shape, 1.0, True, 0.0, 1))
    np.testing.assert_array_equal(w, np.ones(bb_shape, np.float32))
    fused, w = P.fuse_group([img], [a], bb_min, bb_shape, decon_blending=[((1, 1, 0), (4, 4, 3))])
    np.testing.assert_array_equal(w, P.transform_blending(img.shape, (1, 1, 0), (4, 4, 3), a, bb_min, bb_shape))


def test_fuse_group_duplicate_view_and_weight_sum():
    img = RNG.uniform(2, 60, (12, 14, 16)).astype(np.float32)
    a = affine(rot(0, 15), [0.5, 1.0, -1.0])
    bb_min, bb_shape = (-2, -2, -2), (16, 18, 20)
    blend = [((1, 1, 1), (5, 5, 3))]
    one, w1 = P.fuse_group([img], [a], bb_min, bb_shape, fusion_blending=blend, decon_blending=blend)
    two, w2 = P.fuse_group([img, img], [a, a], bb_min, bb_shape, fusion_blending=blend * 2, decon_blending=blend * 2)
    np.testing.assert_array_equal(two, one)          # (v w + v w) / 2w = v
    np.testing.assert_array_equal(w2, 2 * w1)        # the summed decon weights


def test_fuse_group_average_closed_forms():
    # constant views c1 (box 0..9) and c2 shifted by +4 in x; no fusion blending = weight 1 everywhere, so
    # outside a view its 0 joins the average (the reference's ConstantRandomAccessible branch)
    c1, c2 = np.float32(10.0), np.float32(30.0)
    img1 = np.full((10, 10, 10), c1, np.float32)
    img2 = np.full((10, 10, 10), c2, np.float32)
    a1, a2 = IDENT, affine(np.eye(3), [4.0, 0.0, 0.0])
    fused, w = P.fuse_group([img1, img2], [a1, a2], (0, 0, 0), (10, 10, 14))
    assert fused[5, 5, 6] == np.float32((10.0 + 30.0) / 2)   # strictly inside both
    assert fused[5, 5, 11] == np.float32(30.0 / 2)           # inside view 2 only
    assert fused[5, 5, 0] == 0.0                             # on view 1's face (not strictly inside) and outside 2
    np.testing.assert_array_equal(w, np.full((10, 10, 14), 2.0, np.float32))
    # with fusion blending a view outside its image has weight 0 and is skipped
    fb = [((0, 0, 0), (0, 0, 0))] * 2
    fused, _ = P.fuse_group([img1, img2], [a1, a2], (0, 0, 0), (10, 10, 14), fusion_blending=fb)
    assert fused[5, 5, 11] == np.float32(30.0)
    assert fused[5, 5, 6] == np.float32(20.0)
    # intensity adjustment (float)(v * m00 + m01) before resampling, then the min clamp
    fused, _ = P.fuse_group([img1], [a1], (0, 0, 0), (10, 10, 10), intensity_adjust=[(2.0, 5.0)])
    assert fused[5, 5, 5] == np.float32(25.0)
    fused, _ = P.fuse_group([img1], [a1], (0, 0, 0), (10, 10, 10), intensity_adjust=[(0.0, -3.0)])
    assert fused[5, 5, 5] == np.float32(1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("interp", [1, 0])
def test_transform_view_matches_oracle(interp):
    from mvr import hip
    img = RNG.uniform(0, 50, (40, 36, 44)).astype(np.float32)
    a = affine(rot(1, 33) @ rot(2, -12) @ np.diag([1.0, 1.0, 2.2]), [10.3, -4.1, 2.7])
    bb_min, bb_shape = (-5, -12, 0), (70, 48, 60)
    g = hip.transform_view(img, a, bb_min, bb_shape, min_value=1.0, outside=0.0, interpolation=interp)
    o = P.transform_view(img, a, bb_min, bb_shape, min_value=1.0, outside=0.0, interpolation=interp)
    tx, ty, tz = P._input_coords(a, bb_min, bb_shape)
    # voxels whose input position lies within 1e-9 of a face could go either way in FP64 vs the kernel
    edge = np.zeros(bb_shape, bool)
    for t, n in ((tx, 44), (ty, 36), (tz, 40)):
        edge |= (np.abs(t) < 1e-9) | (np.abs(t - (n - 1)) < 1e-9)
    inside = (o != 0) & ~edge
    assert np.array_equal((g != 0) & ~edge, inside)
    assert inside.sum() > 10000
    e = float(np.linalg.norm(g[inside] - o[inside]) / np.linalg.norm(o[inside]))
    print(f"transform_view interp={interp}: rel-L2 {e:.2e}, {inside.sum()} voxels inside")
    assert e <= 1e-6


@pytest.mark.gpu
def test_transform_blending_matches_oracle():
    from mvr import hip
    a = affine(rot(0, 21) @ np.diag([1.0, 1.0, 1.8]), [3.3, -2.0, 5.5])
    img_shape, bb_min, bb_shape = (30, 50, 60), (-4, -20, 0), (64, 60, 72)
    g = hip.transform_blending(img_shape, (2.0, 1.0, 0.0), (15.0, 12.0, 8.0), a, bb_min, bb_shape)
    o = P.transform_blending(img_shape, (2.0, 1.0, 0.0), (15.0, 12.0, 8.0), a, bb_min, bb_shape)
    assert (o > 0).sum() > 10000
    assert np.abs(g - o).max() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("osem,smooth", [(1.0, False), (2.5, False), (1.0, True)])
def test_normalize_weights_matches_oracle(osem, smooth):
    from mvr import hip
    w = RNG.uniform(0, 1.3, (4, 20, 22, 24)).astype(np.float32)
    w[:, :3] = 0
    g = hip.normalize_weights(w, osem_speedup=osem, smooth=smooth)
    o = P.normalize_weights(w, osem_speedup=osem, smooth=smooth)
    assert np.abs(g - o).max() <= 1e-6


# ---------------------------------------------------------------- PSF preparation
def _gauss_psf(shape, sig):
    z, y, x = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    return np.exp(-0.5 * ((x / sig[0]) ** 2 + (y / sig[1]) ** 2 + (z / sig[2]) ** 2)).astype(np.float32)


def test_psf_transform_identity_and_quarter_turn():
    psf = _gauss_psf((9, 11, 13), (1.0, 2.0, 1.5)) + RNG.uniform(0, 0.01, (9, 11, 13)).astype(np.float32)
    p64 = psf.astype(np.float64)
    norm = ((p64 - p64.min()) / (p64.max() - p64.min())).astype(np.float32)
    out = P.transform_psf(psf, IDENT)
    assert out.shape == psf.shape
    np.testing.assert_array_equal(out, norm)
    # x' = -y, y' = x: +x turns onto +y, a quarter turn in the xy plane (rot90 from axis x towards axis y)
    q = affine([[0, -1, 0], [1, 0, 0], [0, 0, 1]], [5, 2, 0])
    out = P.transform_psf(psf, q)
    assert out.shape == (9, 13, 11)
    np.testing.assert_array_equal(out, np.rot90(norm, k=1, axes=(2, 1)))


def test_psf_average_and_same_size():
    from mvr import prep
    a = np.zeros((3, 5, 5), np.float32)
    a[1, 2, 2] = 4
    b = np.zeros((3, 3, 7), np.float32)
    b[1, 1, 3] = 2
    avg = prep.average_psf([a, b])
    assert avg.shape == (3, 3, 5) and avg[1, 1, 2] == 3.0 and avg.sum() == 3.0
    big = prep.make_same_size(np.arange(27, dtype=np.float32).reshape(3, 3, 3) + 1, (5, 5, 5))
    assert big[2, 2, 2] == 14 and big[0, 0, 0] == 1 and big[1, 1, 1] == 1
    small = prep.make_same_size(big, (3, 3, 3))
    np.testing.assert_array_equal(small, np.arange(27, dtype=np.float32).reshape(3, 3, 3) + 1)


@pytest.mark.gpu
def test_transform_psf_matches_oracle():
    from mvr import hip
    psf = _gauss_psf((21, 19, 17), (1.5, 1.2, 3.0))
    a = affine(rot(1, 40) @ rot(2, 15) @ np.diag([1.0, 1.0, 1.7]), [3.0, -1.0, 2.0])
    g = hip.transform_psf(psf, a)
    o = P.transform_psf(psf, a)
    assert g.shape == o.shape and all(s % 2 == 1 for s in g.shape)
    e = float(np.linalg.norm(g - o) / np.linalg.norm(o))
    print(f"transform_psf {psf.shape} -> {g.shape}: rel-L2 {e:.2e}")
    assert e <= 1e-6


@pytest.mark.gpu
def test_prepare_views_pipeline():
    from mvr import prep
    phantom = RNG.uniform(5, 50, (40, 44, 48)).astype(np.float32)
    affines = [affine(rot(1, ang) @ np.diag([1.0, 1.0, 1.4]), [4.0, -3.0, 6.0]) for ang in (0, 35, 70)]
    box_min, box_shape = (-6, -8, 0), (64, 52, 60)
    imgs, ws = prep.prepare_views([phantom] * 3, affines, box_min, box_shape, blending_range_xyz=(10, 10, 6),
                                  blending_border_xyz=(1, 1, 0), osem_speedup=1.0)
    oi = np.stack([P.transform_view(phantom, a, box_min, box_shape, 1.0, True, 0.0, 1) for a in affines])
    ow = P.normalize_weights(np.stack([P.transform_blending(phantom.shape, (1, 1, 0), (10, 10, 6), a, box_min,
                                                            box_shape) for a in affines]))
    assert float(np.linalg.norm(imgs - oi) / np.linalg.norm(oi)) <= 1e-6
    assert np.abs(ws - ow).max() <= 1e-6
    assert ws.sum(0).max() <= 1.0 + 1e-6


@pytest.mark.gpu
def test_fuse_group_matches_oracle():
    from mvr import hip
    rng = np.random.default_rng(9)
    imgs = [rng.uniform(0, 50, (30, 34, 38)).astype(np.float32) for _ in range(3)]
    affs = [affine(rot(1, ang) @ np.diag([1.0, 1.0, 1.6]), [2.0 + ang / 20, -3.0, 1.0]) for ang in (0, 40, 80)]
    adj = [(1.0, 0.0), (1.3, -2.0), (0.8, 4.0)]
    fb = [(1, 1, 0, 8, 8, 5)] * 3
    db = [(0, 0, 0, 12, 12, 6)] * 3
    bb_min, bb_shape = (-10, -6, -4), (50, 44, 56)
    g, gw = hip.fuse_group(imgs, affs, bb_min, bb_shape, adj, fb, db)
    o, ow = P.fuse_group(imgs, affs, bb_min, bb_shape, adj, [((1, 1, 0), (8, 8, 5))] * 3,
                         [((0, 0, 0), (12, 12, 6))] * 3)
    edge = np.zeros(bb_shape, bool)   # input positions within 1e-9 of a face may go either way in FP64
    for a in affs:
        tx, ty, tz = P._input_coords(a, bb_min, bb_shape)
        for t, n in ((tx, 38), (ty, 34), (tz, 30)):
            edge |= (np.abs(t) < 1e-9) | (np.abs(t - (n - 1)) < 1e-9)
    ok = ~edge
    assert (o[ok] > 0).sum() > 20000
    e = float(np.linalg.norm(g[ok] - o[ok]) / np.linalg.norm(o[ok]))
    print(f"fuse_group 3 views: rel-L2 {e:.2e}, weights max |diff| {np.abs(gw - ow).max():.1e}")
    assert e <= 1e-6
    assert np.abs(gw - ow).max() <= 1e-6


@pytest.mark.gpu
def test_prepare_groups_pipeline():
    from mvr import prep
    rng = np.random.default_rng(10)
    ph = rng.uniform(5, 50, (30, 32, 36)).astype(np.float32)
    groups = [[{"img": ph, "affine": affine(rot(1, ang), [3.0, -2.0, 1.0])},
               {"img": ph, "affine": affine(rot(1, ang + 180), [3.0, -2.0, 1.0]), "intensity": (1.1, -1.0)}]
              for ang in (0, 60)]
    bb_min, bb_shape = (-20, -8, -20), (40, 48, 44)
    imgs, ws = prep.prepare_groups(groups, bb_min, bb_shape, blending_range_fusion=(6, 6, 4),
                                   blending_border_fusion=(0, 0, 0), blending_range_decon=(10, 10, 5),
                                   blending_border_decon=(0, 0, 0))
    want_i, want_w = [], []
    for views in groups:
        adj = [v.get("intensity", (1.0, 0.0)) for v in views]
        fi, fw = P.fuse_group([v["img"] for v in views], [v["affine"] for v in views], bb_min, bb_shape, adj,
                              [((0, 0, 0), (6, 6, 4))] * 2, [((0, 0, 0), (10, 10, 5))] * 2)
        want_i.append(fi)
        want_w.append(fw)
    want_w = P.normalize_weights(np.stack(want_w))
    assert float(np.linalg.norm(imgs - np.stack(want_i)) / np.linalg.norm(np.stack(want_i))) <= 1e-6
    assert np.abs(ws - want_w).max() <= 1e-6
