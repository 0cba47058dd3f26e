"""GPU parity: libnli.so kernels (through the C-ABI) against the CPU oracle on the
same seeded inputs.

Bars: bit-exact for the gradient, the distance transform, the pyramid (images and
masks), every mask, the patch distances at a fixed NNF, NL-means and NL-medians (on
random and on PatchMatch fields); NL-Poisson/Depth and
the CG solution within 1e-5 (absolute, images in [0, 1]; relative for depth);
confidence within one float ulp (the exponential is a different libm).
PatchMatch is randomised with a different search schedule, so it is checked by
invariants and by its NNF energy against the serial reference run (GPU energy at
most 1 % above it).
"""
import numpy as np
import pytest

from conftest import bitexact, golden

pytestmark = pytest.mark.gpu


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


# ---------------------------------------------------------------- numerics
@pytest.mark.parametrize("W,H,C", [(37, 29, 3), (64, 48, 1), (33, 65, 4), (256, 200, 3)])
def test_gradient(ctx, oracle, syn, W, H, C):
    img = syn.texture(W, H, 3, 3, W)[..., :1] if C == 1 else syn.texture(W, H, 3, 3, W)
    if C == 4:
        img = np.concatenate([img, syn.depth_planes(W, H, 5, 1)[..., None]], axis=2)
    assert bitexact(ctx.gradient(img), oracle.gradient(img))


@pytest.mark.parametrize("W,H,seed", [(37, 29, 0), (128, 96, 1), (301, 207, 2)])
def test_distance_transform_and_confidence(ctx, oracle, syn, W, H, seed):
    m = syn.ellipse_holes(W, H, 4, 0.25, seed, border=1)
    assert bitexact(ctx.distance_transform(m), oracle.distance_transform(m))
    c_gpu, c_cpu = ctx.confidence(m, 5.0, 0.1), oracle.confidence(m, 5.0, 0.1)
    assert ulps(c_gpu, c_cpu).max() <= 1
    assert bitexact(ctx.confidence(m, 0.0, 0.3), oracle.confidence(m, 0.0, 0.3))


@pytest.mark.parametrize("W,H,seed", [(102, 102, 1), (97, 61, 2), (64, 129, 3), (7, 5, 4)])
def test_poisson_fill_on_device_is_bitexact(ctx, oracle, syn, W, H, seed):
    """The coarsest-level fill (simpois.c) on the device: anti-diagonal Gauss-Seidel
    wavefronts reproduce the sequential in-place sweeps bit for bit."""
    img = syn.texture(W, H, 3, 3, seed)
    hole = syn.ellipse_holes(W, H, 3, 0.2, seed, border=1) if min(W, H) > 16 else np.zeros((H, W), np.uint8)
    if min(W, H) <= 16:
        hole[1:H - 1, 2:W - 1] = 1
    hole[0, :W // 3] = 1                                       # touches the image border
    out = ctx.poisson_fill(img, hole)
    assert bitexact(out, oracle.simpois(img, hole))


def test_golden_numerics(ctx):
    g = golden("numerics")
    img, hole = g["img"], g["hole"]
    rate = float(g["rate_0p1_11"])
    assert bitexact(ctx.gradient(img), g["gradient"])
    assert bitexact(ctx.distance_transform(hole), g["dt"])
    oi, om = ctx.downsample_with_mask(img, hole, rate)
    assert bitexact(oi, g["ds_img"]) and bitexact(om, g["ds_mask"])
    assert bitexact(ctx.downsample(img, rate), g["ds_plain"])
    assert bitexact(ctx.downsample_mask(hole, rate), g["ds_maskonly"])
    assert bitexact(ctx.upsample(img, 50, 41), g["up"])
    assert bitexact(ctx.extended_domain(hole, 9), g["ext9"])


def test_pyramid_chain_1024(ctx, oracle, syn):
    """The 11-level pyramid of the 1024x1024 workload, level by level."""
    img = syn.texture(1024, 1024, 3, 8, 1)
    hole = syn.ellipse_holes(1024, 1024, 6, 0.15, 1)
    rate = oracle.subsampling_rate(0.1, 11)
    for level in range(10):
        a = ctx.downsample_with_mask(img, hole, rate)
        b = oracle.downsample_with_mask(img, hole, rate)
        assert bitexact(a[0], b[0]) and bitexact(a[1], b[1]), level
        up_g = ctx.upsample(b[0], img.shape[1], img.shape[0])
        assert bitexact(up_g, oracle.upsample(b[0], img.shape[1], img.shape[0])), level
        img, hole = b


@pytest.mark.parametrize("factor", [0.5, 0.7197, 0.83])
def test_plain_downsampling(ctx, oracle, syn, factor):
    img = syn.texture(131, 97, 3, 3, 3)
    m = syn.ellipse_holes(131, 97, 3, 0.3, 3)
    assert bitexact(ctx.downsample(img, factor), oracle.downsample(img, factor))
    assert bitexact(ctx.downsample_mask(m, factor), oracle.downsample_mask(m, factor))
    a = ctx.downsample_with_mask(img, m, factor, 0.4, True)
    b = oracle.downsample_with_mask(img, m, factor, 0.4, True)
    assert bitexact(a[0], b[0]) and bitexact(a[1], b[1])


# ---------------------------------------------------------------- masks
@pytest.mark.parametrize("P", [3, 7, 9, 11])
def test_masks(ctx, oracle, syn, P):
    hole = syn.stripe_holes(160, 120, 3, 9, P)       # stripes touch the image border
    assert bitexact(ctx.extended_domain(hole, P), oracle.extended_domain(hole, P))
    for erode in (False, True):
        e, t, s = ctx.scale_masks(hole, P, erode)
        e2, t2, s2, empty = oracle.scale_masks(hole, P, erode)
        assert bitexact(e, e2) and bitexact(t, t2) and bitexact(s, s2)


def test_empty_source_is_reported(ctx, nli):
    hole = np.ones((40, 40), np.uint8)
    with pytest.raises(nli.EmptySourceError):
        ctx.scale_masks(hole, 7, True)


# ---------------------------------------------------------------- distances
def test_patch_distance_golden(ctx):
    g = golden("field64x48")
    img, img4, nnf, tgt = g["img"], g["img4"], g["nnf"], g["tgt"]
    for kind in ["l2", "l1", "l2combined"]:
        assert bitexact(ctx.patch_distance(kind, img, img, nnf, tgt, 9, 10000.0, 0.3, 0.7), g[f"dist_{kind}"]), kind
    assert bitexact(ctx.patch_distance("l2rgbd", img4, img4, nnf, tgt, 9, 10000.0, 0.3, 0.7), g["dist_l2rgbd"])


@pytest.mark.parametrize("kind", ["l2", "l1", "l2combined", "l2rgbd"])
@pytest.mark.parametrize("C", [1, 2, 3, 4])
def test_patch_distance_random_fields(ctx, oracle, syn, kind, C):
    if kind == "l2rgbd" and C != 4:
        pytest.skip("RGBD needs 4 channels")
    W, H, P, r = 90, 70, 7, 3
    img = syn.texture(W, H, 3, 3, C)
    if C == 4:
        img = np.concatenate([img, syn.depth_planes(W, H, 5, C)[..., None]], axis=2)
    img = np.ascontiguousarray(img[..., :C])
    rng = np.random.default_rng(C)
    tm = np.zeros((H, W), np.uint8)
    tm[r:H - r, r:W - r] = rng.integers(0, 2, (H - 2 * r, W - 2 * r))
    nnf = np.stack([rng.integers(r, W - r, (H, W)), rng.integers(r, H - r, (H, W))], axis=2).astype(np.int32)
    dp = oracle.dist_params(kind, P, 2.5, 0.3, 0.6)
    got = ctx.patch_distance(kind, img, img, nnf, tm, P, 2.5, 0.3, 0.6)
    assert bitexact(got, oracle.patch_distance_field(dp, img, img, nnf, tm))


def test_rgbd_needs_four_channels(ctx, syn):
    img = syn.texture(40, 40, 3, 3, 1)
    nnf = np.zeros((40, 40, 2), np.int32)
    with pytest.raises(ValueError, match="Channels do not match"):
        ctx.patch_distance("l2rgbd", img, img, nnf, np.zeros((40, 40), np.uint8))


# ---------------------------------------------------------------- PatchMatch
def energy(oracle, dp, img, nnf, tgt):
    d = oracle.patch_distance_field(dp, img, img, nnf, tgt)
    sel = tgt.astype(bool)
    assert np.all(nnf[sel][:, 0] >= 0)
    return float(np.sum(d[sel].astype(np.float64)))


@pytest.mark.parametrize("kind,lam", [("l2combined", 0.05), ("l2", 0.0), ("l1", 0.0)])
def test_patchmatch_energy_vs_serial_reference(ctx, oracle, syn, kind, lam):
    W = H = 160
    img = syn.texture(W, H, 3, 8, 1)
    hole = syn.ellipse_holes(W, H, 6, 0.15, 1)
    ext, tgt, src, _ = oracle.scale_masks(hole, 9)
    dp = oracle.dist_params(kind, 9, 10000.0, lam)
    e_gpu, e_cpu = [], []
    for seed in range(1, 7):
        n_gpu, d_gpu = ctx.patchmatch(kind, img, src, img, tgt, None, seed=seed, lam=lam, with_dist=True)
        n_cpu = oracle.patchmatch(dp, img, src, img, tgt, None, seed=seed)
        sel = tgt.astype(bool)
        assert np.all(n_gpu[~sel] == -1)
        assert np.all(src[n_gpu[sel][:, 1], n_gpu[sel][:, 0]] == 1)       # entries are source points
        # stored distances are the reference's distances of the entries, bit for bit
        assert bitexact(d_gpu[sel], oracle.patch_distance_field(dp, img, img, n_gpu, tgt)[sel])
        e_gpu.append(energy(oracle, dp, img, n_gpu, tgt))
        e_cpu.append(energy(oracle, dp, img, n_cpu, tgt))
    assert_energy_within_1pct(f"{kind} 160x160", e_gpu, e_cpu)


def ratio_and_se(g, c):
    """Ratio of the mean energies and its standard error (delta method, independent runs)."""
    g, c = np.asarray(g, np.float64), np.asarray(c, np.float64)
    ratio = g.mean() / c.mean()
    se = ratio * np.sqrt(g.var(ddof=1) / len(g) / g.mean() ** 2 + c.var(ddof=1) / len(c) / c.mean() ** 2)
    return ratio, se


def assert_energy_within_1pct(label, g, c):
    """The bar: mean GPU energy at most 1 % above the reference's, resolved by the data
    (the ratio plus two standard errors stays within 1.01)."""
    ratio, se = ratio_and_se(g, c)
    print(f"{label}: energy GPU/serial reference = {ratio:.4f} (standard error {se:.4f}, {len(g)} and {len(c)} "
          f"seeds; seed spread GPU {np.std(g, ddof=1) / np.mean(g):.3f}, reference "
          f"{np.std(c, ddof=1) / np.mean(c):.3f})")
    assert ratio <= 1.01
    assert ratio + 2 * se <= 1.01


@pytest.mark.parametrize("kind,C,patch,lam,lam_d", [
    ("l2combined", 3, 7, 0.05, 0.0),      # candidate-pool sweep, other patch sides
    ("l2combined", 3, 11, 0.5, 0.0),
    ("l2combined", 1, 5, 0.05, 0.0),      # lockstep sweep: channel counts / patch sides
    ("l2", 2, 3, 0.0, 0.0),               # without a specialised pool kernel
    ("l1", 4, 7, 0.0, 0.0),
    ("l2combined", 3, 13, 0.05, 0.0),
])
def test_patchmatch_energy_more_kinds(ctx, oracle, syn, kind, C, patch, lam, lam_d):
    """Mean NNF energy over 12 seeds within 1 % of the serial reference's mean over the
    same seeds (oracle restatement of patch_match.cpp:304-512, pinned bit-exact against
    the reference); stored distances bit-exact."""
    W, H = 144, 120
    img = np.ascontiguousarray(syn.texture(W, H, 3, 6, 2)[..., :C] if C < 4 else np.concatenate(
        [syn.texture(W, H, 3, 6, 2), syn.depth_planes(W, H, 5, 2)[..., None]], axis=2))
    hole = syn.ellipse_holes(W, H, 5, 0.15, 2)
    ext, tgt, src, _ = oracle.scale_masks(hole, patch)
    dp = oracle.dist_params(kind, patch, 10000.0, lam, lam_d)
    e_gpu, e_cpu = [], []
    sel = tgt.astype(bool)
    for seed in range(1, 13):
        n_gpu, d_gpu = ctx.patchmatch(kind, img, src, img, tgt, None, seed=seed, patch=patch, lam=lam, lam_d=lam_d,
                                      with_dist=True)
        n_cpu = oracle.patchmatch(dp, img, src, img, tgt, None, seed=seed)
        assert np.all(n_gpu[~sel] == -1)
        assert np.all(src[n_gpu[sel][:, 1], n_gpu[sel][:, 0]] == 1)
        assert bitexact(d_gpu[sel], oracle.patch_distance_field(dp, img, img, n_gpu, tgt)[sel])
        e_gpu.append(energy(oracle, dp, img, n_gpu, tgt))
        e_cpu.append(energy(oracle, dp, img, n_cpu, tgt))
    assert_energy_within_1pct(f"{kind} C={C} P={patch} lambda {lam}/{lam_d}", e_gpu, e_cpu)


@pytest.mark.parametrize("key,lam,lam_d", [("rgbd144_l10", 1.0, 0.0), ("rgbd144_l0306", 0.3, 0.6)])
def test_patchmatch_energy_rgbd_vs_reference(ctx, oracle, key, lam, lam_d):
    """L2RGBD against the reference's own serial PatchMatch (built in place, run under
    srand(seed) for seeds 1-2048: tests/golden/make_pm_energy.py).  One call from a
    random field is a lottery on this image -- whole regions either find a good match
    early and propagate it, or do not; the reference's own energy varies by 10-13 %
    from seed to seed -- so the mean over 2048 seeds is compared.  Energies are sums of
    the GPU's stored distances, which equal the reference's distances bit for bit
    (checked on the first seeds)."""
    g = golden("pm_energy")
    img, src, tgt = g["rgbd144_img"], g["rgbd144_src"], g["rgbd144_tgt"]
    ref = g[key]
    sel = tgt.astype(bool)
    dp = oracle.dist_params("l2rgbd", 9, 10000.0, lam, lam_d)
    e_gpu = []
    for seed in range(1, len(ref) + 1):
        n, d = ctx.patchmatch("l2rgbd", img, src, img, tgt, None, seed=seed, patch=9, lam=lam, lam_d=lam_d,
                              with_dist=True)
        if seed <= 4:
            assert np.all(src[n[sel][:, 1], n[sel][:, 0]] == 1)
            assert bitexact(d[sel], oracle.patch_distance_field(dp, img, img, n, tgt)[sel])
        e_gpu.append(float(np.sum(d[sel].astype(np.float64))))
    assert_energy_within_1pct(f"l2rgbd 144x120 lambda {lam}/{lam_d}", e_gpu, ref)


def test_patchmatch_energy_1024(ctx, oracle, syn):
    """L2Combined on the 1024x1024 nlpoisson workload (lambda 0.05, patch 9), 16 seeds,
    against the reference's own serial PatchMatch (tests/golden/make_pm_energy.py)."""
    g = golden("pm_energy")
    ref = g["l2c1024_l005"]
    w = syn.config_1024(1)
    _, tgt, src, _ = oracle.scale_masks(w.mask, 9)
    sel = tgt.astype(bool)
    dp = oracle.dist_params("l2combined", 9, 10000.0, 0.05)
    e_gpu = []
    for seed in range(1, len(ref) + 1):
        n, d = ctx.patchmatch("l2combined", w.image, src, w.image, tgt, None, seed=seed, patch=9, lam=0.05,
                              with_dist=True)
        if seed == 1:
            assert bitexact(d[sel], oracle.patch_distance_field(dp, w.image, w.image, n, tgt)[sel])
        e_gpu.append(float(np.sum(d[sel].astype(np.float64))))
    assert_energy_within_1pct("l2combined 1024x1024 lambda 0.05", e_gpu, ref)


@pytest.mark.parametrize("row_scan", [0, 1])
@pytest.mark.parametrize("kind,C,patch", [("l2combined", 3, 9), ("l2combined", 3, 11), ("l2", 3, 7),
                                           ("l1", 3, 9), ("l2rgbd", 4, 9)])
def test_pool_tile_arrangements_give_identical_fields(ctx, syn, nli, kind, C, patch, row_scan):
    """The autotuner chooses among the tile arrangements (1 or 2 target tiles per
    workgroup, or one tile with two waves per tile row where the row scan is off); the
    fields and distances must not depend on the choice (with the row scan automatic, and
    forced on for every distance kind: forced arrangement 3 then runs as 1)."""
    W, H = 150, 118
    img = syn.texture(W, H, 3, 5, 9)
    if C == 4:
        img = np.concatenate([img, syn.depth_planes(W, H, 5, 9)[..., None]], axis=2)
    hole = syn.ellipse_holes(W, H, 4, 0.15, 9)
    ext, tgt, src = ctx.scale_masks(hole, patch, True)
    res = []
    assert ctx.lib.nli_pm_row_scan(row_scan) == 0
    try:
        for tiles in (1, 2, 3, 0):
            assert ctx.lib.nli_pm_tiles_per_block(tiles) == 0
            n, d = ctx.patchmatch(kind, img, src, img, tgt, None, seed=4, patch=patch, lam=0.3, lam_d=0.6,
                                  with_dist=True)
            n2, d2 = ctx.patchmatch(kind, img, src, img, tgt, n, seed=5, patch=patch, lam=0.3, lam_d=0.6,
                                    with_dist=True)
            res.append((n, d, n2, d2))
    finally:
        ctx.lib.nli_pm_tiles_per_block(0)
        ctx.lib.nli_pm_row_scan(0)
    for r in res[1:]:
        for u, v in zip(res[0], r):
            assert np.array_equal(u, v)


def test_patchmatch_given_field_only_improves(ctx, oracle, syn):
    W = H = 96
    img = syn.texture(W, H, 3, 5, 4)
    hole = syn.ellipse_holes(W, H, 3, 0.15, 4)
    ext, tgt, src, _ = oracle.scale_masks(hole, 7)
    dp = oracle.dist_params("l2combined", 7, 10000.0, 0.05)
    init = oracle.patchmatch(dp, img, src, img, tgt, None, seed=5, iters=2)
    d0 = oracle.patch_distance_field(dp, img, img, init, tgt)
    n, d = ctx.patchmatch("l2combined", img, src, img, tgt, init, seed=3, patch=7, lam=0.05, with_dist=True)
    sel = tgt.astype(bool)
    assert np.all(d[sel] <= d0[sel])
    # zero sweeps keep every valid initial entry and its distance
    n0, dd0 = ctx.patchmatch("l2combined", img, src, img, tgt, init, seed=3, patch=7, lam=0.05, iters=0,
                             with_dist=True)
    assert np.array_equal(n0, init) and bitexact(dd0[sel], d0[sel])


@pytest.mark.parametrize("W,H", [(24, 40000), (65600, 24)])
def test_patchmatch_tall_and_wide_images(ctx, oracle, syn, W, H):
    """The candidate-pool sweep packs source coordinates into 16 bits each and holds at
    most 16 random-search windows.  Rows 32768 and beyond (24 x 40000: pool sweep) must
    unpack unsigned, and an image 65600 pixels wide (17 windows, a coordinate beyond 16
    bits) must take the general sweep: every entry a source point, distances bit-exact."""
    img = syn.texture(W, H, 3, 3, 11)
    hole = np.zeros((H, W), np.uint8)
    if H > W:
        hole[33000:33012, 8:16] = 1
        hole[100:104, 10:14] = 1
    else:
        hole[8:16, 40000:40040] = 1
        hole[10:13, 100:104] = 1
    ext, tgt, src, _ = oracle.scale_masks(hole, 9)
    dp = oracle.dist_params("l2combined", 9, 10000.0, 0.05)
    n, d = ctx.patchmatch("l2combined", img, src, img, tgt, None, seed=2, lam=0.05, iters=4, with_dist=True)
    sel = tgt.astype(bool)
    assert np.all(n[~sel] == -1)
    xs, ys = n[sel][:, 0], n[sel][:, 1]
    assert np.all((xs >= 0) & (xs < W) & (ys >= 0) & (ys < H))
    assert np.all(src[ys, xs] == 1)
    assert bitexact(d[sel], oracle.patch_distance_field(dp, img, img, n, tgt)[sel])


def test_patchmatch_rgbd_and_invalid_init(ctx, oracle, syn):
    W, H = 120, 90
    img = np.concatenate([syn.texture(W, H, 3, 3, 7), syn.depth_planes(W, H, 5, 7)[..., None]], axis=2)
    hole = syn.stripe_holes(W, H, 2, 8, 7)
    ext, tgt, src, _ = oracle.scale_masks(hole, 9)
    bad = np.full((H, W, 2), -1, np.int32)         # every entry invalid: random re-initialisation
    n = ctx.patchmatch("l2rgbd", img, src, img, tgt, bad, seed=1, lam=1.0, lam_d=0.0)
    sel = tgt.astype(bool)
    assert np.all(src[n[sel][:, 1], n[sel][:, 0]] == 1)


# ---------------------------------------------------------------- updates
def field_case(oracle, syn, W=96, H=80, P=9, seed=1, kind="l2combined", C=3):
    img = syn.texture(W, H, C, 3, seed)
    hole = syn.ellipse_holes(W, H, 4, 0.15, seed, border=6)
    ext, tgt, src, _ = oracle.scale_masks(hole, P)
    nnf = oracle.patchmatch(oracle.dist_params(kind, P, 10000.0, 0.05), img, src, img, tgt, None, seed=seed)
    conf = oracle.confidence(hole, 5.0, 0.1)
    return img, hole, ext, nnf, conf


def test_nlmeans_bitexact(ctx, oracle, syn):
    img, hole, ext, nnf, conf = field_case(oracle, syn)
    orig = img.copy()
    img = img + np.float32(0.01)                       # image differs from orig inside the scale loop
    a, ta = ctx.update("nlmeans", img, orig, hole, ext, nnf, conf)
    b, tb = oracle.update("nlmeans", img, orig, hole, ext, nnf, conf)
    assert bitexact(a, b)
    assert abs(ta - tb) <= 1e-12 * max(1.0, tb)
    a, _ = ctx.update("nlmeans", img, orig, hole, ext, nnf, None)     # empty confidence
    b, _ = oracle.update("nlmeans", img, orig, hole, ext, nnf, None)
    assert bitexact(a, b)


@pytest.mark.parametrize("seed,P,C", [(1, 9, 3), (2, 7, 3), (3, 9, 1), (4, 11, 3), (5, 7, 2)])
def test_nlmedians_real_fields_bitexact(ctx, oracle, syn, seed, P, C):
    """NL-medians on PatchMatch fields (coherent: many contributors share a colour), bit for
    bit: equal colours among different ones are ordered as the reference's std::sort
    orders them (the weights leave the float total in that order)."""
    img, hole, ext, nnf, conf = field_case(oracle, syn, P=P, seed=seed, kind="l1", C=C)
    a, ta = ctx.update("nlmedians", img, img, hole, ext, nnf, conf, patch=P)
    b, tb = oracle.update("nlmedians", img, img, hole, ext, nnf, conf, patch=P)
    assert bitexact(a, b)
    assert ta == tb


def test_nlmedians_signed_zeros_and_ties(ctx, oracle, syn):
    """Colours quantised to a few levels, zeros stored as -0.0 or +0.0 at random: the
    reference's comparison (a.color < b.color) holds -0.0 and +0.0 equal, so they tie and
    std::sort's order decides which one the median lands on; the result must carry the
    same sign bit as the oracle's, bit for bit."""
    img, hole, ext, nnf, conf = field_case(oracle, syn, kind="l1", seed=6)
    q = (np.round((img - np.float32(0.5)) * np.float32(4)) / np.float32(4)).astype(np.float32)
    rng = np.random.default_rng(6)
    q[(q == 0) & (rng.random(q.shape) < 0.5)] = np.float32(-0.0)
    zeros = q[q == 0]
    assert np.signbit(zeros).any() and (~np.signbit(zeros)).any()
    a, ta = ctx.update("nlmedians", q, q, hole, ext, nnf, conf)
    b, tb = oracle.update("nlmedians", q, q, hole, ext, nnf, conf)
    assert bitexact(a, b)
    assert ta == tb
    sel = hole.astype(bool)
    assert np.signbit(a[sel][a[sel] == 0]).any()   # some medians land on a -0.0


def test_nlmedians(ctx, oracle, syn):
    img, hole, ext, nnf, conf = field_case(oracle, syn, kind="l1")
    a, ta = ctx.update("nlmedians", img, img, hole, ext, nnf, conf)
    b, tb = oracle.update("nlmedians", img, img, hole, ext, nnf, conf)
    assert bitexact(a, b)
    assert ta == tb
    rng = np.random.default_rng(3)                      # a random field has no value ties
    sp = np.argwhere(oracle.scale_masks(hole, 9)[2])
    rnd = np.full_like(nnf, -1)
    t = np.argwhere(oracle.scale_masks(hole, 9)[1])
    pick = sp[rng.integers(0, len(sp), len(t))]
    rnd[t[:, 0], t[:, 1], 0] = pick[:, 1]
    rnd[t[:, 0], t[:, 1], 1] = pick[:, 0]
    a, _ = ctx.update("nlmedians", img, img, hole, ext, rnd, conf, patch=7)
    b, _ = oracle.update("nlmedians", img, img, hole, ext, rnd, conf, patch=7)
    assert bitexact(a, b)


@pytest.mark.parametrize("C,P", [(1, 3), (4, 5), (2, 7), (3, 11), (1, 13), (3, 15)])
def test_nl_updates_patch_sides(ctx, oracle, syn, C, P):
    """NL-means (bit-exact) and NL-medians on a random field (no value ties: bit-exact)
    for every window-size specialisation of the 16-lanes-per-pixel kernels."""
    img, hole, ext, nnf, conf = field_case(oracle, syn, W=80, H=72, P=P, seed=3, C=C)
    orig = img.copy()
    img = img + np.float32(0.01)
    a, ta = ctx.update("nlmeans", img, orig, hole, ext, nnf, conf, patch=P)
    b, tb = oracle.update("nlmeans", img, orig, hole, ext, nnf, conf, patch=P)
    assert bitexact(a, b)
    assert abs(ta - tb) <= 1e-12 * max(1.0, tb)
    rng = np.random.default_rng(P)
    _, tgt, src, _ = oracle.scale_masks(hole, P)
    sp = np.argwhere(src)
    t = np.argwhere(tgt)
    rnd = np.full_like(nnf, -1)
    pick = sp[rng.integers(0, len(sp), len(t))]
    rnd[t[:, 0], t[:, 1], 0] = pick[:, 1]
    rnd[t[:, 0], t[:, 1], 1] = pick[:, 0]
    a, ta = ctx.update("nlmedians", img, img, hole, ext, rnd, conf, patch=P)
    b, tb = oracle.update("nlmedians", img, img, hole, ext, rnd, conf, patch=P)
    assert bitexact(a, b)
    assert abs(ta - tb) <= 1e-12 * max(1.0, tb)


@pytest.mark.parametrize("lam,tol,limit", [(0.05, 1e-4, 100), (0.5, 1e-6, 1000), (1.0, 1e-6, 50), (0.0, 1e-6, 300)])
def test_nlpoisson(ctx, oracle, syn, lam, tol, limit):
    img, hole, ext, nnf, conf = field_case(oracle, syn)
    orig = img.copy()
    a, ta, st = ctx.update("nlpoisson", img, orig, hole, ext, nnf, conf, lam=lam, cg_tol=tol, cg_limit=limit,
                           stats=True)
    b, tb, so = oracle.update("nlpoisson", img, orig, hole, ext, nnf, conf, lam=lam, cg_tol=tol, cg_limit=limit,
                              stats=True)
    err = np.max(np.abs(a.astype(np.float64) - b))
    print(f"lambda {lam}: CG iterations GPU {list(st.cg_iterations[:3])} oracle {list(so.cg_iterations[:3])}, "
          f"max |diff| {err:.2e}")
    assert err <= 1e-5
    for c in range(3):
        assert abs(st.cg_iterations[c] - so.cg_iterations[c]) <= 1
        assert abs(st.cg_nr0[c] - so.cg_nr0[c]) <= 1e-9 * so.cg_nr0[c]
    assert abs(ta - tb) <= 1e-4 * max(tb, 1e-12)


@pytest.mark.parametrize("C,P", [(1, 5), (2, 13), (4, 7), (3, 11), (3, 3), (1, 15), (4, 9)])
def test_nlpoisson_channels_and_patches(ctx, oracle, syn, C, P):
    """The coefficient kernel's channel-count and patch-row specialisations (and its
    generic form for P = 3, 13, 15) against the oracle, FP64 CG."""
    img, hole, ext, nnf, conf = field_case(oracle, syn, W=88, H=76, P=P, seed=2, C=C)
    orig = img.copy()
    img = img + np.float32(0.02)
    a, ta, st = ctx.update("nlpoisson", img, orig, hole, ext, nnf, conf, patch=P, lam=0.05, cg_tol=1e-6,
                           cg_limit=1000, stats=True)
    b, tb, so = oracle.update("nlpoisson", img, orig, hole, ext, nnf, conf, patch=P, lam=0.05, cg_tol=1e-6,
                              cg_limit=1000, stats=True)
    err = np.max(np.abs(a.astype(np.float64) - b))
    print(f"C={C} P={P}: CG iterations GPU {list(st.cg_iterations[:C])} oracle {list(so.cg_iterations[:C])}, "
          f"max |diff| {err:.2e}")
    assert err <= 1e-5
    for c in range(C):
        assert abs(st.cg_iterations[c] - so.cg_iterations[c]) <= 1
        assert abs(st.cg_nr0[c] - so.cg_nr0[c]) <= 1e-9 * so.cg_nr0[c]
    assert abs(ta - tb) <= 1e-4 * max(tb, 1e-12)


def test_nlpoisson_large_hole_bitexact_coefficients(ctx, oracle, syn):
    """Above 32 k hole points the update switches to the thread-per-point coefficient
    kernel; against the oracle at 37 k points (NNF from the GPU PatchMatch)."""
    W, H = 520, 460
    img = syn.texture(W, H, 3, 3, 6)
    hole = syn.ellipse_holes(W, H, 5, 0.16, 6, border=6)
    assert hole.sum() >= 32768
    ext, tgt, src, _ = oracle.scale_masks(hole, 9)
    nnf = ctx.patchmatch("l2combined", img, src, img, tgt, None, seed=6, lam=0.05)
    conf = oracle.confidence(hole, 5.0, 0.1)
    orig = img.copy()
    img = img + np.float32(0.02)
    a, ta, st = ctx.update("nlpoisson", img, orig, hole, ext, nnf, conf, lam=0.05, cg_tol=1e-6, cg_limit=1000,
                           stats=True)
    b, tb, so = oracle.update("nlpoisson", img, orig, hole, ext, nnf, conf, lam=0.05, cg_tol=1e-6, cg_limit=1000,
                              stats=True)
    err = np.max(np.abs(a.astype(np.float64) - b))
    print(f"{int(hole.sum())} hole points: CG iterations GPU {list(st.cg_iterations[:3])} oracle "
          f"{list(so.cg_iterations[:3])}, max |diff| {err:.2e}")
    assert err <= 1e-5
    for c in range(3):
        assert abs(st.cg_iterations[c] - so.cg_iterations[c]) <= 1
        assert abs(st.cg_nr0[c] - so.cg_nr0[c]) <= 1e-9 * so.cg_nr0[c]


def test_nlpoisson_golden(ctx):
    g = golden("field64x48")
    a, ta = ctx.update("nlpoisson", g["img"], g["img"], g["hole"], g["ext"], g["pm_l2c_seed42"], g["conf"],
                       lam=0.05, cg_tol=1e-4, cg_limit=100)
    assert np.max(np.abs(a - g["upd_nlpoisson"])) <= 1e-5
    a, _ = ctx.update("nlmeans", g["img"], g["img"], g["hole"], g["ext"], g["pm_l2c_seed42"], g["conf"])
    assert bitexact(a, g["upd_nlmeans"])


def test_nldepth(ctx, oracle, syn):
    W, H = 96, 80
    img = np.concatenate([syn.texture(W, H, 3, 3, 5), syn.depth_planes(W, H, 5, 5)[..., None]], axis=2)
    hole = syn.stripe_holes(W, H, 2, 6, 5)
    ext, tgt, src, _ = oracle.scale_masks(hole, 9)
    nnf = oracle.patchmatch(oracle.dist_params("l2rgbd", 9, 10000.0, 1.0, 0.0), img, src, img, tgt, None, seed=2)
    conf = oracle.confidence(hole, 5.0, 0.1)
    a, ta = ctx.update("nldepth", img, img, hole, ext, nnf, conf, lam=1.0, lam_d=0.0)
    b, tb = oracle.update("nldepth", img, img, hole, ext, nnf, conf, lam=1.0, lam_d=0.0)
    assert np.max(np.abs(a[..., :3] - b[..., :3])) <= 1e-5
    assert np.max(np.abs(a[..., 3] - b[..., 3])) <= 1e-5 * np.max(np.abs(b[..., 3]))


def test_mixed_precision_cg(ctx, oracle, syn):
    """FP32 CG with FP64 dots and periodic FP64 residual replacement, against the FP64 oracle."""
    img, hole, ext, nnf, conf = field_case(oracle, syn, W=128, H=112)
    b, tb = oracle.update("nlpoisson", img, img, hole, ext, nnf, conf, lam=0.05, cg_tol=1e-8, cg_limit=400)
    a, ta, st = ctx.update("nlpoisson", img, img, hole, ext, nnf, conf, lam=0.05, cg_tol=1e-8, cg_limit=400,
                           cg_precision=1, cg_refine_every=8, stats=True)
    err = np.max(np.abs(a.astype(np.float64) - b))
    print(f"mixed precision: max |x - x_fp64| = {err:.2e}, iterations {list(st.cg_iterations[:3])}")
    assert err <= 2e-4


# ---------------------------------------------------------------- whole process()
def psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return 10 * np.log10(1.0 / mse)


def assert_psnr_within_0p1db(label, g, r):
    """The bar: the GPU's mean hole PSNR at most 0.1 dB below the reference's, resolved
    by the data (the lower end of the difference's two-standard-error interval stays
    above -0.1 dB, and that interval is at most 0.1 dB wide on either side)."""
    g, r = np.asarray(g, np.float64), np.asarray(r, np.float64)
    diff = g.mean() - r.mean()
    se = np.sqrt(g.var(ddof=1) / len(g) + r.var(ddof=1) / len(r))
    print(f"{label}: PSNR GPU {g.mean():.3f} dB ({len(g)} seeds, sd {g.std(ddof=1):.3f}), reference "
          f"{r.mean():.3f} dB ({len(r)} seeds, sd {r.std(ddof=1):.3f}); difference {diff:+.3f} dB, "
          f"standard error {se:.3f} dB")
    assert 2 * se <= 0.1
    assert diff - 2 * se >= -0.1


# GPU seeds per case: enough that two standard errors of the difference stay within
# 0.1 dB given the seed-to-seed spreads (reference: 0.25, 0.41, 0.67 dB)
PROCESS_SEEDS = {"nlmeans256": 40, "nlmedians192": 112, "nlpoisson128": 400, "rgb_nlpoisson128": 400,
                 "small256": 48}


def process_case(name):
    import importlib
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    return importlib.import_module("make_process_psnr").case_inputs(name)


@pytest.mark.parametrize("name", ["nlmeans256", "nlmedians192", "nlpoisson128", "small256"])
def test_process_psnr_vs_reference_run(ctx, nli, name):
    """Whole multiscale runs: the GPU reaches the reconstruction quality (hole PSNR) of
    the reference's own process() (serial build, srand(seed), seeds 1-512 or 1-1024:
    tests/golden/make_process_psnr.py) within 0.1 dB, on means over many seeds
    (PatchMatch is randomised differently; one run's PSNR varies by 0.25-0.7 dB from
    seed to seed on these images).  small256 is the 256x256 nlmeans workload (centred
    64x64 hole, patch 9, 10 iterations)."""
    ref = golden("process_psnr")[name]
    img, hole, kw, peak = process_case(name)
    sel = hole.astype(bool)
    g = []
    for seed in range(1, PROCESS_SEEDS[name] + 1):
        out, st = ctx.inpaint(img, hole, nli.Inpainting(seed=seed, **kw))
        if seed == 1:
            assert np.array_equal(out[~sel], img[~sel])
        g.append(psnr(out[sel], img[sel]))
    assert_psnr_within_0p1db(name, g, ref)


def test_process_psnr_1024_vs_reference_runs(ctx, nli, syn):
    """The 1024x1024 nlpoisson workload (images synthetic.config_1024(seed=1+k), the bench's
    parameters with FP64 CG): whole GPU runs against whole runs of the reference's own
    process() (serial build, srand(seed): tests/golden/make_process_psnr_1024.py, about 35
    minutes a run).  Per image the mean hole PSNR over 4 GPU seeds is compared with the
    mean over the reference's seeds (6 seeds an image, 47 runs).  One run varies by
    0.3-0.4 dB from seed to seed at this size, which leaves a standard error near 0.06 dB
    over the images: the bar is one-sided here (the GPU is not worse: the mean difference
    must not lie more than two standard errors below -0.1 dB); it is resolved in both
    directions at the smaller sizes above."""
    g = golden("process_psnr_1024")
    images, ref = g["images"], g["psnr"]   # psnr[image, seed]; nan where not run
    n_ref = np.sum(np.isfinite(ref), axis=1)
    keep = n_ref > 0
    assert keep.sum() >= 4, "too few reference runs in the fixture"
    kw = dict(method="nlpoisson", patch=9, sigma=10000.0, lam=0.05, iters=200, tol=0.1, scales=11, coarse=0.1,
              init="poisson", cg_tol=1e-4, cg_limit=100)
    diffs = []
    for k, r in zip(images[keep], ref[keep]):
        w = syn.config_1024(seed=1 + int(k))
        sel = w.mask.astype(bool)
        gp = []
        for seed in range(1, 5):
            out, _ = ctx.inpaint(w.image, w.mask, nli.Inpainting(seed=seed, **kw))
            gp.append(psnr(out[sel], w.truth[sel]))
        diffs.append(np.mean(gp) - np.nanmean(r))
    d = np.array(diffs)
    se = d.std(ddof=1) / np.sqrt(len(d))
    print(f"1024x1024: GPU - reference hole PSNR per image {np.round(d, 3)}; mean {d.mean():+.3f} dB, "
          f"SE {se:.3f} dB over {len(d)} images ({int(n_ref.sum())} reference runs)")
    assert d.mean() - 2 * se >= -0.1


def test_process_psnr_rgb_pipeline_vs_reference_run(ctx, nli):
    """The reference program's colour flow (rgb_to_lab, process, lab_to_rgb) on a 0-255
    RGB image, nlpoisson: nli_inpaint_rgb against the reference's own conversions and
    process() (compiled in place), hole PSNR on the 0-255 scale."""
    ref = golden("process_psnr")["rgb_nlpoisson128"]
    rgb, hole, kw, peak = process_case("rgb_nlpoisson128")
    sel = hole.astype(bool)
    g = []
    for seed in range(1, PROCESS_SEEDS["rgb_nlpoisson128"] + 1):
        out, _, _ = ctx.inpaint_rgb(rgb, hole, nli.Inpainting(seed=seed, **kw))
        mse = float(np.mean((out[sel].astype(np.float64) - rgb[sel]) ** 2))
        g.append(10 * np.log10(255.0 ** 2 / mse))
    assert_psnr_within_0p1db("rgb_nlpoisson128", g, ref)


def test_process_rgbd_vs_reference_runs(ctx, nli):
    """nldepth with the 1920x1080 RGB-D workload's weights (lambda_rgb 1, lambda_d 0) on a
    128x96 RGB + depth image with stripe holes: whole GPU runs against whole runs of the
    reference's own process() (serial build, srand(seed), seeds 1-256:
    tests/golden/make_process_psnr.py).  The colour channels' hole PSNR and the depth
    channel's (20 log10(10 m / RMSE)) must each reach the reference's within 0.1 dB on
    means over many seeds (one run varies by 0.2 dB in colour and 0.3 dB in depth from
    seed to seed).  Known pixels are unchanged."""
    g = golden("process_psnr")
    ref_c, ref_d = g["rgbd_nldepth128"], g["rgbd_nldepth128_depth"]
    img, hole, kw, _ = process_case("rgbd_nldepth128")
    sel = hole.astype(bool)
    gc, gd = [], []
    for seed in range(1, RGBD_PROCESS_SEEDS + 1):
        out, st = ctx.inpaint(img, hole, nli.Inpainting(seed=seed, **kw))
        if seed == 1:
            assert np.array_equal(out[~sel], img[~sel])
        gc.append(psnr(out[sel][:, :3], img[sel][:, :3]))
        gd.append(psnr(out[sel][:, 3] / 10.0, img[sel][:, 3] / 10.0))
    assert_psnr_within_0p1db("rgbd_nldepth128 colour", gc, ref_c)
    assert_psnr_within_0p1db("rgbd_nldepth128 depth", gd, ref_d)


# GPU seeds of the RGB-D case: two standard errors of the depth difference stay near
# 0.065 dB with the reference's 256 seeds (seed-to-seed spread 0.32 dB)
RGBD_PROCESS_SEEDS = 160


def test_batch_equals_single_images(ctx, syn, nli):
    imgs, masks = [], []
    for k, (W, H) in enumerate([(96, 80), (130, 70), (64, 64), (110, 120), (80, 96)]):
        imgs.append(syn.texture(W, H, 3, 3, k))
        masks.append(syn.ellipse_holes(W, H, 3, 0.15, k))
    p = nli.Inpainting(method="nlpoisson", patch=7, iters=5, scales=3, coarse=0.4, cg_tol=1e-4, cg_limit=50)
    outs, stats = ctx.inpaint_batch(imgs, masks, p, workers=3)
    for im, m, o, st in zip(imgs, masks, outs, stats):
        single, st1 = ctx.inpaint(im, m, p)
        assert np.array_equal(o, single)
        assert st.iterations == st1.iterations


@pytest.mark.parametrize("method,precision", [("nlpoisson", 1), ("nlpoisson", 0), ("nldepth", 1)])
def test_lockstep_batch_equals_single_images(ctx, syn, nli, method, precision):
    """The lockstep groups (one launch per kernel for all images of a group, the stopping
    rule on the device): every image equals its single run bit for bit, with the same
    iterations per scale and CG iterations, for uneven sizes and groups of 1 to 4 images;
    and equals the concurrent per-image drivers."""
    C = 4 if method == "nldepth" else 3
    imgs, masks = [], []
    for k, (W, H) in enumerate([(96, 80), (130, 70), (64, 64), (110, 120), (80, 96), (150, 101), (72, 90)]):
        im = syn.texture(W, H, 3, 3, k)
        if C == 4:
            im = np.concatenate([im, syn.depth_planes(W, H, 5, k)[..., None]], axis=2)
        imgs.append(im)
        masks.append(syn.stripe_holes(W, H, 2, 6, k) if C == 4 else syn.ellipse_holes(W, H, 3, 0.15, k))
    kw = dict(method=method, patch=9, iters=12, scales=3, coarse=0.5, cg_tol=1e-4, cg_limit=60,
              cg_precision=precision, cg_refine_every=8 if precision else 0, seed=3)
    if C == 4:
        kw.update(lambda_rgb=1.0, lambda_d=0.0)
    p = nli.Inpainting(**kw)
    singles = [ctx.inpaint(im, m, p) for im, m in zip(imgs, masks)]
    try:
        for mode, workers in [(1, 2), (1, 7), (2, 3)]:
            assert ctx.lib.nli_batch_mode(mode) == 0
            outs, stats = ctx.inpaint_batch(imgs, masks, p, workers=workers)
            for k, (o, st) in enumerate(zip(outs, stats)):
                single, st1 = singles[k]
                assert np.array_equal(o, single), (mode, workers, k)
                assert st.iterations == st1.iterations, (mode, workers, k)
                assert st.cg_iterations_total == st1.cg_iterations_total, (mode, workers, k)
    finally:
        ctx.lib.nli_batch_mode(0)


def test_uneven_mixed_batch_equals_single_images(ctx, syn, nli):
    """Per-image parameters (the uneven batch: mixed sizes, nlmeans / nlmedians, patch
    7 or 9): each image equals its own single run, whatever the worker and order."""
    ws = syn.config_uneven(count=8, seed=5, min_side=72, max_side=200)
    params = []
    for w in ws:
        q = dict(w.params)
        q["iters"] = 6
        q["scales"] = 4
        q["coarse"] = max(0.35, 40.0 / min(w.image.shape[:2]))
        params.append(nli.Inpainting(**q))
    outs, stats = ctx.inpaint_batch([w.image for w in ws], [w.mask for w in ws], params, workers=3)
    for w, q, o, st in zip(ws, params, outs, stats):
        single, st1 = ctx.inpaint(w.image, w.mask, q)
        assert np.array_equal(o, single), w.name
        assert st.iterations == st1.iterations
        sel = w.mask.astype(bool)
        assert np.array_equal(o[~sel], w.image[~sel])


def test_uneven_workload_full_size(ctx, oracle, syn, nli):
    """The uneven batch at full size: 64 images of sides 128-1024, nlmeans / nlmedians,
    patch 7 or 9, 8 scales, per-image parameters, on concurrent drivers.  Every image:
    known pixels unchanged, finite values, the same iterations and output as its own
    single run for a few images.  On the three smallest images of each method, the hole PSNR against the
    oracle's run of the reference algorithm (OpenMP PatchMatch) on the same input: one
    run each, so the bar is the run-to-run spread (1.5 dB per image, 0.5 dB on the
    mean); the 0.1 dB bar is resolved on means over many seeds in
    test_process_psnr_vs_reference_run."""
    import oracle as O
    ws = syn.config_uneven(64)
    params = [nli.Inpainting(**w.params) for w in ws]
    outs, stats = ctx.inpaint_batch([w.image for w in ws], [w.mask for w in ws], params, workers=8)
    sizes = [w.image.shape[0] * w.image.shape[1] for w in ws]
    for w, o in zip(ws, outs):
        sel = w.mask.astype(bool)
        assert np.all(np.isfinite(o)), w.name
        assert np.array_equal(o[~sel], w.image[~sel]), w.name
    for k in np.argsort(sizes)[::16]:
        single, st1 = ctx.inpaint(ws[k].image, ws[k].mask, params[k])
        assert np.array_equal(outs[k], single) and stats[k].iterations == st1.iterations, ws[k].name
    g, o = [], []
    order = list(np.argsort(sizes))
    subset = [k for k in order if ws[k].params["method"] == "nlmeans"][:3] + \
        [k for k in order if ws[k].params["method"] == "nlmedians"][:3]
    for k in subset:
        w = ws[k]
        sel = w.mask.astype(bool)
        q = {key: v for key, v in w.params.items() if key != "seed"}
        ref, _ = oracle.process(w.image, w.mask, O.InpaintParams(seed=w.params["seed"], pm_threads=16, **q))
        g.append(psnr(outs[k][sel], w.truth[sel]))
        o.append(psnr(ref[sel], w.truth[sel]))
        print(f"{w.name}: hole PSNR GPU {g[-1]:.3f} dB, reference algorithm {o[-1]:.3f} dB")
        assert g[-1] >= o[-1] - 1.5
    print(f"mean over the six: GPU {np.mean(g):.3f} dB, reference algorithm {np.mean(o):.3f} dB")
    assert np.mean(g) >= np.mean(o) - 0.5


def test_inpaint_errors(ctx, nli):
    img = np.full((40, 40, 3), 0.5, np.float32)
    with pytest.raises(ValueError, match="Mask should not be empty"):
        ctx.inpaint(img, np.zeros((40, 40), np.uint8), nli.Inpainting(scales=1))
    with pytest.raises(nli.EmptySourceError):
        ctx.inpaint(img, np.ones((40, 40), np.uint8), nli.Inpainting(patch=7, scales=1, coarse=1.0))
    with pytest.raises(ValueError, match="Channels do not match"):
        ctx.inpaint(img, np.pad(np.ones((10, 10), np.uint8), 15), nli.Inpainting(method="nldepth", scales=1))
