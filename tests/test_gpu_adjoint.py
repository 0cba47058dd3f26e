"""The NL-Poisson patch energy, its hand-written adjoint and its gradient on the GPU.

* the extraction / aggregation pair passes the dot-product test <E u, w> = <u, E^T w>
  on both sides (target gather, source scatter);
* the energy equals an FP64 brute-force evaluation of its definition;
* on hole pixels the gradient equals twice the residual of the reference's NL-Poisson
  system, built by the oracle's restatement of the reference coefficients
  (patch_non_local_poisson.cpp:139-356): this ties the energy to the reference;
* FP64 central finite differences of the energy agree with the gradient at 256 random
  pixels (hole, extended domain, source pixels, image border).  E is quadratic in each
  pixel, so the central difference is exact up to rounding;
* after the NL-Poisson update with the CG driven to 1e-8 relative residual, the
  gradient on the hole pixels vanishes (the update minimises the energy).
The last two also run at 2048x2048, patch 11, ~20 % ellipse holes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def setup_scene(ctx, syn, W, H, P, fraction, seed, holes=5):
    img = syn.texture(W, H, 3, 4, seed)
    hole = syn.ellipse_holes(W, H, holes, fraction, seed)
    img = np.where(hole[..., None].astype(bool), 0.5, img).astype(np.float32)
    ext, tgt, src = ctx.scale_masks(hole, P, True)
    nnf = ctx.patchmatch("l2combined", img, src, img, tgt, None, seed=seed, patch=P, lam=0.05)
    conf = ctx.confidence(hole)
    return img, hole, ext, nnf, conf


def brute_energy(img, ext, nnf, conf, w, P, lam):
    H, W, C = img.shape
    u = img.astype(np.float64)
    gx = np.zeros_like(u)
    gy = np.zeros_like(u)
    gx[:, :-1] = u[:, 1:] - u[:, :-1]
    gy[:-1, :] = u[1:, :] - u[:-1, :]
    r = P // 2
    e = 0.0
    lam1 = float(np.float32(1) - np.float32(lam))
    lam = float(np.float32(lam))
    for py, px in zip(*np.nonzero(ext)):
        fx, fy = nnf[py, px]
        for oy in range(-r, r + 1):
            for ox in range(-r, r + 1):
                tx, ty, sx, sy = px + ox, py + oy, fx + ox, fy + oy
                if not (0 <= tx < W and 0 <= ty < H and 0 <= sx < W and 0 <= sy < H):
                    continue
                wt = float(np.float32(conf[py, px]) * np.float32(w[oy + r, ox + r]))
                d0 = u[ty, tx] - u[sy, sx]
                dx = gx[ty, tx] - gx[sy, sx]
                dy = gy[ty, tx] - gy[sy, sx]
                e += wt * float(np.sum(lam * d0 * d0 + lam1 * (dx * dx + dy * dy)))
    return e


def test_extract_aggregate_adjoint(ctx):
    rng = np.random.default_rng(7)
    W, H, C, P = 48, 40, 3, 5
    ext = np.zeros((H, W), np.uint8)
    ext[8:30, 10:34] = 1
    ext[0:4, 40:48] = 1                                  # touches the image border
    nnf = np.stack([rng.integers(-3, W + 2, (H, W)), rng.integers(-3, H + 2, (H, W))], axis=-1).astype(np.int32)
    nnf[12, 12] = (-1, -1)                               # a margin point's empty entry
    u = rng.standard_normal((H, W, C))
    for side in (0, 1):
        pt = ctx.patch_extract(u, ext, nnf, patch=P, side=side)
        w = rng.standard_normal(pt.shape)
        back = ctx.adjoint_aggregate(w, ext, nnf, patch=P, side=side)
        lhs = float(np.sum(pt * w))
        rhs = float(np.sum(u * back))
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (side, lhs, rhs)
        # extraction by definition
        e0 = list(zip(*np.nonzero(ext)))[5]
        py, px = e0
        fx, fy = nnf[py, px]
        for k, (oy, ox) in enumerate([(oy, ox) for oy in range(-2, 3) for ox in range(-2, 3)]):
            tx, ty, sx, sy = px + ox, py + oy, fx + ox, fy + oy
            ok = 0 <= tx < W and 0 <= ty < H and 0 <= sx < W and 0 <= sy < H
            want = (u[ty, tx] if side == 0 else u[sy, sx]) if ok else 0.0
            assert np.array_equal(pt[5, k], want * np.ones(C))


def test_energy_matches_definition(ctx, syn, oracle):
    W, H, P, lam = 40, 32, 5, 0.3
    img, hole, ext, nnf, conf = setup_scene(ctx, syn, W, H, P, 0.1, 3, holes=2)
    w = oracle.gaussian_weights(P, P, 10000.0, 10000.0)
    e = ctx.patch_energy(img, ext, nnf, conf, patch=P, lam=lam)
    ref = brute_energy(img, ext, nnf, conf, w, P, lam)
    assert abs(e - ref) <= 1e-12 * abs(ref)


def residual_expected(oracle, img, hole, ext, nnf, conf, P, lam):
    """2 (a1 u - f1 + f2 - L_a2 u) from the oracle's restatement of the reference."""
    a1, a2, f1, f2 = oracle.poisson_coefficients(img, ext, nnf, conf, patch=P, lam=lam)
    H, W, C = img.shape
    out = np.zeros((H, W, C))
    for c in range(C):
        u = img[..., c].astype(np.float64)
        L = oracle.anisotropic_laplacian(u, a2, hole)
        out[..., c] = 2 * (a1 * u - f1[c] + f2[c] - L)
    return out


def test_gradient_is_twice_the_reference_residual(ctx, syn, oracle):
    W, H, P, lam = 96, 80, 7, 0.05
    img, hole, ext, nnf, conf = setup_scene(ctx, syn, W, H, P, 0.12, 4)
    g = ctx.energy_gradient(img, ext, nnf, conf, patch=P, lam=lam)
    want = residual_expected(oracle, img, hole, ext, nnf, conf, P, lam)
    sel = hole.astype(bool)
    err = np.abs(g[sel] - want[sel]).max()
    scale = np.abs(want[sel]).max()
    print(f"max |grad - 2 residual| = {err:.3e} (max |2 residual| = {scale:.3e})")
    # the oracle's F uses float forward differences (gradient.cpp), the energy exact ones
    assert err <= 1e-6 * scale


def sample_pixels(rng, hole, ext, nnf, n):
    H, W = hole.shape
    hy, hx = np.nonzero(hole)
    ey, ex = np.nonzero(ext & ~hole.astype(bool))
    k = n // 4
    pick = [(hx[i], hy[i]) for i in rng.integers(0, len(hx), k)]
    pick += [(ex[i], ey[i]) for i in rng.integers(0, len(ex), k)]
    # source pixels some centre copies from
    cy, cx = np.nonzero(ext)
    for i in rng.integers(0, len(cx), k):
        fx, fy = nnf[cy[i], cx[i]]
        pick.append((int(np.clip(fx + rng.integers(-2, 3), 0, W - 1)), int(np.clip(fy + rng.integers(-2, 3), 0, H - 1))))
    while len(pick) < n:                                 # anywhere, borders included
        pick.append((int(rng.choice([0, W - 1, rng.integers(0, W)])), int(rng.choice([0, H - 1, rng.integers(0, H)]))))
    return pick


def fd_check(ctx, img, hole, ext, nnf, conf, P, lam, n, seed):
    rng = np.random.default_rng(seed)
    g = ctx.energy_gradient(img, ext, nnf, conf, patch=P, lam=lam)
    delta = 1e-2
    worst = 0.0
    gmax = np.abs(g).max()
    for (x, y) in sample_pixels(rng, hole, ext, nnf, n):
        c = int(rng.integers(0, img.shape[2]))
        ep = ctx.patch_energy(img, ext, nnf, conf, patch=P, lam=lam, touch=(x, y, c), delta=delta)
        em = ctx.patch_energy(img, ext, nnf, conf, patch=P, lam=lam, touch=(x, y, c), delta=-delta)
        fd = (ep - em) / (2 * delta)
        worst = max(worst, abs(fd - g[y, x, c]) / max(abs(g[y, x, c]), 1e-3 * gmax))
    return worst


def test_gradient_vs_fp64_finite_differences(ctx, syn):
    W, H, P, lam = 160, 128, 9, 0.05
    img, hole, ext, nnf, conf = setup_scene(ctx, syn, W, H, P, 0.15, 5)
    worst = fd_check(ctx, img, hole, ext, nnf, conf, P, lam, 256, 1)
    print(f"worst relative |FD - gradient| over 256 pixels: {worst:.3e}")
    assert worst <= 1e-9


def stationarity(ctx, img, hole, ext, nnf, conf, P, lam, limit):
    sel = hole.astype(bool)
    g0 = ctx.energy_gradient(img, ext, nnf, conf, patch=P, lam=lam)
    out, _, st = ctx.update("nlpoisson", img, img, hole, ext, nnf, conf, patch=P, lam=lam, cg_tol=1e-16,
                            cg_limit=limit, stats=True)
    g1 = ctx.energy_gradient(out, ext, nnf, conf, patch=P, lam=lam)
    e0 = ctx.patch_energy(img, ext, nnf, conf, patch=P, lam=lam)
    e1 = ctx.patch_energy(out, ext, nnf, conf, patch=P, lam=lam)
    ratio = np.linalg.norm(g1[sel]) / np.linalg.norm(g0[sel])
    rel_res = [np.sqrt(st.cg_nr[c] / st.cg_nr0[c]) for c in range(3)]
    return ratio, e0, e1, rel_res, list(st.cg_iterations[:3])


def test_update_is_a_stationary_point(ctx, syn):
    W, H, P, lam = 160, 128, 9, 0.05
    img, hole, ext, nnf, conf = setup_scene(ctx, syn, W, H, P, 0.15, 6)
    ratio, e0, e1, rel, its = stationarity(ctx, img, hole, ext, nnf, conf, P, lam, 4000)
    print(f"|grad| after/before on the hole: {ratio:.3e}; energy {e0:.6e} -> {e1:.6e}; "
          f"relative residuals {rel}; CG iterations {its}")
    assert max(rel) <= 1e-8
    assert e1 < e0
    assert ratio <= 1e-5


def test_config_2048_patch11(ctx, syn):
    W = H = 2048
    P, lam = 11, 0.05
    img, hole, ext, nnf, conf = setup_scene(ctx, syn, W, H, P, 0.20, 11, holes=6)
    worst = fd_check(ctx, img, hole, ext, nnf, conf, P, lam, 256, 2)
    ratio, e0, e1, rel, its = stationarity(ctx, img, hole, ext, nnf, conf, P, lam, 20000)
    print(f"2048^2 P=11: worst FD error {worst:.3e}; |grad| after/before {ratio:.3e}; relative residuals {rel}; "
          f"CG iterations {its}; energy {e0:.6e} -> {e1:.6e}")
    assert worst <= 1e-9
    assert max(rel) <= 1e-8
    assert ratio <= 1e-5
