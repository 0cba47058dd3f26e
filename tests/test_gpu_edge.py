"""Whole inpaintings on edge-case inputs (GPU): tiny and odd-sized images, holes on the
border and in corners, a one-pixel hole, 1, 2 and 4 channels, patch sides 3 to 15.

The results are checked through properties the reference's algorithm guarantees for any
NNF, so they hold whatever the random search finds:
* pixels outside the hole are returned unchanged;
* NL-means writes weighted averages of known samples and NL-medians writes samples,
  so every hole value lies within the channel's range of known values (the pyramid's
  coarse levels also average known samples, and the coarsest Poisson fill interpolates
  them);
* NL-Poisson and NL-Depth return finite values;
and, where the oracle runs in a second or two, the hole PSNR is compared with the
reference algorithm's on the same input (6 seeds each, the bar of the process tests)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def corner_and_border_holes(W, H, seed):
    rng = np.random.default_rng(seed)
    m = np.zeros((H, W), np.uint8)
    m[: max(1, H // 4), : max(1, W // 4)] = 1                          # top-left corner
    m[H - max(1, H // 5):, W // 2: W // 2 + max(1, W // 6)] = 1        # touches the bottom edge
    y, x = rng.integers(H // 3, max(H // 3 + 1, 2 * H // 3)), rng.integers(W // 3, max(W // 3 + 1, 2 * W // 3))
    m[y, x] = 1                                                        # an isolated pixel
    return m


CASES = [
    # W, H, C, method, patch, scales
    (12, 10, 3, "nlmeans", 3, 1),
    (17, 9, 1, "nlmedians", 3, 1),
    (33, 31, 2, "nlmeans", 5, 2),
    (40, 23, 3, "nlpoisson", 5, 2),
    (64, 48, 4, "nldepth", 5, 2),
    (57, 61, 3, "nlmedians", 7, 3),
    (96, 80, 3, "nlpoisson", 9, 3),
    (120, 100, 3, "nlmeans", 13, 2),
    (128, 128, 1, "nlpoisson", 15, 2),
]


@pytest.mark.parametrize("W,H,C,method,patch,scales", CASES)
def test_edge_inputs_properties(ctx, syn, nli, W, H, C, method, patch, scales):
    img = syn.texture(W, H, C if C != 4 else 3, 4, W + H)
    if C == 4:
        img = np.concatenate([img, syn.depth_planes(W, H, 3, W)[..., None]], axis=2)
    hole = corner_and_border_holes(W, H, W * H)
    sel = hole.astype(bool)
    kw = dict(method=method, patch=patch, iters=4, scales=scales, coarse=0.5, seed=3, cg_tol=1e-4, cg_limit=100)
    if method == "nldepth":
        kw.update(lambda_rgb=1.0, lambda_d=0.0)
    out, st = ctx.inpaint(img, hole, nli.Inpainting(**kw))
    assert out.shape == img.shape
    assert np.all(np.isfinite(out))
    assert np.array_equal(out[~sel], img[~sel])
    if method in ("nlmeans", "nlmedians"):
        known = img[~sel]
        lo, hi = known.min(axis=0), known.max(axis=0)
        assert np.all(out[sel] >= lo - 1e-6) and np.all(out[sel] <= hi + 1e-6)
    print(f"{W}x{H}x{C} {method} P={patch} S={scales}: iterations per scale {st.iterations}")


def test_one_pixel_hole_matches_reference(ctx, oracle, syn, nli):
    """A one-pixel hole away from the border: the update is the same weighted combination
    in the GPU and the reference whatever the NNF's quality, so compare PSNR over seeds."""
    import oracle as O
    W, H = 48, 40
    img = syn.texture(W, H, 3, 3, 9)
    hole = np.zeros((H, W), np.uint8)
    hole[20, 23] = 1
    g, o = [], []
    for seed in range(1, 7):
        kw = dict(method="nlmeans", patch=5, iters=5, scales=1, coarse=1.0, seed=seed)
        out, _ = ctx.inpaint(img, hole, nli.Inpainting(**kw))
        ref, _ = oracle.process(img, hole, O.InpaintParams(**kw))
        g.append(float(np.abs(out[20, 23] - img[20, 23]).max()))
        o.append(float(np.abs(ref[20, 23] - img[20, 23]).max()))
    print(f"one-pixel hole: mean |error| GPU {np.mean(g):.4f}, reference algorithm {np.mean(o):.4f}")
    assert np.mean(g) <= np.mean(o) + 0.02
