"""The C++ plugin classes (include/nli_plugins.hpp) composed as the reference's main():
HipPatchMatch and the HipPatchNonLocal* updaters plugged into ImageInpainting, every
step a C-ABI call on host buffers (libnli_plugins.so).  Its output must equal the
device-resident driver's bit for bit: the two run the same kernels on the same data
in the same order with the same per-call PatchMatch seeds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rgbd(syn, W, H, seed):
    return np.concatenate([syn.texture(W, H, 3, 3, seed), syn.depth_planes(W, H, 5, seed)[..., None]], axis=2)


CASES = [
    ("nlmeans", dict(patch=7, iters=6, scales=3, coarse=0.4), 96, 80),
    ("nlmedians", dict(patch=7, iters=4, scales=3, coarse=0.4), 90, 72),
    ("nlpoisson", dict(patch=9, iters=5, scales=3, coarse=0.4, cg_tol=1e-4, cg_limit=60), 100, 84),
    ("nlpoisson", dict(patch=7, iters=4, scales=2, coarse=0.5, cg_tol=1e-4, cg_limit=60, cg_precision=1,
                       cg_refine_every=8, init="avg"), 80, 64),
    ("nldepth", dict(patch=7, iters=3, scales=2, coarse=0.5, lambda_rgb=1.0, lambda_d=0.0), 96, 72),
]


@pytest.mark.parametrize("method,kw,W,H", CASES, ids=[f"{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_plain_dropin_equals_device_driver(ctx, syn, nli, method, kw, W, H):
    img = rgbd(syn, W, H, 5) if method == "nldepth" else syn.texture(W, H, 3, 3, 5)
    hole = (syn.stripe_holes(W, H, 3, 5, 5) if method == "nldepth" else syn.ellipse_holes(W, H, 4, 0.12, 5))
    p = nli.Inpainting(method=method, seed=3, **kw)
    plain, st_p = ctx.inpaint_plain(img, hole, p)
    dev, st_d = ctx.inpaint(img, hole, p)
    sel = hole.astype(bool)
    assert np.array_equal(plain[~sel], img[~sel])
    assert st_p.iterations == st_d.iterations
    assert np.array_equal(plain, dev)


def test_plain_dropin_reports_empty_source(ctx, nli):
    img = np.full((40, 40, 3), 0.5, np.float32)
    with pytest.raises(nli.EmptySourceError, match="Empty source mask"):
        ctx.inpaint_plain(img, np.ones((40, 40), np.uint8), nli.Inpainting(patch=7, scales=1, coarse=1.0))


def test_plain_dropin_rejects_rgbd_distance_on_rgb(ctx, nli):
    img = np.full((40, 40, 3), 0.5, np.float32)
    with pytest.raises(nli.NliError):
        ctx.inpaint_plain(img, np.pad(np.ones((10, 10), np.uint8), 15), nli.Inpainting(method="nldepth", scales=1))


def test_kernel_timing_all_and_sampled(ctx, syn, nli):
    """nli_profile_enable(_sampled): a PatchMatch call's sweeps are timed as one span (one
    event pair, its sweeps counted as launches).  With every call timed, every sweep of
    the run is counted; timing one call in 4 counts the sweeps of every 4th call (the
    first, the fifth, ...); the algorithmic work of the timed sweeps is positive."""
    img = syn.texture(160, 128, 3, 4, 2)
    hole = syn.ellipse_holes(160, 128, 3, 0.12, 2)
    p = nli.Inpainting(method="nlpoisson", patch=9, iters=3, scales=2, coarse=0.5, cg_tol=1e-4, cg_limit=50, seed=2)
    ctx.inpaint(img, hole, p)   # autotuned once, so both runs below launch the same sweeps
    ctx.profile(["pm_sweep"])
    ctx.inpaint(img, hole, p)
    full = ctx.profile_read("pm_sweep")
    ctx.profile(["pm_sweep"], period=4)
    ctx.inpaint(img, hole, p)
    part = ctx.profile_read("pm_sweep")
    ctx.profile(False)
    print(f"sweeps timed: all {full['launches']}, one in 4 {part['launches']}; "
          f"average {full['ms'] / full['launches'] * 1e3:.1f} vs {part['ms'] / part['launches'] * 1e3:.1f} us")
    K = p.pm_iters   # sweeps per PatchMatch call
    assert full["launches"] > 8 and full["work"] > 0 and full["launches"] % K == 0
    calls = full["launches"] // K
    assert part["launches"] == (calls + 3) // 4 * K
    assert part["work"] > 0
    with pytest.raises(ValueError):
        ctx.profile(["pm_sweep"], period=0)
