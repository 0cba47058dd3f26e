"""The colour path around process() on the GPU (main.cpp:115-309, io_utility.cpp:268-318).

* nli_rgb_to_lab / nli_lab_to_rgb against the oracle (which is bit-exact with the
  reference's own IOUtility code, tests/test_oracle_color.py).  The GPU evaluates the
  reference's powf as pow in double rounded to float: the correctly rounded value, which
  glibc's powf misses by one unit in the last place for a few inputs in ten thousand.
  The bar: at least 99.5 % of the values bit-identical, and every value within
  - rgb_to_lab: 2e-4 (L*a*b* units): one last-place difference of a cube root, scaled
    by 500 for a*;
  - lab_to_rgb: 1e-3 (0-255 units, a thousandth of an 8-bit level): one last-place
    difference of a cube in X, Y or Z (up to about 1) moves the linear RGB value by up
    to about 3e-7 through the matrix, and the gamma curve's slope, 255 * 1.055 / 2.4 *
    r^(-0.58), multiplies that by up to 2000 for dark colours;
  - lab_to_rgb out of the sRGB gamut (results up to about 1000 in magnitude):
    1e-3 + 1e-5 |value|;
* pass-through of images the reference does not convert, the 4-channel form, the
  device-pointer forms, and a 1024 x 1024 round trip;
* nli_inpaint_rgb: equal, bit for bit, to the composition rgb_to_lab -> nli_inpaint ->
  lab_to_rgb and to the plain drop-in's IOUtility composition (nli_plain_inpaint_rgb),
  with and without a depth plane; main.cpp's parameter rules and errors.  (The hole PSNR
  of the whole RGB pipeline against the reference's own is in test_gpu_parity.py.)"""
import numpy as np
import pytest

from conftest import bitexact
from test_oracle_color import colour_cases, lab_cases

pytestmark = pytest.mark.gpu

LAB_TOL = 2e-4
RGB_TOL = 1e-3
SAME = 0.995


def compare(name, got, want, atol, rtol=0.0):
    assert got.shape == want.shape
    same = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    diff = np.abs(got.astype(np.float64) - want)
    bound = atol + rtol * np.abs(want.astype(np.float64))
    k = int(np.argmax(diff / bound))
    print(f"{name}: {same * 100:.3f} % bit-identical, max |difference| {diff.max():.3g}; "
          f"closest to the bar: {diff.flat[k]:.3g} at value {float(want.flat[k]):.6g}")
    assert same >= SAME and np.all(diff <= bound), (same, diff.flat[k], float(want.flat[k]))


def test_rgb_to_lab_matches_oracle(ctx, oracle):
    for seed in (1, 2):
        img = colour_cases(seed)
        compare(f"rgb_to_lab seed {seed}", ctx.rgb_to_lab(img), oracle.rgb_to_lab(img), LAB_TOL)


def test_lab_to_rgb_matches_oracle(ctx, oracle):
    lab = lab_cases(3)
    compare("lab_to_rgb (in and out of gamut)", ctx.lab_to_rgb(lab), oracle.lab_to_rgb(lab), RGB_TOL, rtol=1e-5)
    real = oracle.rgb_to_lab(colour_cases(4))
    compare("lab_to_rgb (converted colours)", ctx.lab_to_rgb(real), oracle.lab_to_rgb(real), RGB_TOL)


def test_eight_bit_inputs_exhaustive_per_channel(ctx, oracle):
    """Every 8-bit value in every channel position (the gamma expansion sees only these
    256 values from 8-bit images), crossed with a few values in the other channels."""
    v = np.arange(256, dtype=np.float32)
    rows = []
    for other in (0.0, 37.0, 128.0, 255.0):
        for c in range(3):
            px = np.full((256, 3), other, np.float32)
            px[:, c] = v
            rows.append(px)
    img = np.stack(rows)
    compare("rgb_to_lab 8-bit", ctx.rgb_to_lab(img), oracle.rgb_to_lab(img), LAB_TOL)


def test_pass_through_four_channels_and_odd_sizes(ctx, oracle):
    rng = np.random.default_rng(5)
    for C in (1, 2, 4):
        img = rng.uniform(0, 255, (13, 7, C)).astype(np.float32)
        assert bitexact(ctx.rgb_to_lab(img), img)
    for C in (1, 2, 5):
        img = rng.uniform(-50, 100, (5, 9, C)).astype(np.float32)
        assert bitexact(ctx.lab_to_rgb(img), img)
    lab4 = np.concatenate([lab_cases(6)[:17, :123], rng.uniform(0, 10, (17, 123, 1)).astype(np.float32)], axis=2)
    out = ctx.lab_to_rgb(lab4)
    assert out.shape == (17, 123, 3)
    assert bitexact(out, ctx.lab_to_rgb(np.ascontiguousarray(lab4[..., :3])))
    compare("lab_to_rgb 4 channels", out, oracle.lab_to_rgb(lab4), RGB_TOL, rtol=1e-5)
    one = np.array([[[255.0, 0.0, 0.0]]], np.float32)
    compare("rgb_to_lab 1 pixel", ctx.rgb_to_lab(one), oracle.rgb_to_lab(one), LAB_TOL)


def test_table_edges(ctx, oracle):
    """rgb_to_lab takes the gamma expansion of the integral values 0..255 from a table
    computed by the same function; -0.0, values just off an integer, 255 and above, and
    negative values are computed directly.  All of them against the oracle."""
    v = np.array([-0.0, 0.0, 1.0, np.nextafter(np.float32(1.0), np.float32(2.0)), 10.0, 10.5, 11.0,
                  np.nextafter(np.float32(255.0), np.float32(0.0)), 255.0, 256.0, 300.0, -1.0, -255.0],
                 np.float32)
    img = np.stack(np.meshgrid(v, v, v, indexing="ij"), axis=-1).reshape(len(v), -1, 3)
    got, want = ctx.rgb_to_lab(img), oracle.rgb_to_lab(img)
    compare("rgb_to_lab table edges", got, want, LAB_TOL)
    sel = (img == 0).all(axis=2) & np.signbit(img).any(axis=2)
    assert np.array_equal(np.signbit(got[sel]), np.signbit(want[sel]))   # -0.0 keeps its sign path


def test_device_forms_equal_host_forms(ctx):
    img = colour_cases(7)
    H, W, _ = img.shape
    d_in = ctx.to_device(img)
    d_lab = ctx.dev_alloc(img.nbytes)
    d_back = ctx.dev_alloc(img.nbytes)
    try:
        ctx.rgb_to_lab_dev(d_in, d_lab, W, H)
        ctx.lab_to_rgb_dev(d_lab, d_back, W, H)
        lab = ctx.from_device(d_lab, img.shape, np.float32)
        back = ctx.from_device(d_back, img.shape, np.float32)
    finally:
        for p in (d_in, d_lab, d_back):
            ctx.dev_free(p)
    assert bitexact(lab, ctx.rgb_to_lab(img))
    assert bitexact(back, ctx.lab_to_rgb(lab))


def test_round_trip_1024(ctx, syn):
    img = (syn.texture(1024, 1024, 3, 8, 1) * 255.0).astype(np.float32)
    back = ctx.lab_to_rgb(ctx.rgb_to_lab(img))
    err = float(np.abs(back - img).max())
    print(f"1024x1024 rgb -> lab -> rgb: max |difference| {err:.3g} (0-255 scale)")
    assert err < 0.01


def rgb_case(syn, W=96, H=80, seed=3):
    rgb = (syn.texture(W, H, 3, 4, seed) * 255.0).astype(np.float32)
    hole = syn.ellipse_holes(W, H, 3, 0.12, seed)
    return rgb, hole


def test_inpaint_rgb_equals_composition_and_plain(ctx, syn, nli):
    rgb, hole = rgb_case(syn)
    p = nli.Inpainting(method="nlpoisson", patch=7, iters=4, scales=3, coarse=0.4, cg_tol=1e-4, cg_limit=60, seed=4)
    out, dout, st = ctx.inpaint_rgb(rgb, hole, p)
    assert dout is None
    lab, st1 = ctx.inpaint(ctx.rgb_to_lab(rgb), hole, p)
    assert bitexact(out, ctx.lab_to_rgb(lab))
    assert st.iterations == st1.iterations
    plain, _, st2 = ctx.inpaint_rgb_plain(rgb, hole, p)
    assert bitexact(out, plain)
    assert st2.iterations == st.iterations
    sel = hole.astype(bool)
    assert np.abs(out[~sel] - rgb[~sel]).max() < 0.01   # known pixels: a round trip only


def test_inpaint_rgb_with_depth(ctx, syn, nli):
    W, H = 120, 96
    rgb = (syn.texture(W, H, 3, 3, 4) * 255.0).astype(np.float32)
    depth = syn.depth_planes(W, H, 5, 4)
    hole = syn.stripe_holes(W, H, 3, 6, 4)
    p = nli.Inpainting(method="nldepth", patch=7, iters=3, scales=3, coarse=0.4, lambda_rgb=1.0, lambda_d=0.0, seed=2)
    out, dout, st = ctx.inpaint_rgb(rgb, hole, p, depth=depth)
    lab4 = np.concatenate([ctx.rgb_to_lab(rgb), depth[..., None]], axis=2)
    ref, _ = ctx.inpaint(lab4, hole, p)
    assert bitexact(out, ctx.lab_to_rgb(ref))
    assert bitexact(dout, np.ascontiguousarray(ref[..., 3]))
    plain, pdepth, _ = ctx.inpaint_rgb_plain(rgb, hole, p, depth=depth)
    assert bitexact(out, plain) and bitexact(dout, pdepth)
    assert np.all(np.isfinite(dout))


def test_inpaint_png_files_end_to_end(ctx, syn, nli, tmp_path):
    """main.cpp:115-312 with files: PNG input, mask and depth written by a second encoder
    (PIL), Context.inpaint_png, the output PNGs read back by PIL.  They hold exactly
    inpaint_rgb's results taken to bytes as iio takes them (uint8)clamp(v + 0.5)."""
    Image = pytest.importorskip("PIL.Image")
    W, H = 120, 96
    rgb8 = np.rint(syn.texture(W, H, 3, 3, 11) * 255.0).astype(np.uint8)
    depth8 = np.rint(syn.depth_planes(W, H, 4, 11) * 255.0 / max(1e-6, syn.depth_planes(W, H, 4, 11).max()))
    depth8 = depth8.astype(np.uint8)
    hole = syn.ellipse_holes(W, H, 3, 0.12, 11)
    Image.fromarray(rgb8, "RGB").save(tmp_path / "in.png")
    Image.fromarray((hole > 0).astype(np.uint8) * 255, "L").convert("1").save(tmp_path / "mask.png")
    Image.fromarray(depth8, "L").save(tmp_path / "depth.png")

    def to_bytes(v):
        t = 0.5 + np.asarray(v, np.float64)
        return np.trunc(np.where(t > 0, np.where(t < 255, t, 255), 0)).astype(np.uint8)

    for method, dname in (("nlpoisson", None), ("nldepth", "depth.png")):
        p = nli.Inpainting(method=method, patch=7, iters=3, scales=2, coarse=0.5, lambda_rgb=1.0, lambda_d=0.0,
                           seed=4)
        out_name, dout_name = str(tmp_path / f"out_{method}.png"), str(tmp_path / f"dout_{method}.png")
        st = ctx.inpaint_png(str(tmp_path / "in.png"), str(tmp_path / "mask.png"), out_name, p,
                             depth_name=str(tmp_path / dname) if dname else None, depth_output_name=dout_name)
        assert st.scales == 2
        d = depth8.astype(np.float32) if dname else None
        ref, dref, _ = ctx.inpaint_rgb(rgb8.astype(np.float32), (hole > 0).astype(np.uint8), p, depth=d)
        got = np.asarray(Image.open(out_name))
        assert bitexact(got, to_bytes(ref))
        if dname:
            assert bitexact(np.asarray(Image.open(dout_name)), to_bytes(dref))
        print(f"{method}: {W}x{H} PNG in, PNG out; {int((hole > 0).sum())} hole pixels")
    with pytest.raises(ValueError, match="same size"):
        Image.fromarray(rgb8[:-1], "RGB").save(tmp_path / "small.png")
        ctx.inpaint_png(str(tmp_path / "small.png"), str(tmp_path / "mask.png"), str(tmp_path / "x.png"), p)
    Image.fromarray(np.zeros((H, W), np.uint8), "L").save(tmp_path / "black.png")
    with pytest.raises(ValueError, match="should not be empty"):
        ctx.inpaint_png(str(tmp_path / "in.png"), str(tmp_path / "black.png"), str(tmp_path / "x.png"), p)


def test_inpaint_rgb_device_form(ctx, syn, nli):
    """nli_inpaint_rgb_dev on device buffers equals nli_inpaint_rgb on host buffers, with
    and without a depth plane."""
    W, H = 112, 90
    rgb = np.rint(syn.texture(W, H, 3, 3, 8) * 255.0).astype(np.float32)
    depth = syn.depth_planes(W, H, 4, 8)
    hole = syn.ellipse_holes(W, H, 3, 0.12, 8)
    for method, d in (("nlpoisson", None), ("nldepth", depth)):
        p = nli.Inpainting(method=method, patch=7, iters=3, scales=2, coarse=0.5, lambda_rgb=1.0, lambda_d=0.0,
                           cg_tol=1e-4, cg_limit=60, seed=5)
        out, dout, st = ctx.inpaint_rgb(rgb, hole, p, depth=d)
        ptrs = [ctx.to_device(rgb), ctx.to_device(hole), ctx.dev_alloc(rgb.nbytes)]
        if d is not None:
            ptrs += [ctx.to_device(d), ctx.dev_alloc(d.nbytes)]
        try:
            st_dev = ctx.inpaint_rgb_dev(ptrs[0], ptrs[1], ptrs[2], W, H, p,
                                         *((ptrs[3], ptrs[4]) if d is not None else ()))
            got = ctx.from_device(ptrs[2], rgb.shape, np.float32)
            gotd = ctx.from_device(ptrs[4], d.shape, np.float32) if d is not None else None
        finally:
            for q in ptrs:
                ctx.dev_free(q)
        assert bitexact(got, out) and st_dev.iterations == st.iterations
        if d is not None:
            assert bitexact(gotd, dout)


def test_inpaint_rgb_parameter_rules_and_errors(ctx, oracle, syn, nli):
    rgb, hole = rgb_case(syn, 80, 64, 5)
    # coarse = 0: min(1.5 * patch / largest distance to the hole boundary, 1) (main.cpp:166-178)
    p = nli.Inpainting(method="nlmeans", patch=5, iters=2, scales=3, coarse=0.0, seed=1)
    _, _, st = ctx.inpaint_rgb(rgb, hole, p)
    max_dist = float(oracle.distance_transform(hole).max())
    rate = nli.subsampling_rate(min(np.float32(1.5) * np.float32(5) / np.float32(max_dist), 1.0), 3)
    w = [80]
    for _ in range(2):
        w.append(ctx.sample_size(w[-1], rate))
    assert [s[0] for s in st.sizes] == w
    _, _, stp = ctx.inpaint_rgb_plain(rgb, hole, p)
    assert stp.iterations == st.iterations
    # a coarsest rate of 1 runs one scale (main.cpp:200-203)
    _, _, st1 = ctx.inpaint_rgb(rgb, hole, nli.Inpainting(method="nlmeans", patch=5, iters=2, scales=6, coarse=1.0))
    assert st1.scales == 1
    with pytest.raises(ValueError, match="Mask should not be empty"):
        ctx.inpaint_rgb(rgb, np.zeros_like(hole), p)
    with pytest.raises(ValueError, match="depth"):
        ctx.inpaint_rgb(rgb, hole, nli.Inpainting(method="nldepth", patch=5, iters=2, scales=2, coarse=0.5))
    with pytest.raises(ValueError, match="confidence_decay_time"):
        ctx.inpaint_rgb(rgb, hole, nli.Inpainting(method="nlmeans", patch=5, iters=2, scales=2, conf_t=0.0))
    with pytest.raises(ValueError, match="confidence_decay_time"):
        ctx.inpaint_rgb_plain(rgb, hole, nli.Inpainting(method="nlmeans", patch=5, iters=2, scales=2, conf_t=0.0))
