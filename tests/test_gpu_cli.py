"""The command-line program on the GPU: the reference README's example call
(README.md:44-46) end to end on the reference's sample images, and the default nlmeans
run against the reference algorithm's colour pipeline (oracle).

* The program's result files equal, byte for byte, those of Context.inpaint_png (the
  device-resident driver) with the same parameters and seed: the program runs the plugin
  classes (one C-ABI call per reference call), whose results equal the device driver's.
* -showpyr writes every level of the output pyramid, -shownnf the NNF illustration.
* Default nlmeans run: the hole of the program's result against the oracle's
  rgb_to_lab -> process -> lab_to_rgb on the same files.  The photograph has no ground
  truth in the hole and the two runs draw different random numbers, so the bar is loose:
  outside the hole both keep the input (up to the 8-bit round trip), inside the two fills
  agree to a PSNR of at least 20 dB on the 0-255 scale."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "non-local-inpainting_amd", "Inpainting")
DATA = os.path.join(ROOT, "tests", "golden", "data")
IMG, MASK, DEPTH = (os.path.join(DATA, n) for n in ("img.png", "mask.png", "depth.png"))


def cli(cwd, *args, timeout=600):
    if not os.path.exists(CLI):
        pytest.skip("Inpainting not built (make -C non-local-inpainting_amd)")
    print("Inpainting", " ".join(args), flush=True)
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def test_readme_example_equals_inpaint_png(ctx, nli, tmp_path):
    cli(str(tmp_path), IMG, MASK, "output.png", "-depth", DEPTH, "-patch", "9", "-method", "nldepth",
        "-lambdar", "1.0", "-lambdad", "0.0", "-seed", "3")
    out = tmp_path / "output.png_nldepth_sc11_0.1_initnone_ps9_10000_conf5_0.1.png"
    assert out.exists() and (tmp_path / "inpainted_depth.png").exists()
    p = nli.Inpainting(method="nldepth", patch=9, lambda_rgb=1.0, lambda_d=0.0, seed=3)
    ctx.inpaint_png(IMG, MASK, str(tmp_path / "mine.png"), p, depth_name=DEPTH,
                    depth_output_name=str(tmp_path / "mine_depth.png"))
    assert out.read_bytes() == (tmp_path / "mine.png").read_bytes()
    assert (tmp_path / "inpainted_depth.png").read_bytes() == (tmp_path / "mine_depth.png").read_bytes()


def test_pyramid_and_nnf_dumps(nli, tmp_path):
    cli(str(tmp_path), IMG, MASK, "res", "-scales", "4", "-iters", "5", "-showpyr", "level", "-shownnf", "nnf.png",
        "-seed", "1")
    sizes = []
    for n in range(4):
        f = tmp_path / f"level{n}.png"
        assert f.exists()
        sizes.append(nli.read_rgb_image(str(f)).shape[:2])
    assert sizes[0] == (468, 625) and all(a[0] > b[0] and a[1] > b[1] for a, b in zip(sizes, sizes[1:]))
    nnf = nli.read_rgb_image(str(tmp_path / "nnf.png"))
    res = nli.read_rgb_image(str(tmp_path / "res_nlmeans_sc4_0.1_initnone_ps7_10000_conf5_0.1.png"))
    assert nnf.shape == res.shape
    mask = nli.read_mono_image(MASK) > 0
    edge = mask[1:, 1:] != mask[1:, :-1]
    assert np.all(nnf[1:, 1:][edge] == 0)                     # the hole boundary is drawn black
    assert np.array_equal(nnf[~mask], res[~mask]) or np.mean(nnf[~mask] == res[~mask]) > 0.99


def test_default_run_against_reference_algorithm(oracle, nli, tmp_path):
    import oracle as O
    cli(str(tmp_path), IMG, MASK, "res", "-seed", "2")
    res = nli.read_rgb_image(str(tmp_path / "res_nlmeans_sc11_0.1_initnone_ps7_10000_conf5_0.1.png"))
    rgb = nli.read_rgb_image(IMG)
    mask = (nli.read_mono_image(MASK) > 0).astype(np.uint8)
    p = O.InpaintParams(method="nlmeans", patch=7, seed=2, pm_threads=16)
    ref = oracle.lab_to_rgb(oracle.process(oracle.rgb_to_lab(rgb), mask, p)[0])
    ref8 = np.clip(np.floor(0.5 + ref.astype(np.float64)), 0, 255)
    sel = mask.astype(bool)
    assert np.max(np.abs(res[~sel] - rgb[~sel])) <= 1            # 8-bit round trip of known pixels
    assert np.max(np.abs(ref8[~sel] - rgb[~sel])) <= 1
    mse = float(np.mean((res[sel].astype(np.float64) - ref8[sel]) ** 2))
    psnr = 10 * np.log10(255.0 ** 2 / mse)
    print(f"README images, default nlmeans: hole PSNR between the program's fill and the reference "
          f"algorithm's fill {psnr:.2f} dB")
    assert psnr >= 20.0
