"""The command-line program (non-local-inpainting_amd/Inpainting) against main.cpp's
interface, on the CPU: usage, option parsing (pick_option, main.cpp:43-61), defaults, the
parameter checks and their messages (main.cpp:120-217) and the output file name
(main.cpp:287-303), through -dryrun, which stops before any GPU work."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "non-local-inpainting_amd", "Inpainting")
DATA = os.path.join(ROOT, "tests", "golden", "data")
IMG, MASK, DEPTH = (os.path.join(DATA, n) for n in ("img.png", "mask.png", "depth.png"))


def run(*args):
    if not os.path.exists(CLI):
        pytest.skip("Inpainting not built (make -C non-local-inpainting_amd)")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_usage():
    r = run("only", "two")
    assert r.returncode == 1
    assert r.stderr.startswith("USAGE:\n")
    assert " -method \tmethod name [nlmeans/nlmedians/nlpoisson] (nlmeans)\n" in r.stderr
    assert " -patch  \tpatch side (7)\n" in r.stderr
    assert " -coarse \tcoarsest rate (0.1)\n" in r.stderr
    assert " -psigma \tGaussian patch weights (10000)\n" in r.stderr
    r = run("a", "b", "-patch", "5", "-method", "nlpoisson")    # options are removed before the count
    assert r.returncode == 1 and "patch side (5)" in r.stderr and "(nlpoisson)" in r.stderr


@pytest.mark.parametrize("args,name", [
    ([], "out_nlmeans_sc11_0.1_initnone_ps7_10000_conf5_0.1.png"),
    (["-method", "nlpoisson", "-patch", "9"], "out_nlpoisson_l0.05_sc11_0.1_initnone_ps9_10000_conf5_0.1.png"),
    (["-init", "avg", "-lambda", "0.3", "-conft", "2.5", "-confa", "0.2"],
     "out_nlmeans_sc11_0.1_initavg_ps7_10000_conf2.5_0.2.png"),
    (["-init", "black", "-coarse", "0.123456789", "-scales", "5", "-psigma", "1.5"],
     "out_nlmeans_sc5_0.123457_init000_ps7_1.5_conf5_0.1.png"),
    (["-coarse", "1", "-method", "nlmedians"], "out_nlmedians_sc1_1_initnone_ps7_10000_conf5_0.1.png"),
    (["-depth", DEPTH, "-patch", "9", "-method", "nldepth", "-lambdar", "1.0", "-lambdad", "0.0"],
     "out_nldepth_sc11_0.1_initnone_ps9_10000_conf5_0.1.png"),
])
def test_output_name(args, name):
    # options may come before, between or after the positional arguments
    r = run(IMG, "-dryrun", MASK, "out", *args)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == name
    if "-coarse" in args and args[args.index("-coarse") + 1] == "1":
        assert "the number of scales is set to 1" in r.stdout


@pytest.mark.parametrize("args,message", [
    (["-patch", "0"], "ERROR: patch_side needs to be a positive integer."),
    (["-coarse", "1.5"], "ERROR: coarsest rate should be between 0 and 1."),
    (["-iters", "-1"], "ERROR: inpainting_iterations cannot be negative."),
    (["-conft", "0"], "ERROR: confidence_decay_time needs to be positive."),
    (["-lambda", "1.2"], "ERROR: lambda should be between 0 and 1."),
    (["-scales", "0"], "ERROR: number of scales must be at least 1."),
    (["-init", "grey"], "ERROR: Unknown initialization type"),
    (["-method", "nlsomething"], "ERROR: Unknown method name"),
    (["-method", "nldepth"], "Specify depth input using -depth [filename]"),
])
def test_parameter_errors(args, message):
    r = run(IMG, MASK, "out", "-dryrun", *args)
    assert r.returncode == 1
    assert message in r.stderr


def test_input_errors(tmp_path, nli):
    small = str(tmp_path / "small.png")
    black = str(tmp_path / "black.png")
    nli.write_rgb_image(small, np.zeros((10, 12, 3), np.float32))
    nli.write_rgb_image(black, np.zeros((468, 625, 3), np.float32))
    r = run(IMG, small, "out", "-dryrun")
    assert r.returncode == 1 and "ERROR: Input image and mask should be of the same size" in r.stderr
    r = run(IMG, black, "out", "-dryrun")
    assert r.returncode == 1 and "ERROR: Mask should not be empty (all black)" in r.stderr
