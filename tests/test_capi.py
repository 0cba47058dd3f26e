"""The C-ABI library on the CPU: it loads, exports every entry point include/nli.h
declares, and its host-side functions agree with the oracle.  No kernel runs here."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from conftest import ROOT, bitexact


def declared_symbols(header="nli.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nli_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(nli):
    lib = nli.load_library()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(nli.EXPORTS) == names


def test_plugin_library_exports_its_declared_symbols(nli):
    lib = nli.load_plugins()
    names = declared_symbols("nli_plugins.hpp")
    own = sorted([n for n in names if n.startswith("nli_plain")] + declared_symbols("nli_io.h"))
    assert len(own) == 14
    assert own == sorted(nli.PLUGIN_EXPORTS)
    missing = [n for n in own if not hasattr(lib, n)]
    assert not missing, missing


def test_plain_inpaint_rejects_null_context(nli):
    lib = nli.load_plugins()
    p = nli.Inpainting().to_struct()
    img = np.zeros((8, 8, 3), np.float32)
    mask = np.zeros((8, 8), np.uint8)
    out = np.zeros_like(img)
    rc = lib.nli_plain_inpaint(None, ct.byref(p), img.ctypes.data_as(ct.c_void_p), mask.ctypes.data_as(ct.c_void_p),
                               8, 8, 3, out.ctypes.data_as(ct.c_void_p), None)
    assert rc == nli.NLI_E_ARGUMENT


def test_version(nli):
    assert nli.load_library().nli_version() == 1


def test_rate_and_sizes_match_oracle(nli, oracle):
    lib = nli.load_library()
    for ratio, scales in [(0.1, 11), (0.1, 4), (0.1, 8), (0.37, 5)]:
        a = nli.subsampling_rate(ratio, scales)
        b = oracle.subsampling_rate(ratio, scales)
        assert np.float32(a) == np.float32(b)
        for size in (1024, 1920, 1080, 813, 129, 17):
            assert lib.nli_sample_size(size, ct.c_float(b)) == oracle.sample_size(size, b)


def test_poisson_fill_matches_oracle(nli, oracle, syn):
    img = syn.texture(45, 38, 3, 3, 6)
    hole = syn.ellipse_holes(45, 38, 3, 0.25, 6, border=2)
    masked = img.copy()
    masked[hole.astype(bool)] = 0.0
    assert bitexact(nli.poisson_fill(masked, hole), oracle.simpois(masked, hole))


def gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(gpu_present(), reason="a GPU is visible")
def test_context_without_gpu_fails_cleanly(nli):
    with pytest.raises(nli.NliError):
        nli.Context(0)


def test_null_context_is_rejected(nli):
    lib = nli.load_library()
    assert lib.nli_synchronize(None) == nli.NLI_E_ARGUMENT
    assert lib.nli_poisson_fill(None, None, 0, 0, 0) == nli.NLI_E_ARGUMENT


def test_process_wide_knobs_check_their_range(nli):
    """The sweep's tile arrangement (0 tuned, 1-3 forced), the row scan (0-2) and the batch
    mode (0-2) are process-wide settings: out-of-range values are refused, in-range ones
    accepted (no GPU work is done)."""
    lib = nli.load_library()
    try:
        for v in (0, 1, 2, 3):
            assert lib.nli_pm_tiles_per_block(v) == nli.NLI_OK
        for v in (-1, 4):
            assert lib.nli_pm_tiles_per_block(v) == nli.NLI_E_ARGUMENT
        for v in (0, 1, 2):
            assert lib.nli_pm_row_scan(v) == nli.NLI_OK
        assert lib.nli_pm_row_scan(3) == nli.NLI_E_ARGUMENT
        for v in (0, 1, 2):
            assert lib.nli_batch_mode(v) == nli.NLI_OK
        for v in (-1, 3):
            assert lib.nli_batch_mode(v) == nli.NLI_E_ARGUMENT
    finally:
        lib.nli_pm_tiles_per_block(0)
        lib.nli_pm_row_scan(0)
        lib.nli_batch_mode(0)
