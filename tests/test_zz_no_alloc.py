"""nli_patch_distance_dev_noalloc: the allocation-free form of nli_patch_distance
(device pointers, scratch in a static per-device array), against nli_patch_distance
itself, bit for bit; two calls in a row with the second smaller (the array keeps the
previous call's contents); images beyond the scratch are refused."""
import ctypes as ct

import numpy as np
import pytest

from conftest import bitexact

pytestmark = pytest.mark.gpu

NLI_E_ARGUMENT = -1


def noalloc_distance(ctx, nli, kind, src, tgt, nnf, tmask, patch=9, sigma=10000.0, lam=0.3, lam_d=0.6):
    sH, sW, C = src.shape
    tH, tW, _ = tgt.shape
    d_src = ctx.to_device(src)
    d_tgt = d_src if tgt is src else ctx.to_device(tgt)
    d_nnf = ctx.to_device(nnf)
    d_tm = ctx.to_device(tmask)
    d_out = ctx.dev_alloc(tW * tH * 4)
    try:
        dp = nli.DistParams(nli.DIST[kind], patch, sigma, lam, lam_d)
        rc = ctx.lib.nli_patch_distance_dev_noalloc(ctx.h, ct.byref(dp), ct.c_void_p(d_src), sW, sH,
                                                    ct.c_void_p(d_tgt), tW, tH, C, ct.c_void_p(d_nnf),
                                                    ct.c_void_p(d_tm), ct.c_void_p(d_out))
        msg = ctx.lib.nli_last_error(ctx.h)   # before the frees below clear it
        out = ctx.from_device(d_out, (tH, tW), np.float32) if rc == 0 else None
    finally:
        for p in {d_src, d_tgt, d_nnf, d_tm, d_out}:
            ctx.dev_free(p)
    return rc, out, msg


def field(syn, W, H, C, P, seed, tW=None, tH=None):
    tW, tH = tW or W, tH or H
    img = syn.texture(W, H, 3, 3, seed)
    if C == 4:
        img = np.concatenate([img, syn.depth_planes(W, H, 5, seed)[..., None]], axis=2)
    img = np.ascontiguousarray(img[..., :C])
    r = P // 2
    rng = np.random.default_rng(seed)
    tm = np.zeros((tH, tW), np.uint8)
    tm[r:tH - r, r:tW - r] = rng.integers(0, 2, (tH - 2 * r, tW - 2 * r))
    nnf = np.stack([rng.integers(r, W - r, (tH, tW)), rng.integers(r, H - r, (tH, tW))], axis=2).astype(np.int32)
    return img, tm, nnf


@pytest.mark.parametrize("kind,C", [("l2combined", 3), ("l2rgbd", 4), ("l2", 3), ("l1", 1)])
def test_noalloc_equals_patch_distance_two_calls(ctx, nli, syn, kind, C):
    for (W, H, P, seed) in [(300, 220, 9, 1), (97, 61, 7, 2)]:           # the second call is smaller
        img, tm, nnf = field(syn, W, H, C, P, seed)
        rc, got, msg = noalloc_distance(ctx, nli, kind, img, img, nnf, tm, patch=P)
        assert rc == 0, msg
        want = ctx.patch_distance(kind, img, img, nnf, tm, P, 10000.0, 0.3, 0.6)
        assert bitexact(got, want), (kind, W, H)


def test_noalloc_distinct_target(ctx, nli, syn):
    """Source and target images differ (both gradients live in the scratch)."""
    src, _, _ = field(syn, 140, 110, 3, 9, 3)
    tgt, tm, nnf = field(syn, 120, 90, 3, 9, 4)
    nnf[..., 0] = np.clip(nnf[..., 0], 4, 140 - 5)
    nnf[..., 1] = np.clip(nnf[..., 1], 4, 110 - 5)
    rc, got, msg = noalloc_distance(ctx, nli, "l2combined", src, tgt, nnf, tm)
    assert rc == 0, msg
    assert bitexact(got, ctx.patch_distance("l2combined", src, tgt, nnf, tm, 9, 10000.0, 0.3, 0.6))


def test_noalloc_refuses_images_beyond_the_scratch(ctx, nli):
    """2 C W H floats of gradient beyond 32 MiB: refused before any kernel runs (the
    buffers passed are of the stated size all the same)."""
    W = H = 2048
    img = np.zeros((H, W, 4), np.float32)
    tm = np.zeros((H, W), np.uint8)
    nnf = np.full((H, W, 2), -1, np.int32)
    rc, _, msg = noalloc_distance(ctx, nli, "l2rgbd", img, img, nnf, tm)
    assert rc == NLI_E_ARGUMENT
    assert b"no-allocation scratch" in msg
