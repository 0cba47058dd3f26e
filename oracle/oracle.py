"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference core
built in place (_ref/libnli_ref*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The
product (the non-local-inpainting_amd package and libnli.so) never imports it.

Arrays follow the reference containers: images float32 (H, W, C) HWC, masks
uint8 (H, W), NNF int32 (H, W, 2) holding (x, y), (-1, -1) = none.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnli_ref.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libnli_ref_omp.so")
REF_COLOR_SO = os.path.join(HERE, "_ref", "libnli_ref_color.so")

DIST = {"l2": 0, "l1": 1, "l2combined": 2, "l2rgbd": 3}
UPD = {"nlmeans": 0, "nlmedians": 1, "nlpoisson": 2, "nldepth": 3}
INIT = {"black": 0, "avg": 1, "none": 2, "poisson": 3}

_f32p = ct.POINTER(ct.c_float)
_f64p = ct.POINTER(ct.c_double)
_u8p = ct.POINTER(ct.c_uint8)
_i32p = ct.POINTER(ct.c_int32)


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


class DistParams(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("patch", ct.c_int), ("sigma", ct.c_float),
                ("lambda_", ct.c_float), ("lambda_d", ct.c_float)]


class PMParams(ct.Structure):
    _fields_ = [("iterations", ct.c_int), ("shots", ct.c_int), ("window", ct.c_int)]


class UpdParams(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("patch", ct.c_int), ("sigma", ct.c_float),
                ("lambda_", ct.c_float), ("lambda_d", ct.c_float), ("cg_tol", ct.c_float),
                ("cg_limit", ct.c_int)]


class CGStats(ct.Structure):
    _fields_ = [("channels_solved", ct.c_int), ("cg_iterations", ct.c_int * 4),
                ("cg_nr0", ct.c_double * 4), ("cg_nr", ct.c_double * 4)]


class ProcessParams(ct.Structure):
    _fields_ = [("method", ct.c_int), ("patch", ct.c_int), ("sigma", ct.c_float),
                ("lambda_", ct.c_float), ("lambda_rgb", ct.c_float), ("lambda_d", ct.c_float),
                ("iters", ct.c_int), ("tol", ct.c_float), ("scales", ct.c_int),
                ("coarse", ct.c_float), ("conf_t", ct.c_float), ("conf_a", ct.c_float),
                ("init", ct.c_int), ("pm_iters", ct.c_int), ("pm_shots", ct.c_int),
                ("pm_window", ct.c_int), ("cg_tol", ct.c_float), ("cg_limit", ct.c_int),
                ("seed", ct.c_uint), ("pm_threads", ct.c_int), ("max_iters_per_scale", ct.c_int)]


class ProcessStats(ct.Structure):
    _fields_ = [("iterations", ct.c_int * 32), ("width", ct.c_int * 32), ("height", ct.c_int * 32),
                ("seconds_pyramid", ct.c_double), ("seconds_patchmatch", ct.c_double),
                ("seconds_update", ct.c_double), ("seconds_other", ct.c_double),
                ("seconds_iterations", ct.c_double * 32), ("seconds_propagate", ct.c_double * 32)]


@dataclass
class InpaintParams:
    """Parameters of one inpainting run; defaults are the reference CLI's
    (main.cpp:68-82, 220-233, 254, 266) except the patch side (9)."""
    method: str = "nlmeans"
    patch: int = 9
    sigma: float = 10000.0
    lam: float = 0.05
    lambda_rgb: float = 0.05
    lambda_d: float = 0.05
    iters: int = 200
    tol: float = 0.1
    scales: int = 11
    coarse: float = 0.1
    conf_t: float = 5.0
    conf_a: float = 0.1
    init: str = "poisson"
    pm_iters: int = 10
    pm_shots: int = 30
    pm_window: int = -1
    cg_tol: float = 1e-6
    cg_limit: int = 1000
    seed: int = 1
    pm_threads: int = 0
    max_iters_per_scale: int = 0

    def to_struct(self) -> ProcessParams:
        return ProcessParams(UPD[self.method], self.patch, self.sigma, self.lam, self.lambda_rgb,
                             self.lambda_d, self.iters, self.tol, self.scales, self.coarse,
                             self.conf_t, self.conf_a, INIT[self.init], self.pm_iters,
                             self.pm_shots, self.pm_window, self.cg_tol, self.cg_limit,
                             self.seed, self.pm_threads, self.max_iters_per_scale)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Oracle:
    """The clean-room CPU restatement."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = ct.CDLL(path, mode=ct.RTLD_LOCAL)
        L = self.lib
        L.nlo_patch_distance.restype = ct.c_float
        L.nlo_update.restype = ct.c_double
        L.nlo_subsampling_rate.restype = ct.c_float
        L.nlo_sample_size.restype = ct.c_int

    # ---- numerics ----
    def gaussian_weights(self, sx, sy, sigx, sigy):
        out = np.zeros(sx * sy, np.float32)
        self.lib.nlo_gaussian_weights(sx, sy, ct.c_float(sigx), ct.c_float(sigy), _p(out, _f32p))
        return out.reshape(sy, sx)

    def gradient(self, img):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, 2 * C), np.float32)
        self.lib.nlo_gradient(_p(img, _f32p), W, H, C, _p(out, _f32p))
        return out

    def distance_transform(self, mask):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        out = np.zeros((H, W), np.float32)
        self.lib.nlo_distance_transform(_p(mask, _u8p), W, H, _p(out, _f32p))
        return out

    def confidence(self, mask, decay=5.0, asym=0.1):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        out = np.zeros((H, W), np.float32)
        self.lib.nlo_confidence(_p(mask, _u8p), W, H, ct.c_float(decay), ct.c_float(asym), _p(out, _f32p))
        return out

    def subsampling_rate(self, ratio, scales):
        return float(self.lib.nlo_subsampling_rate(ct.c_float(ratio), scales))

    def sample_size(self, size, factor):
        return int(self.lib.nlo_sample_size(size, ct.c_float(factor)))

    def downsample_with_mask(self, img, mask, factor, threshold=0.4, mask_val=False):
        img = _c(img, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = img.shape
        sw, sh = self.sample_size(W, factor), self.sample_size(H, factor)
        oi = np.zeros((sh, sw, C), np.float32)
        om = np.zeros((sh, sw), np.uint8)
        rc = self.lib.nlo_downsample_with_mask(_p(img, _f32p), _p(mask, _u8p), W, H, C, ct.c_float(factor),
                                               ct.c_float(threshold), int(mask_val), _p(oi, _f32p), _p(om, _u8p))
        if rc != 0:
            raise ValueError("invalid factor")
        return oi, om

    def downsample(self, img, factor):
        img = _c(img, np.float32)
        H, W, C = img.shape
        sw, sh = self.sample_size(W, factor), self.sample_size(H, factor)
        out = np.zeros((sh, sw, C), np.float32)
        if self.lib.nlo_downsample(_p(img, _f32p), W, H, C, ct.c_float(factor), _p(out, _f32p)) != 0:
            raise ValueError("invalid factor")
        return out

    def downsample_mask(self, mask, factor, threshold=0.4):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        sw, sh = self.sample_size(W, factor), self.sample_size(H, factor)
        out = np.zeros((sh, sw), np.uint8)
        if self.lib.nlo_downsample_mask(_p(mask, _u8p), W, H, ct.c_float(factor), ct.c_float(threshold),
                                        _p(out, _u8p)) != 0:
            raise ValueError("invalid factor")
        return out

    def upsample(self, img, W2, H2):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H2, W2, C), np.float32)
        self.lib.nlo_upsample(_p(img, _f32p), W, H, C, W2, H2, _p(out, _f32p))
        return out

    # ---- colour (io_utility.cpp:268-318) ----
    def rgb_to_lab(self, img):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, C), np.float32)
        self.lib.nlo_rgb_to_lab(_p(img, _f32p), W, H, C, _p(out, _f32p))
        return out

    def lab_to_rgb(self, img):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, 3 if C in (3, 4) else C), np.float32)
        self.lib.nlo_lab_to_rgb(_p(img, _f32p), W, H, C, _p(out, _f32p))
        return out

    # ---- masks ----
    def extended_domain(self, hole, P):
        hole = _c(hole, np.uint8)
        H, W = hole.shape
        out = np.zeros((H, W), np.uint8)
        self.lib.nlo_extended_domain(_p(hole, _u8p), W, H, P, _p(out, _u8p))
        return out

    def scale_masks(self, hole, P, erode=True):
        """(extended, target, source, source_empty) as inpaint_internal builds them."""
        hole = _c(hole, np.uint8)
        H, W = hole.shape
        e, t, s = (np.zeros((H, W), np.uint8) for _ in range(3))
        empty = self.lib.nlo_scale_masks(_p(hole, _u8p), W, H, P, int(erode), _p(e, _u8p), _p(t, _u8p),
                                         _p(s, _u8p))
        return e, t, s, bool(empty)

    def scale_nnf(self, lnnf, lhole, uW, uH):
        lnnf = _c(lnnf, np.int32)
        lhole = _c(lhole, np.uint8)
        lH, lW = lhole.shape
        out = np.zeros((uH, uW, 2), np.int32)
        self.lib.nlo_scale_nnf(_p(lnnf, _i32p), _p(lhole, _u8p), lW, lH, uW, uH, _p(out, _i32p))
        return out

    # ---- distances / PatchMatch ----
    @staticmethod
    def dist_params(kind, patch=9, sigma=10000.0, lam=0.5, lam_d=0.5):
        return DistParams(DIST[kind] if isinstance(kind, str) else kind, patch, sigma, lam, lam_d)

    def patch_distance_field(self, dp, src, tgt, nnf, tmask):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        nnf = _c(nnf, np.int32)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        out = np.zeros((tH, tW), np.float32)
        self.lib.nlo_patch_distance_field(ct.byref(dp), _p(src, _f32p), sW, sH, _p(tgt, _f32p), tW, tH, C,
                                          _p(nnf, _i32p), _p(tmask, _u8p), _p(out, _f32p))
        return out

    def patchmatch(self, dp, src, smask, tgt, tmask, init=None, seed=1, iters=10, shots=30, window=-1,
                   with_metrics=False):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        smask = _c(smask, np.uint8)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        pm = PMParams(iters, shots, window)
        nnf = np.zeros((tH, tW, 2), np.int32)
        dist = np.zeros((tH, tW), np.float32)
        energy = np.zeros(max(iters, 1), np.float64)
        init = None if init is None else _c(init, np.int32)
        rc = self.lib.nlo_patchmatch(ct.byref(dp), ct.byref(pm), _p(src, _f32p), _p(smask, _u8p), sW, sH,
                                     _p(tgt, _f32p), _p(tmask, _u8p), tW, tH, C, _p(init, _i32p),
                                     ct.c_uint(seed), _p(nnf, _i32p), _p(dist, _f32p), _p(energy, _f64p))
        if rc != 0:
            raise ValueError("patchmatch failed")
        return (nnf, dist, energy[:iters]) if with_metrics else nnf

    def patchmatch_omp(self, dp, src, smask, tgt, tmask, init=None, seed=1, threads=0, iters=10, shots=30,
                       window=-1):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        smask = _c(smask, np.uint8)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        pm = PMParams(iters, shots, window)
        nnf = np.zeros((tH, tW, 2), np.int32)
        init = None if init is None else _c(init, np.int32)
        self.lib.nlo_patchmatch_omp(ct.byref(dp), ct.byref(pm), _p(src, _f32p), _p(smask, _u8p), sW, sH,
                                    _p(tgt, _f32p), _p(tmask, _u8p), tW, tH, C, _p(init, _i32p),
                                    ct.c_uint(seed), threads, _p(nnf, _i32p))
        return nnf

    # ---- updaters ----
    def update(self, kind, image, orig, domain, ext, nnf, conf, patch=9, sigma=10000.0, lam=0.05, lam_d=0.05,
               cg_tol=1e-6, cg_limit=1000, stats=False):
        """Returns (new image, total difference[, CGStats])."""
        img = np.array(image, dtype=np.float32, copy=True, order="C")
        orig = _c(orig, np.float32)
        domain = _c(domain, np.uint8)
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        conf = None if conf is None else _c(conf, np.float32)
        H, W, C = img.shape
        up = UpdParams(UPD[kind], patch, sigma, lam, lam_d, cg_tol, cg_limit)
        st = CGStats()
        total = self.lib.nlo_update(ct.byref(up), _p(img, _f32p), _p(orig, _f32p), _p(domain, _u8p),
                                    _p(ext, _u8p), _p(nnf, _i32p), _p(conf, _f32p), W, H, C, ct.byref(st))
        return (img, total, st) if stats else (img, total)

    def poisson_coefficients(self, image, ext, nnf, conf, patch=9, sigma=10000.0, lam=0.05):
        image = _c(image, np.float32)
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        conf = None if conf is None else _c(conf, np.float32)
        H, W, C = image.shape
        a1 = np.zeros((H, W)); a2 = np.zeros((H, W))
        f1 = np.zeros((C, H, W)); f2 = np.zeros((C, H, W))
        self.lib.nlo_poisson_coefficients(patch, ct.c_float(sigma), ct.c_float(lam), _p(image, _f32p),
                                          _p(ext, _u8p), _p(nnf, _i32p), _p(conf, _f32p), W, H, C,
                                          _p(a1, _f64p), _p(a2, _f64p), _p(f1, _f64p), _p(f2, _f64p))
        return a1, a2, f1, f2

    def anisotropic_laplacian(self, u, a, mask):
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        a = _c(a, np.float64)
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        self.lib.nlo_anisotropic_laplacian(_p(u, _f64p), _p(a, _f64p), _p(mask, _u8p), W, H)
        return u

    def conjugate_gradient(self, x0, a1, a2, b, mask, tol, limit):
        x = np.array(x0, dtype=np.float64, copy=True, order="C")
        a1 = _c(a1, np.float64); a2 = _c(a2, np.float64); b = _c(b, np.float64)
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        nr0 = ct.c_double(); nr = ct.c_double()
        it = self.lib.nlo_conjugate_gradient(_p(x, _f64p), _p(a1, _f64p), _p(a2, _f64p), _p(b, _f64p),
                                             _p(mask, _u8p), W, H, ct.c_float(tol), limit, ct.byref(nr0),
                                             ct.byref(nr))
        return x, it, nr0.value, nr.value

    def simpois(self, img, mask):
        """Coarsest-level Poisson fill: img (H, W, C), mask (H, W) 1 = hole."""
        img = _c(img, np.float32)
        H, W, C = img.shape
        planar = np.ascontiguousarray(np.transpose(img, (2, 0, 1))).copy()
        planar[:, mask.astype(bool)] = np.nan
        out = np.zeros_like(planar)
        self.lib.nlo_simpois(_p(out, _f32p), _p(planar, _f32p), W, H, C)
        return np.ascontiguousarray(np.transpose(out, (1, 2, 0)))

    def process(self, image, mask, params: InpaintParams):
        image = _c(image, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = image.shape
        out = np.zeros_like(image)
        st = ProcessStats()
        p = params.to_struct()
        rc = self.lib.nlo_process(ct.byref(p), _p(image, _f32p), _p(mask, _u8p), W, H, C, _p(out, _f32p),
                                  ct.byref(st))
        if rc == -2:
            raise RuntimeError("ERROR: Empty source mask")
        if rc != 0:
            raise ValueError(f"nlo_process failed ({rc})")
        return out, st


class Reference:
    """The reference's own sources compiled in place (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = ct.CDLL(path, mode=ct.RTLD_LOCAL)
        L = self.lib
        L.ref_update.restype = ct.c_double
        L.ref_subsampling_rate.restype = ct.c_float

    def gaussian_weights(self, sx, sy, sigx, sigy):
        out = np.zeros(sx * sy, np.float32)
        self.lib.ref_gaussian_weights(sx, sy, ct.c_float(sigx), ct.c_float(sigy), _p(out, _f32p))
        return out.reshape(sy, sx)

    def gradient(self, img):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, 2 * C), np.float32)
        self.lib.ref_gradient(_p(img, _f32p), W, H, C, _p(out, _f32p))
        return out

    def distance_transform(self, mask):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        out = np.zeros((H, W), np.float32)
        self.lib.ref_distance_transform(_p(mask, _u8p), W, H, _p(out, _f32p))
        return out

    def confidence(self, mask, decay=5.0, asym=0.1):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        out = np.zeros((H, W), np.float32)
        self.lib.ref_confidence(_p(mask, _u8p), W, H, ct.c_float(decay), ct.c_float(asym), _p(out, _f32p))
        return out

    def subsampling_rate(self, ratio, scales):
        return float(self.lib.ref_subsampling_rate(ct.c_float(ratio), scales))

    def downsample_with_mask(self, img, mask, factor, threshold=0.4, mask_val=False):
        img = _c(img, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = img.shape
        sw = int(self.lib.ref_sample_size(W, ct.c_float(factor)))
        sh = int(self.lib.ref_sample_size(H, ct.c_float(factor)))
        oi = np.zeros((sh, sw, C), np.float32)
        om = np.zeros((sh, sw), np.uint8)
        rc = self.lib.ref_downsample_with_mask(_p(img, _f32p), _p(mask, _u8p), W, H, C, ct.c_float(factor),
                                               ct.c_float(threshold), int(mask_val), _p(oi, _f32p), _p(om, _u8p))
        if rc != 0:
            raise ValueError("invalid factor")
        return oi, om

    def downsample(self, img, factor):
        img = _c(img, np.float32)
        H, W, C = img.shape
        sw = int(self.lib.ref_sample_size(W, ct.c_float(factor)))
        sh = int(self.lib.ref_sample_size(H, ct.c_float(factor)))
        out = np.zeros((sh, sw, C), np.float32)
        self.lib.ref_downsample(_p(img, _f32p), W, H, C, ct.c_float(factor), _p(out, _f32p))
        return out

    def downsample_mask(self, mask, factor, threshold=0.4):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        sw = int(self.lib.ref_sample_size(W, ct.c_float(factor)))
        sh = int(self.lib.ref_sample_size(H, ct.c_float(factor)))
        out = np.zeros((sh, sw), np.uint8)
        self.lib.ref_downsample_mask(_p(mask, _u8p), W, H, ct.c_float(factor), ct.c_float(threshold),
                                     _p(out, _u8p))
        return out

    def upsample(self, img, W2, H2):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H2, W2, C), np.float32)
        self.lib.ref_upsample(_p(img, _f32p), W, H, C, W2, H2, _p(out, _f32p))
        return out

    def extended_domain(self, hole, P):
        hole = _c(hole, np.uint8)
        H, W = hole.shape
        out = np.zeros((H, W), np.uint8)
        self.lib.ref_extended_domain(_p(hole, _u8p), W, H, P, _p(out, _u8p))
        return out

    def patch_distance_field(self, kind, src, tgt, nnf, tmask, patch=9, sigma=10000.0, lam=0.5, lam_d=0.5):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        nnf = _c(nnf, np.int32)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        out = np.zeros((tH, tW), np.float32)
        rc = self.lib.ref_patch_distance_field(DIST[kind], patch, ct.c_float(sigma), ct.c_float(lam),
                                               ct.c_float(lam_d), _p(src, _f32p), sW, sH, _p(tgt, _f32p), tW,
                                               tH, C, _p(nnf, _i32p), _p(tmask, _u8p), _p(out, _f32p))
        if rc != 0:
            raise RuntimeError("L2 RGBD Distance: Channels do not match.")
        return out

    def patchmatch(self, kind, src, smask, tgt, tmask, init=None, seed=1, patch=9, sigma=10000.0, lam=0.5,
                   lam_d=0.5, iters=10, shots=30, window=-1):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        smask = _c(smask, np.uint8)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        nnf = np.zeros((tH, tW, 2), np.int32)
        init = None if init is None else _c(init, np.int32)
        rc = self.lib.ref_patchmatch(DIST[kind], patch, ct.c_float(sigma), ct.c_float(lam), ct.c_float(lam_d),
                                     iters, shots, window, _p(src, _f32p), _p(smask, _u8p), sW, sH,
                                     _p(tgt, _f32p), _p(tmask, _u8p), tW, tH, C, _p(init, _i32p),
                                     ct.c_uint(seed), _p(nnf, _i32p))
        if rc != 0:
            raise ValueError("patchmatch returned an empty field")
        return nnf

    def update(self, kind, image, orig, domain, ext, nnf, conf, patch=9, sigma=10000.0, lam=0.05, lam_d=0.05,
               cg_tol=1e-6, cg_limit=1000):
        img = np.array(image, dtype=np.float32, copy=True, order="C")
        orig = _c(orig, np.float32)
        domain = _c(domain, np.uint8)
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        conf = None if conf is None else _c(conf, np.float32)
        H, W, C = img.shape
        total = self.lib.ref_update(UPD[kind], patch, ct.c_float(sigma), ct.c_float(lam), ct.c_float(lam_d),
                                    ct.c_float(cg_tol), cg_limit, _p(img, _f32p), _p(orig, _f32p),
                                    _p(domain, _u8p), _p(ext, _u8p), _p(nnf, _i32p), _p(conf, _f32p), W, H, C)
        return img, total

    def process(self, image, mask, params: InpaintParams):
        image = _c(image, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = image.shape
        out = np.zeros_like(image)
        p = params
        rc = self.lib.ref_process(UPD[p.method], p.patch, ct.c_float(p.sigma), ct.c_float(p.lam),
                                  ct.c_float(p.lambda_rgb), ct.c_float(p.lambda_d), p.iters, ct.c_float(p.tol),
                                  p.scales, ct.c_float(p.coarse), ct.c_float(p.conf_t), ct.c_float(p.conf_a),
                                  INIT[p.init], p.pm_iters, p.pm_shots, p.pm_window, ct.c_float(p.cg_tol),
                                  p.cg_limit, ct.c_uint(p.seed), _p(image, _f32p), _p(mask, _u8p), W, H, C,
                                  _p(out, _f32p))
        if rc == -2:
            raise RuntimeError("ERROR: Empty source mask")
        if rc != 0:
            raise ValueError("ref_process failed")
        return out


class ReferenceColor:
    """The reference's IOUtility::rgb_to_lab / lab_to_rgb (io_utility.cpp:268-318),
    compiled in place (oracle/_ref/libnli_ref_color.so); 3-channel images only."""

    def __init__(self, path: str = REF_COLOR_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = ct.CDLL(path, mode=ct.RTLD_LOCAL)

    def _run(self, fn, img):
        img = _c(img, np.float32)
        H, W, C = img.shape
        if C != 3:
            raise ValueError("3-channel images only")
        out = np.zeros_like(img)
        fn(_p(img, _f32p), W, H, _p(out, _f32p))
        return out

    def rgb_to_lab(self, img):
        return self._run(self.lib.ref_rgb_to_lab, img)

    def lab_to_rgb(self, img):
        return self._run(self.lib.ref_lab_to_rgb, img)


def nnf_energy(oracle: Oracle, dp: DistParams, img, nnf, tmask) -> float:
    """Sum over target points of the patch distance to their NNF entry."""
    d = oracle.patch_distance_field(dp, img, img, nnf, tmask)
    sel = tmask.astype(bool) & (nnf[..., 0] >= 0)
    return float(np.sum(d[sel].astype(np.float64)))
