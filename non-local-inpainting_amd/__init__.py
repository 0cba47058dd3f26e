"""Python binding of libnli, the MI355X (gfx950) implementation of the inner
inpainting loop of Tom-Zheng/non-local-inpainting.

All computation happens in ``libnli.so`` (hand-written HIP kernels behind the
C-ABI of ``include/nli.h``); this module only moves numpy arrays across that
boundary with ctypes.  There is no CPU fallback: if the library is missing or
no GPU is visible, the calls raise.

Arrays follow the reference containers: images float32 (H, W, C) HWC, masks
uint8 (H, W) with 1 = in the set, NNF int32 (H, W, 2) holding (x, y) with
(-1, -1) = no entry.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnli.so")
PLUGINS_PATH = os.path.join(HERE, "libnli_plugins.so")

NLI_OK = 0
NLI_E_ARGUMENT = -1
NLI_E_DEVICE = -2
NLI_E_EMPTY_SOURCE = -3
NLI_E_INTERNAL = -4
NLI_E_IO = -5

DIST = {"l2": 0, "l1": 1, "l2combined": 2, "l2rgbd": 3}
UPD = {"nlmeans": 0, "nlmedians": 1, "nlpoisson": 2, "nldepth": 3}
INIT = {"black": 0, "avg": 1, "none": 2, "poisson": 3}
CG_FP64, CG_MIXED = 0, 1

# every symbol include/nli.h declares
EXPORTS = [
    "nli_ctx_create", "nli_ctx_destroy", "nli_last_error", "nli_synchronize", "nli_version",
    "nli_dev_alloc", "nli_dev_free", "nli_copy_h2d", "nli_copy_d2h", "nli_copy_d2d",
    "nli_gradient", "nli_distance_transform", "nli_confidence", "nli_sample_size",
    "nli_subsampling_rate", "nli_downsample_with_mask", "nli_downsample", "nli_downsample_mask",
    "nli_upsample", "nli_extended_domain", "nli_scale_masks", "nli_scale_nnf", "nli_poisson_fill_dev", "nli_patch_distance",
    "nli_patchmatch", "nli_patchmatch_dev", "nli_update", "nli_update_dev", "nli_poisson_fill",
    "nli_inpaint", "nli_inpaint_dev", "nli_inpaint_params_default", "nli_profile_enable",
    "nli_patch_energy", "nli_energy_gradient", "nli_patch_extract", "nli_adjoint_aggregate",
    "nli_pm_tiles_per_block", "nli_pm_row_scan", "nli_patch_distance_dev_noalloc", "nli_batch_mode",
    "nli_profile_read", "nli_inpaint_batch", "nli_inpaint_batch_dev", "nli_inpaint_batch_each",
    "nli_inpaint_batch_each_dev", "nli_rgb_to_lab", "nli_lab_to_rgb", "nli_rgb_to_lab_dev",
    "nli_lab_to_rgb_dev", "nli_inpaint_rgb", "nli_inpaint_rgb_dev", "nli_profile_enable_sampled",
]
# symbols of libnli_plugins.so (include/nli_plugins.hpp)
PLUGIN_EXPORTS = ["nli_plain_inpaint", "nli_plain_inpaint_rgb", "nli_plain_last_error",
                  # include/nli_io.h
                  "nli_image_info", "nli_read_image", "nli_read_rgb_image", "nli_read_mono_image",
                  "nli_write_image", "nli_write_rgb_image", "nli_write_mono_image", "nli_io_last_error",
                  "nli_pgm_info", "nli_read_pgm_image", "nli_write_pgm_image"]
KERNELS = {"pm_sweep": 0, "pm_init": 1, "cg_iteration": 2, "poisson_coef": 3, "nl_gather": 4, "color": 6}


class NliError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code


class EmptySourceError(NliError):
    pass


class DistParams(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("patch", ct.c_int), ("sigma", ct.c_float),
                ("lambda_", ct.c_float), ("lambda_d", ct.c_float)]


class PMParams(ct.Structure):
    _fields_ = [("iterations", ct.c_int), ("shots", ct.c_int), ("window", ct.c_int),
                ("first_step", ct.c_int)]


class UpdParams(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("patch", ct.c_int), ("sigma", ct.c_float),
                ("lambda_", ct.c_float), ("lambda_d", ct.c_float), ("cg_tol", ct.c_float),
                ("cg_limit", ct.c_int), ("cg_precision", ct.c_int), ("cg_refine_every", ct.c_int)]


class CGStats(ct.Structure):
    _fields_ = [("channels", ct.c_int), ("cg_iterations", ct.c_int * 4),
                ("cg_nr0", ct.c_double * 4), ("cg_nr", ct.c_double * 4)]


class InpaintParams(ct.Structure):
    _fields_ = [("method", ct.c_int), ("patch", ct.c_int), ("sigma", ct.c_float),
                ("lambda_", ct.c_float), ("lambda_rgb", ct.c_float), ("lambda_d", ct.c_float),
                ("iters", ct.c_int), ("tol", ct.c_float), ("scales", ct.c_int),
                ("coarse", ct.c_float), ("conf_t", ct.c_float), ("conf_a", ct.c_float),
                ("init", ct.c_int), ("pm_iters", ct.c_int), ("pm_shots", ct.c_int),
                ("pm_window", ct.c_int), ("cg_tol", ct.c_float), ("cg_limit", ct.c_int),
                ("cg_precision", ct.c_int), ("cg_refine_every", ct.c_int), ("seed", ct.c_uint64)]


class EnergyParams(ct.Structure):
    _fields_ = [("patch", ct.c_int), ("sigma", ct.c_float), ("lambda_", ct.c_float)]


class InpaintStats(ct.Structure):
    _fields_ = [("scales", ct.c_int), ("iterations", ct.c_int * 32), ("width", ct.c_int * 32),
                ("height", ct.c_int * 32), ("hole_pixels", ct.c_int * 32),
                ("target_pixels", ct.c_int * 32), ("cg_iterations_total", ct.c_int),
                ("seconds_total", ct.c_double), ("seconds_patchmatch", ct.c_double),
                ("seconds_update", ct.c_double), ("seconds_other", ct.c_double)]


_f32p = ct.POINTER(ct.c_float)
_u8p = ct.POINTER(ct.c_uint8)
_i32p = ct.POINTER(ct.c_int32)
_lib = None


def load_library(path: str = LIB_PATH) -> ct.CDLL:
    """Load libnli.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run __graft_entry__.build() "
                          "(make -C non-local-inpainting_amd)")
    lib = ct.CDLL(path)
    lib.nli_last_error.restype = ct.c_char_p
    lib.nli_subsampling_rate.restype = ct.c_float
    lib.nli_subsampling_rate.argtypes = [ct.c_float, ct.c_int]
    lib.nli_sample_size.argtypes = [ct.c_int, ct.c_float]
    lib.nli_ctx_create.argtypes = [ct.c_int, ct.POINTER(ct.c_void_p)]
    lib.nli_ctx_destroy.argtypes = [ct.c_void_p]
    lib.nli_last_error.argtypes = [ct.c_void_p]
    _lib = lib
    return lib


_plugins = None


def load_plugins() -> ct.CDLL:
    """Load libnli_plugins.so, the C++ plugin classes composed as the reference's main()
    (raises if it has not been built)."""
    global _plugins
    if _plugins is not None:
        return _plugins
    load_library()
    if not os.path.exists(PLUGINS_PATH):
        raise ImportError(f"{PLUGINS_PATH} is not built: run __graft_entry__.build()")
    lib = ct.CDLL(PLUGINS_PATH)
    lib.nli_plain_last_error.restype = ct.c_char_p
    lib.nli_io_last_error.restype = ct.c_char_p
    lib.nli_read_image.argtypes = [ct.c_char_p, _f32p, ct.c_long]
    lib.nli_read_rgb_image.argtypes = [ct.c_char_p, _f32p, ct.c_long]
    lib.nli_read_mono_image.argtypes = [ct.c_char_p, _f32p, ct.c_long]
    lib.nli_read_pgm_image.argtypes = [ct.c_char_p, _f32p, ct.c_long]
    _plugins = lib
    return lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


@dataclass
class Inpainting:
    """Parameters of ImageInpainting::process with its method pairing
    (main.cpp:245-266).  Defaults are the reference CLI's, patch side 9."""
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
    cg_precision: int = CG_FP64
    cg_refine_every: int = 0
    seed: int = 1

    def to_struct(self) -> InpaintParams:
        return InpaintParams(UPD[self.method], self.patch, self.sigma, self.lam, self.lambda_rgb,
                             self.lambda_d, self.iters, self.tol, self.scales, self.coarse,
                             self.conf_t, self.conf_a, INIT[self.init], self.pm_iters, self.pm_shots,
                             self.pm_window, self.cg_tol, self.cg_limit, self.cg_precision,
                             self.cg_refine_every, self.seed)


@dataclass
class RunStats:
    scales: int
    iterations: list
    sizes: list
    hole_pixels: list
    target_pixels: list
    cg_iterations_total: int
    seconds_total: float
    seconds_patchmatch: float
    seconds_update: float
    extra: dict = field(default_factory=dict)


class Context:
    """One device context (stream + device buffers) of libnli."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = ct.c_void_p()
        rc = self.lib.nli_ctx_create(device, ct.byref(h))
        if rc != NLI_OK:
            raise NliError(rc, f"nli_ctx_create(device={device}) failed: no usable GPU")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nli_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc == NLI_OK:
            return
        msg = self.lib.nli_last_error(self.h).decode(errors="replace")
        if rc == NLI_E_EMPTY_SOURCE:
            raise EmptySourceError(rc, msg)
        if rc == NLI_E_ARGUMENT:
            raise ValueError(msg)
        raise NliError(rc, msg)

    # ---- numerics ----
    def gradient(self, img):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, 2 * C), np.float32)
        self._check(self.lib.nli_gradient(self.h, _ptr(img, _f32p), W, H, C, _ptr(out, _f32p)))
        return out

    def distance_transform(self, mask):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        out = np.zeros((H, W), np.float32)
        self._check(self.lib.nli_distance_transform(self.h, _ptr(mask, _u8p), W, H, _ptr(out, _f32p)))
        return out

    def confidence(self, mask, decay=5.0, asym=0.1):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        out = np.zeros((H, W), np.float32)
        self._check(self.lib.nli_confidence(self.h, _ptr(mask, _u8p), W, H, ct.c_float(decay),
                                            ct.c_float(asym), _ptr(out, _f32p)))
        return out

    def sample_size(self, size, factor):
        return int(self.lib.nli_sample_size(size, factor))

    def downsample_with_mask(self, img, mask, factor, threshold=0.4, mask_val=False):
        img = _c(img, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = img.shape
        sw, sh = self.sample_size(W, factor), self.sample_size(H, factor)
        oi = np.zeros((sh, sw, C), np.float32)
        om = np.zeros((sh, sw), np.uint8)
        self._check(self.lib.nli_downsample_with_mask(self.h, _ptr(img, _f32p), _ptr(mask, _u8p), W, H, C,
                                                      ct.c_float(factor), ct.c_float(threshold), int(mask_val),
                                                      _ptr(oi, _f32p), _ptr(om, _u8p)))
        return oi, om

    def downsample(self, img, factor):
        img = _c(img, np.float32)
        H, W, C = img.shape
        sw, sh = self.sample_size(W, factor), self.sample_size(H, factor)
        out = np.zeros((sh, sw, C), np.float32)
        self._check(self.lib.nli_downsample(self.h, _ptr(img, _f32p), W, H, C, ct.c_float(factor),
                                            _ptr(out, _f32p)))
        return out

    def downsample_mask(self, mask, factor, threshold=0.4):
        mask = _c(mask, np.uint8)
        H, W = mask.shape
        sw, sh = self.sample_size(W, factor), self.sample_size(H, factor)
        out = np.zeros((sh, sw), np.uint8)
        self._check(self.lib.nli_downsample_mask(self.h, _ptr(mask, _u8p), W, H, ct.c_float(factor),
                                                 ct.c_float(threshold), _ptr(out, _u8p)))
        return out

    def upsample(self, img, W2, H2):
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H2, W2, C), np.float32)
        self._check(self.lib.nli_upsample(self.h, _ptr(img, _f32p), W, H, C, W2, H2, _ptr(out, _f32p)))
        return out

    # ---- masks ----
    def extended_domain(self, hole, patch):
        hole = _c(hole, np.uint8)
        H, W = hole.shape
        out = np.zeros((H, W), np.uint8)
        self._check(self.lib.nli_extended_domain(self.h, _ptr(hole, _u8p), W, H, patch, _ptr(out, _u8p)))
        return out

    def scale_masks(self, hole, patch, erode=True):
        """(extended, target, source) of one scale; raises EmptySourceError when the
        eroded source is empty (erode=True only)."""
        hole = _c(hole, np.uint8)
        H, W = hole.shape
        e, t, s = (np.zeros((H, W), np.uint8) for _ in range(3))
        self._check(self.lib.nli_scale_masks(self.h, _ptr(hole, _u8p), W, H, patch, int(erode), _ptr(e, _u8p),
                                             _ptr(t, _u8p), _ptr(s, _u8p)))
        return e, t, s

    # ---- distances / PatchMatch ----
    def patch_distance(self, kind, src, tgt, nnf, tmask, patch=9, sigma=10000.0, lam=0.5, lam_d=0.5):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        nnf = _c(nnf, np.int32)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        dp = DistParams(DIST[kind], patch, sigma, lam, lam_d)
        out = np.zeros((tH, tW), np.float32)
        self._check(self.lib.nli_patch_distance(self.h, ct.byref(dp), _ptr(src, _f32p), sW, sH, _ptr(tgt, _f32p),
                                                tW, tH, C, _ptr(nnf, _i32p), _ptr(tmask, _u8p), _ptr(out, _f32p)))
        return out

    def patchmatch(self, kind, src, smask, tgt, tmask, init=None, seed=1, patch=9, sigma=10000.0, lam=0.5,
                   lam_d=0.5, iters=10, shots=30, window=-1, first_step=0, with_dist=False):
        src = _c(src, np.float32)
        tgt = _c(tgt, np.float32)
        smask = _c(smask, np.uint8)
        tmask = _c(tmask, np.uint8)
        sH, sW, C = src.shape
        tH, tW, _ = tgt.shape
        dp = DistParams(DIST[kind], patch, sigma, lam, lam_d)
        pm = PMParams(iters, shots, window, first_step)
        nnf = np.zeros((tH, tW, 2), np.int32)
        dist = np.zeros((tH, tW), np.float32) if with_dist else None
        init = None if init is None else _c(init, np.int32)
        self._check(self.lib.nli_patchmatch(self.h, ct.byref(dp), ct.byref(pm), _ptr(src, _f32p), _ptr(smask, _u8p),
                                            sW, sH, _ptr(tgt, _f32p), _ptr(tmask, _u8p), tW, tH, C,
                                            _ptr(init, _i32p), ct.c_uint64(seed), _ptr(nnf, _i32p),
                                            _ptr(dist, _f32p)))
        return (nnf, dist) if with_dist else nnf

    # ---- updates ----
    def update(self, kind, image, orig, domain, ext, nnf, conf, patch=9, sigma=10000.0, lam=0.05, lam_d=0.05,
               cg_tol=1e-6, cg_limit=1000, cg_precision=CG_FP64, cg_refine_every=0, stats=False):
        """AImageUpdating::update: returns (new image, total difference[, CGStats])."""
        img = np.array(image, dtype=np.float32, copy=True, order="C")
        orig = _c(orig, np.float32)
        domain = _c(domain, np.uint8)
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        conf = None if conf is None else _c(conf, np.float32)
        H, W, C = img.shape
        up = UpdParams(UPD[kind], patch, sigma, lam, lam_d, cg_tol, cg_limit, cg_precision, cg_refine_every)
        total = ct.c_double()
        st = CGStats()
        self._check(self.lib.nli_update(self.h, ct.byref(up), _ptr(img, _f32p), _ptr(orig, _f32p),
                                        _ptr(domain, _u8p), _ptr(ext, _u8p), _ptr(nnf, _i32p), _ptr(conf, _f32p),
                                        W, H, C, ct.byref(total), ct.byref(st)))
        return (img, total.value, st) if stats else (img, total.value)

    # ---- kernel timing ----
    def profile(self, on=True, period=1):
        """Time kernels with HIP events: True = all, False = none, or an iterable of
        kernel names (keys of KERNELS); one launch in `period` of each is timed.  Resets
        the totals."""
        if on is True or on is False:
            mask = -1 if on else 0
        else:
            mask = 0
            for k in on:
                mask |= 1 << KERNELS[k]
        self._check(self.lib.nli_profile_enable_sampled(self.h, mask, int(period)))

    def profile_read(self, kernel: str) -> dict:
        n = ct.c_int()
        ms, work, units = ct.c_double(), ct.c_double(), ct.c_double()
        self._check(self.lib.nli_profile_read(self.h, KERNELS[kernel], ct.byref(n), ct.byref(ms), ct.byref(work),
                                              ct.byref(units)))
        return dict(launches=n.value, ms=ms.value, work=work.value, units=units.value)

    # ---- device buffers ----
    def dev_alloc(self, nbytes: int) -> int:
        p = ct.c_void_p()
        self._check(self.lib.nli_dev_alloc(self.h, ct.c_uint64(nbytes), ct.byref(p)))
        return p.value

    def dev_free(self, p: int):
        self._check(self.lib.nli_dev_free(self.h, ct.c_void_p(p)))

    def to_device(self, arr) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.dev_alloc(arr.nbytes)
        self._check(self.lib.nli_copy_h2d(self.h, ct.c_void_p(p), arr.ctypes.data_as(ct.c_void_p),
                                          ct.c_uint64(arr.nbytes)))
        return p

    def from_device(self, p: int, shape, dtype):
        out = np.empty(shape, dtype)
        self._check(self.lib.nli_copy_d2h(self.h, out.ctypes.data_as(ct.c_void_p), ct.c_void_p(p),
                                          ct.c_uint64(out.nbytes)))
        return out

    # ---- the NL-Poisson patch energy and its adjoint ----
    def patch_energy(self, image, ext, nnf, conf=None, patch=9, sigma=10000.0, lam=0.05, touch=None, delta=0.0):
        """E(u) for a fixed NNF; touch=(x, y, c) sums only the terms reading that pixel,
        evaluated with its value + delta (FP64)."""
        image = _c(image, np.float32)
        H, W, C = image.shape
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        conf = None if conf is None else _c(conf, np.float32)
        px, py, pc = (int(v) for v in touch) if touch is not None else (-1, -1, -1)
        e = ct.c_double()
        p = EnergyParams(patch, sigma, lam)
        self._check(self.lib.nli_patch_energy(self.h, ct.byref(p), _ptr(image, _f32p), _ptr(ext, _u8p),
                                              _ptr(nnf, _i32p), _ptr(conf, _f32p), W, H, C, px, py, pc,
                                              ct.c_double(delta), ct.byref(e)))
        return e.value

    def energy_gradient(self, image, ext, nnf, conf=None, patch=9, sigma=10000.0, lam=0.05):
        image = _c(image, np.float32)
        H, W, C = image.shape
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        conf = None if conf is None else _c(conf, np.float32)
        g = np.zeros((H, W, C), np.float64)
        p = EnergyParams(patch, sigma, lam)
        self._check(self.lib.nli_energy_gradient(self.h, ct.byref(p), _ptr(image, _f32p), _ptr(ext, _u8p),
                                                 _ptr(nnf, _i32p), _ptr(conf, _f32p), W, H, C,
                                                 g.ctypes.data_as(ct.POINTER(ct.c_double))))
        return g

    def patch_extract(self, u, ext, nnf, patch=9, side=0):
        u = _c(u, np.float64)
        H, W, C = u.shape
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        ne = int(ext.astype(bool).sum())
        out = np.zeros((ne, patch * patch, C), np.float64)
        n = ct.c_int()
        self._check(self.lib.nli_patch_extract(self.h, patch, u.ctypes.data_as(ct.POINTER(ct.c_double)),
                                               _ptr(ext, _u8p), _ptr(nnf, _i32p), W, H, C, side,
                                               out.ctypes.data_as(ct.POINTER(ct.c_double)),
                                               ct.c_uint64(out.size), ct.byref(n)))
        assert n.value == ne
        return out

    def adjoint_aggregate(self, patches, ext, nnf, patch=9, side=0):
        patches = _c(patches, np.float64)
        ext = _c(ext, np.uint8)
        nnf = _c(nnf, np.int32)
        H, W = ext.shape
        C = patches.shape[-1]
        out = np.zeros((H, W, C), np.float64)
        self._check(self.lib.nli_adjoint_aggregate(self.h, patch, patches.ctypes.data_as(ct.POINTER(ct.c_double)),
                                                   patches.shape[0], _ptr(ext, _u8p), _ptr(nnf, _i32p), W, H, C,
                                                   side, out.ctypes.data_as(ct.POINTER(ct.c_double))))
        return out

    def poisson_fill(self, img, mask):
        """Coarsest-level Poisson fill on the device (bit-identical to the host fill)."""
        img = _c(img, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = img.shape
        d_img = self.to_device(img)
        d_mask = self.to_device(mask)
        try:
            self._check(self.lib.nli_poisson_fill_dev(self.h, ct.c_void_p(d_img), ct.c_void_p(d_mask), W, H, C))
            return self.from_device(d_img, img.shape, np.float32)
        finally:
            self.dev_free(d_img)
            self.dev_free(d_mask)

    def copy_d2d(self, dst: int, src: int, nbytes: int):
        self._check(self.lib.nli_copy_d2d(self.h, ct.c_void_p(dst), ct.c_void_p(src), ct.c_uint64(nbytes)))

    def inpaint_dev(self, d_in: int, d_mask: int, d_out: int, W: int, H: int, C: int, params: "Inpainting"):
        """The device-resident process(): inputs and output already in device memory."""
        p = params.to_struct()
        st = InpaintStats()
        self._check(self.lib.nli_inpaint_dev(self.h, ct.byref(p), ct.c_void_p(d_in), ct.c_void_p(d_mask), W, H, C,
                                             ct.c_void_p(d_out), ct.byref(st)))
        return stats_of(st)

    # ---- batches ----
    def inpaint_batch(self, images, masks, params, workers: int = 0):
        """Inpaints a list of images (uneven sizes allowed) concurrently; params is one
        Inpainting for all images or a list with one per image.  Returns
        (outputs, [RunStats])."""
        images = [_c(im, np.float32) for im in images]
        masks = [_c(m, np.uint8) for m in masks]
        n = len(images)
        C = images[0].shape[2] if n else 3
        outs = [np.zeros_like(im) for im in images]
        PF = ct.POINTER(ct.c_float)
        PU = ct.POINTER(ct.c_uint8)
        ins = (PF * n)(*[_ptr(im, _f32p) for im in images])
        ms = (PU * n)(*[_ptr(m, _u8p) for m in masks])
        os_ = (PF * n)(*[_ptr(o, _f32p) for o in outs])
        Ws = (ct.c_int * n)(*[im.shape[1] for im in images])
        Hs = (ct.c_int * n)(*[im.shape[0] for im in images])
        st = (InpaintStats * n)()
        if isinstance(params, (list, tuple)):
            assert len(params) == n
            ps = (InpaintParams * n)(*[q.to_struct() for q in params])
            self._check(self.lib.nli_inpaint_batch_each(self.h, ps, n, ins, ms, Ws, Hs, C, os_, workers, st))
        else:
            p = params.to_struct()
            self._check(self.lib.nli_inpaint_batch(self.h, ct.byref(p), n, ins, ms, Ws, Hs, C, os_, workers, st))
        return outs, [stats_of(st[i]) for i in range(n)]

    def inpaint_batch_dev(self, d_ins, d_masks, d_outs, Ws, Hs, C, params, workers: int = 0):
        """Device-resident batch: lists of device pointers (from to_device / dev_alloc);
        params as in inpaint_batch."""
        n = len(d_ins)
        V = ct.c_void_p
        st = (InpaintStats * n)()
        args = ((V * n)(*d_ins), (V * n)(*d_masks), (ct.c_int * n)(*Ws), (ct.c_int * n)(*Hs), C, (V * n)(*d_outs),
                workers, st)
        if isinstance(params, (list, tuple)):
            assert len(params) == n
            ps = (InpaintParams * n)(*[q.to_struct() for q in params])
            self._check(self.lib.nli_inpaint_batch_each_dev(self.h, ps, n, *args))
        else:
            p = params.to_struct()
            self._check(self.lib.nli_inpaint_batch_dev(self.h, ct.byref(p), n, *args))
        return [stats_of(st[i]) for i in range(n)]

    # ---- full process ----
    def inpaint(self, image, mask, params: Inpainting):
        image = _c(image, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = image.shape
        out = np.zeros_like(image)
        p = params.to_struct()
        st = InpaintStats()
        self._check(self.lib.nli_inpaint(self.h, ct.byref(p), _ptr(image, _f32p), _ptr(mask, _u8p), W, H, C,
                                         _ptr(out, _f32p), ct.byref(st)))
        return out, stats_of(st)


    # ---- the colour path around process() (main.cpp:117, 146, 307-309) ----
    def rgb_to_lab(self, img):
        """IOUtility::rgb_to_lab (io_utility.cpp:268-289): 0-255 RGB to L*a*b*; images
        that do not have 3 channels come back unchanged."""
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, C), np.float32)
        self._check(self.lib.nli_rgb_to_lab(self.h, _ptr(img, _f32p), W, H, C, _ptr(out, _f32p)))
        return out

    def lab_to_rgb(self, img):
        """IOUtility::lab_to_rgb (io_utility.cpp:292-318): L*a*b* (3 or 4 channels; the
        fourth is not converted) to 0-255 RGB; other channel counts come back unchanged."""
        img = _c(img, np.float32)
        H, W, C = img.shape
        out = np.zeros((H, W, 3 if C in (3, 4) else C), np.float32)
        self._check(self.lib.nli_lab_to_rgb(self.h, _ptr(img, _f32p), W, H, C, _ptr(out, _f32p)))
        return out

    def rgb_to_lab_dev(self, d_in: int, d_out: int, W: int, H: int, C: int = 3):
        self._check(self.lib.nli_rgb_to_lab_dev(self.h, ct.c_void_p(d_in), W, H, C, ct.c_void_p(d_out)))

    def lab_to_rgb_dev(self, d_in: int, d_out: int, W: int, H: int, C: int = 3):
        self._check(self.lib.nli_lab_to_rgb_dev(self.h, ct.c_void_p(d_in), W, H, C, ct.c_void_p(d_out)))

    def inpaint_rgb(self, rgb, mask, params: Inpainting, depth=None):
        """main.cpp:115-309 without the file I/O: 0-255 RGB (and an optional depth plane)
        in, Lab inpainting, 0-255 RGB (and depth) out.  Returns (rgb, depth or None, stats)."""
        rgb = _c(rgb, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = rgb.shape
        if C != 3:
            raise ValueError("rgb must be H x W x 3")
        out = np.zeros_like(rgb)
        d_in = _c(depth, np.float32) if depth is not None else None
        d_out = np.zeros((H, W), np.float32) if depth is not None else None
        p = params.to_struct()
        st = InpaintStats()
        self._check(self.lib.nli_inpaint_rgb(self.h, ct.byref(p), _ptr(rgb, _f32p), _ptr(d_in, _f32p),
                                             _ptr(mask, _u8p), W, H, _ptr(out, _f32p), _ptr(d_out, _f32p),
                                             ct.byref(st)))
        return out, d_out, stats_of(st)

    def inpaint_rgb_dev(self, d_rgb: int, d_mask: int, d_out: int, W: int, H: int, params: "Inpainting",
                        d_depth: int = 0, d_depth_out: int = 0):
        """inpaint_rgb on device buffers (d_depth / d_depth_out 0: no depth plane)."""
        p = params.to_struct()
        st = InpaintStats()
        V = ct.c_void_p
        self._check(self.lib.nli_inpaint_rgb_dev(self.h, ct.byref(p), V(d_rgb), V(d_depth or None), V(d_mask), W, H,
                                                 V(d_out), V(d_depth_out or None), ct.byref(st)))
        return stats_of(st)

    def inpaint_png(self, input_name: str, mask_name: str, output_name: str, params: Inpainting,
                    depth_name: str = None, depth_output_name: str = None):
        """main.cpp:115-312 with files: IOUtility::read_rgb_image (input), read_mono_image
        (mask, taken where positive; depth), inpaint_rgb, write_rgb_image (output).  The
        depth plane is written as a grey PNG (write_mono_image); the reference passes the
        one-channel plane to write_rgb_image, which reads channels 1 and 2 from the next
        pixels and past the end of the plane.  Returns the run's statistics."""
        rgb = read_rgb_image(input_name)
        mask_image = read_mono_image(mask_name)
        if rgb.shape[:2] != mask_image.shape:
            raise ValueError("ERROR: Input image and mask should be of the same size")
        mask = (mask_image > 0).astype(np.uint8)
        if not mask.any():
            raise ValueError("ERROR: Mask should not be empty (all black)")
        depth = read_mono_image(depth_name) if depth_name else None
        out, d_out, stats = self.inpaint_rgb(rgb, mask, params, depth=depth)
        write_rgb_image(output_name, out)
        if depth is not None and depth_output_name:
            write_mono_image(depth_output_name, d_out)
        return stats

    def inpaint_rgb_plain(self, rgb, mask, params: Inpainting, depth=None):
        """main.cpp:115-309 through the C++ plugin classes (IOUtility, the plain drop-in);
        returns (rgb, depth or None, stats) like inpaint_rgb."""
        lib = load_plugins()
        rgb = _c(rgb, np.float32)
        mask = _c(mask, np.uint8)
        H, W, _ = rgb.shape
        out = np.zeros_like(rgb)
        d_in = _c(depth, np.float32) if depth is not None else None
        d_out = np.zeros((H, W), np.float32) if depth is not None else None
        p = params.to_struct()
        st = InpaintStats()
        rc = lib.nli_plain_inpaint_rgb(self.h, ct.byref(p), _ptr(rgb, _f32p), _ptr(d_in, _f32p), _ptr(mask, _u8p),
                                       W, H, _ptr(out, _f32p), _ptr(d_out, _f32p), ct.byref(st))
        if rc != NLI_OK:
            msg = lib.nli_plain_last_error().decode(errors="replace")
            if rc == NLI_E_EMPTY_SOURCE:
                raise EmptySourceError(rc, msg)
            if rc == NLI_E_ARGUMENT:
                raise ValueError(msg)
            raise NliError(rc, msg)
        return out, d_out, stats_of(st)

    def inpaint_plain(self, image, mask, params: Inpainting):
        """The plain drop-in: the reference's main() composition of the C++ plugin classes
        (HipPatchMatch, HipPatchNonLocal*, ImageInpainting), one C-ABI call per reference
        call on host buffers."""
        lib = load_plugins()
        image = _c(image, np.float32)
        mask = _c(mask, np.uint8)
        H, W, C = image.shape
        out = np.zeros_like(image)
        p = params.to_struct()
        st = InpaintStats()
        rc = lib.nli_plain_inpaint(self.h, ct.byref(p), _ptr(image, _f32p), _ptr(mask, _u8p), W, H, C,
                                   _ptr(out, _f32p), ct.byref(st))
        if rc != NLI_OK:
            msg = lib.nli_plain_last_error().decode(errors="replace")
            if rc == NLI_E_EMPTY_SOURCE:
                raise EmptySourceError(rc, msg)
            raise NliError(rc, msg)
        return out, stats_of(st)


# ---- image files (include/nli_io.h; io_utility.cpp:119-220 over 3rdparty/iio/iio.c) ----
def _io_check(rc: int):
    if rc == NLI_OK:
        return
    msg = load_plugins().nli_io_last_error().decode(errors="replace")
    if rc == NLI_E_ARGUMENT:
        raise ValueError(msg)
    if rc == NLI_E_IO:
        raise OSError(msg)
    raise NliError(rc, msg)


def image_info(path: str):
    """(W, H, channels, bit_depth) of a PNG as the reader returns it."""
    lib = load_plugins()
    v = [ct.c_int() for _ in range(4)]
    _io_check(lib.nli_image_info(os.fsencode(path), *[ct.byref(x) for x in v]))
    return tuple(x.value for x in v)


def _read(fn, path: str, channels):
    W, H, C, _ = image_info(path)
    C = C if channels is None else channels
    out = np.zeros((H, W, C), np.float32)
    _io_check(fn(os.fsencode(path), _ptr(out, _f32p), out.size))
    return out


def read_image(path: str):
    """iio_read_image_float_vec: (H, W, C) float32, palette and low-bit grey expanded,
    tRNS as alpha, 8-bit samples 0..255 and 16-bit samples 0..65535."""
    return _read(load_plugins().nli_read_image, path, None)


def read_rgb_image(path: str):
    """IOUtility::read_rgb_image (io_utility.cpp:143-168): (H, W, 3) float32."""
    return _read(load_plugins().nli_read_rgb_image, path, 3)


def read_mono_image(path: str):
    """IOUtility::read_mono_image (io_utility.cpp:119-141): (H, W) float32; 8-bit colour
    becomes (uint8)(.299 R + .587 G + .114 B)."""
    return _read(load_plugins().nli_read_mono_image, path, 1)[:, :, 0]


def _write(fn, path: str, img):
    img = _c(img, np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    H, W, C = img.shape
    _io_check(fn(os.fsencode(path), _ptr(img, _f32p), W, H, C))


def write_image(path: str, img):
    """iio_save_image_float_vec with a .png name: 8-bit samples (uint8)clamp(v + 0.5)."""
    _write(load_plugins().nli_write_image, path, img)


def write_rgb_image(path: str, img):
    """IOUtility::write_rgb_image (io_utility.cpp:171-190): channels 0..2 as an RGB PNG."""
    _write(load_plugins().nli_write_rgb_image, path, img)


def write_mono_image(path: str, img):
    """IOUtility::write_mono_image (io_utility.cpp:192-220): one channel, or the luminance
    0.2126 R + 0.7152 G + 0.0722 B of three."""
    _write(load_plugins().nli_write_mono_image, path, img)


def read_pgm_image(path: str):
    """IOUtility::read_pgm_image (io_utility.cpp:26-74): (H, W) float32 from a P2 or P5
    file, with the reference's reading rules (include/nli_io.h)."""
    lib = load_plugins()
    W, H = ct.c_int(), ct.c_int()
    _io_check(lib.nli_pgm_info(os.fsencode(path), ct.byref(W), ct.byref(H)))
    out = np.zeros((H.value, W.value), np.float32)
    _io_check(lib.nli_read_pgm_image(os.fsencode(path), _ptr(out, _f32p), out.size))
    return out


def write_pgm_image(path: str, img):
    """IOUtility::write_pgm_image (io_utility.cpp:80-117): P5, one byte per pixel."""
    _write(load_plugins().nli_write_pgm_image, path, img)


def batch_mode(mode: str):
    """How batches run, process-wide: "concurrent" (one driver per image; the default) or
    "lockstep" (groups stepping their images together, where the parameters allow it)."""
    rc = load_library().nli_batch_mode({"lockstep": 1, "concurrent": 2}[mode])
    if rc != NLI_OK:
        raise ValueError(mode)


def stats_of(st: InpaintStats) -> RunStats:
    s = st.scales
    return RunStats(s, list(st.iterations[:s]), [(st.width[i], st.height[i]) for i in range(s)],
                    list(st.hole_pixels[:s]), list(st.target_pixels[:s]), st.cg_iterations_total,
                    st.seconds_total, st.seconds_patchmatch, st.seconds_update)


def poisson_fill(img, mask):
    """Coarsest-level Poisson fill (host) of the masked pixels of img (H, W, C)."""
    lib = load_library()
    out = np.array(img, dtype=np.float32, copy=True, order="C")
    mask = _c(mask, np.uint8)
    H, W, C = out.shape
    rc = lib.nli_poisson_fill(_ptr(out, _f32p), _ptr(mask, _u8p), W, H, C)
    if rc != NLI_OK:
        raise NliError(rc, "nli_poisson_fill failed")
    return out


def subsampling_rate(ratio: float, scales: int) -> float:
    return float(load_library().nli_subsampling_rate(ratio, scales))
