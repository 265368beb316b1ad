"""Python host mirror of the reference correlator interface over libxcorr.so.

Names, argument meaning and error behaviour follow 0_GEN/Correlation.h:
`corr_patches` is CorrPatches (Correlation.cpp:901-1170), `corr_images` is
CorrImagesR/S (:3144-3290), `fft_2d`/`ift_2d` are FFT_2D/IFT_2D.  The library
is loaded from the package's lib/ directory; if it is missing the import of
this module raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import XcBucketInfo, XcCntRule, XcPair, XcPoint, XcResult, XcRgnRule, XcServeCfg, XcServeStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("XCORR_LIB") or os.path.join(_HERE, "lib", "libxcorr.so")  # XCORR_LIB: another build

_dp = C.POINTER(C.c_double)
_pp = C.POINTER(XcPoint)
_lib = None


class XcorrError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {msg}")


def load():
    """Load libxcorr.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libxcorr.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.xc_version.restype = C.c_char_p
    L.xc_last_error.restype = C.c_char_p
    L.xc_last_error.argtypes = [C.c_void_p]
    L.xc_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.xc_ctx_destroy.argtypes = [C.c_void_p]
    L.xc_fft_size.argtypes = [C.c_int, C.c_int]
    L.xc_fft_size_sqr.argtypes = [C.c_int] * 4
    L.xc_corr_patches_f64.argtypes = [C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(XcRgnRule), C.POINTER(XcCntRule), _dp, _dp, _dp,
                                      C.POINTER(XcResult)]
    L.xc_corr_images_f64.argtypes = [C.c_void_p, C.c_int, _pp, _dp, C.c_int, _pp, _dp, C.c_int,
                                     C.POINTER(XcRgnRule), C.POINTER(XcCntRule), C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(XcResult)]
    for name, args in _OPTIONAL.items():
        f = getattr(L, name, None)
        if f is not None:
            f.argtypes = args[0]
            f.restype = args[1]
    _lib = L
    return L


_OPTIONAL = {
    "xc_corr_patch_to_image_f64": ([C.c_void_p, _pp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, _dp, _dp, _dp, C.POINTER(XcResult)], C.c_int),
    "xc_corr_patches_noalloc_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(XcRgnRule), C.POINTER(XcCntRule), _dp, _dp, _dp, C.POINTER(XcResult)],
                                    C.c_int),
    "xc_fft2d_r2c_f64": ([C.c_void_p, _dp, C.c_int, C.c_int, _dp, C.POINTER(C.c_int)], C.c_int),
    "xc_ift2d_c2r_f64": ([C.c_void_p, _dp, C.c_int, C.c_int, _dp], C.c_int),
    "xc_tile_put": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                     C.POINTER(C.c_void_p)], C.c_int),
    "xc_tile_release": ([C.c_void_p, C.c_void_p], C.c_int),
    "xc_tile_bytes": ([C.c_void_p], C.c_int64),
    "xc_plan_create": ([C.c_void_p, C.POINTER(XcPair), C.c_int, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "xc_plan_launch": ([C.c_void_p], C.c_int),
    "xc_plan_wait": ([C.c_void_p, C.POINTER(XcResult)], C.c_int),
    "xc_plan_run": ([C.c_void_p, C.POINTER(XcResult)], C.c_int),
    "xc_plan_stage_times": ([C.c_void_p, C.c_int, _dp, _dp], C.c_int),
    "xc_plan_model_bytes": ([C.c_void_p], C.c_double),
    "xc_plan_destroy": ([C.c_void_p], C.c_int),
    "xc_plan_buckets": ([C.c_void_p, C.POINTER(XcBucketInfo), C.c_int], C.c_int),
    "xc_plan_set_shapes": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "xc_plan_sweep_waves": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)], C.c_int),
    "xc_plan_set_sweep_waves": ([C.c_void_p, C.c_int], C.c_int),
    "xc_server_create": ([C.c_void_p, C.POINTER(XcServeCfg), C.POINTER(C.c_void_p)], C.c_int),
    "xc_server_submit": ([C.c_void_p, C.POINTER(XcPair), C.POINTER(C.c_uint64)], C.c_int),
    "xc_server_poll": ([C.c_void_p, C.c_uint64, C.POINTER(XcResult)], C.c_int),
    "xc_server_cancel": ([C.c_void_p, C.c_uint64], C.c_int),
    "xc_server_wait": ([C.c_void_p, C.c_uint64, C.POINTER(XcResult), _dp], C.c_int),
    "xc_server_stats": ([C.c_void_p, C.POINTER(XcServeStats)], C.c_int),
    "xc_server_destroy": ([C.c_void_p], C.c_int),
    "xc_batch_corr_f32": ([C.c_void_p, C.POINTER(XcPair), C.c_int, C.POINTER(XcResult)], C.c_int),
    "xc_sweep_angles_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.c_int, C.POINTER(XcRgnRule),
                             C.POINTER(XcCntRule), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(XcResult)],
                            C.c_int),
    "xc_rfromangle_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_double, C.c_int, C.c_int64,
                           C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.XcAngleRec)], C.c_int),
    "xc_angle_scan_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                           C.c_double, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                           C.POINTER(abi.XcAngleRec), C.POINTER(abi.XcAngleRec), C.c_int, C.POINTER(C.c_int)],
                          C.c_int),
    "xc_angle_scan_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                           C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                           C.POINTER(abi.XcAngleRec), C.POINTER(abi.XcAngleRec), C.c_int, C.POINTER(C.c_int)],
                          C.c_int),
    "xc_peak_hunt_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_double, C.c_int, C.c_int64,
                          C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)],
                         C.c_int),
    "xc_rfromangle_affine_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_double, _dp, _dp, C.c_int,
                                  C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.XcAngleRec)], C.c_int),
    "xc_pretweaks_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, C.c_int,
                          C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
                         C.c_int),
    "xc_finish_at_full_res_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int64,
                                   C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)], C.c_int),
    "xc_thm_params_default": ([C.POINTER(abi.XcThmParams)], None),
    "xc_thm_scan_with_tweaks_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.POINTER(abi.XcThmParams),
                                     C.c_double, C.c_double, C.c_double, C.POINTER(abi.XcAngleRec)], C.c_int),
    "xc_thm_use_prior_angles_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.POINTER(abi.XcThmParams),
                                     C.c_double, C.c_double, C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)], C.c_int),
    "xc_thm_denovo_best_angle_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.POINTER(abi.XcThmParams),
                                      C.c_double, C.c_double, C.c_double, C.POINTER(abi.XcAngleRec),
                                      C.POINTER(C.c_int)], C.c_int),
    "xc_thm_post_tweaks_f64": ([C.c_void_p, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.POINTER(abi.XcThmParams),
                                C.POINTER(abi.XcAngleRec)], C.c_int),
    "xc_pixpair_prep_f64": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                             _dp, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "xc_pixpair_prep_dev_f64": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "xc_fourier_match_f64": ([C.c_void_p, _pp, _dp, _dp, C.c_int, C.c_int, _dp, C.POINTER(abi.XcFourierQa)],
                             C.c_int),
    "xc_thm_scan_with_tweaks_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(abi.XcThmParams), C.c_double,
                                     C.c_double, C.c_double, C.POINTER(abi.XcAngleRec)], C.c_int),
    "xc_thm_use_prior_angles_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(abi.XcThmParams), C.c_double,
                                     C.c_double, C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)], C.c_int),
    "xc_thm_denovo_best_angle_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(abi.XcThmParams), C.c_double,
                                      C.c_double, C.c_double, C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)],
                                     C.c_int),
    "xc_thm_post_tweaks_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(abi.XcThmParams),
                                C.POINTER(abi.XcAngleRec)], C.c_int),
    "xc_decimate_vector_f64": ([C.c_void_p, _pp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)],
                               C.c_int),
    "xc_make_thumbs_f64": ([C.c_void_p, _pp, _dp, C.c_int, C.c_int, C.c_int, _pp, _dp, C.c_int, C.c_int, C.c_int,
                           C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                           C.POINTER(C.c_int)], C.c_int),
    "xc_make_thumb_raster_dev_f64": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "xc_crop_f64": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, _pp, C.c_int, _pp, C.c_int, C.c_double,
                     C.c_double, C.c_double, C.c_double, C.c_int, C.c_int64, _pp, _dp, C.POINTER(abi.XcSubimage),
                     _pp, _dp, C.POINTER(abi.XcSubimage), C.POINTER(C.c_int)], C.c_int),
    "xc_crop_nocr_f64": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_int, _dp, _dp, _pp, C.POINTER(abi.XcSubimage), C.POINTER(abi.XcSubimage),
                          C.POINTER(C.c_int)], C.c_int),
    "xc_finish_at_full_res_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.c_int64,
                                   C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)], C.c_int),
    "xc_thmpair_create": ([C.c_char_p], C.c_int),
    "xc_thmpair_append": ([C.c_char_p, C.POINTER(abi.XcThmPair)], C.c_int),
    "xc_thmpair_read_all": ([C.c_char_p, C.POINTER(abi.XcThmPair), C.c_int], C.c_int),
    "xc_thmpair_find": ([C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.XcThmPair)], C.c_int),
    "xc_thmpair_prior_angle": ([C.POINTER(abi.XcThmPair), C.c_int, _dp], C.c_int),
    "xc_peak_hunt_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int64, C.c_int, C.c_int,
                          C.c_int, C.c_int, C.POINTER(abi.XcAngleRec), C.POINTER(C.c_int)], C.c_int),
    "xc_sweep_affine_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.c_int, C.POINTER(abi.XcRgnRule),
                             C.POINTER(abi.XcCntRule), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(abi.XcResult)],
                            C.c_int),
    "xc_pretweaks_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, _dp, _dp, _dp, C.c_int, C.c_int64, C.c_int,
                          C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
}


def _pts(p):
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 2)
    return p, p.ctypes.data_as(_pp)


def _vals(v):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    return v, v.ctypes.data_as(_dp)


def _ref(x):
    return C.byref(x) if x is not None else None


class Context:
    """One HIP stream on one device (xc_ctx)."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        st = self.L.xc_ctx_create(device, C.byref(h))
        if st != abi.XC_OK:
            raise XcorrError(st, "xc_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            self.L.xc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self.L.xc_last_error(self.h).decode()

    def _check(self, st, what, ok=(abi.XC_OK, abi.XC_NO_PEAK, abi.XC_TOO_SMALL)):
        if st not in ok:
            raise XcorrError(st, f"{what}: {self.error()}")
        return st

    def corr_patches(self, ip1, iv1, ip2, iv2, Ox=0, Oy=0, radius=4096, rgn=None, cnt=None, dx0=0.0, dy0=0.0):
        """CorrPatches. Returns (status, r, dx, dy, XcResult); dx,dy keep dx0,dy0 on XC_NO_PEAK."""
        p1, pp1 = _pts(ip1)
        v1, pv1 = _vals(iv1)
        p2, pp2 = _pts(ip2)
        v2, pv2 = _vals(iv2)
        dx, dy, r = C.c_double(dx0), C.c_double(dy0), C.c_double(0)
        res = XcResult()
        st = self.L.xc_corr_patches_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), Ox, Oy, radius,
                                        _ref(rgn), _ref(cnt), C.byref(dx), C.byref(dy), C.byref(r), C.byref(res))
        self._check(st, "xc_corr_patches_f64", ok=(abi.XC_OK, abi.XC_NO_PEAK, abi.XC_E_ODD_R))
        return st, r.value, dx.value, dy.value, res

    def corr_patches_noalloc(self, ip1, iv1, ip2, iv2, Ox=0, Oy=0, radius=4096, rgn=None, cnt=None, dx0=0.0,
                             dy0=0.0):
        """CorrPatches through the entry point that allocates nothing (its workspace is the library's fixed
        per-GPU scratch).  Same returns as corr_patches."""
        p1, pp1 = _pts(ip1)
        v1, pv1 = _vals(iv1)
        p2, pp2 = _pts(ip2)
        v2, pv2 = _vals(iv2)
        dx, dy, r = C.c_double(dx0), C.c_double(dy0), C.c_double(0)
        res = XcResult()
        st = self.L.xc_corr_patches_noalloc_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), Ox, Oy, radius,
                                                _ref(rgn), _ref(cnt), C.byref(dx), C.byref(dy), C.byref(r),
                                                C.byref(res))
        self._check(st, "xc_corr_patches_noalloc_f64", ok=(abi.XC_OK, abi.XC_NO_PEAK, abi.XC_E_ODD_R))
        return st, r.value, dx.value, dy.value, res

    def corr_images(self, mode, ip1, iv1, ip2, iv2, rgn=None, cnt=None, mincor=0.0, nbmaxht=0.0,
                    Ox=0, Oy=0, Rx=-1, Ry=-1):
        """CorrImagesR / CorrImagesS / CorrImagesF (mode XC_MODE_R / XC_MODE_S / XC_MODE_F; nbmaxht is F mode's
        neighbour height limit). Returns (status, r, dx, dy, XcResult)."""
        p1, pp1 = _pts(ip1)
        v1, pv1 = _vals(iv1)
        p2, pp2 = _pts(ip2)
        v2, pv2 = _vals(iv2)
        dx, dy, r = C.c_double(0), C.c_double(0), C.c_double(0)
        res = XcResult()
        st = self.L.xc_corr_images_f64(self.h, mode, pp1, pv1, len(v1), pp2, pv2, len(v2), _ref(rgn), _ref(cnt),
                                       mincor, nbmaxht, Ox, Oy, Rx, Ry, C.byref(dx), C.byref(dy), C.byref(r),
                                       C.byref(res))
        self._check(st, "xc_corr_images_f64")
        return st, r.value, dx.value, dy.value, res


    def fft_2d(self, a):
        """FFT_2D: a[Nslow][Nfast] real -> [Nslow][Nfast/2+1] complex, unnormalised."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        Ny, Nx = a.shape
        out = np.zeros((Ny, Nx // 2 + 1), np.complex128)
        st = self.L.xc_fft2d_r2c_f64(self.h, a.ctypes.data_as(_dp), Nx, Ny, out.ctypes.data_as(_dp), None)
        self._check(st, "xc_fft2d_r2c_f64", ok=(abi.XC_OK,))
        return out

    def ift_2d(self, f, Nfast):
        """IFT_2D: unnormalised c2r of f[Nslow][Nfast/2+1]."""
        f = np.ascontiguousarray(f, dtype=np.complex128)
        Ny = f.shape[0]
        out = np.zeros((Ny, Nfast))
        st = self.L.xc_ift2d_c2r_f64(self.h, f.ctypes.data_as(_dp), Nfast, Ny, out.ctypes.data_as(_dp))
        self._check(st, "xc_ift2d_c2r_f64", ok=(abi.XC_OK,))
        return out

    def corr_patch_to_image(self, ip1, iv1, i2, Ox=0, Oy=0, radius=4096, bfilter=False):
        """CorrPatchToImage. Returns (status, r, dx, dy, XcResult)."""
        p1, pp1 = _pts(ip1)
        v1, pv1 = _vals(iv1)
        i2 = np.ascontiguousarray(i2, dtype=np.float64).ravel()
        dx, dy, r = C.c_double(0), C.c_double(0), C.c_double(0)
        res = XcResult()
        st = self.L.xc_corr_patch_to_image_f64(self.h, pp1, pv1, len(v1), i2.ctypes.data_as(_dp), i2.size, Ox, Oy,
                                               radius, int(bool(bfilter)), C.byref(dx), C.byref(dy), C.byref(r),
                                               C.byref(res))
        self._check(st, "xc_corr_patch_to_image_f64", ok=(abi.XC_OK, abi.XC_NO_PEAK, abi.XC_E_ODD_R))
        return st, r.value, dx.value, dy.value, res

    # ---- tiles and batched plans ----

    def tile_put(self, img, mask=None, normalize=True, device_ptr=None, dtype=None, stride=0, tables32=False,
                 tables=None):
        """Upload (or adopt, with device_ptr) one image as a device-resident tile.

        tables: the summed-area tables' form -- None (exact integer sums for an unmasked uint8 tile of at most
        2^20 pixels, else FP64 (S, S^2) pairs), "f64" (FP64 pairs always) or "f32" (FP32 pairs; also
        tables32=True)."""
        flags = abi.XC_TILE_NORMALIZE if normalize else 0
        if tables32 or tables == "f32":
            flags |= abi.XC_TILE_F32_TABLES
        elif tables == "f64":
            flags |= abi.XC_TILE_F64_TABLES
        elif tables is not None:
            raise ValueError(f"tables: {tables!r}")
        if device_ptr is not None:
            flags |= abi.XC_TILE_DEVICE_SRC
            h, w = img
            data = C.c_void_p(device_ptr)
            mptr = C.c_void_p(mask) if mask is not None else None
            dt = dtype
        else:
            img = np.ascontiguousarray(img)
            h, w = img.shape
            dt = {np.dtype(np.uint8): abi.XC_U8, np.dtype(np.uint16): abi.XC_U16,
                  np.dtype(np.float32): abi.XC_F32, np.dtype(np.float64): abi.XC_F64}[img.dtype]
            data = img.ctypes.data_as(C.c_void_p)
            mptr = None
            if mask is not None:
                mask = np.ascontiguousarray(mask, dtype=np.uint8)
                assert mask.shape == (h, w)
                mptr = mask.ctypes.data_as(C.c_void_p)
        t = C.c_void_p()
        st = self.L.xc_tile_put(self.h, data, dt, w, h, stride, mptr, flags, C.byref(t))
        self._check(st, "xc_tile_put", ok=(abi.XC_OK,))
        return Tile(self, t, w, h)

    def sweep_angles(self, a, b, angles, rgn=None, cnt=None, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """CThmScan angle sweep: CorrImagesS of A rotated by each angle against B (RFromAngle)."""
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        res = (XcResult * ang.size)()
        st = self.L.xc_sweep_angles_f32(self.h, a.h, b.h, ang.ctypes.data_as(_dp), ang.size, _ref(rgn), _ref(cnt),
                                        Ox, Oy, Rx, Ry, res)
        self._check(st, "xc_sweep_angles_f32", ok=(abi.XC_OK,))
        return list(res)

    # ---- CThmScan angle search (RFromAngle, AngleScanMaxR / AngleScanConstXY, PeakHunt) ----

    def rfromangle(self, deg, ap, av, bp, bv, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """RFromAngle in FP64 on point lists. Returns XcAngleRec (A, R, X, Y)."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        rec = abi.XcAngleRec()
        st = self.L.xc_rfromangle_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), deg, olap1d, req_area,
                                      Ox, Oy, Rx, Ry, C.byref(rec))
        self._check(st, "xc_rfromangle_f64", ok=(abi.XC_OK,))
        return rec

    def angle_scan(self, mode, ap, av, bp, bv, center, hlfwid, step, favor_center, olap1d, req_area,
                   Ox=0, Oy=0, Rx=-1, Ry=-1):
        """AngleScanMaxR / AngleScanConstXY in FP64 on point lists. Returns (best, [recs])."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        best, n = abi.XcAngleRec(), C.c_int()
        cap = 4096
        recs = (abi.XcAngleRec * cap)()
        st = self.L.xc_angle_scan_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), mode, center, hlfwid, step,
                                      int(bool(favor_center)), olap1d, req_area, Ox, Oy, Rx, Ry, C.byref(best),
                                      recs, cap, C.byref(n))
        self._check(st, "xc_angle_scan_f64", ok=(abi.XC_OK,))
        return best, list(recs[:min(n.value, cap)])

    def angle_scan_tiles(self, mode, a, b, center, hlfwid, step, favor_center, olap1d, req_area,
                         Ox=0, Oy=0, Rx=-1, Ry=-1):
        """AngleScanMaxR / AngleScanConstXY over resident tiles, all angles in one FP32 sweep."""
        best, n = abi.XcAngleRec(), C.c_int()
        cap = 4096
        recs = (abi.XcAngleRec * cap)()
        st = self.L.xc_angle_scan_f32(self.h, a.h, b.h, mode, center, hlfwid, step, int(bool(favor_center)), olap1d,
                                      req_area, Ox, Oy, Rx, Ry, C.byref(best), recs, cap, C.byref(n))
        self._check(st, "xc_angle_scan_f32", ok=(abi.XC_OK,))
        return best, list(recs[:min(n.value, cap)])

    def peak_hunt(self, best, ap, av, bp, bv, hlfwid, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """PeakHunt in FP64 on point lists from `best` (XcAngleRec). Returns (best, k)."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        b, k = abi.XcAngleRec(best.A, best.R, best.X, best.Y), C.c_int()
        st = self.L.xc_peak_hunt_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), hlfwid, olap1d, req_area,
                                     Ox, Oy, Rx, Ry, C.byref(b), C.byref(k))
        self._check(st, "xc_peak_hunt_f64", ok=(abi.XC_OK,))
        return b, k.value

    def peak_hunt_tiles(self, best, a, b, hlfwid, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """PeakHunt over resident tiles, each probe an FP32 sweep of one angle. Returns (best, k)."""
        bb, k = abi.XcAngleRec(best.A, best.R, best.X, best.Y), C.c_int()
        st = self.L.xc_peak_hunt_f32(self.h, a.h, b.h, hlfwid, olap1d, req_area, Ox, Oy, Rx, Ry, C.byref(bb),
                                     C.byref(k))
        self._check(st, "xc_peak_hunt_f32", ok=(abi.XC_OK,))
        return bb, k.value

    def sweep_affine_tiles(self, a, b, maps, rgn=None, cnt=None, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """CorrImagesS of A mapped by each 2 x 2 map {t0, t1, t3, t4} against B, one FP32 sweep."""
        m = np.ascontiguousarray(maps, dtype=np.float64).reshape(-1, 4)
        res = (abi.XcResult * len(m))()
        st = self.L.xc_sweep_affine_f32(self.h, a.h, b.h, m.ctypes.data_as(_dp), len(m),
                                        None if rgn is None else C.byref(rgn), None if cnt is None else C.byref(cnt),
                                        Ox, Oy, Rx, Ry, res)
        self._check(st, "xc_sweep_affine_f32", ok=(abi.XC_OK,))
        return list(res)

    def pretweaks_tiles(self, a, b, deg, Tdfm, Tptwk, best_r, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """Pretweaks over resident tiles, each round one FP32 sweep. Returns (Tptwk[6], best_r, changed, n_evals)."""
        D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
        W = np.array(Tptwk if Tptwk is not None else [1, 0, 0, 0, 1, 0], dtype=np.float64)
        r, ch, ne = C.c_double(best_r), C.c_int(), C.c_int()
        st = self.L.xc_pretweaks_f32(self.h, a.h, b.h, deg, None if D is None else D.ctypes.data_as(_dp),
                                     W.ctypes.data_as(_dp), C.byref(r), olap1d, req_area, Ox, Oy, Rx, Ry,
                                     C.byref(ch), C.byref(ne))
        self._check(st, "xc_pretweaks_f32", ok=(abi.XC_OK,))
        return W, r.value, ch.value, ne.value

    def rfromangle_affine(self, deg, ap, av, bp, bv, Tdfm, Ttw, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """RFromAngle with deformation and pretweak affines (t[6] each, None = identity)."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
        W = None if Ttw is None else np.ascontiguousarray(Ttw, dtype=np.float64)
        rec = abi.XcAngleRec()
        st = self.L.xc_rfromangle_affine_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), deg,
                                             None if D is None else D.ctypes.data_as(_dp),
                                             None if W is None else W.ctypes.data_as(_dp), olap1d, req_area,
                                             Ox, Oy, Rx, Ry, C.byref(rec))
        self._check(st, "xc_rfromangle_affine_f64", ok=(abi.XC_OK,))
        return rec

    def pretweaks(self, ap, av, bp, bv, deg, Tdfm, Tptwk, best_r, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
        """Pretweaks in FP64. Returns (Tptwk[6], best_r, changed, n_evals)."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
        W = np.array(Tptwk if Tptwk is not None else [1, 0, 0, 0, 1, 0], dtype=np.float64)
        r, ch, ne = C.c_double(best_r), C.c_int(), C.c_int()
        st = self.L.xc_pretweaks_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), deg,
                                     None if D is None else D.ctypes.data_as(_dp), W.ctypes.data_as(_dp),
                                     C.byref(r), olap1d, req_area, Ox, Oy, Rx, Ry, C.byref(ch), C.byref(ne))
        self._check(st, "xc_pretweaks_f64", ok=(abi.XC_OK,))
        return W, r.value, ch.value, ne.value

    def finish_at_full_res(self, best, ap, av, bp, bv, Tdfm, Tptwk, olap1d, req_area):
        """FinishAtFullRes in FP64 on the full-resolution point lists. Returns (best, used)."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
        W = None if Tptwk is None else np.ascontiguousarray(Tptwk, dtype=np.float64)
        b, u = abi.XcAngleRec(best.A, best.R, best.X, best.Y), C.c_int()
        st = self.L.xc_finish_at_full_res_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2),
                                              None if D is None else D.ctypes.data_as(_dp),
                                              None if W is None else W.ctypes.data_as(_dp), olap1d, req_area,
                                              C.byref(b), C.byref(u))
        self._check(st, "xc_finish_at_full_res_f64", ok=(abi.XC_OK,))
        return b, u.value

    def _thm(self, name, ap, av, bp, bv, params, *args, ok=False):
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        best = abi.XcAngleRec()
        extra = [C.byref(best)]
        okv = C.c_int()
        if ok:
            extra.append(C.byref(okv))
        st = getattr(self.L, name)(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), C.byref(params), *args, *extra)
        self._check(st, name, ok=(abi.XC_OK,))
        return (best, okv.value) if ok else best

    def thm_scan_with_tweaks(self, ap, av, bp, bv, params, center, hlfwid, step):
        """AngleScanWithTweaks in FP64; params (an XcThmParams) is updated in place. Returns best."""
        return self._thm("xc_thm_scan_with_tweaks_f64", ap, av, bp, bv, params, center, hlfwid, step)

    def thm_use_prior_angles(self, ap, av, bp, bv, params, ang0, hfangpr):
        """UsePriorAngles in FP64. Returns (best, ok)."""
        return self._thm("xc_thm_use_prior_angles_f64", ap, av, bp, bv, params, ang0, hfangpr, ok=True)

    def thm_denovo_best_angle(self, ap, av, bp, bv, params, ang0, hfangdn, step):
        """DenovoBestAngle in FP64. Returns (best, ok)."""
        return self._thm("xc_thm_denovo_best_angle_f64", ap, av, bp, bv, params, ang0, hfangdn, step, ok=True)

    def thm_post_tweaks(self, ap, av, bp, bv, params, best):
        """PostTweaks in FP64 from the record best. Returns the new record."""
        p1, pp1 = _pts(ap)
        v1, pv1 = _vals(av)
        p2, pp2 = _pts(bp)
        v2, pv2 = _vals(bv)
        b = abi.XcAngleRec(best.A, best.R, best.X, best.Y)
        st = self.L.xc_thm_post_tweaks_f64(self.h, pp1, pv1, len(v1), pp2, pv2, len(v2), C.byref(params), C.byref(b))
        self._check(st, "xc_thm_post_tweaks_f64", ok=(abi.XC_OK,))
        return b

    def pixpair_prep(self, ras, order=2, dog=False, r1=3, r2=6):
        """PixPair::Load's raster preparation of one 8-bit picture (host arrays). Returns a dict with vf,
        vfflt (with dog), vs and vsflt (when downsampled), ws, hs and scl."""
        ras = np.ascontiguousarray(ras, dtype=np.uint8)
        h, w = ras.shape
        ws, hs, scl = C.c_int(), C.c_int(), C.c_int()
        probe = w, h
        while probe[0] > 2048 or probe[1] > 2048:
            probe = probe[0] // 2, probe[1] // 2
        vf = np.empty((h, w))
        vfflt = np.empty((h, w)) if dog else None
        down = probe != (w, h)
        vs = np.empty((probe[1], probe[0])) if down else None
        vsflt = np.empty((probe[1], probe[0])) if down and dog else None
        ptr = lambda a: None if a is None else a.ctypes.data_as(_dp)
        st = self.L.xc_pixpair_prep_f64(self.h, ras.ctypes.data, w, h, order, int(bool(dog)), r1, r2, ptr(vf),
                                        ptr(vfflt), ptr(vs), ptr(vsflt), C.byref(ws), C.byref(hs), C.byref(scl))
        self._check(st, "xc_pixpair_prep_f64", ok=(abi.XC_OK,))
        return {"vf": vf, "vfflt": vfflt, "vs": vs, "vsflt": vsflt, "ws": ws.value, "hs": hs.value, "scl": scl.value}

    def fourier_match(self, pts, av, bv, wvlen=25):
        """FourierMatch in FP64 (point list and the two value lists). Returns an XcFourierQa (qa.dot is the
        metric)."""
        p, pp = _pts(pts)
        a, pa = _vals(av)
        b, pb = _vals(bv)
        d, qa = C.c_double(), abi.XcFourierQa()
        st = self.L.xc_fourier_match_f64(self.h, pp, pa, pb, len(a), wvlen, C.byref(d), C.byref(qa))
        self._check(st, "xc_fourier_match_f64", ok=(abi.XC_OK,))
        return qa

    def thm_tiles(self, which, a, b, params, *args, best=None):
        """The CThmScan sequence over resident tiles in FP32: which is "scan_with_tweaks" (center, hlfwid, step),
        "use_prior_angles" (ang0, hfangpr), "denovo_best_angle" (ang0, hfangdn, step) or "post_tweaks" (from
        best).  Returns best, or (best, ok) for the two that report success."""
        name = f"xc_thm_{which}_f32"
        b0 = abi.XcAngleRec() if best is None else abi.XcAngleRec(best.A, best.R, best.X, best.Y)
        okv = C.c_int()
        with_ok = which in ("use_prior_angles", "denovo_best_angle")
        extra = [C.byref(b0)] + ([C.byref(okv)] if with_ok else [])
        st = getattr(self.L, name)(self.h, a.h, b.h, C.byref(params), *args, *extra)
        self._check(st, name, ok=(abi.XC_OK,))
        return (b0, okv.value) if with_ok else b0

    def decimate_vector(self, pts, vals, w, h, lbin):
        """DecimateVector on host lists. Returns (points (k, 2), values (k,))."""
        p = np.array(pts, dtype=np.float64).reshape(-1, 2).copy()
        v = np.array(vals, dtype=np.float64).copy()
        k = C.c_int()
        st = self.L.xc_decimate_vector_f64(self.h, p.ctypes.data_as(_pp), v.ctypes.data_as(_dp), len(v), w, h, lbin,
                                           C.byref(k))
        self._check(st, "xc_decimate_vector_f64", ok=(abi.XC_OK,))
        return p[:k.value], v[:k.value]

    def make_thumbs(self, ap, av, aw, ah, bp, bv, bw, bh, decfactor, olap1d, req_area):
        """MakeThumbs on both sides. Returns (status, ap, av, bp, bv, olap1d, req_area)."""
        a_p = np.array(ap, dtype=np.float64).reshape(-1, 2).copy()
        a_v = np.array(av, dtype=np.float64).copy()
        b_p = np.array(bp, dtype=np.float64).reshape(-1, 2).copy()
        b_v = np.array(bv, dtype=np.float64).copy()
        ol, rq, ka, kb = C.c_int(olap1d), C.c_int64(req_area), C.c_int(), C.c_int()
        st = self.L.xc_make_thumbs_f64(self.h, a_p.ctypes.data_as(_pp), a_v.ctypes.data_as(_dp), len(a_v), aw, ah,
                                       b_p.ctypes.data_as(_pp), b_v.ctypes.data_as(_dp), len(b_v), bw, bh, decfactor,
                                       C.byref(ol), C.byref(rq), C.byref(ka), C.byref(kb))
        self._check(st, "xc_make_thumbs_f64", ok=(abi.XC_OK, abi.XC_TOO_SMALL, abi.XC_NO_PEAK))
        return st, a_p[:ka.value], a_v[:ka.value], b_p[:kb.value], b_v[:kb.value], ol.value, rq.value

    def thumb_raster_dev(self, ras, dtype, w, h, stride, lbin, out_ptr):
        """Thumbnail of a device picture (pointer ras) into the device array out_ptr; returns the natural sd."""
        sd = C.c_double()
        st = self.L.xc_make_thumb_raster_dev_f64(self.h, ras, dtype, w, h, stride, lbin, out_ptr, C.byref(sd))
        self._check(st, "xc_make_thumb_raster_dev_f64", ok=(abi.XC_OK,))
        return sd.value

    def crop(self, av_ptr, bv_ptr, ws, hs, apts, bpts, tx, ty, xyconf, scl, olap1d, olap2d):
        """CThmUtil::Crop of two ws x hs FP64 device pictures (pointers) with host point lists.
        Returns (status, whole, (ap, av, a), (bp, bv, b)); a, b are XcSubimage records."""
        apts = np.ascontiguousarray(apts, dtype=np.float64).reshape(-1, 2)
        bpts = np.ascontiguousarray(bpts, dtype=np.float64).reshape(-1, 2)
        ap, bp = np.zeros_like(apts), np.zeros_like(bpts)
        avo, bvo = np.zeros(len(apts)), np.zeros(len(bpts))
        a, b, whole = abi.XcSubimage(), abi.XcSubimage(), C.c_int(-1)
        st = self.L.xc_crop_f64(self.h, av_ptr, bv_ptr, ws, hs, apts.ctypes.data_as(_pp), len(apts),
                                bpts.ctypes.data_as(_pp), len(bpts), tx, ty, xyconf, scl, olap1d, olap2d,
                                ap.ctypes.data_as(_pp), avo.ctypes.data_as(_dp), C.byref(a), bp.ctypes.data_as(_pp),
                                bvo.ctypes.data_as(_dp), C.byref(b), C.byref(whole))
        self._check(st, "xc_crop_f64", ok=(abi.XC_OK, abi.XC_TOO_SMALL))
        return st, whole.value, (ap[:a.n], avo[:a.n], a), (bp[:b.n], bvo[:b.n], b)

    def crop_nocr(self, av_ptr, bv_ptr, ws, hs, tx, ty, xyconf, scl, olap1d, points=True):
        """CThmUtil::Crop_NoCR of two ws x hs FP64 device pictures. Returns (whole, points or None, (av, a),
        (bv, b))."""
        avo, bvo = np.zeros(ws * hs), np.zeros(ws * hs)
        p = np.zeros((ws * hs, 2)) if points else None
        a, b, whole = abi.XcSubimage(), abi.XcSubimage(), C.c_int(-1)
        st = self.L.xc_crop_nocr_f64(self.h, av_ptr, bv_ptr, ws, hs, tx, ty, xyconf, scl, olap1d,
                                     avo.ctypes.data_as(_dp), bvo.ctypes.data_as(_dp),
                                     p.ctypes.data_as(_pp) if points else None, C.byref(a), C.byref(b),
                                     C.byref(whole))
        self._check(st, "xc_crop_nocr_f64", ok=(abi.XC_OK,))
        return whole.value, (p[:a.n] if points else None), (avo[:a.n], a), (bvo[:b.n], b)

    def finish_at_full_res_tiles(self, best, a, b, Tdfm, Tptwk, olap1d, req_area):
        """FinishAtFullRes over full-resolution resident tiles in FP32. Returns (best, used)."""
        D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
        W = None if Tptwk is None else np.ascontiguousarray(Tptwk, dtype=np.float64)
        bb, u = abi.XcAngleRec(best.A, best.R, best.X, best.Y), C.c_int()
        st = self.L.xc_finish_at_full_res_f32(self.h, a.h, b.h, None if D is None else D.ctypes.data_as(_dp),
                                              None if W is None else W.ctypes.data_as(_dp), olap1d, req_area,
                                              C.byref(bb), C.byref(u))
        self._check(st, "xc_finish_at_full_res_f32", ok=(abi.XC_OK,))
        return bb, u.value

    def plan(self, pairs, refresh_tiles=False, autotune=False):
        return Plan(self, pairs, refresh_tiles, autotune)

    def batch_corr(self, pairs):
        arr = (XcPair * len(pairs))(*pairs)
        res = (XcResult * len(pairs))()
        st = self.L.xc_batch_corr_f32(self.h, arr, len(pairs), res)
        self._check(st, "xc_batch_corr_f32", ok=(abi.XC_OK,))
        return list(res)


class Tile:
    def __init__(self, ctx, handle, w, h):
        self.ctx, self.h, self.w, self.hgt = ctx, handle, w, h

    @property
    def nbytes(self):
        return self.ctx.L.xc_tile_bytes(self.h)

    def release(self):
        if self.h:
            self.ctx.L.xc_tile_release(self.ctx.h, self.h)
            self.h = None


def make_pair(a, b, mode=abi.XC_MODE_NCC, Ox=0, Oy=0, radius=4096, Rx=-1, Ry=-1, rgn=None, cnt=None, angle=0.0):
    """xc_pair for tiles a, b (CorrPatches semantics for XC_MODE_NCC, CorrImages* for R/S)."""
    return XcPair(a=a.h, b=b.h, mode=mode, Ox=Ox, Oy=Oy, radius=radius, Rx=Rx, Ry=Ry, angle_deg=angle,
                  rgn=rgn if rgn is not None else XcRgnRule(), cnt=cnt if cnt is not None else XcCntRule())


class Plan:
    """A fixed batch of pairs, laid out once and replayed as one HIP graph (xc_plan)."""

    def __init__(self, ctx, pairs, refresh_tiles=False, autotune=False):
        self.ctx = ctx
        self.n = len(pairs)
        arr = (XcPair * self.n)(*pairs)
        h = C.c_void_p()
        flags = (abi.XC_PLAN_REFRESH_TILES if refresh_tiles else 0) | (abi.XC_PLAN_AUTOTUNE if autotune else 0)
        st = ctx.L.xc_plan_create(ctx.h, arr, self.n, flags, C.byref(h))
        ctx._check(st, "xc_plan_create", ok=(abi.XC_OK,))
        self.h = h
        self._res = (XcResult * self.n)()

    def run(self, want=True):
        st = self.ctx.L.xc_plan_run(self.h, self._res if want else None)
        self.ctx._check(st, "xc_plan_run", ok=(abi.XC_OK,))
        return list(self._res) if want else None

    def launch(self):
        self.ctx._check(self.ctx.L.xc_plan_launch(self.h), "xc_plan_launch", ok=(abi.XC_OK,))

    def wait(self, want=False):
        st = self.ctx.L.xc_plan_wait(self.h, self._res if want else None)
        self.ctx._check(st, "xc_plan_wait", ok=(abi.XC_OK,))
        return list(self._res) if want else None

    def stage_times(self, iters=3):
        ms = (C.c_double * 5)()
        by = (C.c_double * 5)()
        st = self.ctx.L.xc_plan_stage_times(self.h, iters, ms, by)
        self.ctx._check(st, "xc_plan_stage_times", ok=(abi.XC_OK,))
        return list(ms), list(by)

    @property
    def model_bytes(self):
        return self.ctx.L.xc_plan_model_bytes(self.h)

    def buckets(self):
        """Size buckets with their launch shapes (tuned ones carry col_us / r2c_us)."""
        nb = self.ctx.L.xc_plan_buckets(self.h, None, 0)
        arr = (XcBucketInfo * max(nb, 1))()
        self.ctx.L.xc_plan_buckets(self.h, arr, nb)
        return [{k: v for k, v in arr[i].as_dict().items() if k != "_pad"} for i in range(nb)]

    def set_shapes(self, bucket, col_c=0, col_xcd=1, r2c_rpt=0, r2c_np=0):
        """Set a bucket's launch shapes by hand (xc_plan_set_shapes)."""
        st = self.ctx.L.xc_plan_set_shapes(self.h, bucket, col_c, col_xcd, r2c_rpt, r2c_np)
        self.ctx._check(st, "xc_plan_set_shapes", ok=(abi.XC_OK,))

    def sweep_waves(self):
        """(wavefronts per CU of the sweeps beside the row transforms, tuned whole-plan us or 0)."""
        w, us = C.c_int32(0), C.c_double(0.0)
        st = self.ctx.L.xc_plan_sweep_waves(self.h, C.byref(w), C.byref(us))
        self.ctx._check(st, "xc_plan_sweep_waves", ok=(abi.XC_OK,))
        return w.value, us.value

    def set_sweep_waves(self, waves):
        """Set the sweeps' share by hand (xc_plan_set_sweep_waves)."""
        st = self.ctx.L.xc_plan_set_sweep_waves(self.h, waves)
        self.ctx._check(st, "xc_plan_set_sweep_waves", ok=(abi.XC_OK,))

    def destroy(self):
        if self.h:
            self.ctx.L.xc_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """Serving loop (xc_server): continuous batching of pair requests over resident tiles."""

    def __init__(self, ctx, hbm_budget=8 << 30, max_batch=64, streams=3, max_wait_us=200, deadline_us=0):
        self.ctx = ctx
        cfg = XcServeCfg(hbm_budget=hbm_budget, max_batch=max_batch, streams=streams, max_wait_us=max_wait_us,
                         deadline_us=deadline_us)
        h = C.c_void_p()
        st = ctx.L.xc_server_create(ctx.h, C.byref(cfg), C.byref(h))
        ctx._check(st, "xc_server_create", ok=(abi.XC_OK,))
        self.h = h

    def submit(self, pair):
        t = C.c_uint64()
        st = self.ctx.L.xc_server_submit(self.h, C.byref(pair), C.byref(t))
        if st != abi.XC_OK:
            raise XcorrError(st, "xc_server_submit")
        return t.value

    def wait(self, ticket, raise_errors=True):
        """(XcResult, latency_us); with raise_errors=False an error status (timeout, cancellation, device
        failure) is returned in the record instead of raised."""
        r = XcResult()
        lat = C.c_double()
        st = self.ctx.L.xc_server_wait(self.h, ticket, C.byref(r), C.byref(lat))
        if st != abi.XC_OK and raise_errors:
            raise XcorrError(st, "xc_server_wait")
        return r, lat.value

    def cancel(self, ticket):
        """XC_OK if the queued request was cancelled, XC_E_NOTREADY if it is already on the device or done."""
        return self.ctx.L.xc_server_cancel(self.h, ticket)

    def poll(self, ticket):
        r = XcResult()
        st = self.ctx.L.xc_server_poll(self.h, ticket, C.byref(r))
        return (r if st == abi.XC_OK else None)

    def stats(self):
        s = XcServeStats()
        self.ctx.L.xc_server_stats(self.h, C.byref(s))
        return s.as_dict()

    def close(self):
        if self.h:
            self.ctx.L.xc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fft_size(n1, n2):
    return load().xc_fft_size(n1, n2)


def version():
    return load().xc_version().decode()


# ---- ThmPair tables (host only: no device needed) ----

def thmpair_create(path):
    """WriteThmPairHdr: create or truncate a ThmPair table with its header line."""
    st = load().xc_thmpair_create(os.fsencode(path))
    if st != abi.XC_OK:
        raise OSError(f"xc_thmpair_create({path}): status {st}")


def thmpair_append(path, atl, acr, btl, bcr, err, A, R, T):
    """WriteThmPair: append one entry (skipped when atl < 0 or btl < 0)."""
    e = abi.XcThmPair(atl, acr, btl, bcr, err, 0, A, R, (C.c_double * 6)(*T))
    st = load().xc_thmpair_append(os.fsencode(path), C.byref(e))
    if st != abi.XC_OK:
        raise OSError(f"xc_thmpair_append({path}): status {st}")


def thmpair_read_all(path):
    """ReadAllThmPair: every entry after the header (XcThmPair list)."""
    L = load()
    n = L.xc_thmpair_read_all(os.fsencode(path), None, 0)
    if n < 0:
        raise OSError(f"xc_thmpair_read_all({path}): status {n}")
    out = (abi.XcThmPair * max(n, 1))()
    L.xc_thmpair_read_all(os.fsencode(path), out, n)
    return list(out[:n])


def thmpair_find(path, atl, acr, btl, bcr):
    """ReadThmPair: the first entry of the given tiles, or None."""
    e = abi.XcThmPair()
    st = load().xc_thmpair_find(os.fsencode(path), atl, acr, btl, bcr, C.byref(e))
    if st == abi.XC_OK:
        return e
    if st == abi.XC_NO_PEAK:
        return None
    raise OSError(f"xc_thmpair_find({path}): status {st}")


def thmpair_prior_angle(entries):
    """Mode-Y starting angle: (count of error-free entries used, median angle) or (0, None)."""
    arr = (abi.XcThmPair * max(len(entries), 1))(*entries)
    a = C.c_double(0.0)
    n = load().xc_thmpair_prior_angle(arr, len(entries), C.byref(a))
    return (n, a.value) if n > 0 else (0, None)
