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
from .abi import XcCntRule, XcPair, XcPoint, XcResult, XcRgnRule

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libxcorr.so")

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
    "xc_fft2d_r2c_f64": ([C.c_void_p, _dp, C.c_int, C.c_int, _dp, C.POINTER(C.c_int)], C.c_int),
    "xc_ift2d_c2r_f64": ([C.c_void_p, _dp, C.c_int, C.c_int, _dp], C.c_int),
    "xc_tile_put": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                     C.POINTER(C.c_void_p)], C.c_int),
    "xc_tile_release": ([C.c_void_p, C.c_void_p], C.c_int),
    "xc_tile_bytes": ([C.c_void_p], C.c_int64),
    "xc_batch_corr_f32": ([C.c_void_p, C.POINTER(XcPair), C.c_int, C.POINTER(XcResult)], C.c_int),
    "xc_sweep_angles_f32": ([C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.c_int, C.POINTER(XcRgnRule),
                             C.POINTER(XcCntRule), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(XcResult)],
                            C.c_int),
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

    def corr_images(self, mode, ip1, iv1, ip2, iv2, rgn=None, cnt=None, mincor=0.0, nbmaxht=0.0,
                    Ox=0, Oy=0, Rx=-1, Ry=-1):
        """CorrImagesR (mode=XC_MODE_R) / CorrImagesS (mode=XC_MODE_S). Returns (status, r, dx, dy, XcResult)."""
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


def fft_size(n1, n2):
    return load().xc_fft_size(n1, n2)


def version():
    return load().xc_version().decode()
