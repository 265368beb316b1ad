"""CPU oracle for the FFT tile-pair cross-correlation path -- TEST INFRASTRUCTURE.

ctypes front end of oracle/liboracle.so (oracle/xc_oracle.c), a plain-C FP64
restatement of billkarsh/Alignment_Projects' 0_GEN/Correlation.cpp path.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
it, and only as the checker or as the timed CPU baseline.  The product
(alignment_projects_amd / libxcorr.so) never imports it.

Parity pins (see tests/test_oracle.py): the reference's own harnesses
(testnormcorr.cpp KAT, time_fft.cpp planted shift, fft.cpp DFT check), a
direct DFT and numpy.fft for the transforms, and a brute-force (no FFT)
normalised cross-correlation with the reference's scan order, zone rules and
parabola quirks for the correlators.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from alignment_projects_amd.abi import (XcAngleRec, XcCntRule, XcFourierQa, XcPoint, XcResult, XcRgnRule,
                                        XcSubimage, XcThmParams)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_pp = C.POINTER(XcPoint)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.xo_corr_patches.argtypes = [_pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(XcRgnRule), C.POINTER(XcCntRule), _dp, _dp, _dp,
                                      C.POINTER(XcResult), _dp, C.c_int]
        L.xo_corr_images.argtypes = [C.c_int, _pp, _dp, C.c_int, _pp, _dp, C.c_int,
                                     C.POINTER(XcRgnRule), C.POINTER(XcCntRule), C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                     C.POINTER(XcResult), _dp, _dp, C.POINTER(C.c_uint8), C.c_int]
        L.xo_corr_patch_to_image.argtypes = [_pp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, _dp, _dp, _dp, C.POINTER(XcResult)]
        L.xo_rfromangle.argtypes = [C.c_double, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_int, C.c_long,
                                    C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(XcResult)]
        L.xo_fft2d_r2c.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.xo_ift2d_c2r.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.xo_fft1d.argtypes = [_dp, C.c_int, C.c_int]
        L.xo_stats.argtypes = [_dp, C.c_int, _dp, _dp]
        L.xo_normalize.argtypes = [_dp, C.c_int]
        L.xo_normalize.restype = C.c_double
        L.xo_integrate_image.argtypes = [_dp, _dp, _ip, C.c_int, C.c_int, _dp, C.c_int]
        L.xo_image_from_values_and_points.argtypes = [_dp, C.c_int, C.c_int, _dp, _pp, C.c_int, C.c_int, C.c_int]
        L.xo_bbox_from_points.argtypes = [_pp, C.c_int, _ip]
        L.xo_boxes_from_shifts.argtypes = [C.c_int] * 6 + [_ip, _ip]
        L.xo_parab_peak_fft.argtypes = [_dp, _dp, C.c_int, _dp, C.c_int, C.c_int]
        L.xo_decimate_vector.argtypes = [_pp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
        _sp = C.POINTER(XcSubimage)
        L.xo_crop.argtypes = [_dp, _dp, C.c_int, C.c_int, _pp, C.c_int, _pp, C.c_int, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_int, C.c_long, _pp, _dp, _sp, _pp, _dp, _sp,
                              C.POINTER(C.c_int)]
        L.xo_crop_nocr.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, _dp, _dp, _pp, _sp, _sp, C.POINTER(C.c_int)]
        L.xo_set_fft_provider.argtypes = [C.c_int, C.c_char_p]
        L.xo_fft_provider_name.restype = C.c_char_p
        L.xo_rgn_ok.argtypes = [C.POINTER(XcRgnRule), C.c_int, C.c_int]
        L.xo_linefit_lincor.argtypes = [_dp, _dp, C.c_int]
        L.xo_linefit_lincor.restype = C.c_double
        L.xo_scan_angles.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp, C.c_int]
        L.xo_scan_select.argtypes = [C.c_int, C.POINTER(XcAngleRec), C.c_int, C.POINTER(XcAngleRec)]
        L.xo_angle_scan.argtypes = [C.c_int, _pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_int, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(XcAngleRec), C.POINTER(XcAngleRec), C.c_int, C.POINTER(C.c_int)]
        L.xo_peak_hunt.argtypes = [_pp, _dp, C.c_int, _pp, _dp, C.c_int, C.c_double, C.c_int, C.c_long,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(XcAngleRec), C.POINTER(C.c_int)]
        L.xo_rfromangle_affine.argtypes = [C.c_double, _pp, _dp, C.c_int, _pp, _dp, C.c_int, _dp, _dp, C.c_int,
                                            C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.xo_pretweaks.argtypes = [C.c_double, _pp, _dp, C.c_int, _pp, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_long,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.xo_finish_at_full_res.argtypes = [_pp, _dp, C.c_int, _pp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_long,
                                             C.POINTER(XcAngleRec), C.POINTER(C.c_int)]
        for nm, extra in (("xo_thm_scan_with_tweaks", [C.c_double, C.c_double, C.c_double, C.POINTER(XcAngleRec)]),
                          ("xo_thm_use_prior_angles", [C.c_double, C.c_double, C.POINTER(XcAngleRec),
                                                       C.POINTER(C.c_int)]),
                          ("xo_thm_denovo_best_angle", [C.c_double, C.c_double, C.c_double, C.POINTER(XcAngleRec),
                                                        C.POINTER(C.c_int)]),
                          ("xo_thm_post_tweaks", [C.POINTER(XcAngleRec)])):
            getattr(L, nm).argtypes = [_pp, _dp, C.c_int, _pp, _dp, C.c_int, C.POINTER(XcThmParams)] + extra
        L.xo_leg_poly_create.argtypes = [_dp, C.c_int, C.c_int]
        L.xo_leg_poly_create.restype = None
        L.xo_leg_poly_flatten.argtypes = [_dp, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.xo_make_dog_kernel.argtypes = [_dp, C.c_int, C.c_int]
        L.xo_downsample.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_downsample.restype = None
        L.xo_pixpair_prep.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                      _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.xo_tightest_bbox.argtypes = [_pp, C.c_int]
        L.xo_fourier_match.argtypes = [_pp, _dp, _dp, C.c_int, C.c_int, _dp, C.POINTER(XcFourierQa)]
        L.xo_convolve.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_fpeak.argtypes = [_dp, _dp, C.c_int, C.c_int, C.POINTER(C.c_long), C.c_long, C.c_double, _ip, _ip]
        _lib = L
    return _lib


def _pts(p):
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 2)
    return p, p.ctypes.data_as(_pp)


def _vals(v):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    return v, v.ctypes.data_as(_dp)


def _ref(x):
    return C.byref(x) if x is not None else None


def corr_patches(ip1, iv1, ip2, iv2, Ox=0, Oy=0, radius=4096, rgn=None, cnt=None,
                 dx0=0.0, dy0=0.0, want_rslt=False):
    """CorrPatches (0_GEN/Correlation.cpp:901-1170). Returns (status, r, dx, dy, XcResult[, rslt])."""
    p1, pp1 = _pts(ip1)
    v1, pv1 = _vals(iv1)
    p2, pp2 = _pts(ip2)
    v2, pv2 = _vals(iv2)
    dx, dy, r = C.c_double(dx0), C.c_double(dy0), C.c_double(0)
    res = XcResult()
    rs = None
    if want_rslt:
        b1 = bbox(p1)
        b2 = bbox(p2)
        Nx = fft_size(b1[1] - b1[0] + 1, b2[1] - b2[0] + 1)
        Ny = fft_size(b1[3] - b1[2] + 1, b2[3] - b2[2] + 1)
        rs = np.zeros(Nx * Ny)
    st = lib().xo_corr_patches(pp1, pv1, len(v1), pp2, pv2, len(v2), Ox, Oy, radius,
                               _ref(rgn), _ref(cnt), C.byref(dx), C.byref(dy), C.byref(r), C.byref(res),
                               rs.ctypes.data_as(_dp) if rs is not None else None,
                               rs.size if rs is not None else 0)
    out = (st, r.value, dx.value, dy.value, res)
    if want_rslt:
        out = out + (rs.reshape(res.ny, res.nx) if res.nx else rs,)
    return out


def corr_images(mode, ip1, iv1, ip2, iv2, rgn=None, cnt=None, mincor=0.0, nbmaxht=0.0,
                Ox=0, Oy=0, Rx=-1, Ry=-1, want_images=False):
    """CorrImagesR/S/F (0_GEN/Correlation.cpp:3066-3290). Returns (status, r, dx, dy, XcResult[, R, S, A]);
    in F mode the S slot holds the F image and A has MakeF's guard band cleared."""
    p1, pp1 = _pts(ip1)
    v1, pv1 = _vals(iv1)
    p2, pp2 = _pts(ip2)
    v2, pv2 = _vals(iv2)
    dx, dy, r = C.c_double(0), C.c_double(0), C.c_double(0)
    res = XcResult()
    R = S = A = None
    cap = 0
    if want_images:
        b1, b2 = bbox(p1), bbox(p2)
        wR = (b1[1] - b1[0]) + (b2[1] - b2[0] + 1)
        hR = (b1[3] - b1[2]) + (b2[3] - b2[2] + 1)
        cap = wR * hR
        R, S, A = np.zeros(cap), np.zeros(cap), np.zeros(cap, np.uint8)
    st = lib().xo_corr_images(mode, pp1, pv1, len(v1), pp2, pv2, len(v2), _ref(rgn), _ref(cnt),
                              mincor, nbmaxht, Ox, Oy, Rx, Ry, C.byref(dx), C.byref(dy), C.byref(r),
                              C.byref(res),
                              R.ctypes.data_as(_dp) if R is not None else None,
                              S.ctypes.data_as(_dp) if S is not None else None,
                              A.ctypes.data_as(C.POINTER(C.c_uint8)) if A is not None else None, cap)
    out = (st, r.value, dx.value, dy.value, res)
    if want_images:
        out = out + (R.reshape(hR, wR), S.reshape(hR, wR), A.reshape(hR, wR))
    return out


def convolve(src, K, sym=True, prenorm=True):
    """Convolve (0_GEN/Correlation.cpp:371-457) of src (h x w) with the odd-sized kernel K (hk x wk)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    hs, ws = src.shape
    hk, wk = K.shape
    dst = np.zeros_like(src)
    lib().xo_convolve(dst.ctypes.data_as(_dp), src.ctypes.data_as(_dp), ws, hs, K.ctypes.data_as(_dp), wk, hk,
                      int(bool(sym)), int(bool(prenorm)))
    return dst


def fpeak(F, R, order, nbmaxht):
    """CCorImg::FPeak (0_GEN/Correlation.cpp:2580-2767) over the candidate indices `order`.
    Returns (rx, ry) or None."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    hR, wR = F.shape
    o = np.ascontiguousarray(order, dtype=np.int64)
    rx, ry = C.c_int(), C.c_int()
    ok = lib().xo_fpeak(F.ctypes.data_as(_dp), R.ctypes.data_as(_dp), wR, hR, o.ctypes.data_as(C.POINTER(C.c_long)),
                        o.size, nbmaxht, C.byref(rx), C.byref(ry))
    return (rx.value, ry.value) if ok else None


def corr_patch_to_image(ip1, iv1, i2, Ox=0, Oy=0, radius=4096, bfilter=False):
    """CorrPatchToImage (0_GEN/Correlation.cpp:1186-1355). Returns (status, r, dx, dy, XcResult)."""
    p1, pp1 = _pts(ip1)
    v1, pv1 = _vals(iv1)
    i2 = np.ascontiguousarray(i2, dtype=np.float64).ravel()
    dx, dy, r = C.c_double(0), C.c_double(0), C.c_double(0)
    res = XcResult()
    st = lib().xo_corr_patch_to_image(pp1, pv1, len(v1), i2.ctypes.data_as(_dp), i2.size, Ox, Oy, radius,
                                      int(bool(bfilter)), C.byref(dx), C.byref(dy), C.byref(r), C.byref(res))
    return st, r.value, dx.value, dy.value, res


def rfromangle(deg, ap, av, bp, bv, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
    """CThmScan::RFromAngle (0_GEN/CThmScan.cpp:329-366). Returns (status, R, X, Y, XcResult)."""
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    X, Y, R = C.c_double(0), C.c_double(0), C.c_double(0)
    res = XcResult()
    st = lib().xo_rfromangle(deg, pp1, pv1, len(v1), pp2, pv2, len(v2), olap1d, req_area, Ox, Oy, Rx, Ry,
                             C.byref(X), C.byref(Y), C.byref(R), C.byref(res))
    return st, R.value, X.value, Y.value, res


def fft2d_r2c(a):
    """FFT_2D (0_GEN/Correlation_fft_fftw.cpp:38-56): a[Ny][Nx] -> [Ny][Nx/2+1] complex."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    Ny, Nx = a.shape
    out = np.zeros((Ny, Nx // 2 + 1), np.complex128)
    lib().xo_fft2d_r2c(a.ctypes.data_as(_dp), Nx, Ny, out.ctypes.data_as(_dp))
    return out


def ift2d_c2r(f, Nx):
    """IFT_2D (0_GEN/Correlation_fft_fftw.cpp:67-92): unnormalised c2r."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    Ny = f.shape[0]
    out = np.zeros((Ny, Nx))
    lib().xo_ift2d_c2r(f.ctypes.data_as(_dp), Nx, Ny, out.ctypes.data_as(_dp))
    return out


def fft1d(z, inverse=False):
    z = np.array(z, dtype=np.complex128)
    lib().xo_fft1d(z.ctypes.data_as(_dp), z.size, int(inverse))
    return z


def stats(v):
    v, pv = _vals(v)
    a, s = C.c_double(), C.c_double()
    lib().xo_stats(pv, v.size, C.byref(a), C.byref(s))
    return a.value, s.value


def normalize(v):
    v = np.array(v, dtype=np.float64)
    lib().xo_normalize(v.ctypes.data_as(_dp), v.size)
    return v


def integrate_image(img, wS, hS):
    img = np.ascontiguousarray(img, dtype=np.float64)
    wI = img.shape[1]
    S = np.zeros(wS * hS)
    S2 = np.zeros(wS * hS)
    nz = np.zeros(wS * hS, np.int32)
    lib().xo_integrate_image(S.ctypes.data_as(_dp), S2.ctypes.data_as(_dp), nz.ctypes.data_as(_ip),
                             wS, hS, img.ctypes.data_as(_dp), wI)
    return S.reshape(hS, wS), S2.reshape(hS, wS), nz.reshape(hS, wS)


def image_from_values_and_points(w, h, v, p, xmin, ymin):
    p, pp = _pts(p)
    v, pv = _vals(v)
    out = np.zeros(w * h)
    lib().xo_image_from_values_and_points(out.ctypes.data_as(_dp), w, h, pv, pp, len(v), xmin, ymin)
    return out.reshape(h, w)


def bbox(p):
    p, pp = _pts(p)
    o = (C.c_int * 4)()
    lib().xo_bbox_from_points(pp, len(p), o)
    return tuple(o)


def boxes_from_shifts(w1, h1, w2, h2, x, y):
    b1, b2 = (C.c_int * 4)(), (C.c_int * 4)()
    lib().xo_boxes_from_shifts(w1, h1, w2, h2, x, y, b1, b2)
    return tuple(b1), tuple(b2)


def fft_size(n1, n2):
    return lib().xo_fft_size(n1, n2)


def ceil_pow2(n):
    return lib().xo_ceil_pow2(n)


def decimate_vector(p, v, w, h, lbin):
    p = np.array(p, dtype=np.float64).reshape(-1, 2)
    v = np.array(v, dtype=np.float64)
    k = lib().xo_decimate_vector(p.ctypes.data_as(_pp), v.ctypes.data_as(_dp), len(v), w, h, lbin)
    return p[:k].copy(), v[:k].copy()


def crop(av, bv, apts, bpts, tx, ty, xyconf, scl, olap1d, olap2d):
    """CThmUtil::Crop; av, bv: hs x ws pictures.  Returns (status, whole, (ap, av, a), (bp, bv, b))."""
    av = np.ascontiguousarray(av, dtype=np.float64)
    bv = np.ascontiguousarray(bv, dtype=np.float64)
    hs, ws = av.shape
    apts = np.ascontiguousarray(apts, dtype=np.float64).reshape(-1, 2)
    bpts = np.ascontiguousarray(bpts, dtype=np.float64).reshape(-1, 2)
    ap, bp = np.zeros_like(apts), np.zeros_like(bpts)
    avo, bvo = np.zeros(len(apts)), np.zeros(len(bpts))
    a, b, whole = XcSubimage(), XcSubimage(), C.c_int(-1)
    st = lib().xo_crop(av.ctypes.data_as(_dp), bv.ctypes.data_as(_dp), ws, hs, apts.ctypes.data_as(_pp), len(apts),
                       bpts.ctypes.data_as(_pp), len(bpts), tx, ty, xyconf, scl, olap1d, olap2d,
                       ap.ctypes.data_as(_pp), avo.ctypes.data_as(_dp), C.byref(a), bp.ctypes.data_as(_pp),
                       bvo.ctypes.data_as(_dp), C.byref(b), C.byref(whole))
    return st, whole.value, (ap[:a.n], avo[:a.n], a), (bp[:b.n], bvo[:b.n], b)


def crop_nocr(av, bv, tx, ty, xyconf, scl, olap1d):
    """CThmUtil::Crop_NoCR.  Returns (status, whole, points, (av, a), (bv, b))."""
    av = np.ascontiguousarray(av, dtype=np.float64)
    bv = np.ascontiguousarray(bv, dtype=np.float64)
    hs, ws = av.shape
    p = np.zeros((ws * hs, 2))
    avo, bvo = np.zeros(ws * hs), np.zeros(ws * hs)
    a, b, whole = XcSubimage(), XcSubimage(), C.c_int(-1)
    st = lib().xo_crop_nocr(av.ctypes.data_as(_dp), bv.ctypes.data_as(_dp), ws, hs, tx, ty, xyconf, scl, olap1d,
                            avo.ctypes.data_as(_dp), bvo.ctypes.data_as(_dp), p.ctypes.data_as(_pp), C.byref(a),
                            C.byref(b), C.byref(whole))
    return st, whole.value, p[:a.n], (avo[:a.n], a), (bvo[:b.n], b)


def rgn_ok(rule, olw, olh):
    return bool(lib().xo_rgn_ok(C.byref(rule), olw, olh))


MKL_CANDIDATES = ("/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", "libfftw3.so.3")


def set_fft_provider(which):
    """0: the oracle's radix-2 FFT.  1: the FFTW3 API from an installed library
    (libfftw3 or MKL's FFTW3 wrappers), called as the reference calls it."""
    if which == 0:
        lib().xo_set_fft_provider(0, None)
        return provider_name()
    for path in MKL_CANDIDATES:
        if lib().xo_set_fft_provider(1, path.encode()) == 0:
            return provider_name()
    raise RuntimeError("no library exporting the FFTW3 API was found")


def provider_name():
    return lib().xo_fft_provider_name().decode()


# ---- CThmScan angle search ----

def linefit_lincor(x, y):
    """LineFit's correlation coefficient (0_GEN/Maths.cpp:457-507)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().xo_linefit_lincor(x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), x.size)


def scan_angles(mode, center, hlfwid, step, favor):
    """The angle list of AngleScanMaxR / AngleScanConstXY (0_GEN/CThmScan.cpp:444-470, 543-544)."""
    n = lib().xo_scan_angles(mode, center, hlfwid, step, int(bool(favor)), None, 0)
    a = np.zeros(max(n, 1))
    lib().xo_scan_angles(mode, center, hlfwid, step, int(bool(favor)), a.ctypes.data_as(_dp), n)
    return a[:n]


def scan_select(mode, recs):
    """The selection of AngleScanMaxR / AngleScanConstXY over (A, R, X, Y) records."""
    arr = (XcAngleRec * len(recs))(*[XcAngleRec(*r) for r in recs])
    best = XcAngleRec()
    lib().xo_scan_select(mode, arr, len(recs), C.byref(best))
    return best


def angle_scan(mode, ap, av, bp, bv, center, hlfwid, step, favor, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
    """AngleScanMaxR / AngleScanConstXY (0_GEN/CThmScan.cpp:431-621) over RFromAngle. Returns (best, recs)."""
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    cap = 4096
    recs = (XcAngleRec * cap)()
    best, n = XcAngleRec(), C.c_int()
    lib().xo_angle_scan(mode, pp1, pv1, len(v1), pp2, pv2, len(v2), center, hlfwid, step, int(bool(favor)), olap1d,
                        req_area, Ox, Oy, Rx, Ry, C.byref(best), recs, cap, C.byref(n))
    return best, list(recs[:min(n.value, cap)])


def peak_hunt(best, ap, av, bp, bv, hlfwid, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
    """PeakHunt (0_GEN/CThmScan.cpp:676-732). Returns (best, k)."""
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    b, k = XcAngleRec(best.A, best.R, best.X, best.Y), C.c_int()
    lib().xo_peak_hunt(pp1, pv1, len(v1), pp2, pv2, len(v2), hlfwid, olap1d, req_area, Ox, Oy, Rx, Ry, C.byref(b),
                       C.byref(k))
    return b, k.value


def pretweaks(ap, av, bp, bv, deg, Tdfm, Tptwk, best_r, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
    """CThmScan::Pretweaks (0_GEN/CThmScan.cpp:205-323). Returns (Tptwk[6], best_r, changed, n_evals)."""
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
    W = np.array(Tptwk if Tptwk is not None else [1, 0, 0, 0, 1, 0], dtype=np.float64)
    r, ch, ne = C.c_double(best_r), C.c_int(), C.c_int()
    lib().xo_pretweaks(deg, pp1, pv1, len(v1), pp2, pv2, len(v2), None if D is None else D.ctypes.data_as(_dp),
                       W.ctypes.data_as(_dp), C.byref(r), olap1d, req_area, Ox, Oy, Rx, Ry, C.byref(ch), C.byref(ne))
    return W, r.value, ch.value, ne.value


def rfromangle_affine(deg, ap, av, bp, bv, Tdfm, Ttw, olap1d, req_area, Ox=0, Oy=0, Rx=-1, Ry=-1):
    """RFromAngle with deformation and pretweak affines. Returns (status, R, X, Y)."""
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
    W = None if Ttw is None else np.ascontiguousarray(Ttw, dtype=np.float64)
    X, Y, R = C.c_double(), C.c_double(), C.c_double()
    st = lib().xo_rfromangle_affine(deg, pp1, pv1, len(v1), pp2, pv2, len(v2),
                                    None if D is None else D.ctypes.data_as(_dp),
                                    None if W is None else W.ctypes.data_as(_dp), olap1d, req_area, Ox, Oy, Rx, Ry,
                                    C.byref(X), C.byref(Y), C.byref(R))
    return st, R.value, X.value, Y.value


def finish_at_full_res(best, ap, av, bp, bv, Tdfm, Tptwk, olap1d, req_area):
    """CThmScan::FinishAtFullRes (0_GEN/CThmScan.cpp:810-842). Returns (best, used)."""
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    D = None if Tdfm is None else np.ascontiguousarray(Tdfm, dtype=np.float64)
    W = None if Tptwk is None else np.ascontiguousarray(Tptwk, dtype=np.float64)
    b, u = XcAngleRec(best.A, best.R, best.X, best.Y), C.c_int()
    lib().xo_finish_at_full_res(pp1, pv1, len(v1), pp2, pv2, len(v2), None if D is None else D.ctypes.data_as(_dp),
                                None if W is None else W.ctypes.data_as(_dp), olap1d, req_area, C.byref(b), C.byref(u))
    return b, u.value


def _thm(name, ap, av, bp, bv, params, *args, ok=False, best=None):
    p1, pp1 = _pts(ap)
    v1, pv1 = _vals(av)
    p2, pp2 = _pts(bp)
    v2, pv2 = _vals(bv)
    b = XcAngleRec() if best is None else XcAngleRec(best.A, best.R, best.X, best.Y)
    okv = C.c_int()
    extra = [C.byref(b)] + ([C.byref(okv)] if ok else [])
    getattr(lib(), name)(pp1, pv1, len(v1), pp2, pv2, len(v2), C.byref(params), *args, *extra)
    return (b, okv.value) if ok else b


def thm_scan_with_tweaks(ap, av, bp, bv, params, center, hlfwid, step):
    """AngleScanWithTweaks (0_GEN/CThmScan.cpp:647-672); params updated in place."""
    return _thm("xo_thm_scan_with_tweaks", ap, av, bp, bv, params, center, hlfwid, step)


def thm_use_prior_angles(ap, av, bp, bv, params, ang0, hfangpr):
    """UsePriorAngles (0_GEN/CThmScan.cpp:734-758). Returns (best, ok)."""
    return _thm("xo_thm_use_prior_angles", ap, av, bp, bv, params, ang0, hfangpr, ok=True)


def thm_denovo_best_angle(ap, av, bp, bv, params, ang0, hfangdn, step):
    """DenovoBestAngle (0_GEN/CThmScan.cpp:764-790). Returns (best, ok)."""
    return _thm("xo_thm_denovo_best_angle", ap, av, bp, bv, params, ang0, hfangdn, step, ok=True)


def thm_post_tweaks(ap, av, bp, bv, params, best):
    """PostTweaks (0_GEN/CThmScan.cpp:796-802) from the record best."""
    return _thm("xo_thm_post_tweaks", ap, av, bp, bv, params, best=best)


def leg_poly_create(order, n):
    """LegPolyCreate (0_GEN/Maths.cpp:962-997): (order+1) x n."""
    L = np.zeros((order + 1) * n)
    lib().xo_leg_poly_create(L.ctypes.data_as(_dp), order, n)
    return L.reshape(order + 1, n)


def leg_poly_flatten(ras, order):
    """LegPolyFlatten for 8-bit rasters (0_GEN/Maths.cpp:1006-1086), then Normalize."""
    ras = np.ascontiguousarray(ras, dtype=np.uint8)
    out = np.zeros(ras.shape)
    st = lib().xo_leg_poly_flatten(out.ctypes.data_as(_dp), ras.ctypes.data, ras.shape[1], ras.shape[0], order)
    assert st == 0
    return out


def make_dog_kernel(r1, r2):
    """MakeDoGKernel (0_GEN/CPixPair.cpp:36-80): D x D."""
    D = lib().xo_make_dog_kernel(None, r1, r2)
    K = np.zeros(D * D)
    lib().xo_make_dog_kernel(K.ctypes.data_as(_dp), r1, r2)
    return K.reshape(D, D)


def pixpair_prep(ras, order=2, dog=False, r1=3, r2=6):
    """PixPair::Load's raster stage for one picture (0_GEN/CPixPair.cpp:276-460). Returns the same dict as
    Context.pixpair_prep, or None for the reference's failures."""
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
    st = lib().xo_pixpair_prep(ras.ctypes.data, w, h, order, int(bool(dog)), r1, r2, ptr(vf), ptr(vfflt), ptr(vs),
                               ptr(vsflt), C.byref(ws), C.byref(hs), C.byref(scl))
    if st:
        return None
    return {"vf": vf, "vfflt": vfflt, "vs": vs, "vsflt": vsflt, "ws": ws.value, "hs": hs.value, "scl": scl.value}


def downsample(src, scl):
    """PixPair::Downsample (0_GEN/CPixPair.cpp:90-113)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    hf, wf = src.shape
    ws, hs = wf // scl, hf // scl
    out = np.zeros((hs, ws))
    lib().xo_downsample(out.ctypes.data_as(_dp), src.ctypes.data_as(_dp), wf, scl, ws, hs)
    return out


def tightest_bbox(pts):
    """TightestBBox (0_GEN/Geometry.cpp:187-267): the whole-degree rotation of least box area."""
    p, pp = _pts(pts)
    return lib().xo_tightest_bbox(pp, len(p))


def fourier_match(pts, av, bv, wvlen=25):
    """FourierMatch (0_GEN/Metrics.cpp:517-658). Returns an XcFourierQa."""
    p, pp = _pts(pts)
    a, pa = _vals(av)
    b, pb = _vals(bv)
    d, qa = C.c_double(), XcFourierQa()
    lib().xo_fourier_match(pp, pa, pb, len(a), wvlen, C.byref(d), C.byref(qa))
    return qa
