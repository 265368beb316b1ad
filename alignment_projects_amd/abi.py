"""ctypes mirror of include/xcorr.h: status codes, POD structs and rule builders.

The legality rules are the descriptor form of the reference's EvalType
callbacks (0_GEN/Correlation.h:91); each builder names the callback it
replaces.
"""
import ctypes as C

XC_OK = 0
XC_NO_PEAK = 1
XC_TOO_SMALL = 2
XC_E_NOTREADY = 3
XC_E_ODD_R = -44
XC_E_ARG = -2
XC_E_SHAPE = -3
XC_E_DEVICE = -4
XC_E_NOMEM = -5
XC_E_BUDGET = -6
XC_E_UNSUPPORTED = -7
XC_E_TIMEOUT = -8
XC_E_CANCELLED = -9

STATUS_NAMES = {
    XC_OK: "XC_OK", XC_NO_PEAK: "XC_NO_PEAK", XC_TOO_SMALL: "XC_TOO_SMALL",
    XC_E_NOTREADY: "XC_E_NOTREADY", XC_E_ODD_R: "XC_E_ODD_R", XC_E_ARG: "XC_E_ARG",
    XC_E_SHAPE: "XC_E_SHAPE", XC_E_DEVICE: "XC_E_DEVICE", XC_E_NOMEM: "XC_E_NOMEM",
    XC_E_BUDGET: "XC_E_BUDGET", XC_E_UNSUPPORTED: "XC_E_UNSUPPORTED",
    XC_E_TIMEOUT: "XC_E_TIMEOUT", XC_E_CANCELLED: "XC_E_CANCELLED",
}

XC_MODE_NCC = 0
XC_MODE_R = 1
XC_MODE_S = 2
XC_MODE_F = 3

XC_U8, XC_U16, XC_F32, XC_F64 = 0, 1, 2, 3
XC_TILE_NORMALIZE = 1
XC_TILE_DEVICE_SRC = 2
XC_TILE_F32_TABLES = 4
XC_TILE_F64_TABLES = 8


class XcPoint(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double)]


class XcRgnRule(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("min_w", C.c_int32), ("min_h", C.c_int32),
        ("exact_w", C.c_int32), ("exact_h", C.c_int32),
        ("or_w_ge", C.c_int32), ("or_h_gt", C.c_int32),
        ("table_w", C.c_int32), ("table_h", C.c_int32),
        ("_pad", C.c_int32),
        ("min_area", C.c_int64),
        ("table", C.POINTER(C.c_uint8)),
    ]


class XcCntRule(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("_pad", C.c_int32),
                ("k1", C.c_int64), ("k2", C.c_int64)]


class XcResult(C.Structure):
    _fields_ = [
        ("r", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("peak", C.c_double), ("sub_x", C.c_double), ("sub_y", C.c_double),
        ("status", C.c_int32), ("lag_x", C.c_int32), ("lag_y", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("_pad", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("_")}


class XcAngleRec(C.Structure):
    """CorRec without its affine (CThmScan.h:36-44): angle, score, displacement."""
    _fields_ = [("A", C.c_double), ("R", C.c_double), ("X", C.c_double), ("Y", C.c_double)]


# xc_ang_proc: void (*)(int *Ox, int *Oy, int *Rx, int *Ry, double deg, void *arg)
ANG_PROC = C.CFUNCTYPE(None, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                       C.c_double, C.c_void_p)


class XcThmParams(C.Structure):
    """CThmScan's matcher state for one pair (CThmScan.h:85-97) as the thm entry points read and update it."""
    _fields_ = [("rthresh", C.c_double), ("const_xy", C.c_int32), ("pretweak", C.c_int32), ("olap1D", C.c_int32),
                ("Ox", C.c_int32), ("Oy", C.c_int32), ("Rx", C.c_int32), ("Ry", C.c_int32), ("n_evals", C.c_int32),
                ("req_area", C.c_int64), ("Tdfm", C.c_double * 6), ("Tptwk", C.c_double * 6),
                ("ang_proc", C.c_void_p), ("ang_arg", C.c_void_p)]

    @classmethod
    def default(cls, olap1D=0, req_area=0):
        """CThmScan's constructor values (CThmScan.cpp:160-177)."""
        p = cls()
        p.rthresh, p.const_xy, p.pretweak, p.olap1D, p.req_area = 0.30, 1, 1, olap1D, req_area
        p.Ox, p.Oy, p.Rx, p.Ry = 0, 0, -1, -1
        for i, v in enumerate((1, 0, 0, 0, 1, 0)):
            p.Tdfm[i] = v
            p.Tptwk[i] = v
        return p


class XcFourierQa(C.Structure):
    """FourierMatch's metric and the figures the reference logs beside it (0_GEN/Metrics.cpp:517-658)."""
    _fields_ = [("dot", C.c_double), ("energy", C.c_double), ("msd", C.c_double), ("tot_power", C.c_double),
                ("half_frac", C.c_double), ("half_index", C.c_int32), ("maxf", C.c_int32), ("deg", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("_pad", C.c_int32)]


class XcSubimage(C.Structure):
    """One side of CThmUtil::Crop's overlap (SubI, 1_DMesh/CThmUtil.h:14-20): corner, size, point count."""
    _fields_ = [("Ox", C.c_double), ("Oy", C.c_double), ("w", C.c_int32), ("h", C.c_int32), ("n", C.c_int64)]


class XcThmPair(C.Structure):
    """One ThmPair_za^zb.txt entry (ThmPair, 0_GEN/PipeFiles.h:145-152)."""
    _fields_ = [("atl", C.c_int32), ("acr", C.c_int32), ("btl", C.c_int32), ("bcr", C.c_int32),
                ("err", C.c_int32), ("_pad", C.c_int32), ("A", C.c_double), ("R", C.c_double),
                ("T", C.c_double * 6)]


XC_SCAN_MAXR = 0
XC_SCAN_CONSTXY = 1


class XcPair(C.Structure):
    _fields_ = [
        ("a", C.c_void_p), ("b", C.c_void_p),
        ("mode", C.c_int32), ("Ox", C.c_int32), ("Oy", C.c_int32),
        ("radius", C.c_int32), ("Rx", C.c_int32), ("Ry", C.c_int32),
        ("angle_deg", C.c_double),
        ("rgn", XcRgnRule), ("cnt", XcCntRule),
    ]


def rgn_none():
    """NULL LegalRgn callback: every overlap is legal."""
    return XcRgnRule()


def rgn_big_enough(olap1d, req_area):
    """BigEnough (0_GEN/CThmScan.cpp:28-33): w>=olap1D && h>=olap1D && w*h > reqArea."""
    return XcRgnRule(enabled=1, min_w=olap1d, min_h=olap1d, min_area=req_area)


def rgn_some_area():
    """SomeArea (Retired_Proj/4_TestNormCorr/testnormcorr.cpp:11-14): w > 1 && h > 1."""
    return XcRgnRule(enabled=1, min_w=2, min_h=2, min_area=-1)


def rgn_in_section_legal():
    """InSectionLegal (1_DMesh/InSectionOverlap.cpp:41-47): n > 30000 && (w >= 1500 || h > 1500)."""
    return XcRgnRule(enabled=1, min_area=30000, or_w_ge=1500, or_h_gt=1500)


def rgn_exact(w, h):
    """lr / lr31x31 (1_Mosaic_current/mos.cpp:1270, 0_GEN/Inspect.cpp:631): w == W && h == H."""
    return XcRgnRule(enabled=1, exact_w=w, exact_h=h, min_area=-1)


def rgn_min(min_w, min_h, min_area=-1):
    return XcRgnRule(enabled=1, min_w=min_w, min_h=min_h, min_area=min_area)


def cnt_none():
    """NULL / lc / lcTRUE LegalCnt callbacks: always legal."""
    return XcCntRule()


def cnt_enough_points(req_area):
    """EnoughPoints (0_GEN/CThmScan.cpp:39-42): c1 > reqArea && c2 > reqArea."""
    return XcCntRule(enabled=1, k1=req_area, k2=req_area)


def check(status, what="xcorr call", allow=(XC_OK, XC_NO_PEAK, XC_TOO_SMALL), msg=None):
    if status not in allow:
        name = STATUS_NAMES.get(status, str(status))
        raise RuntimeError(f"{what} failed: {name}" + (f" ({msg})" if msg else ""))
    return status


XC_PLAN_REFRESH_TILES = 1
XC_PLAN_AUTOTUNE = 2


class XcBucketInfo(C.Structure):
    _fields_ = [("nh", C.c_int32), ("ny", C.c_int32), ("whiten", C.c_int32), ("pairs", C.c_int32),
                ("col_c", C.c_int32), ("col_xcd", C.c_int32), ("col_us", C.c_double),
                ("r2c_rpt", C.c_int32), ("r2c_np", C.c_int32), ("r2c_us", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class XcServeCfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_batch", C.c_int32), ("streams", C.c_int32),
                ("max_wait_us", C.c_int32), ("deadline_us", C.c_int32)]


class XcServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("batches", C.c_int64), ("spectra_hits", C.c_int64), ("spectra_misses", C.c_int64),
                ("arena_bytes", C.c_int64), ("arena_peak", C.c_int64), ("timed_out", C.c_int64),
                ("cancelled", C.c_int64), ("failed", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}
