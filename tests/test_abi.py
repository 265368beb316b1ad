"""The C-ABI library loads on a machine without a GPU and exports every symbol its headers declare."""
import ctypes
import glob
import os
import re

from alignment_projects_amd import abi, xcorr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xc_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = xcorr.load()
    names = declared(os.path.join(ROOT, "include", "xcorr.h"))
    assert "xc_corr_patches_f64" in names and "xc_plan_create" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # sizes of the POD structs the header defines (x86-64 / gfx950 share the layout)
    assert ctypes.sizeof(abi.XcPoint) == 16
    assert ctypes.sizeof(abi.XcRgnRule) == 56
    assert ctypes.sizeof(abi.XcCntRule) == 24
    assert ctypes.sizeof(abi.XcResult) == 72
    assert ctypes.sizeof(abi.XcBucketInfo) == 32
    assert ctypes.sizeof(abi.XcServeCfg) == 24
    assert ctypes.sizeof(abi.XcServeStats) == 64
    assert ctypes.sizeof(abi.XcAngleRec) == 32
    assert ctypes.sizeof(abi.XcThmPair) == 6 * 4 + 2 * 8 + 6 * 8
    assert ctypes.sizeof(abi.XcThmParams) == 8 + 8 * 4 + 8 + 12 * 8
    assert ctypes.sizeof(abi.XcFourierQa) == 5 * 8 + 6 * 4
    assert ctypes.sizeof(abi.XcSubimage) == 32
    assert ctypes.sizeof(abi.XcPair) == 2 * 8 + 6 * 4 + 8 + 56 + 24


def test_version_and_sizes_without_device():
    assert "gfx950" in xcorr.version()
    assert xcorr.fft_size(512, 512) == 1024
    L = xcorr.load()
    assert L.xc_fft_size_sqr(5, 3, 7, 2) == 16


def test_no_device_is_an_error_not_a_crash():
    import pytest
    L = xcorr.load()
    h = ctypes.c_void_p()
    st = L.xc_ctx_create(0, ctypes.byref(h))
    if st == abi.XC_OK:
        L.xc_ctx_destroy(h)
        pytest.skip("a device is present")
    assert st == abi.XC_E_DEVICE


def test_compat_library_exports_reference_signatures():
    import subprocess
    lib = os.path.join(ROOT, "alignment_projects_amd", "lib", "libaln_corr.so")
    syms = subprocess.run(["nm", "-D", "-C", "--defined-only", lib], capture_output=True, text=True,
                          check=True).stdout
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "aln_corr_compat.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b([A-Z][A-Za-z0-9_]+)\s*\(", text)) - {"Point"})
    assert {"CorrPatches", "CorrPatchToImage", "CorrImagesF", "CorrImagesR", "CorrImagesS", "FFT_2D", "IFT_2D"} <= set(names)
    missing = [n for n in names if f" {n}(" not in syms]
    assert not missing, missing


def test_thm_params_default_matches_the_library():
    """xc_thm_params_default (host only) writes the same block, field by field, as the Python default."""
    L = xcorr.load()
    a, b = abi.XcThmParams(), abi.XcThmParams.default()
    a.n_evals = 7
    L.xc_thm_params_default(ctypes.byref(a))
    for name, _ in abi.XcThmParams._fields_:
        va, vb = getattr(a, name), getattr(b, name)
        assert (list(va) == list(vb)) if hasattr(va, "__len__") else va == vb, name
