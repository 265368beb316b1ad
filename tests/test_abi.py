"""The C-ABI library loads on a machine without a GPU and exports every symbol its headers declare."""
import ctypes
import glob
import os
import re
import subprocess

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


# each POD struct of the header and its ctypes mirror
MIRRORS = {"xc_point": "XcPoint", "xc_rgn_rule": "XcRgnRule", "xc_cnt_rule": "XcCntRule", "xc_result": "XcResult",
           "xc_pair": "XcPair", "xc_bucket_info": "XcBucketInfo", "xc_angle_rec": "XcAngleRec",
           "xc_thm_params": "XcThmParams", "xc_subimage": "XcSubimage", "xc_fourier_qa": "XcFourierQa",
           "xc_thmpair": "XcThmPair", "xc_serve_cfg": "XcServeCfg", "xc_serve_stats": "XcServeStats"}


def _header_layouts(tmp_path):
    """sizeof and offsetof of every field of the mirrored structs, as the C compiler lays out the header."""
    hdr = open(os.path.join(ROOT, "include", "xcorr.h")).read()
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "xcorr.h"', "int main(void) {"]
    for cname, pyname in MIRRORS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in getattr(abi, pyname)._fields_:
            if re.search(rf"\b{f}\b", hdr):
                lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    lay = {}
    for ln in out.split("\n"):
        if ln:
            c, f, v = ln.split()
            lay[(c, f)] = int(v)
    return lay


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the C compiler's size and field offsets for every struct of the header."""
    lay = _header_layouts(tmp_path)
    for cname, pyname in MIRRORS.items():
        cls = getattr(abi, pyname)
        assert ctypes.sizeof(cls) == lay[(cname, "size")], (cname, ctypes.sizeof(cls), lay[(cname, "size")])
        for f, _ in cls._fields_:
            if (cname, f) in lay:
                assert getattr(cls, f).offset == lay[(cname, f)], (cname, f)
    assert lay[("xc_serve_stats", "failed")] == 80 and lay[("xc_serve_cfg", "deadline_us")] == 20


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
