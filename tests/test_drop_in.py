"""The C++ drop-in covers 0_GEN/Correlation.h: every function the header declares is either defined by
libaln_corr.so or kept from the reference's own object by the link recipe (kept_from_reference.txt), and the
recipe's script (localize_replaced.sh) makes a link take the drop-in's definitions while keeping the others.
Host code only: runs without a device."""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT = os.path.join(ROOT, "alignment_projects_amd", "compat")
LIB = os.path.join(ROOT, "alignment_projects_amd", "lib")
FIXTURE = os.path.join(ROOT, "tests", "golden", "correlation_h_functions.json")


def _declared():
    return json.load(open(FIXTURE))["functions"]


def _kept():
    names = []
    for ln in open(os.path.join(COMPAT, "kept_from_reference.txt")):
        ln = ln.split("#", 1)[0].strip()
        if ln:
            names.append(ln)
    return names


def _exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(LIB, "libaln_corr.so")], check=True,
                         capture_output=True, text=True).stdout
    return {m.group(1) for m in re.finditer(r" [TW] ([A-Za-z_][A-Za-z0-9_]*)\(", out)}


def test_fixture_matches_the_reference_header():
    hdr = "/root/reference/0_GEN/Correlation.h"
    if not os.path.exists(hdr):
        pytest.skip("reference checkout not present")
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(ROOT, "tests", "golden",
                                                                     "make_correlation_h_functions.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    assert mk.declared_functions(hdr) == _declared()


def test_every_declared_function_is_defined_or_kept():
    declared, kept, exported = _declared(), _kept(), _exported()
    assert len(declared) == 16
    missing = [f for f in declared if f not in exported and f not in kept]
    assert not missing, missing
    assert not [f for f in kept if f in exported], "a kept function is also defined by the drop-in"
    assert not [f for f in kept if f not in declared]
    for f in ("Convolve", "CorrPatches", "CorrPatchToImage", "CorrImagesF", "CorrImagesR", "CorrImagesS", "FFT_2D",
              "IFT_2D", "FFTSize", "FFTSizeSQR", "ParabPeakFFT"):
        assert f in exported, f


def test_localized_object_links_with_the_drop_in(tmp_path):
    """An object standing in for Correlation.o that defines a replaced function (FFTSize, returning 0) and a
    kept one (RVectors): after the script, FFTSize is local and RVectors global; a program linked with the
    object and the drop-in gets the drop-in's FFTSize (CeilPow2(n1 + n2 - 1)) and the object's RVectors."""
    src = tmp_path / "corr.cpp"
    src.write_text(
        "#include <vector>\n"
        "int FFTSize(int, int) { return 0; }\n"
        "double RVectors(const std::vector<double> &a, const std::vector<double> &) { return a.size(); }\n")
    obj = tmp_path / "corr.o"
    subprocess.run(["g++", "-c", "-fPIC", str(src), "-o", str(obj)], check=True)
    r = subprocess.run(["sh", os.path.join(COMPAT, "localize_replaced.sh"), str(obj),
                        os.path.join(LIB, "libaln_corr.so")], check=True, capture_output=True, text=True)
    assert "made local 1 functions" in r.stdout
    syms = subprocess.run(["nm", "-C", str(obj)], check=True, capture_output=True, text=True).stdout
    assert re.search(r" t FFTSize\(int, int\)", syms) and re.search(r" T RVectors\(", syms), syms
    main = tmp_path / "main.cpp"
    main.write_text(
        "#include <stdio.h>\n#include <vector>\n"
        "int FFTSize(int, int);\n"
        "double RVectors(const std::vector<double> &, const std::vector<double> &);\n"
        "int main() { std::vector<double> v(3); printf(\"%d %g\\n\", FFTSize(513, 512), RVectors(v, v)); }\n")
    exe = tmp_path / "main"
    subprocess.run(["g++", str(main), str(obj), "-L" + LIB, "-laln_corr", "-lxcorr", "-Wl,-rpath," + LIB, "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["1024", "3"], out
