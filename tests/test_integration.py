"""The ADDA-side binding (integration/hipmatvec.c) against the reference's own headers: it must compile with warnings
as errors next to the MatVec declaration that src/iterative.c:137-138 re-declares (a mismatch of the signature of
src/matvec.c:128-133 is then a compile error), and against src/fft.h's prototypes of the functions it replaces.
Skipped where the reference checkout is absent (the GPU box)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")


def _matvec_decl():
    """The MatVec declaration of src/iterative.c (text between 'void MatVec(' and the closing ');')."""
    with open(os.path.join(REF, "iterative.c")) as f:
        txt = f.read()
    m = re.search(r"void MatVec\s*\([^;]*\);", txt)
    assert m, "MatVec declaration not found in src/iterative.c"
    return m.group(0)


def test_hipmatvec_compiles_against_reference_headers(tmp_path):
    decl = tmp_path / "iterative_decl.h"
    decl.write_text('#include "types.h"\n#include "timing.h"\n#include <stdbool.h>\n' + _matvec_decl() + "\n")
    for hip_macro in ([], ["-DHIP"]):
        cmd = ["gcc", "-std=gnu99", "-fsyntax-only", "-Wall", "-Wextra", "-Wno-unused-parameter", "-Werror",
               "-DFFT_TEMPERTON", "-I" + REF, "-I" + os.path.join(ROOT, "include"), "-include", str(decl)] + \
            hip_macro + [os.path.join(ROOT, "integration", "hipmatvec.c")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_hipmatvec_signature_mismatch_is_caught(tmp_path):
    """The check above has teeth: a MatVec declared with another signature does not compile."""
    decl = tmp_path / "bad_decl.h"
    decl.write_text('#include "types.h"\n#include "timing.h"\n#include <stdbool.h>\n'
                    + _matvec_decl().replace("bool her", "int her") + "\n")
    cmd = ["gcc", "-std=gnu99", "-fsyntax-only", "-DFFT_TEMPERTON", "-I" + REF, "-I" + os.path.join(ROOT, "include"),
           "-include", str(decl), os.path.join(ROOT, "integration", "hipmatvec.c")]
    assert subprocess.run(cmd, capture_output=True).returncode != 0


def test_hip_flavor_links():
    """ADDA built with the binding in place of matvec.c and fft.c (integration/Makefile) links against
    libadda_hip.so: every symbol the rest of ADDA needs from the two replaced files is provided."""
    exe = os.path.join(ROOT, "oracle", "_ref", "adda_hip_flavor")
    r = subprocess.run(["make", "-f", os.path.join("integration", "Makefile"), "-j8", "REF=/root/reference"],
                       cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert os.path.exists(exe)
    nm = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
    for sym in ("MatVec", "InitDmatrix", "Free_FFT_Dmat", "fftFit", "IterativeSolver"):
        assert re.search(r" T %s$" % sym, nm, re.M), sym
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libadda_hip.so" in ldd and "not found" not in ldd
