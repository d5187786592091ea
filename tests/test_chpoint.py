"""ADDA's iteration checkpoint format (SaveIterChpoint / LoadIterChpoint, src/iterative.c:165-278) on the CPU: the
reference's own checkpoint files (tests/golden/chp_*/chp.0, written by its SPARSE build after 4 iterations of each
solver on a 64-dipole cube) are read by adda_chp_read and written back byte for byte by adda_chp_write; the reference's
errors for another method, another vector size and a longer file. No GPU: only the host library is called."""
import os
import shutil

import numpy as np
import pytest

import golden_io as G

METHODS = ["bcgs2", "bicg", "bicgstab", "cgnr", "csym", "qmr", "qmr2"]
NROWS = 3 * 64  # local_nRows of the 4x4x4 cube


@pytest.fixture(scope="module")
def A():
    import adda_amd
    return adda_amd


@pytest.mark.parametrize("method", METHODS)
def test_reference_checkpoint_round_trip(A, method, tmp_path):
    src = G.path("chp_" + method, "chp.0")
    ind, ns, nv = A.chp_layout(method)
    # the file is exactly the header, the method's scalars and 3 + nv vectors (no padding between fields)
    assert os.path.getsize(src) == 4 + 8 + 4 + 4 + 3 * 8 + 8 * ns + 16 * NROWS * (3 + nv)
    with open(src, "rb") as f:
        assert int(np.frombuffer(f.read(4), np.int32)[0]) == ind
    head, sc, vecs = A.chp_read(os.path.dirname(src), method, NROWS)
    # -maxiter 4 ends the loop at niter = 5; resid_scale = 1/||S.Einc||^2, prev_err = sqrt(resid_scale*|r_4|^2)
    assert head["niter"] == 5 and head["resid_scale"] > 0
    assert np.all(np.isfinite(sc)) and np.all(np.isfinite(vecs)) and np.abs(vecs[0]).max() > 0
    out = tmp_path / "chp"
    A.chp_write(str(out), method, head, sc, vecs, run_dir="part")
    with open(src, "rb") as f1, open(out / "chp.0", "rb") as f2:
        assert f1.read() == f2.read()
    with open(out / "chp.log") as f:
        assert f.read() == "Info about the run, which produced the checkpoint, can be found in ../part"


def test_reference_checkpoint_errors(A, tmp_path):
    d = os.path.dirname(G.path("chp_qmr", "chp.0"))
    with pytest.raises(RuntimeError, match="different iterative method"):
        A.chp_read(d, "bicgstab", NROWS)
    with pytest.raises(RuntimeError, match="different vector size"):
        A.chp_read(d, "qmr", NROWS + 3)
    shutil.copy(os.path.join(d, "chp.0"), tmp_path / "chp.0")
    with open(tmp_path / "chp.0", "ab") as f:
        f.write(b"\0")
    with pytest.raises(RuntimeError, match="too long"):
        A.chp_read(str(tmp_path), "qmr", NROWS)
    with pytest.raises(RuntimeError, match="cannot open"):
        A.chp_read(str(tmp_path / "missing"), "qmr", NROWS)
