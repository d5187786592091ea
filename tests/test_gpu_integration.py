"""ADDA itself, built with its FFT MatVec replaced by the HIP flavor (integration/hipmatvec.c over libadda_hip.so,
integration/Makefile -> oracle/_ref/adda_hip_flavor), run on the GPU: the reference's own IterativeSolver, all seven
Krylov methods, InitCC, GenerateB and post-processing call our MatVec through the drop-in boundary, and reproduce the
reference's SPARSE goldens (comp2exec tolerances: CrossSec abs 1e-14 / rel 1e-8, mueller abs 1e-9 / rel 1e-5).
The executable is built in the build container (where the reference sources are); skipped when it is absent."""
import os
import subprocess

import numpy as np
import pytest

import golden_io as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "adda_hip_flavor")


def _run(tmp_path, args, name):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/adda_hip_flavor not built (needs the reference sources)")
    run = tmp_path / name
    cmd = [EXE] + args + ["-dir", str(run)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=tmp_path,
                       env=dict(os.environ, ADDA_HIP_GPU="0"))
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return run


def _crosssec(path):
    with open(path) as f:
        return {l.split("=")[0].strip(): float(l.split("=")[1]) for l in f if "=" in l}


@pytest.mark.parametrize("method,golden", [("qmr", "c1_qmr"), ("bicgstab", "c1_bicgstab"), ("cgnr", "c1_cgnr"),
                                           ("bicg", "c1_bicg"), ("csym", "c1_csym"), ("bcgs2", "c1_bcgs2"),
                                           ("qmr2", "c1_qmr2")])
def test_adda_solvers_on_hip_matvec(hipdev, tmp_path, method, golden):
    """ADDA's solvers over the GPU MatVec, sphere 16^3 (grid 32^3), against the SPARSE build's results."""
    run = _run(tmp_path, ["-shape", "read", G.path("geom", "sphere16.geom"), "-dpl", "15", "-m", "1.5", "0", "-iter",
                          method, "-eps", "10"], method)
    for pol in ("Y", "X"):
        a, b = G.crosssec(golden, pol), _crosssec(run / ("CrossSec-" + pol))
        assert G.numeric_close([b["Cext"], b["Cabs"]], [a["Cext"], a["Cabs"]], 1e-14, 1e-8)[0], pol
    if method == "qmr":
        ours = np.loadtxt(run / "mueller", skiprows=1)
        ref = G.mueller("c1_qmr")
        assert ours.shape == ref.shape
        assert G.numeric_close(ours, ref, 1e-9, 1e-5)[0]


def test_adda_surface_on_hip_matvec(hipdev, tmp_path):
    """ADDA's surface mode (-surf, image dipole) over the GPU MatVec with its substrate term: Cext, Cabs and DipPol
    of the reference's SPARSE -surf run."""
    run = _run(tmp_path, ["-shape", "read", G.path("geom", "sphere8.geom"), "-dpl", "15", "-m", "1.5", "0.01",
                          "-int_surf", "img", "-iter", "qmr", "-eps", "10", "-store_dip_pol", "-surf", "2.0", "1.5",
                          "0.1", "-prop", "0", "0", "-1"], "surf")
    for pol in ("Y", "X"):
        a, b = G.crosssec("surf_img_above", pol), _crosssec(run / ("CrossSec-" + pol))
        assert G.numeric_close([b["Cext"], b["Cabs"]], [a["Cext"], a["Cabs"]], 1e-14, 1e-8)[0], pol
    ours = np.loadtxt(run / "DipPol-Y", skiprows=1)
    ref = np.loadtxt(G.path("surf_img_above", "DipPol-Y"), skiprows=1)
    assert G.numeric_close(ours, ref, 1e-12, 1e-6)[0]
