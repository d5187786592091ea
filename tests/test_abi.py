"""CPU checks of the C ABI: every function declared in include/*.h is exported by the built libraries, the host-side
logic that needs no GPU (particle builder, polarizability, grid sizing) agrees with the oracle/test restatements, and
the product path fails loudly without a GPU (no CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

from cases import make_shape, oracle, Case, PROP_Z, POL_Y

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(adda_\w+)\s*\(", txt)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("adda_hip.h", "libadda_hip.so"), ("adda.h", "libadda.so")])
def test_every_declared_symbol_is_exported(header, lib):
    L = ctypes.CDLL(os.path.join(ROOT, "a-dda_amd", lib))
    names = _declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_fails_loudly_without_gpu():
    import adda_amd
    if adda_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    pos, box = make_shape("sphere", 8)
    with pytest.raises(RuntimeError, match="not available"):
        adda_amd.Operator(pos, box, (16, 16, 16), np.zeros(len(pos), np.int32), 1, 1.0, 0.4)


def test_fftfit_matches_oracle_rule():
    import adda_amd
    L = adda_amd.hip()
    for x in range(2, 700):
        assert L.adda_hip_fftfit(x, 1) == oracle().or_fftfit(x, 1, 1)


def test_fastgrid_rule():
    """The speed-oriented grid: a power of two at most 25 % above fftFit's size, else fftFit's size."""
    import adda_amd
    L = adda_amd.hip()
    assert [L.adda_hip_fastgrid(x, 1) for x in (104, 256, 640, 96, 120, 32)] == [128, 256, 640, 96, 128, 32]
    for x in range(2, 2100):
        n, g = L.adda_hip_fftfit(x, 1), L.adda_hip_fastgrid(x, 1)
        assert g >= x and g % 2 == 0
        assert g == n or (g & (g - 1) == 0 and 4 * g <= 5 * n), (x, n, g)
    assert adda_amd.fast_grid((64, 52, 32)) == (128, 128, 64)
    # divisors (ADDA passes process counts): an odd factor rules out the power of two, a factor other than 2, 3, 5
    # any size (an error, not an endless search)
    assert L.adda_hip_fastgrid(104, 3) == L.adda_hip_fftfit(104, 3) == 108
    assert L.adda_hip_fastgrid(100, 5) == L.adda_hip_fftfit(100, 5) == 100
    assert L.adda_hip_fastgrid(120, 4) == 128 and L.adda_hip_fftfit(120, 4) == 120
    assert L.adda_hip_fftfit(100, 7) < 0 and L.adda_hip_fastgrid(100, 7) < 0


@pytest.mark.parametrize("shape,grid,ay,az,nd", [("sphere", 16, 1, 1, 2176), ("ellipsoid", 64, 0.8, 0.5, None),
                                                  ("box", 10, 0.7, 0.4, None), ("sphere", 128, 1, 1, 1099136)])
def test_particle_builder_matches_restatement(shape, grid, ay, az, nd):
    import adda_amd
    pos, mat, box = adda_amd.make_particle(shape, grid, ay, az)
    ref, rbox = make_shape(shape, grid, ay, az)
    assert tuple(box) == tuple(rbox)
    assert np.array_equal(pos, ref)
    if nd:
        assert len(pos) == nd  # 2176: sample log; 1099136: config 2


def test_couple_constant_matches_oracle():
    import adda_amd
    pos, box = make_shape("sphere", 8)
    for m in (1.5, 1.33 + 0.001j, 2.0 + 0.5j):
        c = Case(pos, box, m, dpl=13.3)
        for polrel in (0, 1, 2):
            for prop, pol in ((PROP_Z, POL_Y), (np.array([0.6, 0, 0.8]), np.array([0, 0.8, 0.6]))):
                cc, _ = c.cc(prop, pol, polrel)
                v = adda_amd.couple_constant(m, polrel, c.wavenum, c.gridspace, prop, pol)
                assert abs(v - cc[0, 0]) <= 1e-15 * abs(v)


def test_volume_ratio():
    import adda_amd
    L = adda_amd.host()
    assert abs(L.adda_volume_ratio(b"sphere", 1, 1) - np.pi / 6) < 1e-15
    assert abs(L.adda_volume_ratio(b"ellipsoid", 0.8, 0.5) - np.pi / 6 * 0.4) < 1e-15
