"""Regenerates the golden fixtures under tests/golden/ with ADDA's own SPARSE-mode executable.

The executable is built from the reference sources by oracle/Makefile.ref into oracle/_ref/adda_sparse (direct-sum
MatVec, same Green's tensor, polarizability, Krylov solvers and post-processing as the FFT path). This script runs it
in this container only; the fixtures it writes are plain text data (geometry, CrossSec-*, mueller, DipPol-*, log
residual lines) and are committed with the exact command line of every run in cmd.txt.

    python tests/golden/make_golden.py [name ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from cases import ADJ_DELTA, adj8_materials, make_shape, orientations, read_shape  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "adda_sparse")


def write_geom(path, pos, mat=None):
    with open(path, "w") as f:
        if mat is None:
            for p in pos:
                f.write("%d %d %d\n" % tuple(p))
        else:
            f.write("Nmat=%d\n" % (int(mat.max()) + 1))
            for p, m in zip(pos, mat):
                f.write("%d %d %d %d\n" % (p[0], p[1], p[2], m + 1))


def geometries():
    g = {}
    pos, _ = make_shape("sphere", 16)
    g["sphere16"] = (pos, None)
    pos, _ = make_shape("ellipsoid", 12, 0.8, 0.5)
    g["ell12"] = (pos, None)
    pos, _ = make_shape("sphere", 8)
    g["sphere8"] = (pos, None)
    pos, _ = make_shape("box", 8, 0.75, 0.5)
    rng = np.random.default_rng(20261203)
    g["box8mm"] = (pos, rng.integers(0, 3, len(pos)))
    pos, _ = make_shape("box", 8)
    g["adj8"] = (pos, adj8_materials(len(pos))[1])
    pos, _ = make_shape("box", 4)
    g["box4"] = (pos, None)
    return g




RUNS = {
    # name: (geometry, args, files to keep)
    "c1_qmr": ("sphere16", "-dpl 15 -m 1.5 0 -iter qmr -eps 10", ["CrossSec-Y", "CrossSec-X", "mueller"]),
    "c1_qmr_e5": ("sphere16", "-dpl 15 -m 1.5 0 -iter qmr -eps 5", ["CrossSec-Y", "log"]),
    "c1_bicgstab": ("sphere16", "-dpl 15 -m 1.5 0 -iter bicgstab -eps 10", ["CrossSec-Y", "CrossSec-X"]),
    "c1_cgnr": ("sphere16", "-dpl 15 -m 1.5 0 -iter cgnr -eps 10", ["CrossSec-Y", "CrossSec-X"]),
    "c1_bicg": ("sphere16", "-dpl 15 -m 1.5 0 -iter bicg -eps 10", ["CrossSec-Y", "CrossSec-X"]),
    "c1_csym": ("sphere16", "-dpl 15 -m 1.5 0 -iter csym -eps 10", ["CrossSec-Y", "CrossSec-X"]),
    "c1_bcgs2": ("sphere16", "-dpl 15 -m 1.5 0 -iter bcgs2 -eps 10", ["CrossSec-Y", "CrossSec-X", "log"]),
    "c1_qmr2": ("sphere16", "-dpl 15 -m 1.5 0 -iter qmr2 -eps 10", ["CrossSec-Y", "CrossSec-X", "log"]),
    "ell12_orient": ("ell12", "-dpl 12 -m 1.33 0.05 -orient 30 40 50 -iter qmr -eps 10 -ntheta 19 -store_dip_pol",
                     ["CrossSec-Y", "CrossSec-X", "mueller", "DipPol-Y", "DipPol-X"]),
    "sphere8_cm": ("sphere8", "-dpl 10 -m 1.1 0.1 -pol cm -iter qmr -eps 10 -ntheta 19 -store_dip_pol",
                   ["CrossSec-Y", "CrossSec-X", "mueller", "DipPol-Y"]),
    "box8_multimat": ("box8mm", "-dpl 14 -m 1.2 0.01 1.5 0.1 1.8 0 -iter bicgstab -eps 10 -ntheta 19 -store_dip_pol",
                      ["CrossSec-Y", "CrossSec-X", "DipPol-Y"]),
}
# particle above a substrate, image-dipole reflection (-int_surf img): incidence from above and from below (the default
# direction), a perfect reflector, oblique incidence from above, and from below beyond the critical angle (evanescent
# transmitted wave)
SURF = "-dpl 15 -m 1.5 0.01 -int_surf img -iter qmr -eps 10 -store_dip_pol"
RUNS.update({
    "surf_img_above": ("sphere8", SURF + " -surf 2.0 1.5 0.1 -prop 0 0 -1", ["CrossSec-Y", "CrossSec-X", "DipPol-Y"]),
    "surf_img_below": ("sphere8", SURF + " -surf 2.0 1.5 0.1", ["CrossSec-Y", "CrossSec-X", "DipPol-Y"]),
    "surf_inf_above": ("sphere8", SURF + " -surf 2.0 inf -prop 0 0 -1", ["CrossSec-Y", "CrossSec-X", "DipPol-Y"]),
    "surf_img_oblique": ("box8mm", "-dpl 14 -m 1.2 0.01 1.5 0.1 1.8 0 -int_surf img -iter qmr -eps 10 -store_dip_pol "
                         "-surf 1.7 2.0 0.5 -prop 0.5 0.3 -1", ["CrossSec-Y", "CrossSec-X", "DipPol-Y", "DipPol-X"]),
    "surf_img_tir": ("sphere8", SURF + " -surf 1.9 1.5 0 -prop 0.8 0 0.6",
                     ["CrossSec-Y", "CrossSec-X", "DipPol-Y", "DipPol-X"]),
})
# orientations of the batched-RHS workload (ellipsoid 1:0.8:0.5, m = 1.33+0.001i, default dpl 10|m|), first four
for _i, (_a, _b, _g) in enumerate(orientations(4)):
    RUNS["ell12_c3o%d" % _i] = ("ell12", "-dpl 13.3 -m 1.33 0.001 -orient %.6f %.6f %.6f -iter qmr -eps 10" % (_a, _b, _g),
                               ["CrossSec-Y", "CrossSec-X"])


def adj8_runs():
    """name -> (m list, voxel or None, (dRe, dIm))"""
    pos, _ = make_shape("box", 8)
    m, mat, vox = adj8_materials(len(pos))
    runs = {"adj8_base": (m, None, (0, 0))}
    for j in vox:
        for tag, d in (("rp", (ADJ_DELTA, 0)), ("rm", (-ADJ_DELTA, 0)), ("ip", (0, ADJ_DELTA)), ("im", (0, -ADJ_DELTA))):
            runs["adj8_v%d_%s" % (j, tag)] = (m, int(j), d)
    return runs


def run_adj(name):
    m, j, (dr, di) = adj8_runs()[name]
    pos, _ = make_shape("box", 8)
    p, _ = read_shape(pos)
    _, mat, _ = adj8_materials(len(pos))
    mats = list(m)
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        gfile = os.path.join(HERE, "geom", "adj8.geom")
        if j is not None:
            mat = mat.copy()
            mats.append(m[mat[j]] + dr + 1j * di)
            mat[j] = 14
            gfile = os.path.join(tmp, "adj8_v%d.geom" % j)
            write_geom(gfile, p, mat)
        margs = " ".join("%.10g %.10g" % (z.real, z.imag) for z in mats)
        args = "-dpl 18 -m %s -iter qmr -eps 12" % margs
        cmd = [EXE, "-shape", "read", gfile] + args.split() + ["-dir", os.path.join(tmp, "run")]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, cwd=tmp)
        out = os.path.join(HERE, "adj8", name)
        os.makedirs(out, exist_ok=True)
        for f in ("CrossSec-Y", "CrossSec-X"):
            shutil.copy(os.path.join(tmp, "run", f), os.path.join(out, f))
        with open(os.path.join(out, "cmd.txt"), "w") as f:
            geo = "tests/golden/geom/adj8.geom" if j is None else \
                "adj8_v%d.geom (tests/golden/geom/adj8.geom with voxel %d moved to material 15)" % (j, j)
            f.write("oracle/_ref/adda_sparse -shape read %s %s -dir run\n" % (geo, args))
        return name
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# Iteration checkpoints (SaveIterChpoint, src/iterative.c:165-220): for each solver the uninterrupted run and a run
# stopped by -maxiter 4 whose final state -chpoint 0s saves (type "always"); -sym enf computes one polarization only,
# which the reference requires for checkpoints (src/param.c:2173).
CHP_METHODS = ["bcgs2", "bicg", "bicgstab", "cgnr", "csym", "qmr", "qmr2"]
CHP_ARGS = "-sym enf -dpl 10 -m 1.5 0.1 -eps 10"


def run_chp(name):
    meth = name[len("chp_"):]
    out = os.path.join(HERE, name)
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        gfile = os.path.join(HERE, "geom", "box4.geom")
        full = [EXE, "-shape", "read", gfile] + CHP_ARGS.split() + ["-iter", meth, "-dir", os.path.join(tmp, "full")]
        part = [EXE, "-shape", "read", gfile] + CHP_ARGS.split() + ["-iter", meth, "-maxiter", "4", "-chpoint", "0s",
                                                                   "-chp_dir", os.path.join(tmp, "chp"), "-dir",
                                                                   os.path.join(tmp, "part")]
        for c in (full, part):
            subprocess.run(c, check=True, stdout=subprocess.DEVNULL, cwd=tmp)
        os.makedirs(out, exist_ok=True)
        shutil.copy(os.path.join(tmp, "full", "CrossSec-Y"), os.path.join(out, "CrossSec-Y"))
        with open(os.path.join(tmp, "full", "log")) as fi, open(os.path.join(out, "log_RE"), "w") as fo:
            for line in fi:
                if line.startswith("RE_") or "matrix-vector" in line:
                    fo.write(line)
        shutil.copy(os.path.join(tmp, "chp", "chp.0"), os.path.join(out, "chp.0"))
        with open(os.path.join(out, "cmd.txt"), "w") as f:
            f.write("oracle/_ref/adda_sparse -shape read tests/golden/geom/box4.geom %s -iter %s -dir full\n"
                    % (CHP_ARGS, meth))
            f.write("oracle/_ref/adda_sparse -shape read tests/golden/geom/box4.geom %s -iter %s -maxiter 4 -chpoint 0s "
                    "-chp_dir chp -dir part   (chp.0 = chp/chp.0; chp/chp.log reads 'Info about the run, which produced "
                    "the checkpoint, can be found in ../part')\n" % (CHP_ARGS, meth))
        return name
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run(name):
    if name.startswith("adj8"):
        return run_adj(name)
    if name.startswith("chp_"):
        return run_chp(name)
    geo, args, keep = RUNS[name]
    out = os.path.join(HERE, name)
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        gfile = os.path.join(HERE, "geom", geo + ".geom")
        cmd = [EXE, "-shape", "read", gfile] + args.split() + ["-dir", os.path.join(tmp, "run")]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, cwd=tmp)
        os.makedirs(out, exist_ok=True)
        for f in keep:
            src = os.path.join(tmp, "run", f)
            if f == "log":  # keep the residual history and solver counters only
                with open(src) as fi, open(os.path.join(out, "log_RE"), "w") as fo:
                    for line in fi:
                        if line.startswith("RE_") or line.startswith("CoupleConstant") or "matrix-vector" in line:
                            fo.write(line)
            else:
                shutil.copy(src, os.path.join(out, f))
        rel = " ".join(os.path.relpath(c, ROOT) if c.startswith(ROOT) else os.path.basename(c) if c.startswith(tmp)
                       else c for c in cmd)
        with open(os.path.join(out, "cmd.txt"), "w") as f:
            f.write("oracle/_ref/adda_sparse -shape read tests/golden/geom/%s.geom %s -dir run\n" % (geo, args))
        return name
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    if not os.path.exists(EXE):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    os.makedirs(os.path.join(HERE, "geom"), exist_ok=True)
    for gname, (pos, mat) in geometries().items():
        p, _ = read_shape(pos)
        write_geom(os.path.join(HERE, "geom", gname + ".geom"), p, mat)
    chp = ["chp_" + m for m in CHP_METHODS]
    names = sys.argv[1:] or list(RUNS) + list(adj8_runs()) + chp
    if names == ["adj8"]:
        names = list(adj8_runs())
    if names == ["chp"]:
        names = chp
    with ThreadPoolExecutor(max_workers=6) as ex:
        for n in ex.map(run, names):
            print("done", n, flush=True)


if __name__ == "__main__":
    main()
