"""Write the golden vectors of tests/golden/ from the reference's own code.

Run in a container that holds the reference source tree (default
/root/reference).  It
  * compiles the reference kernels from their sources (oracle/Makefile,
    target ``ref``) and evaluates rhs4sg_rev (rhs4sg_rev.C:43) for three
    one-sided configurations and addsgd4fort_rev (tests/testil/
    addsgd4fort_rev.C:4) on seeded random inputs (tests/common.py
    ``random_inputs``); only the outputs and checksums of the inputs are
    stored;
  * extracts the exact SBP coefficients from GetStencilCoefficients
    (ew-cfromfort.C:1487-1944) as rational numbers;
  * evaluates the reference metric (tests/testil/metricC.C:127) and
    rhs4sgcurv_rev (rhs4sgcurv_rev.C:34) on a grid under seeded Gaussian
    hills (tests/common.py ``hills_depth``), with and without the top SBP
    rows (curv.npz / curv.json);
  * copies the point-source accuracy fixture
    (pytest/reference/pointsource/pointsource-h0p04/PointSourceErr.txt) and
    the parameters of its deck (pytest/reference/pointsource/pointsource.in);
  * copies the LOH.1 and LOH.2 station records at h = 100 m
    (pytest/reference/loh{1,2}/LOH{1,2}-h100/sta10.txt) and writes LOH.2's
    point-source list from its deck (pytest/reference/loh2/LOH.2-h100.in).

Usage: python tests/golden/make_golden.py [/path/to/reference] [--loh-only]
"""
import json
import os
import re
import shutil
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from common import hills_depth, npts, random_inputs  # noqa: E402
from oracle import oracle as O  # noqa: E402

GRID = (8, 9, 13)
H = 0.1
BETA = 0.02


def sbp_from_reference(ref):
    src = open(os.path.join(ref, "src", "ew-cfromfort.C")).read()
    start = src.index("void EW::GetStencilCoefficients")
    body = src[start:src.index("_sbop[4]", start) + 40]
    num = r"(-?\d+)\.0?/(\d+)\.?0?"
    acof = {}
    for m in re.finditer(r"acof\((\d),(\d),(\d)\)\s*=\s*([^;]+);", body):
        k, q, mm = map(int, m.groups()[:3])
        e = m.group(4).strip()
        if e == "0":
            continue
        r = re.fullmatch(num, e)
        acof[(k, q, mm)] = Fraction(int(r.group(1)), int(r.group(2)))
    d4a, d4b = Fraction(2, 3), Fraction(-1, 12)
    named = {"d4a": d4a, "d4b": d4b, "-d4a": -d4a, "-d4b": -d4b}
    bope = {}
    for m in re.finditer(r"bope\((\d),(\d)\)\s*=\s*([^;]+);", body):
        k, q = map(int, m.groups()[:2])
        e = m.group(3).replace(" ", "")
        if e == "0":
            continue
        if e in named:
            bope[(k, q)] = named[e]
        else:
            r = re.fullmatch(num, e)
            bope[(k, q)] = Fraction(int(r.group(1)), int(r.group(2)))
    gh = re.search(r"ghcof\(1\)\s*=\s*" + num, body)
    sbop = [Fraction(int(a), int(b)) for a, b in re.findall(r"_sbop\[\d\]\s*=\s*(-?\s*\d+)\.0/(\d+)", body)]
    return {
        "acof": [[k, q, m, str(v)] for (k, q, m), v in sorted(acof.items())],
        "bope": [[k, q, str(v)] for (k, q), v in sorted(bope.items())],
        "ghcof1": str(Fraction(int(gh.group(1)), int(gh.group(2)))),
        "sbop": [str(v) for v in sbop],
    }


def sbp_arrays(tab):
    acof = np.zeros(384)
    bope = np.zeros(48)
    ghcof = np.zeros(6)
    for k, q, m, v in tab["acof"]:
        acof[(k - 1) + 6 * (q - 1) + 48 * (m - 1)] = float(Fraction(v))
    for k, q, v in tab["bope"]:
        bope[(k - 1) + 6 * (q - 1)] = float(Fraction(v))
    ghcof[0] = float(Fraction(tab["ghcof1"]))
    return acof, bope, ghcof


def checksums(f):
    return {k: float(np.sum(v)) for k, v in sorted(f.items())}


CURV_GRID = (10, 9, 13)


def curv_golden(lib, acof, bope, ghcof):
    nx, ny, nz = CURV_GRID
    out, meta = {}, {"grid": CURV_GRID, "h": H, "cases": {}}
    for top in (0, 1):
        seed = 41 + top
        box, f = random_inputs(seed, nx, ny, nz)
        _, z = hills_depth(nx, ny, nz, H, seed + 100)
        n = npts(box)
        met, jac = np.zeros(4 * n), np.zeros(n)
        assert O.metric(1, box, H, z, met, jac, lib=lib) == 0
        lu = np.zeros_like(f["u"])
        O.rhs4sgcurv(1, box, f["u"], f["mu"], f["la"], met, jac, lu, [0, 0, 0, 0, top, 0], acof, bope, ghcof,
                     f["strx"], f["stry"], lib=lib)
        name = "rhs4sgcurv_rev_top%d" % top
        out[name] = lu
        meta["cases"][name] = {"seed": seed, "met_sum": float(np.sum(met)), "checksums": checksums(f)}
    np.savez_compressed(os.path.join(HERE, "curv.npz"), **out)
    with open(os.path.join(HERE, "curv.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


def main(ref):
    O.build(ref=True)
    lib = O.reference()
    tab = sbp_from_reference(ref)
    with open(os.path.join(HERE, "sbp_coefficients.json"), "w") as fh:
        json.dump(tab, fh, indent=0)
    acof, bope, ghcof = sbp_arrays(tab)
    nx, ny, nz = GRID
    out = {}
    meta = {"grid": GRID, "h": H, "beta": BETA, "cases": {}}
    for seed, os4, os5 in ((11, 0, 0), (12, 1, 0), (13, 1, 1)):
        box, f = random_inputs(seed, nx, ny, nz)
        lu = np.zeros_like(f["u"])
        O.rhs4sg(1, box, nz, [0, 0, 0, 0, os4, os5], acof, bope, ghcof, lu, f["u"], f["mu"], f["la"], H,
                 f["strx"], f["stry"], f["strz"], lib=lib)
        name = "rhs4sg_rev_os%d%d" % (os4, os5)
        out[name] = lu
        meta["cases"][name] = {"seed": seed, "onesided": [0, 0, 0, 0, os4, os5], "checksums": checksums(f)}
    box, f = random_inputs(21, nx, ny, nz)
    up = f["up"].copy()
    O.addsgd4(1, box, up, f["u"], f["um"], f["rho"], f["dcx"], f["dcy"], f["dcz"], f["strx"], f["stry"],
              f["strz"], f["cox"], f["coy"], f["coz"], BETA, lib=lib)
    out["addsgd4fort_rev"] = up
    meta["cases"]["addsgd4fort_rev"] = {"seed": 21, "checksums": checksums(f)}
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **out)
    curv_golden(lib, acof, bope, ghcof)
    with open(os.path.join(HERE, "kernels.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    # point-source accuracy fixture and its deck parameters
    line = open(os.path.join(ref, "pytest", "reference", "pointsource", "pointsource-h0p04",
                             "PointSourceErr.txt")).read().split()
    deck = open(os.path.join(ref, "pytest", "reference", "pointsource", "pointsource.in")).read()
    ps = {"t": line[0], "linf": line[1], "l2": line[2], "solinf": line[3], "deck": deck}
    with open(os.path.join(HERE, "pointsource_err.json"), "w") as fh:
        json.dump(ps, fh, indent=1)
    print("golden vectors written to", HERE)


def loh_fixtures(ref):
    """Station records of LOH.1/LOH.2 (USGS text, copied as they are) and the
    3200 point sources of the LOH.2 deck as [x, y, z, Mxy, t0]."""
    base = os.path.join(ref, "pytest", "reference")
    for case in ("loh1", "loh2"):
        shutil.copyfile(os.path.join(base, case, case.upper() + "-h100", "sta10.txt"),
                        os.path.join(HERE, case + "_h100_sta10.txt"))
    src = []
    for line in open(os.path.join(base, "loh2", "LOH.2-h100.in")):
        if not line.startswith("source "):
            continue
        kv = dict(t.split("=", 1) for t in line.split()[1:])
        assert kv["type"] == "Gaussian" and float(kv["freq"]) == 16.6667 and float(kv["m0"]) == 1.0
        assert set(kv) == {"x", "y", "z", "Mxy", "t0", "freq", "type", "m0"}
        src.append([float(kv["x"]), float(kv["y"]), float(kv["z"]), float(kv["Mxy"]), float(kv["t0"])])
    with open(os.path.join(HERE, "loh2_h100_sources.json"), "w") as fh:
        json.dump({"deck": "pytest/reference/loh2/LOH.2-h100.in", "type": "Gaussian", "freq": 16.6667, "m0": 1.0,
                   "columns": ["x", "y", "z", "Mxy", "t0"], "sources": src}, fh)
    print("LOH fixtures written:", len(src), "LOH.2 sources")


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    ref_dir = args[0] if args else "/root/reference"
    if "--loh-only" not in sys.argv:
        main(ref_dir)
    loh_fixtures(ref_dir)
