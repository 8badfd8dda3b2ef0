"""The benchmark configurations as parity cases.

C1, the Gaussian point source (121^3, h = 0.05, rho = 1, cs = 1,
cp = sqrt 3, a full moment tensor at the centre with a Gaussian of
sigma = 0.08 (freq 12.5) centred at t0 = 0.32, supergrid of 20 points on
all six faces, CFL 1.3, t = 0.6): the reference program run on that deck
(tests/golden/make_ref_runs.py) gives dt = 0.6/21 and the analytic-solution
errors Linf, L2 and max|u_exact| of tests/golden/c1_pointsource.json.  The
CPU oracle loop and the HIP library (fused sweeps and the reference stage
order) must give the same numbers to 1e-11 relative, and the HIP field must
equal the oracle's to 1e-12.
"""
import json
import os

import numpy as np
import pytest

from common import ROOT, relerr
from sw4lite_amd import lib as L

C1 = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_pointsource.json")))
C1_WANT = (C1["linf"], C1["l2"], C1["solinf"])
C1_GRID = dict(nx=121, ny=121, nz=121, h=0.05, bc=[L.BC_SUPERGRID] * 6, sg_gp=20, rho=1.0, cs=1.0,
               cp=np.sqrt(3.0), tmax=0.6)
C1_SRC = dict(x=3.0, y=3.0, z=3.0, m=[1.0, 0.3, -0.2, 0.8, 0.4, 0.6], t0=0.32, freq=12.5, tfkind=L.TF_GAUSSIAN)
C1_TOL = 1e-11


def test_c1_deck_is_the_workload():
    deck = C1["deck"]
    assert "grid x=6.0 y=6.0 z=6.0 h=0.05" in deck and "supergrid gp=20" in deck and "time t=0.6" in deck
    assert "testpointsource\n" in deck and "cfl=1.3" in deck
    assert "Mxx=1 Myy=0.8 Mzz=0.6 Mxy=0.3 Mxz=-0.2 Myz=0.4 t0=0.32 freq=12.5 type=Gaussian" in deck
    assert C1["box"] == [-1, 123, -1, 123, -1, 123] and C1["nsteps"] == 21
    assert [float(v) for v in C1["printed"][1:]] == pytest.approx(list(C1_WANT), rel=1e-11)


def c1_oracle():
    from oracle.loop import OracleRun
    d = C1_GRID
    mu = d["rho"] * d["cs"] ** 2
    la = d["rho"] * d["cp"] ** 2 - 2 * mu
    run = OracleRun(d["nx"], d["ny"], d["nz"], d["h"], d["bc"], d["rho"], mu, la, [C1_SRC], sg_gp=d["sg_gp"],
                    tmax=d["tmax"])
    return run


def test_c1_cpu_oracle():
    run = c1_oracle()
    assert run.nsteps == C1["nsteps"] and run.dt == pytest.approx(C1["dt"], rel=1e-15)
    run.step(run.nsteps)
    li, l2, xli, _ = run.point_source_error(C1_SRC)
    for g, w in zip((li, l2, xli), C1_WANT):
        assert abs(g - w) <= C1_TOL * abs(w), ((li, l2, xli), C1_WANT)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_c1_gpu(mode):
    from sw4lite_amd import EW
    d = C1_GRID
    print("C1 121^3 Gaussian point source, mode", mode, flush=True)
    e = EW(d["nx"], d["ny"], d["nz"], d["h"], d["bc"], corder=1, sg_gp_thickness=d["sg_gp"])
    try:
        e.set_material_const(d["rho"], d["cs"], d["cp"])
        s = C1_SRC
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(tmax=d["tmax"])
        inf = e.info()
        assert inf.nsteps == C1["nsteps"] and inf.dt == pytest.approx(C1["dt"], rel=1e-15)
        e.timesteploop(mode=mode)
        got = e.point_source_error()
        u, _ = e.get_solution()
    finally:
        e.close()
    print("C1 mode", mode, got, C1_WANT, flush=True)
    for g, w in zip(got, C1_WANT):
        assert abs(g - w) <= C1_TOL * abs(w), (got, C1_WANT)
    if mode == 0:
        run = c1_oracle()
        run.step(run.nsteps)
        assert relerr(u, run.U) < 1e-12
