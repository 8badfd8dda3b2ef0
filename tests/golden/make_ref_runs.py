"""Golden results of whole runs of the reference program, made in the
container that holds the reference sources: the reference is compiled from
them by oracle/Makefile (target ``ref``: oracle/_ref/libsw4ew.so) and each
deck below is run through its own constructor (EW::EW, EW.C:162-224).
Only numbers and small output files are stored.

  c1_pointsource.json   the Gaussian point-source workload (121^3, h = 0.05,
                        rho = 1, cs = 1, cp = sqrt 3, supergrid of 20 points
                        on all six faces, CFL 1.3, t = 0.6): dt, step count
                        and the analytic-solution error of the final field
                        (EW::exactSol + normOfDifference, EW.C:2889-2905),
                        at full precision and as the run's own
                        PointSourceErr.txt.

Usage: python tests/golden/make_ref_runs.py [name ...]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import refew as R  # noqa: E402

C1_DECK = """grid x=6.0 y=6.0 z=6.0 h=0.05
time t=0.6
testpointsource
supergrid gp=20
source x=3.0 y=3.0 z=3.0 Mxx=1 Myy=0.8 Mzz=0.6 Mxy=0.3 Mxz=-0.2 Myz=0.4 t0=0.32 freq=12.5 type=Gaussian
developer cfl=1.3 corder=1
"""


def c1():
    with R.RefEW(C1_DECK) as e:
        box = e.boxes[0]
        t = e.dt * e.nsteps
        li, l2, xi = e.point_source_error(t)
        with open(os.path.join(e.path, "PointSourceErr.txt")) as f:
            printed = f.read().split()
        res = {"deck": C1_DECK, "box": list(box), "dt": e.dt, "nsteps": e.nsteps, "t": t, "linf": li, "l2": l2,
               "solinf": xi, "printed": printed}
    with open(os.path.join(HERE, "c1_pointsource.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("c1:", res["nsteps"], "steps, dt", res["dt"], "linf", li, "l2", l2, "solinf", xi)


RUNS = {"c1": c1}

if __name__ == "__main__":
    O.build(ref=True)
    for name in sys.argv[1:] or list(RUNS):
        RUNS[name]()
