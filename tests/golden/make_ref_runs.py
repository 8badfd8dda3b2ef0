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
  ref_stations/         the reference's receiver files (USGS text and binary
                        SAC, TimeSeries::writeFile) of a small layered deck
                        (STATIONS_DECK: two blocks, one Gaussian moment
                        source, displacement and velocity stations, each in
                        x/y/z and in east/north/up components) after 40 steps;
                        the deck itself is ref_stations/stations.in.
  topo_curvilinear.npz  the stacked grids of the reference's topography
                        deck (decks/curvilinear.in, the reference's
                        pytest/reference/topo/curvilinear.in) before any
                        step: grid sizes, z of the curvilinear grid on its
                        surface and along one column, the grid-point sources
                        of the moment source and the Jacobian at them.  The
                        deck's own receiver files are the reference's fixture
                        (topo_curvilinear/sta0{1,2,3}.txt); the reference
                        program built here reproduces them to 4e-14 of 11.6.
  ref_checkpoint/       the reference's checkpoint file
                        (CheckPoint::write_checkpoint) of a small deck at
                        cycle 6, gzip-compressed, run with the interleaved
                        layout (corder=0) in which the reference extracts
                        the fields it writes (Sarray.C:616-633); the deck is
                        ref_checkpoint/checkpoint.in.

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


STATIONS_DECK = """grid h=100 x=3000 y=3000 z=1700
time steps=40
supergrid gp=10
block vp=6000 vs=3464 rho=2700
block vp=4000 vs=2000 rho=2600 z2=1000
source x=1700 y=1900 z=600 mxy=1e18 mxx=3e17 t0=0.36 freq=16.6667 type=Gaussian
developer cfl=1.3 corder=1
rec x=2100 y=2300 z=0 file=sta10 usgsformat=1 sacformat=1
rec x=2100 y=2300 z=0 file=sta10e nsew=1 usgsformat=1 sacformat=1
rec x=2150 y=1340 z=200 file=stb usgsformat=1 sacformat=1 variables=velocity
rec x=2150 y=1340 z=200 file=stbe sta=STB nsew=1 usgsformat=1 sacformat=1 variables=velocity
"""

CHECKPOINT_DECK = """grid h=100 x=1300 y=1200 z=1100
time steps=9
supergrid gp=4
block vp=6000 vs=3464 rho=2700
block vp=4000 vs=2000 rho=2600 z2=500
source x=600 y=650 z=400 mxy=1e18 t0=0.2 freq=16.6667 type=Gaussian
developer cfl=1.3 corder=0
checkpoint cycle=6 file=ck
"""


def stations():
    import shutil
    out = os.path.join(HERE, "ref_stations")
    os.makedirs(out, exist_ok=True)
    with R.RefEW(STATIONS_DECK) as e:
        for f in sorted(os.listdir(e.path)):
            shutil.copy(os.path.join(e.path, f), os.path.join(out, f))
    with open(os.path.join(out, "stations.in"), "w") as f:
        f.write(STATIONS_DECK)
    print("stations:", sorted(os.listdir(out)))


def checkpoint():
    import gzip
    out = os.path.join(HERE, "ref_checkpoint")
    os.makedirs(out, exist_ok=True)
    with R.RefEW(CHECKPOINT_DECK) as e:
        names = [f for f in os.listdir(e.path) if f.endswith(".sw4checkpoint")]
        for f in names:
            with open(os.path.join(e.path, f), "rb") as src, gzip.open(os.path.join(out, f + ".gz"), "wb") as dst:
                dst.write(src.read())
    with open(os.path.join(out, "checkpoint.in"), "w") as f:
        f.write(CHECKPOINT_DECK)
    print("checkpoint:", names)


def topo():
    """The curvilinear grid of the reference's topography deck
    (pytest/reference/topo/curvilinear.in, copied as decks/curvilinear.in)
    with no time steps: grid sizes, z of the curvilinear grid on its surface
    plane and along the column of receiver sta03, the grid-point sources of
    its moment source (grid 0 Cartesian, 1 curvilinear) and the curvilinear
    Jacobian at each of them."""
    import numpy as np
    deck = open(os.path.join(HERE, "decks", "curvilinear.in")).read()
    deck = "\n".join(ln for ln in deck.splitlines() if not ln.strip().startswith("fileio"))
    deck = deck.replace("time t=0.5", "time steps=0")
    with R.RefEW(deck) as e:
        (hc, zminc, nzc), (hv, _, nzv) = e.grid(0), e.grid(1)
        ib, ie, jb, je, kb, ke = e.boxes[1]
        ni, nj, nk = ie - ib + 1, je - jb + 1, ke - kb + 1
        z = e.array(7, 1).reshape(nk, nj, ni)
        jac = e.array(9, 1).reshape(nk, nj, ni)
        ijkg, f, _, _ = e.point_sources()
        gijk = np.stack([ijkg[:, 3], ijkg[:, 0], ijkg[:, 1], ijkg[:, 2]], axis=1).astype(np.int32)
        jac_at = np.array([jac[k - kb, j - jb, i - ib] if g == 1 else 0.0 for g, i, j, k in gijk])
        np.savez_compressed(os.path.join(HERE, "topo_curvilinear.npz"), h=hc, zmin_cart=zminc, nz_cart=nzc,
                            nz_curv=nzv, box_curv=np.array(e.boxes[1]), z_surface=z[1 - kb], z_column=z[:, 26 - jb, 31 - ib],
                            gijk=gijk, force=f, jac_at=jac_at)
    print("topo:", nzc, nzv, len(gijk), "grid-point sources")


RUNS = {"c1": c1, "stations": stations, "checkpoint": checkpoint, "topo": topo}

if __name__ == "__main__":
    O.build(ref=True)
    for name in sys.argv[1:] or list(RUNS):
        RUNS[name]()
