"""Timing of the reference program itself on the host cores (the
cpu_baseline leg of bench.py).

TEST INFRASTRUCTURE ONLY.  The reference's whole program, compiled from its
own sources into oracle/_ref/libsw4ew.so (oracle/Makefile), runs an input
deck: parse, set-up and its own time-step loop (EW::timesteploop,
EW.C:2339-2928) with its OpenMP C kernels.  Two runs of the same deck that
differ only in the number of steps give the time of the extra steps, without
the set-up.

Run as ``python -m oracle.refbench nx ny nz h steps srcs.json`` it prints
one JSON line; bench.py starts it as a child process under a time limit.
"""
import json
import os
import sys
import time


def workload_deck(nx, ny, nz, h, zmax, sources, nsteps, path):
    """Deck of a slab of the layered workload: cs = 1000 + 2500 z / zmax,
    cp = 1.8 cs (linear in depth, the reference's block gradients,
    MaterialBlock.C:160-168), rho linear between its values at the top and
    bottom of the slab; free surface on top and supergrid of 30 points on
    the other sides (the reference's default sides, EW.C:940-948); moment
    sources with Ricker time functions."""
    g = 2500.0 / zmax
    zb = (nz - 1) * h

    def rho(z):
        return 1700.0 * (1.8 * (1000.0 + g * z) / 1000.0) ** 0.25

    lines = ["fileio path=%s verbose=0" % path,
             "grid nx=%d ny=%d nz=%d h=%.17g" % (nx, ny, nz, h),
             "time steps=%d" % nsteps,
             "supergrid gp=30",
             "block vs=1000 vp=1800 rho=%.17g vsgrad=%.17g vpgrad=%.17g rhograd=%.17g"
             % (rho(0.0), g, 1.8 * g, (rho(zb) - rho(0.0)) / zb),
             "developer cfl=1.3 corder=1"]
    for s in sources:
        m = s["m"]
        lines.append("source x=%.17g y=%.17g z=%.17g Mxx=%.17g Myy=%.17g Mzz=%.17g Mxy=%.17g Mxz=%.17g Myz=%.17g "
                     "t0=%.17g freq=%.17g type=Ricker" % (s["x"], s["y"], s["z"], m[0], m[1], m[2], m[3], m[4],
                                                          m[5], s["t0"], s["freq"]))
    return "\n".join(lines) + "\n"


def host_description():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def time_reference(nx, ny, nz, h, zmax, sources, steps, warm=1):
    """Seconds of `steps` time steps of the reference program on the deck
    (difference of two runs, warm and warm + steps steps)."""
    from oracle import refew
    out = []
    for n in (warm, warm + steps):
        import tempfile
        with tempfile.TemporaryDirectory(prefix="refbench_") as tmp:
            deck = workload_deck(nx, ny, nz, h, zmax, sources, n, os.path.join(tmp, "out"))
            t0 = time.perf_counter()
            e = refew.RefEW(deck, workdir=tmp)
            out.append(time.perf_counter() - t0)
            e.close()
    return out[1] - out[0], out


def main():
    nx, ny, nz = (int(v) for v in sys.argv[1:4])
    h, zmax, steps = float(sys.argv[4]), float(sys.argv[5]), int(sys.argv[6])
    sources = json.load(open(sys.argv[7]))
    dt, runs = time_reference(nx, ny, nz, h, zmax, sources, steps)
    model, ncpu, avail = host_description()
    print(json.dumps({"seconds": dt, "runs": runs, "cpu_model": model, "nproc": ncpu, "affinity": avail,
                      "omp_threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}), flush=True)


if __name__ == "__main__":
    main()
