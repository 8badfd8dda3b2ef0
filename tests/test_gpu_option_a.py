"""GPU: the kernel-level drop-in compiled as C++ (tests/option_a/option_a.cpp,
built by __graft_entry__.build against include/sw4hip.h and libsw4hip.so).

The program keeps the reference's EW member signatures (Force, evalRHS,
evalPredictor, evalCorrector, evalDpDmInTime, addSuperGridDamping,
enforceBC; EW.h:93-127) with bodies that call the C ABI, and steps them in
the reference's stage order (EW.C:2527-2842).  Its solution after 8 steps
must equal the CPU oracle loop's to 1e-12 in both array layouts, with a
free surface and supergrid sides, and with periodic sides over a Dirichlet
bottom."""
import os
import subprocess

import numpy as np
import pytest

from common import ROOT, relerr
from test_gpu_parity import SOURCES, layered_material

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "option_a", "option_a")


@pytest.mark.parametrize("corder,bc", [(1, [2, 2, 2, 2, 0, 2]), (0, [2, 2, 2, 2, 0, 2]), (1, [3, 3, 3, 3, 0, 1])])
def test_option_a_program_matches_oracle(tmp_path, corder, bc):
    from oracle.loop import OracleRun
    assert os.path.exists(EXE), "tests/option_a/option_a is not built (__graft_entry__.build)"
    nx, ny, nz, h, nsteps = 30, 28, 26, 0.08, 8
    rho, mu, la = layered_material(nx, ny, nz, h, 5)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, SOURCES, corder=corder, sg_gp=6, nsteps=nsteps)
    ijk, frc, fq, t0, tf = ref.src
    sg = ref.sg
    parts = [np.array([nx, ny, nz, corder, nsteps, len(fq)] + list(bc) + ref.onesided[4:6] + [0, 0], dtype="<i4"),
             np.array([h, ref.dt, ref.beta], dtype="<f8"),
             np.asarray(ref.wind, dtype="<i4"),
             np.asarray(ref.acof, "<f8"), np.asarray(ref.bope, "<f8"), np.asarray(ref.ghcof, "<f8"),
             np.asarray(ref.sbop, "<f8")]
    parts += [np.asarray(sg[k], "<f8") for k in ("dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz")]
    parts += [np.asarray(a, "<f8") for a in (ref.rho, ref.mu, ref.la)]
    parts += [np.asarray(ijk, "<i4").ravel(), np.asarray(frc, "<f8").ravel(), np.asarray(fq, "<f8"),
              np.asarray(t0, "<f8"), np.asarray(tf, "<i4")]
    parts += [np.asarray(ref.U, "<f8"), np.asarray(ref.Um, "<f8")]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    inp.write_bytes(b"".join(p.tobytes() for p in parts))
    print("option_a: corder", corder, "bc", bc, flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    ref.step(nsteps)
    got = np.frombuffer(out.read_bytes(), "<f8")
    n3 = ref.U.size
    u, um = got[:n3], got[n3:]
    err = relerr(u, ref.U)
    print("option_a: relerr", err, flush=True)
    assert np.max(np.abs(ref.U)) > 0 and err < 1e-12 and relerr(um, ref.Um) < 1e-12
