"""GPU: sw4hip_bcfortsg_adjoint_noalloc, the allocation-free variant of
sw4hip_bcfortsg_adjoint (its scratch is a fixed device array of the library).

Both entry points run the same kernels, so on the same inputs they must give
the same bits: the adjoint field y and the material derivatives gmu, glambda.
Two calls in a row, the second on a smaller grid, show that nothing of the
first call's scratch contents leaks into the second; a plane beyond the
scratch limit is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from common import random_inputs
from sw4lite_amd import lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a visible GPU"
    L.load()
    return _t


def _p(a):
    return C.c_void_p(a.data_ptr()) if hasattr(a, "data_ptr") else a.ctypes.data_as(C.c_void_p)


def _run(torch, fn, nx, ny, nz, bc, seed):
    """One call of fn (an entry point with the sw4hip_bcfortsg_adjoint
    arguments) on seeded inputs; returns y, gmu, glambda on the host."""
    from sw4lite_amd import ew
    h = 0.1
    box, f = random_inputs(seed, nx, ny, nz)
    sbop = np.ascontiguousarray(ew.stencil_coefficients()[3])
    wind, _ = ew.host_boundary_windows(L.grid_desc(nx, ny, nz, h, bc))
    rng = np.random.default_rng(seed + 1)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    y = d(rng.uniform(-1, 1, f["u"].size))
    fwd = d(f["u"])
    mu, la, sx, sy = d(f["mu"]), d(f["la"]), d(f["strx"]), d(f["stry"])
    gmu, gla = d(np.full_like(f["mu"], 0.25)), d(np.full_like(f["la"], -0.5))
    bcc = np.ascontiguousarray(bc, dtype=np.int32)
    w = np.ascontiguousarray(wind, dtype=np.int32)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(fn(*box, _p(w), nx, ny, nz, _p(y), _p(fwd), h, _p(bcc), _p(sbop), _p(mu), _p(la), _p(sx), _p(sy),
               _p(gmu), _p(gla), stream))
    torch.cuda.synchronize()
    return y.cpu().numpy(), gmu.cpu().numpy(), gla.cpu().numpy()


@pytest.mark.parametrize("bc", [[2, 2, 2, 2, 0, 0], [1, 1, 2, 2, 0, 2]])
def test_noalloc_matches_allocating_entry(torch, bc):
    lib = L.load()
    # a larger grid first, then a smaller one: the second call must not see
    # the first call's plane or weights
    for nx, ny, nz, seed in ((40, 37, 18, 501), (16, 15, 14, 502)):
        want = _run(torch, lib.sw4hip_bcfortsg_adjoint, nx, ny, nz, bc, seed)
        got = _run(torch, lib.sw4hip_bcfortsg_adjoint_noalloc, nx, ny, nz, bc, seed)
        for name, a, b in zip(("y", "gmu", "glambda"), want, got):
            ndiff = int(np.count_nonzero(a != b))
            print("noalloc bc", bc, "grid", (nx, ny, nz), name, "points differing:", ndiff, flush=True)
            assert ndiff == 0
        assert np.max(np.abs(want[1] - 0.25)) > 0  # the closure's derivatives were accumulated


def test_noalloc_rejects_planes_beyond_the_limit(torch):
    """(nx+4)(ny+4) above the scratch limit: refused with SW4HIP_ERANGE
    before any kernel runs (the tiny arrays below are never touched)."""
    from sw4lite_amd import ew
    lib = L.load()
    nx, ny, nz, h = 1100, 1100, 14, 0.1
    bc = [2, 2, 2, 2, 0, 2]
    wind, _ = ew.host_boundary_windows(L.grid_desc(nx, ny, nz, h, bc))
    w = np.ascontiguousarray(wind, dtype=np.int32)
    bcc = np.ascontiguousarray(bc, dtype=np.int32)
    sbop = np.ascontiguousarray(ew.stencil_coefficients()[3])
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    box = (-1, nx + 2, -1, ny + 2, -1, nz + 2)
    rc = lib.sw4hip_bcfortsg_adjoint_noalloc(*box, _p(w), nx, ny, nz, _p(t), _p(t), h, _p(bcc), _p(sbop), _p(t),
                                             _p(t), _p(t), _p(t), _p(t), _p(t),
                                             C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 5, "expected SW4HIP_ERANGE"
    assert b"scratch limit" in lib.sw4hip_last_error()
