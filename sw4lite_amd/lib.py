"""ctypes bindings of ``libsw4hip.so`` (declarations in include/sw4hip.h).

The library is loaded from the package directory; nothing here falls back
to another implementation: if the library is missing the import of
:func:`load` raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsw4hip.so")

BC_FREE, BC_DIRICHLET, BC_SUPERGRID, BC_PERIODIC = 0, 1, 2, 3
TF_RICKER, TF_GAUSSIAN, TF_C6SMOOTHBUMP = 0, 1, 2

P = C.c_void_p
I = C.c_int
D = C.c_double
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)


class GridDesc(C.Structure):
    _fields_ = [("nx", I), ("ny", I), ("nz", I), ("h", D), ("zmin", D), ("corder", I),
                ("bc", I * 6), ("sg_gp_thickness", I), ("sg_damping", D), ("cfl", D)]


class DeckReport(C.Structure):
    _fields_ = [("nx", I), ("ny", I), ("nz", I), ("h", D), ("dt", D), ("t", D), ("nsteps", I), ("cycle", I),
                ("nstations", I), ("ncheckpoints", I), ("point_source_test", I), ("linf", D), ("l2", D),
                ("solinf", D), ("ngrids", I), ("nz_curv", I)]


class Topography(C.Structure):
    _fields_ = [("zmax", D), ("order", I), ("zeta_break", D), ("amp", D), ("xc", D), ("yc", D), ("lx", D),
                ("ly", D)]


class Station(C.Structure):
    _fields_ = [("file", C.c_char_p), ("sta", C.c_char_p), ("mode", I), ("nsew", I), ("usgs", I), ("sac", I),
                ("x", D), ("y", D), ("z", D), ("gx", D), ("gy", D), ("gz", D), ("utc", I * 7)]


class EWInfo(C.Structure):
    _fields_ = [("dt", D), ("nsteps", I), ("t", D), ("cycle", I), ("onesided", I * 6),
                ("wind", I * 36), ("npts", C.c_long), ("nsources", I), ("nunique", I)]


_SIGS = {
    "sw4hip_last_error": (C.c_char_p, []),
    "sw4hip_version": (I, []),
    "sw4hip_get_stencil_coefficients": (I, [P, P, P, P]),
    "sw4hip_host_supergrid": (I, [C.POINTER(GridDesc)] + [P] * 9),
    "sw4hip_host_boundary_windows": (I, [C.POINTER(GridDesc), P, P]),
    "sw4hip_host_moment_source": (I, [C.POINTER(GridDesc), D, D, D, P, I, IP, P, P]),
    "sw4hip_rhs4sg": (I, [I] * 8 + [P, P, P, P, P, P, P, P, D, P, P, P, P]),
    "sw4hip_force": (I, [I] * 7 + [D, I, I, P, P, P, P, P, P, P]),
    "sw4hip_predfort": (I, [I] * 7 + [P] * 6 + [D, P]),
    "sw4hip_corrfort": (I, [I] * 7 + [P] * 4 + [D, P]),
    "sw4hip_dpdmtfort": (I, [I] * 6 + [P] * 4 + [D, P]),
    "sw4hip_addsgd4": (I, [I] * 7 + [P] * 13 + [D, P]),
    "sw4hip_bcfortsg": (I, [I] * 7 + [P, I, I, I, P, D, P, P, P, P, P, P, P, P]),
    "sw4hip_rhs4sg_adjoint": (I, [I] * 7 + [P] * 8 + [D, P, P, P, P]),
    "sw4hip_rhs4sg_matgrad": (I, [I] * 7 + [P] * 8 + [D, P, P, P, P, P, P]),
    "sw4hip_addsgd4_adjoint": (I, [I] * 6 + [P] * 13 + [D, P, P, P]),
    "sw4hip_bcfortsg_adjoint": (I, [I] * 6 + [P, I, I, I, P, P, D, P, P, P, P, P, P, P, P, P]),
    "sw4hip_bcfortsg_adjoint_noalloc": (I, [I] * 6 + [P, I, I, I, P, P, D, P, P, P, P, P, P, P, P, P]),
    "sw4hip_ew_reset": (I, [P]),
    "sw4hip_batch_create": (I, [P, I, C.POINTER(P)]),
    "sw4hip_batch_destroy": (I, [P]),
    "sw4hip_batch_set_kchunk": (I, [P, I]),
    "sw4hip_batch_tune": (I, [P, I, P, IP, P]),
    "sw4hip_batch_step": (I, [P, I, I]),
    "sw4hip_batch_gradient": (I, [P, I, I, P, P, P, P, I, I]),
    "sw4hip_batch_release_workspace": (I, [P]),
    "sw4hip_ew_gradient": (I, [P, I, I, P, P, P, P, I]),
    "sw4hip_ew_gradient_comp": (I, [P, I, I, P, P, P, P, I]),
    "sw4hip_batch_gradient_comp": (I, [P, I, I, P, P, P, P, I, I]),
    "sw4hip_ew_gradient_plan": (I, [P, I, I, P, P]),
    "sw4hip_ew_set_gradient_store": (I, [P, I]),
    "sw4hip_ew_release_workspace": (I, [P]),
    "sw4hip_host_moment_source_curv": (I, [C.POINTER(GridDesc), P, D, D, D, P, I, IP, P, P]),
    "sw4hip_host_topography_grid": (I, [C.POINTER(Topography), I, I, D, D, IP, IP, P]),
    "sw4hip_host_topography_receiver": (I, [C.POINTER(Topography), I, I, D, D, D, D, D, I, IP, P]),
    "sw4hip_host_topography_source": (I, [C.POINTER(Topography), I, I, D, D, P, D, D, D, I, P, I, IP, P, P]),
    "sw4hip_ew_set_topography": (I, [P, P, I]),
    "sw4hip_metric": (I, [I] * 7 + [P] * 5 + [P]),
    "sw4hip_rhs4sgcurv": (I, [I] * 7 + [P] * 12 + [P]),
    "sw4hip_addsgd4c": (I, [I] * 7 + [P] * 11 + [D, P]),
    "sw4hip_freesurfcurvisg": (I, [I] * 9 + [P] * 8 + [P]),
    "sw4hip_enforce_cart_topo": (I, [I] * 7 + [P, I, I, P, I, P]),
    "sw4hip_ew_create": (I, [C.POINTER(GridDesc), C.POINTER(P)]),
    "sw4hip_ew_destroy": (I, [P]),
    "sw4hip_ew_set_stream": (I, [P, P]),
    "sw4hip_ew_set_kernel_variant": (I, [P, I, I]),
    "sw4hip_ew_set_material": (I, [P, P, P, P, I]),
    "sw4hip_ew_set_material_const": (I, [P, D, D, D]),
    "sw4hip_ew_add_moment_source": (I, [P, D, D, D, P, D, D, I]),
    "sw4hip_ew_setup": (I, [P, D, I]),
    "sw4hip_ew_get_info": (I, [P, C.POINTER(EWInfo)]),
    "sw4hip_ew_get_supergrid": (I, [P] + [P] * 9),
    "sw4hip_ew_get_sources": (I, [P, P, P, P, P, P, P]),
    "sw4hip_ew_set_solution": (I, [P, P, P, I]),
    "sw4hip_ew_get_solution": (I, [P, P, P, I]),
    "sw4hip_ew_step": (I, [P, I, I]),
    "sw4hip_ew_step_timed": (I, [P, I, P]),
    "sw4hip_ew_step_comp_timed": (I, [P, I, P]),
    "sw4hip_ew_point_source_error": (I, [P, DP, DP, DP]),
    "sw4hip_ew_add_receiver": (I, [P, D, D, D, I, IP]),
    "sw4hip_ew_get_receivers": (I, [P, IP, P]),
    "sw4hip_ew_device_array": (I, [P, I, C.POINTER(P)]),
    "sw4hip_ew_sync": (I, [P]),
    "sw4hip_run_deck": (I, [C.c_char_p, C.c_char_p, I, C.POINTER(DeckReport)]),
    "sw4hip_parse_deck": (I, [C.c_char_p, C.POINTER(DeckReport)]),
    "sw4hip_write_station": (I, [C.POINTER(Station), C.c_char_p, D, D, I, P]),
    "sw4hip_write_checkpoint": (I, [C.c_char_p, D, I, I, I, I, P, P]),
    "sw4hip_read_checkpoint": (I, [C.c_char_p, I, I, I, DP, IP, P, P]),
}

_lib = None


class Sw4HipError(RuntimeError):
    pass


def load():
    """Load libsw4hip.so (raises if it has not been built).

    PyTorch ships its own HIP runtime under the same soname as the system
    one.  Importing torch first makes the library bind to that runtime, so a
    process that uses both holds exactly one HIP runtime; loading the library
    first would map a second runtime that cannot share the device."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise Sw4HipError("libsw4hip.so is not built; run python -m sw4lite_amd.build")
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIBPATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise Sw4HipError("sw4hip error %d: %s" % (rc, load().sw4hip_last_error().decode()))
    return rc


def grid_desc(nx, ny, nz, h, bc, corder=1, zmin=0.0, sg_gp_thickness=20, sg_damping=0.02, cfl=1.3):
    d = GridDesc()
    d.nx, d.ny, d.nz, d.h, d.zmin, d.corder = nx, ny, nz, h, zmin, corder
    for s in range(6):
        d.bc[s] = bc[s]
    d.sg_gp_thickness, d.sg_damping, d.cfl = sg_gp_thickness, sg_damping, cfl
    return d
