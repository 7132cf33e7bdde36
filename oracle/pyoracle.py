"""ctypes wrapper of the CPU restatement (libspark_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import contextlib
import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libspark_oracle.so"

RECON = {"weno5": 0, "mp5": 1, "plm": 2, "fog": 3}
RIEMANN = {"hllc": 0, "hll": 1}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double), ("cfl", ctypes.c_double), ("cvisc", ctypes.c_double),
        ("smlrho", ctypes.c_double), ("smallp", ctypes.c_double), ("smalle", ctypes.c_double),
        ("smallx", ctypes.c_double), ("eintSwitch", ctypes.c_double), ("tiny", ctypes.c_double),
        ("gasConst", ctypes.c_double), ("abar", ctypes.c_double), ("grav", ctypes.c_double * 3),
        ("ndim", ctypes.c_int), ("nguard", ctypes.c_int), ("nstencil", ctypes.c_int), ("rk_order", ctypes.c_int),
        ("recon", ctypes.c_int), ("riemann", ctypes.c_int), ("use_flattening", ctypes.c_int),
        ("use_hybrid_riemann", ctypes.c_int), ("nscalars", ctypes.c_int), ("use_gravity", ctypes.c_int),
        ("nspecies", ctypes.c_int), ("eos", ctypes.c_int),
        ("species_a", ctypes.c_double * 8), ("species_gamma", ctypes.c_double * 8),
    ]


class OrcBlock(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("geo", ctypes.c_void_p), ("geometry", ctypes.c_int)]


_lib = None
_D = ctypes.POINTER(ctypes.c_double)


def load(path):
    """Open a build of the restatement at `path` with every signature set."""
    path = Path(path)
    if not path.exists():
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = ctypes.CDLL(str(path))
    P = ctypes.POINTER(OrcParams)
    B = ctypes.POINTER(OrcBlock)
    pp = ctypes.POINTER(ctypes.c_void_p)
    L.orc_prim2con.argtypes = [_D, _D]
    L.orc_prim2flx.argtypes = [ctypes.c_int, _D, _D]
    L.orc_riemann_hllc.argtypes = [P, ctypes.c_int, _D, _D, ctypes.c_int, _D]
    L.orc_riemann_hllc.restype = ctypes.c_int
    L.orc_riemann_hll.argtypes = [P, ctypes.c_int, _D, _D, _D]
    L.orc_riemann_hll.restype = ctypes.c_int
    L.orc_recon.argtypes = [P, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D]
    L.orc_eos_dens_ei_cell.argtypes = [P, _D, ctypes.c_int]
    L.orc_flattening.argtypes = [P, B, _D, ctypes.POINTER(ctypes.c_int), _D]
    L.orc_get_face_flux.argtypes = [P, B, _D, _D, ctypes.POINTER(ctypes.c_int), _D, _D, _D,
                                    ctypes.POINTER(ctypes.c_int)]
    L.orc_update_soln.argtypes = [P, B, _D, _D, ctypes.POINTER(ctypes.c_int), ctypes.c_double, _D, _D, _D, _D]
    L.orc_eos_wrapped.argtypes = [P, B, _D, ctypes.POINTER(ctypes.c_int)]
    L.orc_shock_detect.argtypes = [P, B, _D]
    L.orc_compute_dt.argtypes = [P, B, _D, ctypes.POINTER(ctypes.c_int)]
    L.orc_compute_dt.restype = ctypes.c_double
    L.orc_stage.argtypes = [P, B, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D, _D,
                            ctypes.POINTER(ctypes.c_int)]
    L.orc_hydro_block.argtypes = [P, B, _D, ctypes.c_double, _D, _D, _D]
    L.orc_hydro_block.restype = ctypes.c_int
    L.orc_correct_fluxes.argtypes = [P, B, _D, ctypes.c_double, _D, _D, _D]
    L.orc_ug_fill_guards.argtypes = [P, B, ctypes.c_int, ctypes.c_int, ctypes.c_int, pp, ctypes.c_int]
    L.orc_ug_advance.argtypes = [P, B, ctypes.c_int, ctypes.c_int, ctypes.c_int, pp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, _D]
    L.orc_ug_advance.restype = ctypes.c_int
    L.orc_batch_advance.argtypes = [P, B, ctypes.c_int, pp, ctypes.c_int, ctypes.c_int, _D]
    L.orc_batch_advance.restype = ctypes.c_int
    L.orc_set_dt_control.argtypes = [ctypes.c_double] * 4
    L.orc_set_dt_control.restype = None
    L.orc_set_geometry.argtypes = [pp, ctypes.c_int]
    L.orc_set_geometry.restype = None
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(path):
    """Route every call of this module to another build of the restatement
    (the planted-mistake builds of the pin tests), then restore."""
    global _lib
    saved = _lib
    _lib = load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def params(cfg: dict) -> OrcParams:
    p = OrcParams()
    for name, _ in OrcParams._fields_:
        if name == "grav":
            g = cfg.get("grav", (0.0, 0.0, 0.0))
            for i in range(3):
                p.grav[i] = float(g[i])
            continue
        if name in ("species_a", "species_gamma"):
            arr = getattr(p, name)
            for i, x in enumerate(cfg.get(name, ())):
                arr[i] = float(x)
            continue
        if name in ("nspecies", "eos"):
            v = cfg.get(name, 0)
            if name == "eos" and isinstance(v, str):
                v = {"gamma": 0, "multigamma": 1}[v]
            setattr(p, name, v)
            continue
        v = cfg[name]
        if name == "recon" and isinstance(v, str):
            v = RECON[v]
        if name == "riemann" and isinstance(v, str):
            v = RIEMANN[v]
        setattr(p, name, v)
    return p


def block(b: dict) -> OrcBlock:
    """b["geo"] (optional): the flat float64 geometry arrays of the block
    (rc, rl, rr over NX, then area and volume over NY x NX) with
    b["geometry"] 1 (cylindrical) or 2 (spherical)."""
    geo = b.get("geo")
    if geo is not None:
        assert geo.dtype == np.float64 and geo.flags["C_CONTIGUOUS"] and geo.ndim == 1
    return OrcBlock(int(b["nx"]), int(b["ny"]), int(b["nz"]), float(b["dx"]), float(b["dy"]), float(b["dz"]),
                    geo.ctypes.data if geo is not None else None, int(b.get("geometry", 0)))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def ptr_array(arrays):
    out = (ctypes.c_void_p * len(arrays))()
    for i, a in enumerate(arrays):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        out[i] = a.ctypes.data
    return out


def hydro_block(cfg, b, U, dt, fb=None):
    P, B = params(cfg), block(b)
    if fb is None:
        return lib().orc_hydro_block(ctypes.byref(P), ctypes.byref(B), dptr(U), dt, None, None, None)
    return lib().orc_hydro_block(ctypes.byref(P), ctypes.byref(B), dptr(U), dt, dptr(fb[0]), dptr(fb[1]),
                                 dptr(fb[2]))


def compute_dt(cfg, b, U):
    P, B = params(cfg), block(b)
    loc = (ctypes.c_int * 3)()
    dt = lib().orc_compute_dt(ctypes.byref(P), ctypes.byref(B), dptr(U), loc)
    return dt, list(loc)


def ug_advance(cfg, b, ug, arrays, nsteps, refill=False, nthreads=0, geo=None):
    """geo (optional): per block the flat geometry arrays (see block())."""
    P, B = params(cfg), block(b)
    dts = np.zeros(max(1, nsteps))
    gp = ptr_array(geo) if geo is not None else None  # kept alive: the oracle holds the pointers
    if gp is not None:
        lib().orc_set_geometry(gp, len(geo))
    try:
        nerr = lib().orc_ug_advance(ctypes.byref(P), ctypes.byref(B), ug[0], ug[1], ug[2], ptr_array(arrays),
                                    nsteps, int(refill), nthreads, dptr(dts))
    finally:
        if gp is not None:
            lib().orc_set_geometry(None, 0)
    return dts[:nsteps], nerr


@contextlib.contextmanager
def dt_control(dt_init=0.0, change_factor=0.0, dt_min=0.0, dt_max=0.0):
    """The driver's dt limits for the advance functions inside the block
    (Driver_computeDt.F90:450-460 and dr_dtInit); reset to the CFL dt after."""
    lib().orc_set_dt_control(float(dt_init), float(change_factor), float(dt_min), float(dt_max))
    try:
        yield
    finally:
        lib().orc_set_dt_control(0.0, 0.0, 0.0, 0.0)


def ug_fill(cfg, b, ug, arrays):
    P, B = params(cfg), block(b)
    nunk = 11 + cfg["nscalars"]
    lib().orc_ug_fill_guards(ctypes.byref(P), ctypes.byref(B), ug[0], ug[1], ug[2], ptr_array(arrays), nunk)


def batch_advance(cfg, blocks, arrays, nsteps, nthreads=0):
    P = params(cfg)
    bl = (OrcBlock * len(blocks))(*[block(b) for b in blocks])
    dts = np.zeros(max(1, nsteps))
    nerr = lib().orc_batch_advance(ctypes.byref(P), bl, len(blocks), ptr_array(arrays), nsteps, nthreads, dptr(dts))
    return dts[:nsteps], nerr


def stage(cfg, b, U0, S, stage_no, ngoff, dt, fb=None):
    P, B = params(cfg), block(b)
    nerr = ctypes.c_int(0)
    if fb is None:
        lib().orc_stage(ctypes.byref(P), ctypes.byref(B), dptr(U0), dptr(S), stage_no, ngoff, dt, None, None, None,
                        ctypes.byref(nerr))
    else:
        lib().orc_stage(ctypes.byref(P), ctypes.byref(B), dptr(U0), dptr(S), stage_no, ngoff, dt, dptr(fb[0]),
                        dptr(fb[1]), dptr(fb[2]), ctypes.byref(nerr))
    return nerr.value


def shock_detect(cfg, b, U):
    P, B = params(cfg), block(b)
    lib().orc_shock_detect(ctypes.byref(P), ctypes.byref(B), dptr(U))


def correct_fluxes(cfg, b, U, dt, fb):
    P, B = params(cfg), block(b)
    lib().orc_correct_fluxes(ctypes.byref(P), ctypes.byref(B), dptr(U), dt, dptr(fb[0]), dptr(fb[1]), dptr(fb[2]))


# ---- single-function entry points (closed-form tests) ----------------------
def prim2con(V):
    V = np.ascontiguousarray(V, dtype=np.float64)
    U = np.zeros(5)
    lib().orc_prim2con(dptr(V), dptr(U))
    return U


def prim2flx(d, V):
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.zeros(5)
    lib().orc_prim2flx(d, dptr(V), dptr(F))
    return F


def hllc(cfg, d, VL, VR, in_shock=False):
    P = params(cfg)
    F = np.zeros(5)
    VL = np.ascontiguousarray(VL, dtype=np.float64)
    VR = np.ascontiguousarray(VR, dtype=np.float64)
    ierr = lib().orc_riemann_hllc(ctypes.byref(P), d, dptr(VL), dptr(VR), int(in_shock), dptr(F))
    return F, ierr


def hll(cfg, d, VL, VR):
    P = params(cfg)
    F = np.zeros(5)
    VL = np.ascontiguousarray(VL, dtype=np.float64)
    VR = np.ascontiguousarray(VR, dtype=np.float64)
    ierr = lib().orc_riemann_hll(ctypes.byref(P), d, dptr(VL), dptr(VR), dptr(F))
    return F, ierr


def recon(cfg, pen, i1, flat=1.0):
    """pen: (nrecon, plen) array; returns (uPlus, uMinus) at cell i1."""
    P = params(cfg)
    pen = np.ascontiguousarray(pen, dtype=np.float64)
    nr, plen = pen.shape
    uP, uM = np.zeros(nr), np.zeros(nr)
    lib().orc_recon(ctypes.byref(P), nr, dptr(pen), i1, plen, flat, dptr(uP), dptr(uM))
    return uP, uM


def eos_cell(cfg, v, full=True):
    P = params(cfg)
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    lib().orc_eos_dens_ei_cell(ctypes.byref(P), dptr(v), int(full))
    return v


class OrcAmr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nb0", ctypes.c_int * 3), ("level", ctypes.POINTER(ctypes.c_int)),
                ("origin", ctypes.POINTER(ctypes.c_int)), ("bc", ctypes.c_int * 3)]


def _amr(n, nb0, levels, origins, bc=(0, 0, 0)):
    lv = np.ascontiguousarray(np.asarray(levels, dtype=np.int32))
    og = np.ascontiguousarray(np.asarray(origins, dtype=np.int32).reshape(-1))
    A = OrcAmr()
    A.n = int(n)
    for d in range(3):
        A.nb0[d] = int(nb0[d])
    A.level = lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    A.origin = og.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    for d in range(3):
        A.bc[d] = int(bc[d])
    return A, (lv, og)


def amr_fill(cfg, blocks, nb0, origins, arrays, depth, bc=(0, 0, 0)):
    P = params(cfg)
    A, keep = _amr(blocks[0]["nx"], nb0, [b["level"] for b in blocks], origins, bc)
    nunk = 11 + cfg["nscalars"]
    nbad = lib().orc_amr_fill_guards(ctypes.byref(P), ctypes.byref(A), len(blocks), ptr_array(arrays), nunk,
                                     int(depth))
    if nbad:
        raise ValueError(f"hierarchy refused: {nbad} blocks have guard cells over a leaf two or more levels finer")


def flux_exchange(cfg, blocks, nb0, origins, fb, fc, bc=(0, 0, 0)):
    """fb / fc: lists of 3 arrays per block (x, y, z registers, flux fastest)."""
    P = params(cfg)
    A, keep = _amr(blocks[0]["nx"], nb0, [b["level"] for b in blocks], origins, bc)
    lib().orc_flux_exchange(ctypes.byref(P), ctypes.byref(A), len(blocks), ptr_array(fb), ptr_array(fc))


def amr_advance(cfg, blocks, nb0, origins, arrays, nsteps, flux_correct=True, nthreads=0, bc=(0, 0, 0)):
    P = params(cfg)
    A, keep = _amr(blocks[0]["nx"], nb0, [b["level"] for b in blocks], origins, bc)
    bl = (OrcBlock * len(blocks))(*[block(b) for b in blocks])
    dts = np.zeros(max(1, nsteps))
    nerr = lib().orc_amr_advance(ctypes.byref(P), ctypes.byref(A), bl, len(blocks), ptr_array(arrays), nsteps,
                                 int(bool(flux_correct)), nthreads, dptr(dts))
    if nerr < 0:
        raise ValueError("hierarchy refused: a guard layer lies over a leaf two or more levels finer")
    return dts[:nsteps], nerr
