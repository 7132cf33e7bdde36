"""ctypes binding of the C ABI declared in include/spark_hip.h.

This is host plumbing for the Python harness (tests, bench).  The Fortran
host side binds the same symbols through ISO_C_BINDING
(flash-x_amd/fortran/spark_hip_mod.F90).  Loading fails loudly when the
HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent          # flash-x_amd/
# SPARK_HIP_LIB selects another build of the library (same ABI)
LIB_PATH = Path(os.environ.get("SPARK_HIP_LIB", str(PKG_ROOT / "libspark_hip.so")))

SPARK_OK = 0
SPARK_NPROP = 11
SPARK_MAX_SCALARS = 8

RECON = {"weno5": 0, "mp5": 1, "plm": 2, "fog": 3}
GEOM = {"cartesian": 0, "cylindrical": 1, "spherical": 2}
BC = {"outflow": 0, "periodic": 1, "reflect": 2}
RIEMANN = {"hllc": 0, "hll": 1}
FORM = {"plain_fp64": 0, "fp32_flux": 1, "batched_fp64": 2}
GUARD = {"telescope": 0, "refill": 1}
EOS = {"gamma": 0, "multigamma": 1}

# canonical variable order of the device state (alphabetical unk order of
# Spark + Gamma, then mass scalars)
VARS = ["dens", "eint", "ener", "gamc", "game", "pres", "shok", "temp", "velx", "vely", "velz"]
VI = {n: i for i, n in enumerate(VARS)}


class SparkParams(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double), ("cfl", ctypes.c_double), ("cvisc", ctypes.c_double),
        ("smlrho", ctypes.c_double), ("smallp", ctypes.c_double), ("smalle", ctypes.c_double),
        ("smallx", ctypes.c_double), ("eintSwitch", ctypes.c_double), ("tiny", ctypes.c_double),
        ("gasConst", ctypes.c_double), ("abar", ctypes.c_double), ("grav", ctypes.c_double * 3),
        ("ndim", ctypes.c_int), ("nguard", ctypes.c_int), ("nstencil", ctypes.c_int),
        ("rk_order", ctypes.c_int), ("recon", ctypes.c_int), ("riemann", ctypes.c_int),
        ("use_flattening", ctypes.c_int), ("use_hybrid_riemann", ctypes.c_int),
        ("nscalars", ctypes.c_int), ("use_gravity", ctypes.c_int), ("flux_correct", ctypes.c_int),
        ("form", ctypes.c_int), ("guard_mode", ctypes.c_int), ("nunk_vars", ctypes.c_int),
        ("var_index", ctypes.c_int * (SPARK_NPROP + SPARK_MAX_SCALARS)),
        ("nspecies", ctypes.c_int), ("eos", ctypes.c_int), ("reserved", ctypes.c_int),
        ("species_a", ctypes.c_double * SPARK_MAX_SCALARS), ("species_gamma", ctypes.c_double * SPARK_MAX_SCALARS),
    ]


class SparkBlockDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("level", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class SparkKernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_longlong), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double), ("flops", ctypes.c_double)]


class SparkError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libspark_hip.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SparkError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        dpp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "spark_params_default": [ctypes.POINTER(SparkParams)],
            "spark_init": [ctypes.c_int],
            "spark_finalize": [],
            "spark_batch_create": [ctypes.POINTER(SparkParams), ctypes.c_int, ctypes.POINTER(SparkBlockDesc),
                                   ctypes.POINTER(vp)],
            "spark_batch_destroy": [vp],
            "spark_pack": [vp, dpp],
            "spark_pack_noalloc": [vp, dpp],
            "spark_unpack": [vp, dpp, ctypes.c_int, ctypes.c_int],
            "spark_set_uniform_grid": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "spark_fill_guardcells": [vp],
            "spark_compute_dt": [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)],
            "spark_hydro_step": [vp, ctypes.c_double, ctypes.c_double],
            "spark_step_begin": [vp],
            "spark_stage": [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int],
            "spark_step_end": [vp],
            "spark_advance": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
            "spark_set_dt_control": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double],
            "spark_set_geometry": [vp, ctypes.c_int, dpp, dpp, dpp, dpp, dpp],
            "spark_get_flux_registers": [vp, dpp, dpp, dpp],
            "spark_set_flux_corrections": [vp, dpp, dpp, dpp],
            "spark_flux_correct": [vp, ctypes.c_double],
            "spark_flux_exchange": [vp],
            "spark_set_amr_grid": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)],
            "spark_synchronize": [vp],
            "spark_get_counters": [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int],
            "spark_last_error": [ctypes.c_char_p, ctypes.c_int],
            "spark_profile": [vp, ctypes.c_int],
            "spark_set_option": [vp, ctypes.c_int, ctypes.c_int],
            "spark_batched_tiles": [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
            "spark_profile_read": [vp, ctypes.POINTER(SparkKernelStat), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def exported_symbols():
    """Names of the extern "C" entry points the header declares."""
    return [
        "spark_params_default", "spark_init", "spark_finalize", "spark_batch_create", "spark_batch_destroy",
        "spark_pack", "spark_unpack", "spark_set_uniform_grid", "spark_fill_guardcells", "spark_compute_dt",
        "spark_hydro_step", "spark_step_begin", "spark_stage", "spark_step_end", "spark_advance",
        "spark_get_flux_registers", "spark_set_flux_corrections", "spark_flux_correct", "spark_synchronize",
        "spark_get_counters", "spark_last_error", "spark_profile", "spark_profile_read", "spark_set_option",
        "spark_batched_tiles", "spark_set_amr_grid", "spark_flux_exchange", "spark_set_dt_control",
        "spark_set_geometry", "spark_pack_noalloc",
    ]


def check(rc: int, what: str = ""):
    if rc != SPARK_OK:
        buf = ctypes.create_string_buffer(1024)
        lib().spark_last_error(buf, 1024)
        raise SparkError(f"{what} failed (status {rc}): {buf.value.decode(errors='replace')}")


def make_params(cfg: dict) -> SparkParams:
    """Build spark_params from a config dict (keys as in the C struct, plus
    string forms for recon / riemann / form / guard_mode)."""
    p = SparkParams()
    check(lib().spark_params_default(ctypes.byref(p)), "spark_params_default")
    for k, v in cfg.items():
        if k == "recon":
            v = RECON[v] if isinstance(v, str) else v
        elif k == "riemann":
            v = RIEMANN[v] if isinstance(v, str) else v
        elif k == "form":
            v = FORM[v] if isinstance(v, str) else v
        elif k == "guard_mode":
            v = GUARD[v] if isinstance(v, str) else v
        if k == "grav":
            for i in range(3):
                p.grav[i] = float(v[i])
            continue
        if k == "var_index":
            for i, x in enumerate(v):
                p.var_index[i] = int(x)
            continue
        if k in ("species_a", "species_gamma"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
            continue
        if k == "eos":
            v = EOS[v] if isinstance(v, str) else v
        if k == "geometry":
            continue  # per-block arrays: Batch.set_geometry
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    if "nunk_vars" not in cfg:
        p.nunk_vars = SPARK_NPROP + p.nscalars
    return p


def init(device: int = 0):
    check(lib().spark_init(device), "spark_init")


def block_descs(blocks):
    """blocks: sequence of dicts with nx, ny, nz, level, dx, dy, dz."""
    arr = (SparkBlockDesc * len(blocks))()
    for i, b in enumerate(blocks):
        arr[i].nx, arr[i].ny, arr[i].nz = int(b["nx"]), int(b["ny"]), int(b["nz"])
        arr[i].level = int(b.get("level", 1))
        arr[i].dx, arr[i].dy, arr[i].dz = float(b["dx"]), float(b["dy"]), float(b["dz"])
    return arr


def _ptrs(arrays):
    out = (ctypes.c_void_p * len(arrays))()
    for i, a in enumerate(arrays):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        out[i] = a.ctypes.data
    return out


class Batch:
    """Device-resident set of blocks (one spark_batch)."""

    def __init__(self, cfg: dict, blocks):
        self.cfg = dict(cfg)
        self.params = make_params(cfg)
        self.blocks = list(blocks)
        self._descs = block_descs(self.blocks)
        h = ctypes.c_void_p()
        check(lib().spark_batch_create(ctypes.byref(self.params), len(self.blocks), self._descs, ctypes.byref(h)),
              "spark_batch_create")
        self.h = h

    def close(self):
        if self.h:
            check(lib().spark_batch_destroy(self.h), "spark_batch_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, arrays):
        check(lib().spark_pack(self.h, _ptrs(arrays)), "spark_pack")

    def pack_noalloc(self, arrays):
        check(lib().spark_pack_noalloc(self.h, _ptrs(arrays)), "spark_pack_noalloc")

    def unpack(self, arrays, interior_only=False, which=0):
        check(lib().spark_unpack(self.h, _ptrs(arrays), int(interior_only), int(which)), "spark_unpack")

    def set_uniform_grid(self, nbx, nby, nbz):
        check(lib().spark_set_uniform_grid(self.h, nbx, nby, nbz), "spark_set_uniform_grid")

    def fill_guardcells(self):
        check(lib().spark_fill_guardcells(self.h), "spark_fill_guardcells")

    def compute_dt(self):
        dt = ctypes.c_double()
        loc = (ctypes.c_int * 5)()
        check(lib().spark_compute_dt(self.h, ctypes.byref(dt), loc), "spark_compute_dt")
        return dt.value, list(loc)

    def hydro_step(self, dt, dt_old=None):
        check(lib().spark_hydro_step(self.h, float(dt), float(dt if dt_old is None else dt_old)),
              "spark_hydro_step")

    def step_begin(self):
        check(lib().spark_step_begin(self.h), "spark_step_begin")

    def stage(self, stage, dt, ngoff):
        check(lib().spark_stage(self.h, int(stage), float(dt), int(ngoff)), "spark_stage")

    def step_end(self):
        check(lib().spark_step_end(self.h), "spark_step_end")

    def advance(self, nsteps):
        dts = (ctypes.c_double * max(1, nsteps))()
        check(lib().spark_advance(self.h, int(nsteps), dts), "spark_advance")
        return np.array(dts[:nsteps])

    def set_geometry(self, geometry, rows=None):
        """geometry GEOM["cartesian"] (0), ["cylindrical"] (1) or
        ["spherical"] (2); rows: per block a dict of the Grid arrays over the
        guard-celled box -- "rc", "rl", "rr" (NX), "area", "vol" (NY x NX)."""
        g = GEOM[geometry] if isinstance(geometry, str) else int(geometry)
        if g == 0:
            check(lib().spark_set_geometry(self.h, 0, None, None, None, None, None), "spark_set_geometry")
            return
        keep = [[np.ascontiguousarray(r[k], dtype=np.float64) for r in rows] for k in ("rc", "area", "vol", "rl", "rr")]
        check(lib().spark_set_geometry(self.h, g, *[_ptrs(k) for k in keep]), "spark_set_geometry")

    def set_dt_control(self, dt_init=0.0, change_factor=0.0, dt_min=0.0, dt_max=0.0):
        """The driver's dt limits for advance() (all 0: the CFL dt)."""
        check(lib().spark_set_dt_control(self.h, float(dt_init), float(change_factor), float(dt_min), float(dt_max)),
              "spark_set_dt_control")

    def synchronize(self):
        check(lib().spark_synchronize(self.h), "spark_synchronize")

    def counters(self):
        c = (ctypes.c_longlong * 4)()
        check(lib().spark_get_counters(self.h, c, 4), "spark_get_counters")
        return list(c)

    def get_flux_registers(self, fbx, fby, fbz):
        check(lib().spark_get_flux_registers(self.h, _ptrs(fbx), _ptrs(fby), _ptrs(fbz)), "spark_get_flux_registers")

    def set_flux_corrections(self, fbx, fby, fbz):
        check(lib().spark_set_flux_corrections(self.h, _ptrs(fbx), _ptrs(fby), _ptrs(fbz)),
              "spark_set_flux_corrections")

    def set_option(self, option: str, value: int):
        code = {"graph": 1, "autotune": 2, "same_level_halos": 3}[option]
        check(lib().spark_set_option(self.h, code, int(value)), "spark_set_option")

    def tiles(self):
        rows = (ctypes.c_int * (6 * 64))()
        n = ctypes.c_int()
        check(lib().spark_batched_tiles(self.h, rows, 64, ctypes.byref(n)), "spark_batched_tiles")
        return [tuple(rows[6 * i:6 * i + 6]) for i in range(min(n.value, 64))]

    def profile(self, enable: bool):
        check(lib().spark_profile(self.h, int(enable)), "spark_profile")

    def profile_read(self):
        stats = (SparkKernelStat * 64)()
        n = ctypes.c_int()
        check(lib().spark_profile_read(self.h, stats, 64, ctypes.byref(n)), "spark_profile_read")
        return [dict(name=stats[i].name.decode(), launches=stats[i].launches, ms=stats[i].total_ms,
                     bytes=stats[i].bytes, flops=stats[i].flops) for i in range(min(n.value, 64))]

    def flux_correct(self, dt):
        check(lib().spark_flux_correct(self.h, float(dt)), "spark_flux_correct")

    def set_amr_grid(self, nb0, origins, bc=(0, 0, 0)):
        """nb0: level-1 blocks per dimension; origins: (nblocks, 3) first
        interior cell of each block in its level's global index space;
        bc: per dimension, BC["outflow"] (0), BC["periodic"] (1) or
        BC["reflect"] (2)."""
        o = np.ascontiguousarray(np.asarray(origins, dtype=np.int32).reshape(-1))
        per = (ctypes.c_int * 3)(*[BC[x] if isinstance(x, str) else int(x) for x in bc])
        check(lib().spark_set_amr_grid(self.h, int(nb0[0]), int(nb0[1]), int(nb0[2]),
                                       o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), per), "spark_set_amr_grid")

    def flux_exchange(self):
        check(lib().spark_flux_exchange(self.h), "spark_flux_exchange")


def gpu_available() -> bool:
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    try:
        import torch  # noqa: F401  (device probe only)
        return torch.cuda.is_available()
    except Exception:
        return False
