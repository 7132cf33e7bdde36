"""The C-ABI library loads and exports every entry point include/spark_hip.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from sparkhip import capi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "spark_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*int\s+(spark_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == sorted(capi.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_params_default_matches_reference_defaults():
    p = capi.make_params({})
    assert (p.gamma, p.cfl, p.cvisc, p.smlrho, p.smalle, p.eintSwitch, p.tiny) == (1.6667, 0.8, 0.1, 1e-10, 1e-10,
                                                                                  0.0, 1e-32)
    assert p.use_flattening == 1 and p.use_hybrid_riemann == 0 and p.rk_order == 2 and p.nguard == 6


def test_batch_create_without_init_is_an_error_not_a_crash():
    import pytest
    with pytest.raises(capi.SparkError):
        capi.Batch({"ndim": 3}, [dict(nx=8, ny=8, nz=8, dx=1.0, dy=1.0, dz=1.0)])
