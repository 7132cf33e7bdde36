"""The Fortran ISO_C_BINDING module matches the C ABI (no device calls).

A small Fortran program built against flash-x_amd/fortran/spark_hip_mod.F90
fills spark_params through spark_params_default and writes what it sees;
the same struct read through ctypes must agree member by member, so the
bind(C) derived types have the C layout.  Every spark_* entry point of the
header has a bind(C) interface in the module.
"""
import re
import shutil
import struct
import subprocess
from pathlib import Path

import pytest

from sparkhip import capi

ROOT = Path(__file__).resolve().parent.parent
FC = "/opt/rocm/bin/amdflang"
MOD = ROOT / "flash-x_amd" / "fortran" / "spark_hip_mod.F90"

PROBE = r"""
program probe
  use iso_c_binding
  use spark_hip
  implicit none
  type(spark_params_t), target :: p
  type(spark_block_desc_t), target :: bd(2)
  integer(c_int) :: rc
  integer :: u
  rc = spark_params_default(p)
  open (newunit=u, file='probe.bin', access='stream', form='unformatted', status='replace')
  write (u) rc, int(c_sizeof(p), c_int), int(c_sizeof(bd(1)), c_int)
  write (u) p
  close (u)
end program probe
"""


def header_symbols():
    text = (ROOT / "include" / "spark_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*int\s+(spark_\w+)\s*\(", text, flags=re.M)))


def test_module_binds_every_header_entry_point():
    text = MOD.read_text()
    bound = sorted(set(re.findall(r'bind\(C,\s*name="(spark_\w+)"\)', text)))
    assert bound == header_symbols()


@pytest.mark.skipif(not Path(FC).exists(), reason="amdflang not present")
def test_params_struct_layout_agrees_with_c(tmp_path):
    shutil.copy(MOD, tmp_path / "spark_hip_mod.F90")
    (tmp_path / "probe.F90").write_text(PROBE)
    lib = capi.LIB_PATH
    subprocess.run([FC, "-O0", "-c", "spark_hip_mod.F90"], cwd=tmp_path, check=True, timeout=120)
    subprocess.run([FC, "-O0", "probe.F90", "spark_hip_mod.o", f"-L{lib.parent}", "-lspark_hip",
                    f"-Wl,-rpath,{lib.parent}", "-o", "probe"], cwd=tmp_path, check=True, timeout=120)
    subprocess.run(["./probe"], cwd=tmp_path, check=True, timeout=60)
    raw = (tmp_path / "probe.bin").read_bytes()
    rc, psize, bsize = struct.unpack_from("<3i", raw, 0)
    assert rc == 0
    import ctypes
    assert psize == ctypes.sizeof(capi.SparkParams)
    assert bsize == ctypes.sizeof(capi.SparkBlockDesc)
    # an unformatted write of a derived type carries its members without the
    # trailing padding of the C struct
    body = raw[12:]
    got = capi.SparkParams.from_buffer_copy(body + bytes(psize - len(body)))
    ref = capi.make_params({})
    for name, _ in capi.SparkParams._fields_:
        a, b = getattr(got, name), getattr(ref, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
