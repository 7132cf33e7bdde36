"""The Fortran driver (flash-x_amd/fortran/spark_driver) runs the library
through the ISO_C_BINDING module on a Flash-X style unk array and matches
the oracle bit for bit (plain FP64 form).

The driver is the Flash-X unsplit loop in miniature: per step
Hydro_computeDt then Hydro, with unk(NUNK_VARS, i, j, k, blockID) handed
over as one c_loc base address per block.  roundtrip=1 packs and unpacks
unk around every step (host unk authoritative between Hydro calls).
"""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from sparkhip import capi, problems
import pyoracle

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "flash-x_amd" / "fortran" / "spark_driver"


def write_problem(path, prob, nsteps, roundtrip):
    c = prob.cfg
    b = prob.blocks[0]
    hdr = [c["ndim"], b["nx"], b["ny"], b["nz"], *prob.ug, c["nguard"], 11 + c["nscalars"], nsteps,
           capi.GUARD[c["guard_mode"]], c["nscalars"], capi.RECON[c["recon"]], capi.RIEMANN[c["riemann"]],
           c["rk_order"], c["nstencil"], c["use_flattening"], capi.FORM[c["form"]], roundtrip]
    with open(path, "wb") as f:
        f.write(struct.pack("<19i", *hdr))
        f.write(struct.pack("<5d", b["dx"], b["dy"], b["dz"], c["gamma"], c["cfl"]))
        for a in prob.arrays:
            f.write(np.ascontiguousarray(a).tobytes())


def read_result(path, prob):
    raw = Path(path).read_bytes()
    nsteps, status = struct.unpack_from("<2i", raw, 0)
    dts = np.frombuffer(raw, dtype=np.float64, count=nsteps, offset=8)
    off = 8 + 8 * nsteps
    out = []
    for a in prob.arrays:
        n = a.size
        out.append(np.frombuffer(raw, dtype=np.float64, count=n, offset=off).reshape(a.shape))
        off += 8 * n
    assert off == len(raw)
    return status, dts, out


@pytest.mark.parametrize("case,roundtrip", [("sod2d", 0), ("sod2d", 1), ("sedov_refill", 0)])
def test_fortran_driver_matches_oracle(tmp_path, case, roundtrip):
    assert DRIVER.exists(), "build the library first (make -C flash-x_amd)"
    if case == "sod2d":
        prob = problems.sod2d(nbx=4, nby=4, n=16, nguard=10)
        nsteps, refill = 6, False
    else:
        prob = problems.sedov3d(nb=2, n=16, nguard=6, guard_mode="refill")
        nsteps, refill = 4, True
    pin, pout = tmp_path / "problem.bin", tmp_path / "result.bin"
    write_problem(pin, prob, nsteps, roundtrip)
    print(f"fortran driver: {case} roundtrip={roundtrip} nsteps={nsteps}", flush=True)
    r = subprocess.run([str(DRIVER), str(pin), str(pout)], capture_output=True, text=True, timeout=90)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    status, dts, out = read_result(pout, prob)
    assert status == 0
    ref = [a.copy() for a in prob.arrays]
    dref, _ = pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, ref, nsteps, refill=refill)
    assert np.array_equal(dts, dref)
    g = prob.cfg["nguard"]
    sl = (slice(None),) * (3 - prob.cfg["ndim"]) + (slice(g, -g),) * prob.cfg["ndim"]
    for a, c in zip(out, ref):
        assert np.array_equal(a[sl], c[sl])
