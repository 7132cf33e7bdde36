"""The benchmark's Sedov workload against the analytic solution the
reference ships: 3D Sedov, 8^3 blocks of 16^3 (128^3 zones), 6 guards
refilled on the device, CFL 0.45, advanced to t = 0.02 (shock radius 27.5
cells) in the plain FP64 and FP32-flux forms.  Bar: shock radius within
1.5 cells of the analytic one and radially binned density within 20 % in
L1 behind the shock, in both forms."""
import numpy as np
import pytest

from sparkhip import capi, problems
from sedov_analytic import radial_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    capi.init(0)
    yield


@pytest.mark.parametrize("form", ["plain_fp64", "fp32_flux"])
def test_sedov128_against_analytic_profile(form):
    prob = problems.sedov3d(nb=8, n=16, nguard=6, guard_mode="refill", cfl=0.45)
    b = capi.Batch(dict(prob.cfg, form=form), prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_uniform_grid(*prob.ug)
        t, steps = 0.0, 0
        while t < 0.02:
            t += float(b.advance(1)[0])
            steps += 1
        out = [a.copy() for a in prob.arrays]
        b.unpack(out)
        cnt = b.counters()
    finally:
        b.close()
    R, front, l1, dx = radial_check(prob, out, t)
    print(f"{form}: t {t:.4f} after {steps} steps, shock {R / dx:.2f} cells, front {front / dx:.2f} cells, "
          f"L1(rho) inside {l1:.3f}", flush=True)
    assert cnt[0] == 0
    assert abs(front - R) / dx < 1.5
    assert l1 < 0.2
