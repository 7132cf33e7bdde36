"""Block hierarchy on the GPU: device guard fill over levels (periodic x/y,
reflecting z) and the fine/coarse flux-register exchange + correction,
against the oracle.

Bar: the plain FP64 form equals the oracle bit for bit (dt sequence and
every interior value) over several steps of a 2-level Rayleigh-Taylor
problem with 8 mass scalars, gravity and flux correction; the batched
FP64 form equals the plain form bit for bit; the FP32-flux form stays
within L1(rho) <= 1e-5 of the oracle.
"""
import numpy as np
import pytest

from sparkhip import capi, problems
import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    capi.init(0)
    yield


def two_level(**kw):
    return problems.rt_amr(nb0=(2, 2, 2), n=8, lmax=2, slabs=((0.5, 1.0),), **kw)


def run(prob, nsteps, form):
    b = capi.Batch(dict(prob.cfg, form=form), prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_amr_grid(prob.info["nb0"], prob.info["origins"], prob.info["bc"])
        dts = b.advance(nsteps)
        out = [a.copy() for a in prob.arrays]
        b.unpack(out)
        cnt = b.counters()
    finally:
        b.close()
    return dts, out, cnt


def interiors(prob, arrays):
    g = prob.cfg["nguard"]
    return [a[g:-g, g:-g, g:-g] for a in arrays]


def test_guard_fill_matches_oracle():
    prob = two_level(nscalars=2)
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_amr_grid(prob.info["nb0"], prob.info["origins"], prob.info["bc"])
        b.fill_guardcells()
        out = [a.copy() for a in prob.arrays]
        b.unpack(out)
    finally:
        b.close()
    ref = [a.copy() for a in prob.arrays]
    pyoracle.amr_fill(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref, depth=prob.cfg["nguard"],
                      bc=prob.info["bc"])
    nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(out, ref))
    print(f"{len(prob.blocks)} blocks, levels {prob.info['leaves_per_level']}: differing values {nd}", flush=True)
    assert nd == 0


@pytest.mark.parametrize("flux_correct", [1, 0])
def test_plain_fp64_amr_steps_match_oracle(flux_correct):
    prob = two_level(nscalars=8, flux_correct=flux_correct)
    nsteps = 3
    print(f"RT 2 levels {prob.info['leaves_per_level']}, flux_correct={flux_correct}, {nsteps} steps", flush=True)
    d, out, cnt = run(prob, nsteps, "plain_fp64")
    ref = [a.copy() for a in prob.arrays]
    dref, nerr = pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref, nsteps,
                                      flux_correct=bool(flux_correct), bc=prob.info["bc"])
    assert nerr == 0 and cnt[0] == 0
    assert np.array_equal(d, dref), (d, dref)
    nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(interiors(prob, out), interiors(prob, ref)))
    print(f"differing interior values {nd}", flush=True)
    assert nd == 0


def test_batched_fp64_amr_bit_exact_with_plain():
    prob = two_level(nscalars=8)
    d64, o64, _ = run(prob, 3, "plain_fp64")
    db, ob, _ = run(prob, 3, "batched_fp64")
    assert np.array_equal(d64, db)
    assert all(np.array_equal(x, y) for x, y in zip(interiors(prob, o64), interiors(prob, ob)))


def test_fp32_flux_amr_vs_oracle():
    prob = two_level(nscalars=8)
    nsteps = 5
    d32, o32, cnt = run(prob, nsteps, "fp32_flux")
    ref = [a.copy() for a in prob.arrays]
    dref, _ = pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref, nsteps,
                                   bc=prob.info["bc"])
    vols = [b["dx"] ** 3 for b in prob.blocks]
    gi, ri = interiors(prob, o32), interiors(prob, ref)
    num = sum(np.sum(np.abs(x[..., 0] - y[..., 0])) * v for x, y, v in zip(gi, ri, vols))
    den = sum(np.sum(np.abs(y[..., 0])) * v for y, v in zip(ri, vols))
    print(f"L1(rho) {num / den:.2e}, dt rel diff {np.max(np.abs(d32 - dref) / dref):.2e}", flush=True)
    assert cnt[0] == 0
    assert num / den <= 1e-5
