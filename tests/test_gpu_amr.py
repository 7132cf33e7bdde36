"""Block hierarchy on the GPU: device guard fill over levels (periodic x/y,
reflecting z) and the fine/coarse flux-register exchange + correction,
against the oracle.

Bar: the plain FP64 form equals the oracle bit for bit (dt sequence and
every interior value) over several steps of a 2-level Rayleigh-Taylor
problem with 8 mass scalars, gravity and flux correction; the batched
FP64 form equals the plain form bit for bit; the FP32-flux form stays
within L1(rho) <= 1e-5 of the oracle.
"""
import importlib.util
import json
from pathlib import Path

import numpy as np
import pytest

from sparkhip import capi, problems
import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def _device():
    capi.init(0)
    yield


def two_level(**kw):
    return problems.rt_amr(nb0=(2, 2, 2), n=8, lmax=2, slabs=((0.5, 1.0),), **kw)


def run(prob, nsteps, form, **options):
    b = capi.Batch(dict(prob.cfg, form=form), prob.blocks)
    try:
        for k, v in options.items():
            b.set_option(k, v)
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


@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 2, 0), (2, 1, 1)])
def test_guard_fill_boundary_codes_match_oracle(bc):
    """Device guard fill over the 3-level hierarchy with each boundary code
    (outflow, periodic, reflecting) on some dimension."""
    prob = three_level(nscalars=2)
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_amr_grid(prob.info["nb0"], prob.info["origins"], bc)
        b.fill_guardcells()
        out = [a.copy() for a in prob.arrays]
        b.unpack(out)
    finally:
        b.close()
    ref = [a.copy() for a in prob.arrays]
    pyoracle.amr_fill(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref, depth=prob.cfg["nguard"],
                      bc=bc)
    nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(out, ref))
    print(f"bc {bc}: differing values {nd}", flush=True)
    assert nd == 0


# hybrid: a pressure jump of 10x across the planes x = 0.5 and x = 1 (= 0,
# periodic), both on block faces of every level, with use_hybridRiemann, so
# the faces there read SHOK across block and level boundaries
SHOCKED = dict(use_hybrid_riemann=1, p_boost=(0.5, 1.0, 10.0))
# the first two mass scalars as species (clamped and renormalised at faces)
SPECIES = dict(nspecies=2, eos="gamma")


@pytest.mark.parametrize("flux_correct,hybrid", [(1, 0), (0, 0), (1, 1), (1, "species")])
def test_plain_fp64_amr_steps_match_oracle(flux_correct, hybrid):
    extra = SHOCKED if hybrid == 1 else SPECIES if hybrid == "species" else {}
    prob = two_level(nscalars=8, flux_correct=flux_correct, **extra)
    nsteps = 3
    print(f"RT 2 levels {prob.info['leaves_per_level']}, flux_correct={flux_correct}, hybrid={hybrid}, "
          f"{nsteps} steps", flush=True)
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


def three_level(**kw):
    # 1 x 1 x 2 level-1 blocks of 16^3, 3 levels: 32 of the 72 leaves (levels
    # 2 and 3) have 26 leaves of their own level around them, so the batched
    # forms read their halos from those neighbours and update only their
    # interior every stage; the other 40 read the guard sectors that have a
    # leaf of their level across from it and telescope over the rest
    return problems.rt_amr(nb0=(1, 1, 2), n=16, lmax=3, **kw)


def same_level_blocks(prob):
    """(leaves whose 26 guard sectors all have a leaf of their own level
    across (periodic wrap, no other domain edge), leaves with some)."""
    info = prob.info
    n = prob.blocks[0]["nx"]
    lv = [b["level"] for b in prob.blocks]
    lmax = max(lv)
    org = np.asarray(info["origins"]).reshape(-1, 3)
    md = [x << (lmax - 1) for x in info["nb0"]]
    slot = -np.ones(md[::-1], dtype=int)
    for i, (o, L) in enumerate(zip(org, lv)):
        sh = lmax - L
        lo = [(int(o[d]) // n) << sh for d in range(3)]
        slot[lo[2]:lo[2] + (1 << sh), lo[1]:lo[1] + (1 << sh), lo[0]:lo[0] + (1 << sh)] = i
    per = [b == problems.BC_PERIODIC for b in info["bc"]]
    full = some = 0
    for o, L in zip(org, lv):
        sh = lmax - L
        nbl = [info["nb0"][d] << (L - 1) for d in range(3)]
        nsame = 0
        for off in np.ndindex(3, 3, 3):
            if off == (1, 1, 1):
                continue
            q = [int(o[d]) // n + off[d] - 1 for d in range(3)]
            ok = True
            for d in range(3):
                if not 0 <= q[d] < nbl[d]:
                    ok = ok and per[d]
                    q[d] %= nbl[d]
            nsame += ok and lv[slot[q[2] << sh, q[1] << sh, q[0] << sh]] == L
        full += nsame == 26
        some += 0 < nsame < 26
    return full, some


def rt_three_level_16():
    """The shape of the 3D Rayleigh-Taylor workload at a twelfth of its size:
    16^3 blocks on three levels of leaves (8 / 96 / 256 on a 2 x 2 x 6
    level-1 grid, z in [0, 3], level 2 where z meets (0.5, 2.5), level 3
    where it meets (1.25, 1.75)), so the hierarchy has two level jumps and
    no guard layer (9 deep at level 1: 18 level-2 cells) reaches past a
    level-2 leaf into level 3; 8 mass scalars, gravity g = -1 along z,
    periodic x/y, reflecting z, fine/coarse flux correction every step."""
    return problems.rt_amr(nb0=(2, 2, 6), n=16, lmax=3, slabs=((0.5, 2.5), (1.25, 1.75)), nscalars=8)


@pytest.fixture(scope="module")
def rt16_oracle():
    prob = rt_three_level_16()
    nsteps = 3
    ref = [a.copy() for a in prob.arrays]
    dref, nerr = pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref, nsteps,
                                      flux_correct=True, bc=prob.info["bc"])
    return prob, nsteps, dref, nerr, ref


@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64"])
def test_three_level_16cubed_steps_match_oracle(form, rt16_oracle):
    """3 steps over the three-level 16^3 hierarchy against the oracle's
    hierarchy driver (guard fill over levels, dt, Hydro per leaf, register
    exchange, correction): dt sequence and every interior value equal."""
    prob, nsteps, dref, nerr, ref = rt16_oracle
    assert prob.cfg["flux_correct"] == 1 and prob.cfg["use_gravity"] == 1
    assert list(prob.info["leaves_per_level"]) == [8, 96, 256]
    assert prob.info["bc"] == (problems.BC_PERIODIC, problems.BC_PERIODIC, problems.BC_REFLECT)
    print(f"RT 3 levels {prob.info['leaves_per_level']}, 16^3 blocks, {form}, {nsteps} steps", flush=True)
    d, out, cnt = run(prob, nsteps, form)
    assert nerr == 0 and cnt[0] == 0
    assert np.array_equal(d, dref), (d, dref)
    nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(interiors(prob, out), interiors(prob, ref)))
    print(f"differing interior values {nd}", flush=True)
    assert nd == 0


def test_three_level_16cubed_fp32_flux_vs_oracle(rt16_oracle):
    """The FP32-flux form over the same 3 steps: L1(rho) and L1(E) within
    1e-5 of the oracle (L1(rho u) is printed: the flow starts at rest, so
    the relative momentum error has a near-zero denominator)."""
    prob, nsteps, dref, _, ref = rt16_oracle
    d32, o32, cnt = run(prob, nsteps, "fp32_flux")
    assert cnt[0] == 0
    vols = [b["dx"] ** 3 for b in prob.blocks]
    gi, ri = interiors(prob, o32), interiors(prob, ref)
    V = capi.VI
    fields = {"rho": lambda a: a[..., V["dens"]], "rhou": lambda a: a[..., V["dens"]] * a[..., V["velx"]],
              "E": lambda a: a[..., V["dens"]] * a[..., V["ener"]]}
    for name, f in fields.items():
        num = sum(np.sum(np.abs(f(x) - f(y))) * v for x, y, v in zip(gi, ri, vols))
        den = sum(np.sum(np.abs(f(y))) * v for y, v in zip(ri, vols))
        print(f"L1({name}) {num / den:.2e}", flush=True)
        if name != "rhou":
            assert num / den <= 1e-5, name
    assert np.max(np.abs(d32 - dref) / dref) <= 1e-5


def _golden_rt16():
    """tests/golden/make_rt16_oracle.py as a module, and its record."""
    spec = importlib.util.spec_from_file_location("make_rt16_oracle", GOLDEN / "make_rt16_oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod, json.loads((GOLDEN / "rt16_oracle_3steps.json").read_text())


@pytest.fixture(scope="module")
def rt16_golden():
    mod, rec = _golden_rt16()
    prob = mod.problem()
    assert list(prob.info["leaves_per_level"]) == rec["leaves_per_level"] == [8, 96, 256]
    # the oracle's hashes hold for exactly these input bytes
    assert mod.input_hash(prob) == rec["input_sha256"], "generated input differs from the golden record's"
    return mod, rec, prob


@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64"])
def test_three_level_16cubed_steps_match_oracle_hashes(form, rt16_golden):
    """3 steps over the three-level 16^3 hierarchy (two level jumps, 8 mass
    scalars, gravity, periodic x/y, reflecting z, flux correction) against
    the oracle's hierarchy driver: the dt sequence equals the oracle's as
    hex floats and every leaf interior hashes to the oracle's SHA-256."""
    mod, rec, prob = rt16_golden
    assert prob.cfg["flux_correct"] == 1 and prob.cfg["use_gravity"] == 1
    assert prob.info["bc"] == (problems.BC_PERIODIC, problems.BC_PERIODIC, problems.BC_REFLECT)
    print(f"RT 3 levels {rec['leaves_per_level']}, 16^3 blocks, {form}, {rec['steps']} steps", flush=True)
    d, out, cnt = run(prob, rec["steps"], form)
    assert cnt[0] == 0
    assert [float(x).hex() for x in d] == rec["dts"], (d, rec["dts"])
    got = mod.interior_hashes(prob, out)
    bad = [i for i, (x, y) in enumerate(zip(got, rec["interior_sha256"])) if x != y]
    print(f"leaves differing from the oracle: {len(bad)}", flush=True)
    assert len(got) == len(rec["interior_sha256"]) and not bad, bad[:10]


def test_three_level_16cubed_fp32_flux_vs_plain(rt16_golden):
    """The FP32-flux form over the same 3 steps against the plain FP64 form
    (which the hash test pins to the oracle): L1(rho) and L1(E) within 1e-5
    (L1(rho u) is printed: the flow starts at rest, so the relative momentum
    error has a near-zero denominator)."""
    mod, rec, prob = rt16_golden
    d64, o64, _ = run(prob, rec["steps"], "plain_fp64")
    d32, o32, cnt = run(prob, rec["steps"], "fp32_flux")
    assert cnt[0] == 0
    vols = [b["dx"] ** 3 for b in prob.blocks]
    gi, ri = interiors(prob, o32), interiors(prob, o64)
    V = capi.VI
    fields = {"rho": lambda a: a[..., V["dens"]], "rhou": lambda a: a[..., V["dens"]] * a[..., V["velx"]],
              "E": lambda a: a[..., V["dens"]] * a[..., V["ener"]]}
    for name, f in fields.items():
        num = sum(np.sum(np.abs(f(x) - f(y))) * v for x, y, v in zip(gi, ri, vols))
        den = sum(np.sum(np.abs(f(y))) * v for y, v in zip(ri, vols))
        print(f"L1({name}) {num / den:.2e}", flush=True)
        if name != "rhou":
            assert num / den <= 1e-5, name
    assert np.max(np.abs(d32 - d64) / d64) <= 1e-5


@pytest.mark.parametrize("case", ["gamma", "hybrid", "species"])
def test_batched_fp64_amr_same_level_halos_bit_exact_with_plain(case):
    prob = three_level(nscalars=8, **{"gamma": {}, "hybrid": SHOCKED, "species": SPECIES}[case])
    full, some = same_level_blocks(prob)
    print(f"RT 3 levels {prob.info['leaves_per_level']}: {full} blocks read all their halos from same-level "
          f"leaves, {some} some of them", flush=True)
    assert full > 0 and some > 0
    d64, o64, _ = run(prob, 3, "plain_fp64")
    db, ob, cnt = run(prob, 3, "batched_fp64")
    assert cnt[0] == 0
    assert np.array_equal(d64, db)
    nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(interiors(prob, o64), interiors(prob, ob)))
    print(f"differing interior values {nd}", flush=True)
    assert nd == 0


@pytest.mark.parametrize("form,hybrid", [("batched_fp64", 0), ("fp32_flux", 0), ("fp32_flux", 1)])
def test_same_level_halos_equal_telescoping(form, hybrid):
    """Reading the halo of a leaf surrounded by same-level leaves from those
    leaves gives the values telescoping computes in its guard cells: the
    batched forms with and without it agree bit for bit."""
    prob = three_level(nscalars=3, **(SHOCKED if hybrid else {}))
    d_on, o_on, _ = run(prob, 3, form)
    d_off, o_off, _ = run(prob, 3, form, same_level_halos=0)
    assert np.array_equal(d_on, d_off)
    nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(interiors(prob, o_on), interiors(prob, o_off)))
    print(f"{form}: differing interior values {nd}", flush=True)
    assert nd == 0


@pytest.mark.parametrize("hybrid", [0, 1])
def test_fp32_flux_amr_same_level_halos_vs_plain(hybrid):
    prob = three_level(nscalars=8, **(SHOCKED if hybrid else {}))
    nsteps = 10
    d64, o64, _ = run(prob, nsteps, "plain_fp64")
    d32, o32, cnt = run(prob, nsteps, "fp32_flux")
    vols = [b["dx"] ** 3 for b in prob.blocks]
    gi, ri = interiors(prob, o32), interiors(prob, o64)
    num = sum(np.sum(np.abs(x[..., 0] - y[..., 0])) * v for x, y, v in zip(gi, ri, vols))
    den = sum(np.sum(np.abs(y[..., 0])) * v for y, v in zip(ri, vols))
    print(f"L1(rho) {num / den:.2e}, dt rel diff {np.max(np.abs(d32 - d64) / d64):.2e}", flush=True)
    assert cnt[0] == 0
    assert num / den <= 1e-5


@pytest.mark.parametrize("form", ["plain_fp64", "fp32_flux"])
def test_amr_mass_conserved_with_flux_correction(form):
    """A periodic domain is closed, so the total mass sum(rho dV) over the
    leaves only changes by rounding when the coarse faces take the
    restricted fine fluxes; without the correction the fine/coarse faces
    leak.  Periodic z and no gravity: at a reflecting wall the two face
    states are mirror images only up to the rounding of the reconstruction
    (FP32 in the FP32-flux form), and under gravity the telescoped guard
    cells below the wall evolve with the same downward gravity as the
    interior, so the wall itself passes mass (as in the reference, which
    fills guard cells once per step)."""
    drift = {}
    for fc in (1, 0):
        prob = three_level(nscalars=2, flux_correct=fc, use_gravity=0)
        prob.info["bc"] = (problems.BC_PERIODIC,) * 3
        g = prob.cfg["nguard"]
        vols = [b["dx"] ** 3 for b in prob.blocks]
        m0 = sum(np.sum(a[g:-g, g:-g, g:-g, 0]) * v for a, v in zip(prob.arrays, vols))
        _, out, cnt = run(prob, 10, form)
        assert cnt[0] == 0
        m1 = sum(np.sum(a[g:-g, g:-g, g:-g, 0]) * v for a, v in zip(out, vols))
        drift[fc] = abs(m1 - m0) / m0
    print(f"{form}: relative mass drift {drift[1]:.2e} with correction, {drift[0]:.2e} without", flush=True)
    assert drift[1] < 1e-13
    assert drift[0] > 100 * drift[1]


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


@pytest.mark.parametrize("form", ["plain_fp64", "fp32_flux"])
def test_graph_replays_match_eager_steps_on_the_hierarchy(form):
    """The captured step over a hierarchy (guard fill with sector masks and
    halo copies, three stages, flux exchange and correction), replayed with
    eager unpacks and synchronisations between replays, reproduces the eager
    steps bit for bit."""
    prob = three_level(nscalars=3)
    results = {}
    for graph in (False, True):
        b = capi.Batch(dict(prob.cfg, form=form), prob.blocks)
        try:
            b.set_option("graph", int(graph))
            b.pack(prob.arrays)
            b.set_amr_grid(prob.info["nb0"], prob.info["origins"], prob.info["bc"])
            seq = []
            for _ in range(4):
                dts = b.advance(1)
                out = [a.copy() for a in prob.arrays]
                b.unpack(out)
                seq.append((dts.copy(), out))
        finally:
            b.close()
        results[graph] = seq
    for s, ((de, oe), (dg, og)) in enumerate(zip(results[False], results[True])):
        assert np.array_equal(de, dg), s
        nd = sum(int(np.count_nonzero(x != y)) for x, y in zip(interiors(prob, oe), interiors(prob, og)))
        assert nd == 0, f"replay {s}: {nd} differing interior values"
