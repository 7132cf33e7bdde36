"""The oracle's block-hierarchy services (guard fill over levels,
flux-register exchange) against closed forms and the single-level fill."""
from pathlib import Path

import numpy as np
import pytest

from sparkhip import problems
import pyoracle


def two_level(nscalars=0, n=8, nguard=6):
    """2 x 2 x 2 level-1 blocks of n^3; the upper layer refined once."""
    return problems.rt_amr(nb0=(2, 2, 2), n=n, lmax=2, slabs=((0.5, 1.0),), nguard=nguard, nscalars=nscalars)


def test_c4_hierarchy_counts_and_balance():
    slabs = ((0.25, 1.75), (0.5, 1.5))
    leaves = problems.amr_leaves((4, 4, 8), 16, 3, lambda L, lo, hi: lo[2] < slabs[L - 1][1] and hi[2] > slabs[L - 1][0],
                                 10)
    levels = np.bincount([L for L, _ in leaves], minlength=4)[1:]
    assert list(levels) == [32, 256, 4096]
    # the leaves tile the domain: volumes add up to the level-1 grid
    assert sum(8.0 ** -(L - 1) for L, _ in leaves) == 4 * 4 * 8


def test_single_level_fill_equals_uniform_fill():
    p = problems.sedov3d(nb=2, n=8, nguard=6)
    a = [x.copy() for x in p.arrays]
    b = [x.copy() for x in p.arrays]
    g = 6
    for y in b:  # clear the guards so the fill must rewrite them
        inner = y[g:-g, g:-g, g:-g].copy()
        y[:] = 0.0
        y[g:-g, g:-g, g:-g] = inner
    pyoracle.ug_fill(p.cfg, p.blocks[0], p.ug, a)
    origins = [[(i % 2) * 8, ((i // 2) % 2) * 8, (i // 4) * 8] for i in range(8)]
    blocks = [dict(b_, level=1) for b_ in p.blocks]
    pyoracle.amr_fill(p.cfg, blocks, (2, 2, 2), origins, b, depth=6)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def linear(X, Y, Z):
    return 1.0 + 2.0 * X + 3.0 * Y + 5.0 * Z


def test_fill_across_levels_restriction_and_prolongation():
    """On a linear field the mean of 8 fine cells (restriction) and the
    limited linear prolongation (MC slopes) both reproduce the field at the
    guard cell's centre; injection would miss it by a quarter coarse cell."""
    p = two_level()
    g, n = p.cfg["nguard"], 8
    arrays = []
    for b, o in zip(p.blocks, p.info["origins"]):
        NZ, NY, NX = problems.gc_shape(p.cfg, b)
        Z, Y, X = problems._centers(p.cfg, b, tuple(o[d] * b["dx"] for d in range(3)), NZ, NY, NX)
        U = np.zeros((NZ, NY, NX, 11))
        U[...] = linear(X, Y, Z)[..., None]
        arrays.append(U)
    ref = [a.copy() for a in arrays]
    pyoracle.amr_fill(p.cfg, p.blocks, p.info["nb0"], p.info["origins"], arrays, depth=g)  # outflow everywhere
    checked_fine = checked_coarse = 0
    for b, o, a, r in zip(p.blocks, p.info["origins"], arrays, ref):
        L = b["level"]
        for (k, j, i) in [(g - 1, g + 3, g + 2), (g - 2, g + 4, g + 5), (g + n, g + 1, g + 5), (g - 3, g + 6, g + 1)]:
            gi = [o[0] + i - g, o[1] + j - g, o[2] + k - g]
            N = 16 * 2 ** (L - 1)
            if not all(2 <= gi[d] < N - 2 for d in range(2)) or not 0 <= gi[2] < N // 2 * 2:
                continue  # stencils clamped at the domain edge are not exact
            zc = (gi[2] + 0.5) * b["dx"]
            if L == 1 and zc > 0.5:
                assert abs(a[k, j, i, 0] - r[k, j, i, 0]) < 1e-13  # restriction from the finer leaf
                checked_fine += 1
            if L == 2 and zc < 0.5:
                assert abs(a[k, j, i, 0] - r[k, j, i, 0]) < 1e-13  # prolongation from the coarser leaf
                checked_coarse += 1
    assert checked_fine > 0 and checked_coarse > 0


def test_prolongation_is_limited_at_a_jump():
    """Across a discontinuity the MC slope vanishes, so fine guard cells
    next to a jump take no value outside the two states."""
    p = two_level()
    g = p.cfg["nguard"]
    arrays = []
    for b, o in zip(p.blocks, p.info["origins"]):
        NZ, NY, NX = problems.gc_shape(p.cfg, b)
        Z, Y, X = problems._centers(p.cfg, b, tuple(o[d] * b["dx"] for d in range(3)), NZ, NY, NX)
        U = np.zeros((NZ, NY, NX, 11))
        U[...] = np.where(Z < 0.45, 1.0, 2.0)[..., None]
        arrays.append(U)
    pyoracle.amr_fill(p.cfg, p.blocks, p.info["nb0"], p.info["origins"], arrays, depth=g)
    for a in arrays:
        assert a[..., 0].min() >= 1.0 and a[..., 0].max() <= 2.0


def test_flux_exchange_constant_registers():
    p = two_level()
    n, nf = 8, 5
    fb, fc = [], []
    for b in p.blocks:
        val = 3.0 if b["level"] == 2 else 1.0
        for d in range(3):
            shp = ((n + (d == 2)) * (n + (d == 1)) * (n + (d == 0)) * nf)
            fb.append(np.full(shp, val))
            fc.append(np.full(shp, np.nan))
    pyoracle.flux_exchange(p.cfg, p.blocks, p.info["nb0"], p.info["origins"], fb, fc)
    for bi, (b, o) in enumerate(zip(p.blocks, p.info["origins"])):
        cz = fc[3 * bi + 2].reshape(n + 1, n, n, nf)
        assert np.all(np.isfinite(fc[3 * bi])) and np.all(np.isfinite(cz))
        if b["level"] == 1:
            # the high-z face of a lower-layer coarse block borders the refined layer
            assert np.all(cz[n] == 2.0) and np.all(cz[:n] == 0.0)
            assert np.all(fc[3 * bi] == 0.0) and np.all(fc[3 * bi + 1] == 0.0)
        else:
            assert np.all(cz == 0.0)


def test_reflecting_and_periodic_boundaries():
    """REFLECT mirrors the interior across the domain face and negates the
    normal velocity; PERIODIC wraps to the opposite side."""
    p = two_level(nscalars=0)
    g = p.cfg["nguard"]
    arrays = []
    rng = np.random.default_rng(3)
    for b in p.blocks:
        arrays.append(rng.random(problems.gc_shape(p.cfg, b) + (11,)))
    pyoracle.amr_fill(p.cfg, p.blocks, p.info["nb0"], p.info["origins"], arrays, depth=g, bc=(1, 1, 2))
    V = {"velx": 8, "vely": 9, "velz": 10}
    for b, o, a in zip(p.blocks, p.info["origins"], arrays):
        if b["level"] != 1 or o[2] != 0:
            continue
        # z = 0 face of the domain: guard layer m mirrors interior layer m
        for m in range(3):
            assert np.array_equal(a[g - 1 - m, g:-g, g:-g, V["velz"]], -a[g + m, g:-g, g:-g, V["velz"]])
            assert np.array_equal(a[g - 1 - m, g:-g, g:-g, V["velx"]], a[g + m, g:-g, g:-g, V["velx"]])
            assert np.array_equal(a[g - 1 - m, g:-g, g:-g, 0], a[g + m, g:-g, g:-g, 0])
        if o[0] == 0:
            # x = 0 face: periodic, the guard layer left of x = 0 is the last interior layer of the x-last block
            other = [c for c, oo, bb in zip(arrays, p.info["origins"], p.blocks)
                     if bb["level"] == 1 and oo[0] == 8 and oo[1] == o[1] and oo[2] == 0][0]
            assert np.array_equal(a[g:-g, g:-g, g - 1, :], other[g:-g, g:-g, -g - 1, :])


def test_flux_correction_conserves_mass():
    """Closed domain (periodic x/y, reflecting z, no gravity): with the
    fine/coarse correction the total mass of the leaves changes only by
    rounding over steps; without it the fine/coarse faces leak."""
    def mass(prob, arrs):
        g = prob.cfg["nguard"]
        return sum(np.sum(a[g:-g, g:-g, g:-g, 0]) * b["dx"] ** 3 for a, b in zip(arrs, prob.blocks))

    drift = {}
    for fc in (1, 0):
        prob = problems.rt_amr(nscalars=0, nb0=(2, 2, 2), n=8, lmax=2, slabs=((0.5, 1.0),), use_gravity=0,
                               flux_correct=fc)
        arr = [a.copy() for a in prob.arrays]
        m0 = mass(prob, arr)
        pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], arr, 2,
                             flux_correct=bool(fc), bc=prob.info["bc"], nthreads=8)
        drift[fc] = abs(mass(prob, arr) - m0) / m0
    assert drift[1] < 1e-14
    assert drift[0] > 1e-7


def test_golden_rt16_record_is_the_oracle_result():
    """tests/golden/rt16_oracle_3steps.json (the three-level 16^3 hierarchy,
    3 steps) is what the oracle gives for what the generator builds: same
    input bytes, same dt sequence, same interior hash for every leaf."""
    import importlib.util
    import json
    gold = Path(__file__).resolve().parent / "golden"
    spec = importlib.util.spec_from_file_location("make_rt16_oracle", gold / "make_rt16_oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rec = json.loads((gold / "rt16_oracle_3steps.json").read_text())
    prob = mod.problem()
    assert mod.input_hash(prob) == rec["input_sha256"]
    ref = [a.copy() for a in prob.arrays]
    dts, nerr = pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref,
                                     rec["steps"], flux_correct=True, bc=prob.info["bc"])
    assert nerr == 0
    assert [float(d).hex() for d in dts] == rec["dts"]
    assert mod.interior_hashes(prob, ref) == rec["interior_sha256"]


def test_fill_refuses_a_guard_layer_over_a_leaf_two_levels_finer():
    """A level-1 block whose 9 guard layers (1.125 in z) reach past the
    level-2 leaves above it into level 3: the oracle's fill has no cell set
    for a leaf two levels finer, so it refuses the hierarchy (as the device
    library does) instead of reading outside the finer block."""
    prob = problems.rt_amr(nb0=(1, 1, 4), n=8, lmax=3, slabs=((1.5, 2.5), (1.9, 2.1)), nguard=10, nscalars=0)
    assert list(prob.info["leaves_per_level"])[0] > 0 and list(prob.info["leaves_per_level"])[2] > 0
    arr = [a.copy() for a in prob.arrays]
    with pytest.raises(ValueError, match="two or more levels finer"):
        pyoracle.amr_fill(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], arr, depth=9,
                          bc=prob.info["bc"])
    with pytest.raises(ValueError, match="two or more levels finer"):
        pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], arr, 1,
                             bc=prob.info["bc"])
    # one layer shallower than a level-2 leaf is tall (8 level-2 cells = 4
    # level-1 cells) the same hierarchy fills
    pyoracle.amr_fill(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], arr, depth=4,
                      bc=prob.info["bc"])
