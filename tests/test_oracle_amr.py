"""The oracle's block-hierarchy services (guard fill over levels,
flux-register exchange) against closed forms and the single-level fill."""
import numpy as np

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


def test_fill_across_levels_restriction_and_injection():
    p = two_level()
    g, n = p.cfg["nguard"], 8
    dx0 = 1.0 / 16
    arrays = []
    for b, o in zip(p.blocks, p.info["origins"]):
        NZ, NY, NX = problems.gc_shape(p.cfg, b)
        Z, Y, X = problems._centers(p.cfg, b, tuple(o[d] * b["dx"] for d in range(3)), NZ, NY, NX)
        U = np.zeros((NZ, NY, NX, 11))
        U[...] = linear(X, Y, Z)[..., None]
        arrays.append(U)
    ref = [a.copy() for a in arrays]
    pyoracle.amr_fill(p.cfg, p.blocks, p.info["nb0"], p.info["origins"], arrays, depth=g)
    checked_fine = checked_coarse = 0
    for b, o, a, r in zip(p.blocks, p.info["origins"], arrays, ref):
        L, dx = b["level"], b["dx"]
        NZ, NY, NX = a.shape[:3]
        for (k, j, i) in [(g - 1, g + 3, g + 2), (g + n, g + 1, g + 5), (g + 2, g - 2, g + n + 1)]:
            gi = [o[0] + i - g, o[1] + j - g, o[2] + k - g]
            nmax = 2 * 16 * 2 ** (L - 1)
            if not all(0 <= gi[d] < (nmax if d == 2 else nmax // 2) for d in range(3)):
                continue
            zc = (gi[2] + 0.5) * dx
            if L == 1 and zc > 0.5:
                # the source is the finer leaf: the mean of 8 cells of a linear field
                assert abs(a[k, j, i, 0] - r[k, j, i, 0]) < 1e-13
                checked_fine += 1
            if L == 2 and zc < 0.5:
                # the source is the coarser leaf: the coarse cell containing it
                c = [(gi[d] // 2 + 0.5) * 2 * dx for d in range(3)]
                assert a[k, j, i, 0] == linear(c[0], c[1], c[2])
                checked_coarse += 1
    assert checked_fine > 0 and checked_coarse > 0


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
