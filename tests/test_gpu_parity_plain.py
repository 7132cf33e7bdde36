"""Plain FP64 form on the GPU against the CPU restatement (oracle).

Bar: the FP64 form agrees with the oracle within 1e-12 relative per
variable on identical inputs; both sides are compiled without FMA
contraction, so the expected result is bit-identical and the tests report
the number of differing values.
"""
import numpy as np
import pytest

from sparkhip import capi, problems
import pyoracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _device():
    capi.init(0)
    yield


def clone(arrays):
    return [a.copy() for a in arrays]


def gpu_advance(prob, arrays, nsteps):
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(arrays)
        if prob.ug is not None:
            b.set_uniform_grid(*prob.ug)
        dts = b.advance(nsteps)
        out = clone(arrays)
        b.unpack(out)
        cnt = b.counters()
    finally:
        b.close()
    return dts, out, cnt


def oracle_advance(prob, arrays, nsteps):
    arrs = clone(arrays)
    if prob.ug is not None:
        dts, nerr = pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, arrs, nsteps,
                                        refill=(prob.cfg["guard_mode"] == "refill"), nthreads=0)
    else:
        dts, nerr = pyoracle.batch_advance(prob.cfg, prob.blocks, arrs, nsteps, nthreads=0)
    return dts, arrs, nerr


def interior_views(prob, arrays, pad=0):
    """Views of the interior (+pad layers) of each block: the cells the
    reference copies back into unk (updateState)."""
    g, nd = prob.cfg["nguard"], prob.cfg["ndim"]
    out = []
    for a in arrays:
        sl = [slice(g - pad, a.shape[2 - d] - g + pad) if d < nd else slice(None) for d in range(3)][::-1]
        out.append(a[tuple(sl)])
    return out


def compare(got, ref, label):
    worst = 0.0
    ndiff = 0
    for g, r in zip(got, ref):
        scale = np.maximum(np.abs(r), 1e-300)
        rel = np.abs(g - r) / scale
        rel[(g == r)] = 0.0
        worst = max(worst, float(np.max(rel)))
        ndiff += int(np.count_nonzero(g != r))
    print(f"[{label}] max rel diff {worst:.3e}, differing values {ndiff}", flush=True)
    assert worst <= RTOL, f"{label}: max relative difference {worst:.3e} > {RTOL}"
    return worst, ndiff


def run_case(prob, nsteps, label):
    print(f"case {label}: {len(prob.blocks)} blocks, {nsteps} steps", flush=True)
    d_g, out_g, cnt = gpu_advance(prob, prob.arrays, nsteps)
    d_o, out_o, nerr = oracle_advance(prob, prob.arrays, nsteps)
    np.testing.assert_allclose(d_g, d_o, rtol=RTOL, atol=0)
    assert cnt[0] == nerr
    return compare(interior_views(prob, out_g), interior_views(prob, out_o), label)


def test_sod2d_full_config_telescope():
    """The full 2D Sod shock tube: 16x16 blocks of 16^2, WENO5+HLLC, RK3, CFL 0.8, 20 steps (NGUARD 10 telescoping)."""
    prob = problems.sod2d(nbx=16, nby=16, n=16, nguard=10)
    run_case(prob, 20, "sod2d 256^2")


def test_sedov3d_small_telescope():
    prob = problems.sedov3d(nb=2, n=16, nguard=10, r_cells=3.5)
    run_case(prob, 4, "sedov3d 2^3x16^3 telescope g10")


def test_sedov3d_small_refill_matches_telescope_oracle():
    """6-guard refill mode on the device reproduces the oracle in the same mode."""
    prob = problems.sedov3d(nb=2, n=16, nguard=6, guard_mode="refill")
    run_case(prob, 4, "sedov3d refill g6")


def test_batched_small_mixed_sizes():
    prob = problems.batched(nblocks=24, nguard=10, sizes=(8, 16, 32), probs=(0.25, 0.625, 0.125), shock_every=4)
    run_case(prob, 3, "batched 24 blocks")


@pytest.mark.parametrize("recon,riemann,rk,ng,nst", [
    ("weno5", "hll", 3, 10, 3),
    ("mp5", "hllc", 3, 10, 3),
    ("plm", "hllc", 2, 4, 2),
    ("plm", "hll", 2, 4, 2),
    ("fog", "hll", 2, 4, 2),
    ("weno5", "hllc", 2, 6, 3),
])
def test_sod2d_variants(recon, riemann, rk, ng, nst):
    prob = problems.sod2d(nbx=4, nby=4, n=16, nguard=ng, recon=recon, riemann=riemann, rk_order=rk, nstencil=nst)
    run_case(prob, 5, f"sod2d {recon}+{riemann} rk{rk}")


def test_hybrid_riemann_and_no_flattening():
    prob = problems.sod2d(nbx=4, nby=4, n=16, nguard=10, use_hybrid_riemann=1, use_flattening=0)
    run_case(prob, 5, "sod2d hybrid, no flattening")


def test_hybrid_riemann_3d_mixed_sizes():
    """3D hybrid HLLC/HLL on blocks of 8^3, 16^3 and 32^3: the shock flags of
    the guard band pick the Riemann solver of faces there, and the 32^3 boxes
    (50 cells wide with the band) span several shockDetect column tiles."""
    prob = problems.batched(nblocks=12, nguard=10, sizes=(8, 16, 32), probs=(0.25, 0.5, 0.25), shock_every=3,
                            use_hybrid_riemann=1)
    assert 32 in [b["nx"] for b in prob.blocks]
    run_case(prob, 3, "batched hybrid 3d mixed sizes")


def test_scalars_and_gravity():
    prob = problems.batched(nblocks=8, nguard=10, sizes=(8, 16), probs=(0.5, 0.5), shock_every=2, nscalars=8,
                            use_gravity=1, grav=(0.0, 0.0, -1.0))
    run_case(prob, 3, "batched ns=8 gravity")


def test_stage_level_entry_points():
    """step_begin + one stage + the stage state match the oracle's stage."""
    prob = problems.sedov3d(nb=1, n=16, nguard=10)
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        dt, _ = b.compute_dt()
        b.step_begin()
        b.stage(1, dt, 6)
        S = clone(prob.arrays)
        b.unpack(S, which=1)
    finally:
        b.close()
    U = prob.arrays[0].copy()
    pyoracle.shock_detect(prob.cfg, prob.blocks[0], U)
    S_ref = U.copy()
    pyoracle.stage(prob.cfg, prob.blocks[0], U, S_ref, 1, 6, dt)
    # stage 1 updates interior +/- 6; the device writes TEMP/GAME only in the
    # last stage (nothing reads them before), so compare what the stage writes
    idx = [capi.VI[v] for v in ("dens", "velx", "vely", "velz", "pres", "eint", "ener", "gamc")]
    compare([interior_views(prob, S, 6)[0][..., idx]], [interior_views(prob, [S_ref], 6)[0][..., idx]], "stage 1")
    dto, _ = pyoracle.compute_dt(prob.cfg, prob.blocks[0], prob.arrays[0])
    assert dt == dto


def test_stage_level_entry_points_two_stages():
    """Stages 1 and 2 through the stage-level entry points: the stage state
    read back after each (stage 2 writes its own buffer on the device, the
    oracle updates one buffer in place) matches the oracle's over the
    region the stage writes."""
    prob = problems.sedov3d(nb=1, n=16, nguard=10)
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        dt, _ = b.compute_dt()
        b.step_begin()
        b.stage(1, dt, 6)
        S1 = clone(prob.arrays)
        b.unpack(S1, which=1)
        b.stage(2, dt, 3)
        S2 = clone(prob.arrays)
        b.unpack(S2, which=1)
    finally:
        b.close()
    U = prob.arrays[0].copy()
    pyoracle.shock_detect(prob.cfg, prob.blocks[0], U)
    S_ref = U.copy()
    idx = [capi.VI[v] for v in ("dens", "velx", "vely", "velz", "pres", "eint", "ener", "gamc")]
    pyoracle.stage(prob.cfg, prob.blocks[0], U, S_ref, 1, 6, dt)
    compare([interior_views(prob, S1, 6)[0][..., idx]], [interior_views(prob, [S_ref], 6)[0][..., idx]], "stage 1")
    pyoracle.stage(prob.cfg, prob.blocks[0], U, S_ref, 2, 3, dt)
    compare([interior_views(prob, S2, 3)[0][..., idx]], [interior_views(prob, [S_ref], 3)[0][..., idx]], "stage 2")


def test_flux_correction_matches_oracle():
    """Flux registers after one step and hy_rk_correctFluxes with given
    correction data match the oracle."""
    prob = problems.batched(nblocks=4, nguard=10, sizes=(8, 16), probs=(0.5, 0.5), shock_every=2, flux_correct=1)
    rng = np.random.default_rng(7)
    nf = 5
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        dt, _ = b.compute_dt()
        b.hydro_step(dt)
        regs = [[np.zeros(((bl["nz"] + (d == 2)) * (bl["ny"] + (d == 1)) * (bl["nx"] + (d == 0)) * nf))
                 for bl in prob.blocks] for d in range(3)]
        b.get_flux_registers(*regs)
        corr = [[rng.normal(scale=1e-3, size=r.shape) for r in regs[d]] for d in range(3)]
        b.set_flux_corrections(*corr)
        b.flux_correct(dt)
        out = clone(prob.arrays)
        b.unpack(out)
    finally:
        b.close()
    ref = clone(prob.arrays)
    for i, bl in enumerate(prob.blocks):
        fb = [np.zeros_like(regs[d][i]) for d in range(3)]
        pyoracle.hydro_block(prob.cfg, bl, ref[i], dt, fb=fb)
        # the boundary faces (the ones the register hand-off reads) match;
        # interior faces are not accumulated on the device and read 0
        got, want = [], []
        for d in range(3):
            e = [bl["nx"], bl["ny"], bl["nz"]]
            e[d] += 1
            g4 = regs[d][i].reshape(e[2], e[1], e[0], nf)
            w4 = fb[d].reshape(e[2], e[1], e[0], nf)
            sl = [slice(None)] * 3
            sl[2 - d] = [0, e[d] - 1]
            got.append(g4[tuple(sl)])
            want.append(w4[tuple(sl)])
            inner = [slice(None)] * 3
            inner[2 - d] = slice(1, e[d] - 1)
            assert not np.any(g4[tuple(inner)]), f"interior registers block {i} dir {d}"
        compare(got, want, f"registers block {i}")
        pyoracle.correct_fluxes(prob.cfg, bl, ref[i], dt, [corr[d][i] for d in range(3)])
    compare(interior_views(prob, out), interior_views(prob, ref), "flux correction")


def test_pack_unpack_roundtrip_with_var_map():
    """A permuted unk variable map and extra unk slots survive pack/unpack."""
    prob = problems.sedov3d(nb=1, n=8, nguard=10)
    nunk = 14
    perm = [13, 0, 5, 2, 9, 1, 7, 11, 3, 6, 10]
    host = []
    for a in prob.arrays:
        h = np.full(a.shape[:-1] + (nunk,), -7.0)
        for v in range(11):
            h[..., perm[v]] = a[..., v]
        host.append(h)
    cfg = dict(prob.cfg, nunk_vars=nunk, var_index=perm)
    b = capi.Batch(cfg, prob.blocks)
    try:
        b.pack(host)
        back = [np.zeros_like(h) for h in host]
        b.unpack(back)
    finally:
        b.close()
    for h, g in zip(host, back):
        for v in range(11):
            assert np.array_equal(h[..., perm[v]], g[..., perm[v]])


@pytest.mark.parametrize("guard_mode", ["telescope", "refill"])
def test_species_multigamma_matches_oracle(guard_mode):
    """Two gases (A = 1, gamma 1.4 | A = 4, gamma 5/3) across the diagonal
    Sod interface with the Multigamma EOS: species faces clamped and
    renormalised, gamc per cell from the mixture; 6 steps."""
    kw = dict(nguard=6, guard_mode="refill") if guard_mode == "refill" else {}
    prob = problems.sod2d_species(nbx=2, nby=2, n=16, **kw)
    d, out, cnt = gpu_advance(prob, prob.arrays, 6)
    dref, ref, nerr = oracle_advance(prob, prob.arrays, 6)
    assert cnt[0] == 0 and nerr == 0
    assert np.array_equal(d, dref), (d, dref)
    compare(interior_views(prob, out), interior_views(prob, ref), f"multigamma {guard_mode}")


def test_batched_forms_reject_what_they_do_not_implement():
    """The FP32 reconstruction takes gamc = gamma (no Multigamma), and the
    species must fit the fused kernel's first scalar chunk."""
    prob = problems.species_tube3d(nb=1, n=16)
    with pytest.raises(capi.SparkError, match="FP32-flux form implements the gamma law only"):
        capi.Batch(dict(prob.cfg, form="fp32_flux"), prob.blocks)
    five = problems.species_tube3d(nb=1, n=16, species_a=(1.0, 4.0, 2.0, 3.0, 5.0),
                                   species_gamma=(1.4, 5.0 / 3.0, 1.4, 1.4, 1.4))
    with pytest.raises(capi.SparkError, match="at most 4 species"):
        capi.Batch(dict(five.cfg, form="batched_fp64"), five.blocks)


def test_advance_with_driver_dt_limits_matches_oracle():
    """spark_set_dt_control: dtinit for the first step, then the CFL dt
    limited to twice the previous one and clamped to dt_max
    (Driver_computeDt.F90:450-460), kept on the device across graph
    replays; the plain form equals the oracle with the same limits."""
    prob = problems.sod2d(nbx=2, nby=2, n=16, nguard=10)
    ctl = dict(dt_init=1e-5, change_factor=2.0, dt_max=1.5e-3)
    nsteps = 12
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_uniform_grid(*prob.ug)
        b.set_dt_control(**ctl)
        d1 = b.advance(5)        # the first advance captures the step graph
        d2 = b.advance(nsteps - 5)
        out = clone(prob.arrays)
        b.unpack(out)
        b.set_dt_control()       # back to the CFL dt
        d3 = b.advance(1)
    finally:
        b.close()
    ref = clone(prob.arrays)
    with pyoracle.dt_control(**ctl):
        dref, nerr = pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, ref, nsteps)
    d = np.concatenate([d1, d2])
    print(f"dts {d}", flush=True)
    assert nerr == 0
    assert np.array_equal(d, dref), (d, dref)
    assert d[0] == 1e-5 and d[1] == 2e-5 and d.max() == 1.5e-3
    compare(interior_views(prob, out), interior_views(prob, ref), "sod2d with driver dt limits")
    ref1 = clone(ref)
    dcfl, _ = pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, ref1, 1)
    assert np.array_equal(d3, dcfl)


@pytest.mark.parametrize("form", ["batched_fp64", "fp32_flux"])
def test_batched_forms_apply_the_driver_dt_limits(form):
    prob = problems.sedov3d(nb=2, n=16, nguard=6, guard_mode="refill")
    ctl = dict(dt_init=1e-6, change_factor=2.0, dt_min=0.0, dt_max=0.0)
    res = {}
    for f in ("plain_fp64", form):
        b = capi.Batch(dict(prob.cfg, form=f), prob.blocks)
        try:
            b.pack(prob.arrays)
            b.set_uniform_grid(*prob.ug)
            b.set_dt_control(**ctl)
            dts = b.advance(6)
            out = clone(prob.arrays)
            b.unpack(out)
        finally:
            b.close()
        res[f] = (dts, out)
    assert list(res[form][0]) == [1e-6 * 2.0 ** k for k in range(6)]
    assert np.array_equal(res["plain_fp64"][0], res[form][0])
    if form == "batched_fp64":
        nd = sum(int(np.count_nonzero(a != c)) for a, c in
                 zip(interior_views(prob, res["plain_fp64"][1]), interior_views(prob, res[form][1])))
        assert nd == 0


@pytest.mark.parametrize("case", ["1d_telescope", "1d_refill", "2d_telescope"])
def test_spherical_geometry_matches_oracle(case):
    """Spherical blast with polar and azimuthal velocity, 1D and 2D (r,
    theta): facM/facP from theta-dependent areas and volumes,
    (rho w^2 + 2 p)/r, -rho u w / r and (rho v)^2 / rho / r scaled by
    dx / dx_sph (geoFacs), the r dtheta width in the CFL step; the plain
    form equals the oracle bit for bit."""
    guard_mode = "refill" if case.endswith("refill") else "telescope"
    kw = dict(nguard=6, guard_mode="refill") if guard_mode == "refill" else {}
    if case.startswith("2d"):
        prob = problems.spherical_blast(nb=(4, 2), n=16, ndim=2, **kw)
    else:
        prob = problems.spherical_blast(nb=(4, 1), n=32, **kw)
    nsteps = 10
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_uniform_grid(*prob.ug)
        b.set_geometry("spherical", prob.info["geo"])
        d = b.advance(nsteps)
        out = clone(prob.arrays)
        b.unpack(out)
        cnt = b.counters()
    finally:
        b.close()
    ref = clone(prob.arrays)
    dref, nerr = pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, ref, nsteps,
                                     refill=guard_mode == "refill", geo=[b["geo"] for b in prob.blocks])
    assert cnt[0] == 0 and nerr == 0
    assert np.array_equal(d, dref), (d, dref)
    compare(interior_views(prob, out), interior_views(prob, ref), f"spherical {case}")


@pytest.mark.parametrize("guard_mode", ["telescope", "refill"])
def test_cylindrical_geometry_matches_oracle(guard_mode):
    """2D cylindrical (r, z) blast with azimuthal swirl: geometric factors
    facM/facP on the radial faces and the sources (rho w^2 + p)/r and
    -rho u w / r (geoFacs, hy_rk_updateSoln.F90-mc:98-175) from the Grid's
    centre radii, face areas and volumes; the plain form equals the oracle
    bit for bit, dt sequence included."""
    kw = dict(nguard=6, guard_mode="refill") if guard_mode == "refill" else {}
    prob = problems.cylindrical_blast(nbx=4, nby=2, n=16, **kw)
    geo = prob.info["geo"]
    nsteps = 8
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        b.pack(prob.arrays)
        b.set_uniform_grid(*prob.ug)
        b.set_geometry("cylindrical", geo)
        d = b.advance(nsteps)
        out = clone(prob.arrays)
        b.unpack(out)
        cnt = b.counters()
    finally:
        b.close()
    ref = clone(prob.arrays)
    dref, nerr = pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, ref, nsteps,
                                     refill=guard_mode == "refill", geo=[b["geo"] for b in prob.blocks])
    assert cnt[0] == 0 and nerr == 0
    assert np.array_equal(d, dref), (d, dref)
    compare(interior_views(prob, out), interior_views(prob, ref), f"cylindrical {guard_mode}")


def test_geometry_is_refused_where_it_is_not_implemented():
    prob = problems.cylindrical_blast(nbx=1, nby=1, n=16)
    rows = prob.info["geo"]
    s3 = problems.sedov3d(nb=1, n=16, nguard=10)
    b = capi.Batch(dict(s3.cfg, form="batched_fp64"), s3.blocks)
    try:
        with pytest.raises(capi.SparkError, match="Cartesian only"):
            b.set_geometry("cylindrical", rows)
    finally:
        b.close()
    b = capi.Batch(s3.cfg, s3.blocks)
    try:
        with pytest.raises(capi.SparkError, match="2D"):
            b.set_geometry("cylindrical", rows)
        with pytest.raises(capi.SparkError, match="1D \\(r\\) or 2D"):
            b.set_geometry("spherical", rows)
    finally:
        b.close()
