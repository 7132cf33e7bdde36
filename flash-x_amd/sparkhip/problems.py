"""Synthetic block sets for the benchmark configurations.

Each generator returns a Problem: the run configuration (spark_params
fields), the block descriptors and one numpy array per block holding
unk(var, i, j, k) over the guard-celled box (C shape (NZ, NY, NX, nunk),
so the variable index is fastest, as in the reference unk layout).

  sod2d      2D Sod tube, interface x + y = 1, 16x16 blocks of 16^2
  sedov3d    3D Sedov blast, 8^3 blocks of 16^3, E = 1 in r < 3.5 cells
  batched    4096 independent 3D blocks of 8^3/16^3/32^3 (2:5:1), levels
             0..4, smooth random Fourier fields plus one planar shock per
             16 blocks, fields evaluated in every cell including guards
  rayleigh_taylor   3-level 16^3 block hierarchy, rho 2 over 1, g = -1,
             8 mass scalars with tanh profiles
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .capi import VI, SPARK_NPROP, BC

BC_OUTFLOW, BC_PERIODIC, BC_REFLECT = BC["outflow"], BC["periodic"], BC["reflect"]

GAS_CONST = 8.3144598e7


@dataclass
class Problem:
    name: str
    cfg: dict
    blocks: list
    arrays: list
    ug: tuple | None = None          # (nbx, nby, nbz) for uniform block grids
    info: dict = field(default_factory=dict)

    @property
    def zones(self) -> int:
        nd = self.cfg["ndim"]
        return int(sum(b["nx"] * (b["ny"] if nd > 1 else 1) * (b["nz"] if nd > 2 else 1) for b in self.blocks))


def gc_shape(cfg, b):
    g = cfg["nguard"]
    nd = cfg["ndim"]
    NX = b["nx"] + 2 * g
    NY = b["ny"] + 2 * g if nd > 1 else 1
    NZ = b["nz"] + 2 * g if nd > 2 else 1
    return NZ, NY, NX


def base_cfg(**kw):
    cfg = dict(gamma=1.4, cfl=0.8, cvisc=0.1, smlrho=1e-10, smallp=1e-10, smalle=1e-10, smallx=1e-10,
               eintSwitch=0.0, tiny=1e-32, gasConst=GAS_CONST, abar=1.0, grav=(0.0, 0.0, 0.0), ndim=3,
               nguard=10, nstencil=3, rk_order=3, recon="weno5", riemann="hllc", use_flattening=1,
               use_hybrid_riemann=0, nscalars=0, use_gravity=0, flux_correct=0, form="plain_fp64",
               guard_mode="telescope")
    cfg.update(kw)
    return cfg


def fill_thermo(U, dens, pres, vx, vy, vz, gamma, abar=1.0, scalars=None):
    """Primitive fields -> a consistent Gamma-law unk state."""
    eint = pres / ((gamma - 1.0) * dens)
    U[..., VI["dens"]] = dens
    U[..., VI["pres"]] = pres
    U[..., VI["velx"]] = vx
    U[..., VI["vely"]] = vy
    U[..., VI["velz"]] = vz
    U[..., VI["eint"]] = eint
    U[..., VI["ener"]] = eint + 0.5 * (vx * vx + vy * vy + vz * vz)
    U[..., VI["gamc"]] = gamma
    U[..., VI["game"]] = gamma
    U[..., VI["temp"]] = eint * (gamma - 1.0) / GAS_CONST * abar
    U[..., VI["shok"]] = 0.0
    if scalars is not None:
        for s, X in enumerate(scalars):
            U[..., SPARK_NPROP + s] = X


def _centers(cfg, b, origin, NZ, NY, NX):
    g = cfg["nguard"]
    nd = cfg["ndim"]
    x = origin[0] + (np.arange(NX) - g + 0.5) * b["dx"]
    y = origin[1] + (np.arange(NY) - (g if nd > 1 else 0) + 0.5) * b["dy"]
    z = origin[2] + (np.arange(NZ) - (g if nd > 2 else 0) + 0.5) * b["dz"]
    return np.meshgrid(z, y, x, indexing="ij")  # Z, Y, X each (NZ, NY, NX)


def sod2d(nbx=16, nby=16, n=16, nguard=10, gamma=1.4, **kw) -> Problem:
    """2D Sod problem with the interface on x + y = 1 (cells straddling it
    take the average of both states, as Sod's Simulation_initBlock does)."""
    cfg = base_cfg(ndim=2, nguard=nguard, gamma=gamma, **kw)
    dx = 1.0 / (nbx * n)
    dy = 1.0 / (nby * n)
    blocks, arrays = [], []
    nunk = SPARK_NPROP + cfg["nscalars"]
    for bj in range(nby):
        for bi in range(nbx):
            b = dict(nx=n, ny=n, nz=1, level=1, dx=dx, dy=dy, dz=1.0)
            NZ, NY, NX = gc_shape(cfg, b)
            Z, Y, X = _centers(cfg, b, (bi * n * dx, bj * n * dy, 0.0), NZ, NY, NX)
            xl, xr = X - 0.5 * dx, X + 0.5 * dx
            lpos = 1.0 - Y
            left = xr <= lpos
            mid = (xl < lpos) & (xr > lpos)
            dens = np.where(left, 1.0, np.where(mid, 0.5 * (1.0 + 0.125), 0.125))
            pres = np.where(left, 1.0, np.where(mid, 0.5 * (1.0 + 0.1), 0.1))
            U = np.zeros((NZ, NY, NX, nunk))
            zero = np.zeros_like(dens)
            fill_thermo(U, dens, pres, zero, zero, zero, gamma)
            blocks.append(b)
            arrays.append(U)
    return Problem("sod2d", cfg, blocks, arrays, ug=(nbx, nby, 1))


def mixture_thermo(X, species_a, species_gamma):
    """abar and gamc of a Multigamma mixture of mass fractions X (..., ns):
    c_v adds over the species (c_v,s = R / (A_s (gamma_s - 1))), so
    gamc = 1 + (R / abar) / c_v with 1/abar = sum X_s / A_s."""
    a = np.asarray(species_a, dtype=float)
    g = np.asarray(species_gamma, dtype=float)
    abar = 1.0 / np.sum(X / a, axis=-1)
    cv_over_r = np.sum(X / (a * (g - 1.0)), axis=-1)
    return abar, 1.0 + 1.0 / (abar * cv_over_r)


def sod2d_species(nbx=4, nby=4, n=16, nguard=10, species_a=(1.0, 4.0), species_gamma=(1.4, 5.0 / 3.0),
                  width=0.02, **kw) -> Problem:
    """Two-gas shock tube on [0,1]^2 with the Multigamma EOS: the Sod states
    (rho, p) = (1, 1) | (0.125, 0.1) across x + y = 1, gas 1 (A = 1, gamma
    1.4) on the left and gas 2 (A = 4, gamma 5/3) on the right, the mass
    fractions blended by a tanh of the given width; gamc, eint and T follow
    the mixture."""
    ns = len(species_a)
    cfg = base_cfg(ndim=2, nguard=nguard, gamma=float(species_gamma[0]), nscalars=ns, nspecies=ns, eos="multigamma",
                   species_a=tuple(species_a), species_gamma=tuple(species_gamma), **kw)
    dx = 1.0 / (nbx * n)
    dy = 1.0 / (nby * n)
    blocks, arrays = [], []
    nunk = SPARK_NPROP + cfg["nscalars"]
    for bj in range(nby):
        for bi in range(nbx):
            b = dict(nx=n, ny=n, nz=1, level=1, dx=dx, dy=dy, dz=1.0)
            NZ, NY, NX = gc_shape(cfg, b)
            Z, Y, X = _centers(cfg, b, (bi * n * dx, bj * n * dy, 0.0), NZ, NY, NX)
            s = (X + Y - 1.0) / np.sqrt(2.0)
            right = 0.5 * (1.0 + np.tanh(s / width))
            dens = 1.0 + (0.125 - 1.0) * right
            pres = 1.0 + (0.1 - 1.0) * right
            Xs = np.stack([1.0 - right, right] + [np.zeros_like(right)] * (ns - 2), axis=-1)
            U = np.zeros((NZ, NY, NX, nunk))
            _fill_mixture(U, dens, pres, Xs, species_a, species_gamma)
            blocks.append(b)
            arrays.append(U)
    return Problem("sod2d_species", cfg, blocks, arrays, ug=(nbx, nby, 1))


def _fill_mixture(U, dens, pres, Xs, species_a, species_gamma, scalars=None):
    """A fluid at rest with a Multigamma mixture Xs (..., ns): eint, gamc,
    game and T from the mixture; further mass scalars after the species."""
    abar, gamc = mixture_thermo(Xs, species_a, species_gamma)
    eint = pres / ((gamc - 1.0) * dens)
    U[..., VI["dens"]] = dens
    U[..., VI["pres"]] = pres
    U[..., VI["eint"]] = eint
    U[..., VI["ener"]] = eint
    U[..., VI["gamc"]] = gamc
    U[..., VI["game"]] = pres / (eint * dens) + 1.0
    U[..., VI["temp"]] = eint * (gamc - 1.0) * abar / GAS_CONST
    ns = Xs.shape[-1]
    for q in range(ns):
        U[..., SPARK_NPROP + q] = Xs[..., q]
    for q, Xq in enumerate(scalars or []):
        U[..., SPARK_NPROP + ns + q] = Xq


def species_tube3d(nb=2, n=16, nguard=10, species_a=(1.0, 4.0), species_gamma=(1.4, 5.0 / 3.0), width=0.03,
                   nscalars=None, eos="multigamma", **kw) -> Problem:
    """Two-gas shock tube on [0,1]^3: the Sod states (rho, p) = (1, 1) |
    (0.125, 0.1) across x + y + z = 1.5, gas 1 on the low side and gas 2 on
    the high side, the mass fractions blended by a tanh of the given width.
    nscalars (default: the species count) adds passive mass scalars after
    the species (smooth profiles in z).  eos "multigamma" takes gamc, eint
    and T from the mixture; "gamma" advects the species under the gamma law
    of gas 1."""
    ns = len(species_a)
    nsc = ns if nscalars is None else nscalars
    cfg = base_cfg(ndim=3, nguard=nguard, gamma=float(species_gamma[0]), nscalars=nsc, nspecies=ns, eos=eos,
                   species_a=tuple(species_a), species_gamma=tuple(species_gamma), **kw)
    if eos == "gamma":
        species_gamma = (float(species_gamma[0]),) * ns
        species_a = (1.0,) * ns
    dx = 1.0 / (nb * n)
    blocks, arrays = [], []
    nunk = SPARK_NPROP + nsc
    for bk in range(nb):
        for bj in range(nb):
            for bi in range(nb):
                b = dict(nx=n, ny=n, nz=n, level=1, dx=dx, dy=dx, dz=dx)
                NZ, NY, NX = gc_shape(cfg, b)
                Z, Y, X = _centers(cfg, b, (bi * n * dx, bj * n * dx, bk * n * dx), NZ, NY, NX)
                right = 0.5 * (1.0 + np.tanh((X + Y + Z - 1.5) / np.sqrt(3.0) / width))
                dens = 1.0 + (0.125 - 1.0) * right
                pres = 1.0 + (0.1 - 1.0) * right
                Xs = np.stack([1.0 - right, right] + [np.zeros_like(right)] * (ns - 2), axis=-1)
                extra = [0.5 * (1.0 + np.tanh((Z - 0.5) / (0.05 * (q + 1)))) for q in range(nsc - ns)]
                U = np.zeros((NZ, NY, NX, nunk))
                _fill_mixture(U, dens, pres, Xs, species_a, species_gamma, extra)
                blocks.append(b)
                arrays.append(U)
    return Problem("species_tube3d", cfg, blocks, arrays, ug=(nb, nb, nb))


def _coords(n, nguard, bi, dx, x0):
    """Cell-centre, left and right edge coordinates over the guard-celled
    index of block bi along one axis, as Grid_getCellCoords.F90:173-200
    computes them: x0 + (K - shift) dx with K the 1-based global cell index
    plus one and shift 1.5 (centre), 2 (left edge, faces) or 1 (right edge)."""
    NX = n + 2 * nguard
    K = (bi * n + 1 + (np.arange(NX) - nguard) + 1).astype(np.float64)
    return x0 + (K - 1.5) * dx, x0 + (K - 2.0) * dx, x0 + (K - 1.0) * dx


def geo_flat(rows):
    """The block's geometry arrays in the layout the oracle reads: rc, rl,
    rr, then area and volume (NY x NX, x fastest)."""
    return np.ascontiguousarray(np.concatenate([rows["rc"], rows["rl"], rows["rr"], rows["area"].ravel(),
                                                rows["vol"].ravel()]))


def cylindrical_rows(n, nguard, bi, nb, dx, dz, r0, ny=1):
    """Grid arrays of block bi along r (2D cylindrical, r from r0) over the
    guard-celled box (ny rows in z): radii (_coords), area 2 pi |r_left| dz
    (Grid_getCellFaceAreas.F90, CYLINDRICAL, 2D) and volume
    2 pi |r_centre| dr dz (Grid_getCellVolumes.F90, CYLINDRICAL, 2D), the
    same for every z row."""
    rc, rl, rr = _coords(n, nguard, bi, dx, r0)
    area = 2.0 * np.pi * np.abs(rl) * dz
    vol = 2.0 * np.pi * np.abs(rc) * dx * dz
    return dict(rc=rc, rl=rl, rr=rr, area=np.tile(area, (ny, 1)), vol=np.tile(vol, (ny, 1)))


def spherical_rows(n, nguard, bi, dx, r0, thf=None):
    """Grid arrays of block bi along r (spherical, r from r0): radii, area
    r_left^2 4 pi and volume dr (r_l^2 + r_l r_r + r_r^2) 4 pi / 3 in 1D;
    with the theta faces thf (NY + 1) of the block's rows, area
    r_left^2 (cos th_j - cos th_j+1) 2 pi and volume
    dr (...) (cos th_j - cos th_j+1) 2 pi / 3 (Grid_getCellFaceAreas.F90,
    Grid_getCellVolumes.F90, SPHERICAL)."""
    rc, rl, rr = _coords(n, nguard, bi, dx, r0)
    cv = dx * (rl * rl + rl * rr + rr * rr)
    if thf is None:
        area = (rl * rl * 4.0 * np.pi)[None, :]
        vol = (cv * 4.0 * np.pi / 3.0)[None, :]
    else:
        dc = (np.cos(thf[:-1]) - np.cos(thf[1:]))[:, None]
        area = (rl * rl)[None, :] * dc * 2.0 * np.pi
        vol = cv[None, :] * dc * 2.0 * np.pi / 3.0
    return dict(rc=rc, rl=rl, rr=rr, area=np.ascontiguousarray(area), vol=np.ascontiguousarray(vol))


def spherical_blast(nb=(4, 1), n=32, nguard=10, ndim=1, r0=0.5, th0=np.pi / 4, p_in=10.0, r_in=0.3,
                    vtheta=0.05, swirl=0.1, vr=0.0, gamma=1.4, **kw) -> Problem:
    """Spherical blast: r in [r0, r0 + 1] (and, in 2D, theta from th0 over
    nb[1] / nb[0] rad), rho = 1, p = p_in for r < r0 + r_in (2D: also theta
    below the middle) and 1 beyond, radial velocity vr, polar velocity
    vtheta (y component) and azimuthal velocity swirl * r (z component).
    info["geo"]: per block the Grid arrays (spherical_rows); each block dict
    carries them flat for the oracle."""
    cfg = base_cfg(ndim=ndim, nguard=nguard, gamma=gamma, geometry="spherical", **kw)
    nbr, nbt = nb[0], (nb[1] if ndim > 1 else 1)
    dx = 1.0 / (nbr * n)
    blocks, arrays, geo = [], [], []
    nunk = SPARK_NPROP + cfg["nscalars"]
    thmid = th0 + 0.5 * nbt * n * dx
    for bj in range(nbt):
        for bi in range(nbr):
            b = dict(nx=n, ny=n if ndim > 1 else 1, nz=1, level=1, dx=dx, dy=dx if ndim > 1 else 1.0, dz=1.0)
            NZ, NY, NX = gc_shape(cfg, b)
            Z, Y, X = _centers(cfg, b, (r0 + bi * n * dx, th0 + bj * n * dx, 0.0), NZ, NY, NX)
            hot = X < r0 + r_in
            if ndim > 1:
                hot = hot & (Y < thmid)
            pres = np.where(hot, p_in, 1.0)
            dens = np.ones_like(pres)
            U = np.zeros((NZ, NY, NX, nunk))
            fill_thermo(U, dens, pres, np.full_like(pres, vr), np.full_like(pres, vtheta), swirl * X, gamma)
            thf = None
            if ndim > 1:
                _, tl, tr = _coords(n, nguard, bj, dx, th0)
                thf = np.append(tl, tr[-1])
            g = spherical_rows(n, nguard, bi, dx, r0, thf)
            blocks.append(dict(b, geo=geo_flat(g), geometry=2))
            arrays.append(U)
            geo.append(g)
    return Problem("spherical_blast", cfg, blocks, arrays, ug=(nbr, nbt, 1), info=dict(geo=geo, r0=r0))


def cylindrical_blast(nbx=4, nby=4, n=16, nguard=10, r0=1.0, p_in=10.0, radius=0.15, swirl=0.1, vr=0.0,
                      gamma=1.4, **kw) -> Problem:
    """2D cylindrical (r, z) blast: r in [r0, r0 + 1], z in [0, nby/nbx],
    rho = 1, p = 1 with p = p_in inside a circle of the given radius at the
    domain centre (in the r-z plane), radial velocity vr, azimuthal velocity
    swirl * r (the z velocity component).  info["geo"]: per block the
    geometry rows (cylindrical_rows) for spark_set_geometry and the
    oracle."""
    cfg = base_cfg(ndim=2, nguard=nguard, gamma=gamma, geometry="cylindrical", **kw)
    dx = 1.0 / (nbx * n)
    blocks, arrays, geo = [], [], []
    nunk = SPARK_NPROP + cfg["nscalars"]
    rmid, zmid = r0 + 0.5, 0.5 * nby * n * dx
    for bj in range(nby):
        for bi in range(nbx):
            b = dict(nx=n, ny=n, nz=1, level=1, dx=dx, dy=dx, dz=1.0)
            NZ, NY, NX = gc_shape(cfg, b)
            Z, Y, X = _centers(cfg, b, (r0 + bi * n * dx, bj * n * dx, 0.0), NZ, NY, NX)
            inside = (X - rmid) ** 2 + (Y - zmid) ** 2 < radius * radius
            pres = np.where(inside, p_in, 1.0)
            dens = np.ones_like(pres)
            U = np.zeros((NZ, NY, NX, nunk))
            fill_thermo(U, dens, pres, np.full_like(pres, vr), np.zeros_like(pres), swirl * X, gamma)
            g = cylindrical_rows(n, nguard, bi, nbx, dx, dx, r0, ny=NY)
            b = dict(b, geo=geo_flat(g), geometry=1)
            blocks.append(b)
            arrays.append(U)
            geo.append(g)
    return Problem("cylindrical_blast", cfg, blocks, arrays, ug=(nbx, nby, 1), info=dict(geo=geo, r0=r0))


def sedov3d(nb=8, n=16, nguard=6, gamma=1.4, E=1.0, r_cells=3.5, p_amb=1e-5, rho=1.0, center=(0.5, 0.5, 0.5),
            **kw) -> Problem:
    """3D Sedov blast on [0,1]^3: E deposited as pressure in the cells whose
    centres lie within r_cells cell widths of `center` (the domain centre by
    default; p scaled so the deposited energy is exactly E)."""
    cfg = base_cfg(ndim=3, nguard=nguard, gamma=gamma, **kw)
    N = nb * n
    dx = 1.0 / N
    r0 = r_cells * dx
    idx = [(np.arange(N) + 0.5) * dx - center[d] for d in range(3)]
    r2 = idx[2][:, None, None] ** 2 + idx[1][None, :, None] ** 2 + idx[0][None, None, :] ** 2
    n_in = int(np.count_nonzero(r2 < r0 * r0))
    p_in = (gamma - 1.0) * E / (n_in * dx ** 3) if n_in else p_amb
    blocks, arrays = [], []
    nunk = SPARK_NPROP + cfg["nscalars"]
    for bk in range(nb):
        for bj in range(nb):
            for bi in range(nb):
                b = dict(nx=n, ny=n, nz=n, level=1, dx=dx, dy=dx, dz=dx)
                NZ, NY, NX = gc_shape(cfg, b)
                Z, Y, X = _centers(cfg, b, (bi * n * dx, bj * n * dx, bk * n * dx), NZ, NY, NX)
                rr = (X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2
                pres = np.where(rr < r0 * r0, p_in, p_amb)
                dens = np.full_like(pres, rho)
                U = np.zeros((NZ, NY, NX, nunk))
                zero = np.zeros_like(pres)
                fill_thermo(U, dens, pres, zero, zero, zero, gamma)
                blocks.append(b)
                arrays.append(U)
    return Problem("sedov3d", cfg, blocks, arrays, ug=(nb, nb, nb), info=dict(p_in=p_in, n_in=n_in, E=E))


class _Fourier:
    """Smooth random field s(x) in [0, 1]: 0.5 + 0.5 sum a_k cos(2 pi k.x + phi_k) / sum |a_k|."""

    def __init__(self, rng, nmodes=6, kmax=3):
        self.k = rng.integers(-kmax, kmax + 1, size=(nmodes, 3)).astype(np.float64)
        self.k[np.all(self.k == 0, axis=1)] = 1.0
        self.a = rng.normal(size=nmodes)
        self.phi = rng.uniform(0, 2 * np.pi, size=nmodes)
        self.norm = np.sum(np.abs(self.a))

    def __call__(self, X, Y, Z):
        """X, Y, Z: (NZ, NY, NX) cell-centre grids of one block.  Each mode is
        evaluated as Re(exp(i phi) ex(x) ey(y) ez(z)) from 1D factors."""
        x, y, z = X[0, 0, :], Y[0, :, 0], Z[:, 0, 0]
        s = np.zeros(X.shape)
        for k, a, ph in zip(self.k, self.a, self.phi):
            ex = np.exp(2j * np.pi * k[0] * x)
            ey = np.exp(2j * np.pi * k[1] * y)
            ez = (a * np.exp(1j * ph)) * np.exp(2j * np.pi * k[2] * z)
            s += ((ez[:, None] * ey[None, :])[:, :, None] * ex[None, None, :]).real
        return 0.5 + 0.5 * s / self.norm


def batched(nblocks=4096, nguard=10, gamma=1.4, seed_geom=20261202, seed_field=20261203, sizes=(8, 16, 32),
            probs=(0.25, 0.625, 0.125), levels=(0, 4), shock_every=16, nscalars=0, **kw) -> Problem:
    """Independent 3D blocks of uneven size and level (frozen guards)."""
    cfg = base_cfg(ndim=3, nguard=nguard, gamma=gamma, nscalars=nscalars, **kw)
    rg = np.random.Generator(np.random.PCG64(seed_geom))
    rf = np.random.Generator(np.random.PCG64(seed_field))
    size_idx = rg.choice(len(sizes), size=nblocks, p=np.asarray(probs) / np.sum(probs))
    lev = rg.integers(levels[0], levels[1] + 1, size=nblocks)
    origins = rg.uniform(0.0, 1.0, size=(nblocks, 3))
    f_rho, f_p = _Fourier(rf), _Fourier(rf)
    f_u = [_Fourier(rf) for _ in range(3)]
    u_scale = 0.3 / np.sqrt(1.0 / 8.0)  # cos-sum std ~ 1/sqrt(8) after normalisation
    blocks, arrays = [], []
    nunk = SPARK_NPROP + nscalars
    nshock = 0
    for i in range(nblocks):
        n = int(sizes[size_idx[i]])
        dx = 2.0 ** (-int(lev[i])) / 16.0
        b = dict(nx=n, ny=n, nz=n, level=int(lev[i]) + 1, dx=dx, dy=dx, dz=dx)
        NZ, NY, NX = gc_shape(cfg, b)
        Z, Y, X = _centers(cfg, b, origins[i], NZ, NY, NX)
        s = f_rho(X, Y, Z)
        dens = 0.5 * 4.0 ** s
        pres = 0.5 + 1.5 * f_p(X, Y, Z)
        vel = [u_scale * (f(X, Y, Z) - 0.5) for f in f_u]
        if shock_every and i % shock_every == 0:
            nshock += 1
            nv = rf.normal(size=3)
            nv /= np.linalg.norm(nv)
            M = rf.uniform(2.0, 5.0)
            c = origins[i] + (0.5 + rf.uniform(-0.25, 0.25)) * n * dx
            post = (nv[0] * (X - c[0]) + nv[1] * (Y - c[1]) + nv[2] * (Z - c[2])) < 0.0
            cs = np.sqrt(gamma * pres / dens)
            r21 = (gamma + 1) * M * M / ((gamma - 1) * M * M + 2)
            p21 = (2 * gamma * M * M - (gamma - 1)) / (gamma + 1)
            up = M * cs * (1.0 - 1.0 / r21)
            vel = [np.where(post, v + nv[d] * up, v) for d, v in enumerate(vel)]
            dens = np.where(post, dens * r21, dens)
            pres = np.where(post, pres * p21, pres)
        scal = None
        if nscalars:
            scal = [0.5 * (1 + np.tanh((X - 0.5) / (0.05 * (q + 1)))) for q in range(nscalars)]
        U = np.empty((NZ, NY, NX, nunk))
        fill_thermo(U, dens, pres, vel[0], vel[1], vel[2], gamma, scalars=scal)
        blocks.append(b)
        arrays.append(U)
    return Problem("batched", cfg, blocks, arrays, info=dict(seeds=(seed_geom, seed_field), nshock=nshock,
                                                             size_counts=np.bincount(size_idx, minlength=len(sizes))))


def amr_leaves(nb0, n, lmax, refine, nguard):
    """Leaves of a block hierarchy: level-1 grid nb0 (blocks per dimension),
    blocks of n^3 cells, block (level L, index b) refined while L < lmax and
    refine(L, lo, hi) is true for its physical box (dx at level 1 is
    1 / (nb0[0] * n)).  Coarse leaves that lie within a finer leaf's guard
    reach (nguard cells) by more than one level are refined further, so
    every leaf under a block's guard cells is within one level of it.
    Returns a sorted list of (level, (bx, by, bz))."""
    dx0 = 1.0 / (nb0[0] * n)

    def box(L, b):
        w = n * dx0 / 2 ** (L - 1)
        return [b[d] * w for d in range(3)], [(b[d] + 1) * w for d in range(3)]

    leaves = set()
    todo = [(1, (i, j, k)) for k in range(nb0[2]) for j in range(nb0[1]) for i in range(nb0[0])]
    while todo:
        L, b = todo.pop()
        lo, hi = box(L, b)
        if L < lmax and refine(L, lo, hi):
            todo += [(L + 1, (2 * b[0] + i, 2 * b[1] + j, 2 * b[2] + k)) for k in (0, 1) for j in (0, 1) for i in (0, 1)]
        else:
            leaves.add((L, b))
    # balance: no leaf under a block's guard reach more than one level coarser
    while True:
        md = [nb0[d] << (lmax - 1) for d in range(3)]
        slot = np.full(md[::-1], -1, dtype=np.int64)
        lv = {}
        for idx, (L, b) in enumerate(sorted(leaves)):
            sh = lmax - L
            slot[b[2] << sh:(b[2] + 1) << sh, b[1] << sh:(b[1] + 1) << sh, b[0] << sh:(b[0] + 1) << sh] = L
        split = set()
        for (L, b) in leaves:
            sh = lmax - L
            lo = [max(0, ((b[d] * n - nguard) << sh) // n) for d in range(3)]
            hi = [min(md[d], ((((b[d] + 1) * n + nguard) << sh) - 1) // n + 1) for d in range(3)]
            region = slot[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
            if region.min() < L - 1:
                for (L2, b2) in leaves:
                    if L2 < L - 1:
                        sh2 = lmax - L2
                        if all(lo[d] < (b2[d] + 1) << sh2 and (b2[d] << sh2) < hi[d] for d in range(3)):
                            split.add((L2, b2))
        if not split:
            break
        for (L, b) in split:
            leaves.discard((L, b))
            leaves.update((L + 1, (2 * b[0] + i, 2 * b[1] + j, 2 * b[2] + k))
                          for k in (0, 1) for j in (0, 1) for i in (0, 1))
    return sorted(leaves)


def rt_amr(nb0=(4, 4, 8), n=16, lmax=3, slabs=((0.25, 1.75), (0.5, 1.5)), nguard=10, nscalars=8, gamma=1.4,
           seed=20261204, width=0.0, p_boost=None, **kw) -> Problem:
    """3D Rayleigh-Taylor on a block hierarchy: heavy fluid (rho = 2) over
    light (rho = 1) across z0 = the middle of the domain, constant gravity
    g = -1 along z as a source term, hydrostatic pressure with p = 2.5 at the
    interface, interface z0 + 0.01 * (cos 2pi x cos 2pi y + seeded random
    modes, normalised), nscalars mass scalars X_k = (1 + tanh((z - zi) /
    w_k)) / 2 with w_k = 0.01 (k + 1).  Level-L blocks are refined where
    their z range meets slabs[L-1].  x, y span [0, 1]; dx = 1 / (nb0[0] n)
    at level 1.  x and y are periodic, z is reflecting.  CFL 0.4: the CFL
    step takes the largest (|u_d| + c) / dx_d over directions, so for a
    flow at rest in 3D the unsplit RK3 / WENO5 update sees about three times
    that Courant number and 0.8 lets a grid-scale odd-even mode grow.
    p_boost = (x_lo, x_hi, f) multiplies the pressure by f for x_lo <= x <
    x_hi (shocks run out of both planes).  Returns a Problem with info nb0 /
    origins / bc for spark_set_amr_grid."""
    cfg = base_cfg(**{**dict(ndim=3, nguard=nguard, gamma=gamma, nscalars=nscalars, use_gravity=1,
                             grav=(0.0, 0.0, -1.0), flux_correct=1, cfl=0.4), **kw})
    dx0 = 1.0 / (nb0[0] * n)
    zlen = nb0[2] * n * dx0
    z0 = 0.5 * zlen

    def refine(L, lo, hi):
        a, b = slabs[L - 1]
        return lo[2] < b and hi[2] > a

    leaves = amr_leaves(nb0, n, lmax, refine, nguard)
    rng = np.random.Generator(np.random.PCG64(seed))
    nm = 4
    kx = rng.integers(1, 4, size=nm)
    ky = rng.integers(1, 4, size=nm)
    amp = rng.normal(size=nm)
    ph = rng.uniform(0, 2 * np.pi, size=nm)

    def zi(X, Y):
        r = sum(a * np.cos(2 * np.pi * (i * X + j * Y) + p) for i, j, a, p in zip(kx, ky, amp, ph))
        pert = np.cos(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.5 * r / np.sum(np.abs(amp))
        return z0 + 0.01 * pert / 1.5

    blocks, arrays, origins = [], [], []
    nunk = SPARK_NPROP + nscalars
    for (L, b) in leaves:
        dx = dx0 / 2 ** (L - 1)
        blk = dict(nx=n, ny=n, nz=n, level=L, dx=dx, dy=dx, dz=dx)
        NZ, NY, NX = gc_shape(cfg, blk)
        Z, Y, X = _centers(cfg, blk, tuple(b[d] * n * dx for d in range(3)), NZ, NY, NX)
        zint = zi(X, Y)
        if width > 0.0:
            # rho = 1.5 + 0.5 tanh(s / w), s = z - zi; p = 2.5 - int_0^s rho
            sz = (Z - zint) / width
            dens = 1.5 + 0.5 * np.tanh(sz)
            pres = 2.5 - (1.5 * (Z - zint) + 0.5 * width * (np.logaddexp(sz, -sz) - np.log(2.0)))
        else:
            heavy = Z > zint
            dens = np.where(heavy, 2.0, 1.0)
            pres = np.where(heavy, 2.5 - 2.0 * (Z - zint), 2.5 + 1.0 * (zint - Z))
        if p_boost is not None:
            pres = np.where((X >= p_boost[0]) & (X < p_boost[1]), p_boost[2] * pres, pres)
        zero = np.zeros_like(dens)
        scal = [0.5 * (1.0 + np.tanh((Z - zint) / (0.01 * (q + 1)))) for q in range(nscalars)] if nscalars else None
        U = np.empty((NZ, NY, NX, nunk))
        fill_thermo(U, dens, pres, zero, zero, zero, gamma, scalars=scal)
        blocks.append(blk)
        arrays.append(U)
        origins.append([b[d] * n for d in range(3)])
    counts = np.bincount([L for L, _ in leaves], minlength=lmax + 1)[1:]
    return Problem("rt_amr", cfg, blocks, arrays,
                   info=dict(nb0=tuple(nb0), origins=np.asarray(origins, dtype=np.int32), leaves_per_level=counts,
                             bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECT), seed=seed))
