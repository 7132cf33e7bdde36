"""Closed-form pins of three parts of the CPU restatement that the whole-
solver checks leave open: the artificial viscosity (Spark/Hydro.ini:285-301),
the mass-scalar fluxes (Spark/Hydro.ini:238-247) and the gravity source of
the update (Spark/hy_rk_updateSoln.F90-mc:179-190, applied at :257-265).

Each check builds a case whose answer follows from the mathematics alone:

* viscosity: first-order (FOG) states in a supersonic pencil make HLLC the
  upwind flux F(V_{i-1}), so one stage's face flux must be
  F(V_{i-1}) + cvisc max(-(u_i - u_{i-1}), 0) (U_{i-1} - U_i), with U the
  conserved cell vector; an expanding pencil adds nothing;
* scalars: in a uniform supersonic flow the mass equation is the linear
  upwind operator, and rho X_s must follow the same RK Taylor polynomial as
  rho, upwinded from the left for u > 0 and from the right for u < 0;
* gravity: a uniform state under constant g has zero flux differences, so
  RK3 integrates m' = rho g, E' = m . g exactly: one step gives
  d(rho u) = rho g dt and dE = m0 . g dt + rho |g|^2 dt^2 / 2, with the
  internal energy unchanged.

The checks are plain functions so that tests/test_oracle_mutants.py can run
them against builds of the restatement with a mistake planted in each part.
"""
import numpy as np

import pyoracle
from sparkhip import problems
from sparkhip.capi import VI

G = 1.4


def _pencil(n, ng, fields, nscalars=0, **kw):
    """One 1D block of n cells over [0, 1] with ng guard cells; fields(x)
    gives (rho, u, v, w, p) and, with scalars, a list of mass fractions."""
    cfg = problems.base_cfg(ndim=1, nguard=ng, nscalars=nscalars, **kw)
    dx = 1.0 / n
    x = (np.arange(n + 2 * ng) - ng + 0.5) * dx
    f = fields(x)
    d, u, v, w, p = f[:5]
    U = np.zeros((1, 1, n + 2 * ng, 11 + nscalars))
    problems.fill_thermo(U, d, p, u, v, w, cfg["gamma"], scalars=f[5] if nscalars else None)
    blk = dict(nx=n, ny=1, nz=1, level=1, dx=dx, dy=1.0, dz=1.0)
    return cfg, blk, U


def _euler_flux(rho, u, v, w, p, rhoe):
    """x-direction Euler flux of a primitive state with internal energy
    density rhoe."""
    E = 0.5 * rho * (u * u + v * v + w * w) + rhoe
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, (E + p) * u])


def _one_stage_flux(cfg, blk, U):
    """x face fluxes of the interior faces from stage 1 of RK2 (weight 1/2
    in the register, an exact scaling)."""
    n, nf = blk["nx"], 5 + cfg["nscalars"]
    S = U.copy()
    fb = [np.zeros((n + 1) * nf), np.zeros(n * nf), np.zeros(n * nf)]
    pyoracle.stage(cfg, blk, U.copy(), S, 1, 0, 1e-4, fb=fb)
    return 2.0 * fb[0].reshape(n + 1, nf)


def check_viscosity(compressive):
    """Supersonic pencil (u between 4 and 6, c < 1.6), FOG states."""
    n, ng = 16, 4
    slope = -2.0 if compressive else 2.0
    fields = lambda x: (1.0 + 0.3 * np.sin(2 * np.pi * x), 5.0 + slope * (x - 0.5), 0.3 * np.cos(3 * x),
                        -0.2 + 0.1 * x, 1.0 + 0.2 * x)
    cfg, blk, U = _pencil(n, ng, fields, recon="fog", nstencil=2, rk_order=2, use_flattening=0)
    F = _one_stage_flux(cfg, blk, U)
    c = U[0, 0]
    rho, u, v, w, p = (c[:, VI[k]] for k in ("dens", "velx", "vely", "velz", "pres"))
    rhoe = rho * c[:, VI["eint"]]
    cons = np.array([rho, rho * u, rho * v, rho * w, 0.5 * rho * (u * u + v * v + w * w) + rhoe])
    lo = ng + np.arange(n + 1) - 1  # the cell on the low side of face i
    hi = lo + 1
    expect = _euler_flux(rho[lo], u[lo], v[lo], w[lo], p[lo], rhoe[lo])
    cv = cfg["cvisc"] * np.maximum(-(u[hi] - u[lo]), 0.0)
    expect = expect + cv * (cons[:, lo] - cons[:, hi])
    if compressive:
        assert np.all(cv > 0.01)
    else:
        assert np.all(cv == 0.0)
    np.testing.assert_allclose(F.T, expect, rtol=1e-13, atol=1e-14)


def check_scalar_upwinding(u0):
    """Uniform supersonic flow u0 (|u0| = 5, p = 1) carrying rho and two
    scalars; FOG + HLLC gives the upwind flux, so after one RK3 step
    rho X_s equals sum_k (dt L)^k / k! applied to rho X_s at the start."""
    n, ng = 32, 6
    X0 = lambda x: 0.5 + 0.3 * np.sin(2 * np.pi * x + 1.0)
    X1 = lambda x: 0.5 * (1.0 + np.tanh((x - 0.5) / 0.1))
    fields = lambda x: (1.0 + 0.2 * np.sin(2 * np.pi * x), np.full_like(x, u0), np.zeros_like(x),
                        np.zeros_like(x), np.ones_like(x), [X0(x), X1(x)])
    cfg, blk, U = _pencil(n, ng, fields, nscalars=2, recon="fog", nstencil=2, rk_order=3)
    rho0 = U[0, 0, :, VI["dens"]].copy()
    m0 = [rho0 * U[0, 0, :, 11 + s] for s in range(2)]
    dt, _ = pyoracle.compute_dt(cfg, blk, U)
    pyoracle.hydro_block(cfg, blk, U, dt)
    c = u0 * dt / blk["dx"]
    if u0 > 0:
        D = lambda r: r - np.concatenate(([np.nan], r[:-1]))       # r_i - r_{i-1}
        L = lambda r: -c * D(r)
    else:
        D = lambda r: np.concatenate((r[1:], [np.nan])) - r        # r_{i+1} - r_i
        L = lambda r: -c * D(r)
    taylor = lambda r: r + L(r) + L(L(r)) / 2 + L(L(L(r))) / 6
    rho = U[0, 0, :, VI["dens"]]
    np.testing.assert_allclose(rho[ng:-ng], taylor(rho0)[ng:-ng], rtol=1e-13)
    for s in range(2):
        got = rho * U[0, 0, :, 11 + s]
        np.testing.assert_allclose(got[ng:-ng], taylor(m0[s])[ng:-ng], rtol=1e-12)


def check_gravity(ndim, grav, vel):
    """Uniform state (rho 1.3, p 0.9, velocity vel) under constant grav,
    one RK3 step of dt = 0.01 (WENO5 + HLLC, flattening on)."""
    n, ng = 8, 10
    rho0, p0, dt = 1.3, 0.9, 0.01
    cfg = problems.base_cfg(ndim=ndim, nguard=ng, use_gravity=1, grav=tuple(grav))
    blk = dict(nx=n, ny=n if ndim > 1 else 1, nz=n if ndim > 2 else 1, level=1, dx=1.0 / n, dy=1.0 / n,
               dz=1.0 / n)
    U = np.zeros(problems.gc_shape(cfg, blk) + (11,))
    one = np.ones(U.shape[:-1])
    problems.fill_thermo(U, rho0 * one, p0 * one, vel[0] * one, vel[1] * one, vel[2] * one, G)
    E0 = rho0 * U[..., VI["ener"]].copy()
    e0 = U[..., VI["eint"]].copy()
    pyoracle.hydro_block(cfg, blk, U, dt)
    sl = tuple(slice(ng, -ng) if d < ndim else slice(None) for d in range(3))[::-1]
    I = U[sl]
    np.testing.assert_array_equal(I[..., VI["dens"]], rho0)
    for d, key in enumerate(("velx", "vely", "velz")):
        np.testing.assert_allclose(rho0 * I[..., VI[key]], rho0 * vel[d] + rho0 * grav[d] * dt, rtol=1e-13,
                                   atol=1e-15)
    dE = sum(rho0 * vel[d] * grav[d] * dt + 0.5 * rho0 * grav[d] ** 2 * dt * dt for d in range(3))
    got = rho0 * I[..., VI["ener"]] - E0[sl]
    np.testing.assert_allclose(got, dE, rtol=1e-9)
    np.testing.assert_allclose(I[..., VI["eint"]], e0[sl], rtol=1e-13)
    np.testing.assert_allclose(I[..., VI["pres"]], p0, rtol=1e-13)


def test_viscosity_compressive_supersonic_pencil():
    check_viscosity(compressive=True)


def test_viscosity_adds_nothing_on_expansion():
    check_viscosity(compressive=False)


def test_scalars_upwinded_rightward_flow():
    check_scalar_upwinding(5.0)


def test_scalars_upwinded_leftward_flow():
    check_scalar_upwinding(-5.0)


def test_gravity_1d_moving_state():
    check_gravity(1, (-1.0, 0.5, 0.25), (0.3, -0.2, 0.1))


def test_gravity_3d_rayleigh_taylor_direction():
    """g = -1 along z, the Rayleigh-Taylor setting, on a state moving up."""
    check_gravity(3, (0.0, 0.0, -1.0), (0.1, -0.05, 0.4))
