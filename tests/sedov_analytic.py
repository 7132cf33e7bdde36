"""Radial comparison of a 3D Sedov run with the analytic profile the
reference ships (tests/golden/sedov_analytic_gamma1.4.txt: gamma 1.4, E = 1,
rho_up = 1, t = 1).  Self-similar scaling: with S = (rho_up / (E t^2))^(1/5)
the profile at time t is rho(r) = rho_1(r S) (rho_up = 1), shock radius
r_0 / S (the header of the reference file)."""
from pathlib import Path

import numpy as np

_TAB = np.loadtxt(Path(__file__).with_name("golden") / "sedov_analytic_gamma1.4.txt")
R0 = 1.03277747


def analytic_rho(r, t, E=1.0, rho_up=1.0):
    S = (rho_up / (E * t * t)) ** 0.2
    x = np.asarray(r) * S
    return np.where(x <= R0, np.interp(x, _TAB[:, 0], _TAB[:, 2]) * rho_up, rho_up), R0 / S


def radial_check(prob, arrays, t):
    """Bins the interior densities of a uniform-grid Sedov problem on [0,1]^3
    in shells of one cell width about the centre.  Returns the analytic shock
    radius, the numerical front (outermost shell above half the numerical
    peak over ambient) and the L1 error of the shell means inside the shock,
    relative to the analytic shell means."""
    g = prob.cfg["nguard"]
    nb = prob.ug[0]
    n = prob.blocks[0]["nx"]
    dx = prob.blocks[0]["dx"]
    r_all, rho_all = [], []
    c = (np.arange(n) + 0.5) * dx
    for i, a in enumerate(arrays):
        bx, by, bz = i % nb, (i // nb) % nb, i // (nb * nb)
        Z, Y, X = np.meshgrid(c + bz * n * dx, c + by * n * dx, c + bx * n * dx, indexing="ij")
        r_all.append(np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2).ravel())
        rho_all.append(a[g:-g, g:-g, g:-g, 0].ravel())
    r = np.concatenate(r_all)
    rho = np.concatenate(rho_all)
    ra, R = analytic_rho(r, t)
    shell = (r / dx).astype(int)
    keep = shell < int(1.4 * R / dx) + 1
    cnt = np.bincount(shell[keep])
    ok = cnt > 0
    rb = (np.bincount(shell[keep], r[keep])[ok] / cnt[ok])
    nb_ = (np.bincount(shell[keep], rho[keep])[ok] / cnt[ok])
    ab = (np.bincount(shell[keep], ra[keep])[ok] / cnt[ok])
    thr = 1.0 + 0.5 * (nb_.max() - 1.0)
    front = rb[np.where(nb_ > thr)[0].max()]
    inside = rb <= R
    l1 = float(np.sum(np.abs(nb_ - ab)[inside]) / np.sum(ab[inside]))
    return R, float(front), l1, dx
