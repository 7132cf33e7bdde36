"""Exact solution of the 1D Riemann problem for an ideal gas.

Iterative pressure solve of Toro, "Riemann Solvers and Numerical Methods
for Fluid Dynamics", ch. 4 (Newton iteration on f(p) = fL + fR + du, then
sampling of the self-similar solution).  Used to pin the CPU restatement's
whole-solver behaviour on the Sod problem.
"""
import numpy as np


def _fk(p, rk, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        f = (p - pk) * np.sqrt(A / (p + B))
        df = np.sqrt(A / (B + p)) * (1.0 - 0.5 * (p - pk) / (B + p))
    else:  # rarefaction
        f = 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rk * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rL, uL, pL, rR, uR, pR, g):
    cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    p = max(1e-8, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (cL + cR))
    for _ in range(100):
        fL, dfL = _fk(p, rL, pL, cL, g)
        fR, dfR = _fk(p, rR, pR, cR, g)
        dp = (fL + fR + uR - uL) / (dfL + dfR)
        pn = max(1e-12, p - dp)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _fk(p, rL, pL, cL, g)
    fR, _ = _fk(p, rR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(x, t, x0, rL, uL, pL, rR, uR, pR, g):
    """Density, velocity, pressure at positions x at time t (interface x0)."""
    ps, us = star_state(rL, uL, pL, rR, uR, pR, g)
    cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    out = np.zeros((3, len(x)))
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pL:  # left shock
                sL = uL - cL * np.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if s <= sL:
                    r, u, p = rL, uL, pL
                else:
                    r = rL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
                    u, p = us, ps
            else:  # left rarefaction
                shl = uL - cL
                csl = cL * (ps / pL) ** ((g - 1) / (2 * g))
                stl = us - csl
                if s <= shl:
                    r, u, p = rL, uL, pL
                elif s >= stl:
                    r, u, p = rL * (ps / pL) ** (1 / g), us, ps
                else:
                    u = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + s)
                    c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - s))
                    r = rL * (c / cL) ** (2 / (g - 1))
                    p = pL * (c / cL) ** (2 * g / (g - 1))
        else:  # right of contact
            if ps > pR:  # right shock
                sR = uR + cR * np.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if s >= sR:
                    r, u, p = rR, uR, pR
                else:
                    r = rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
                    u, p = us, ps
            else:  # right rarefaction
                shr = uR + cR
                csr = cR * (ps / pR) ** ((g - 1) / (2 * g))
                str_ = us + csr
                if s >= shr:
                    r, u, p = rR, uR, pR
                elif s <= str_:
                    r, u, p = rR * (ps / pR) ** (1 / g), us, ps
                else:
                    u = 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + s)
                    c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - s))
                    r = rR * (c / cR) ** (2 / (g - 1))
                    p = pR * (c / cR) ** (2 * g / (g - 1))
        out[:, n] = (r, u, p)
    return out
