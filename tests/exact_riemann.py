"""Exact solution of the 1D Riemann problem of the Euler equations (ideal gas), after Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4: the star pressure by Newton's
method on f_L(p) + f_R(p) + u_R - u_L = 0, then the wave pattern sampled at x/t.  Independent
of the library and the oracle; used to check shock capturing."""
import numpy as np
from scipy.optimize import brentq


def _f(p, rho, pk, g):
    a = np.sqrt(g * pk / rho)
    if p > pk:
        A = 2 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        return (p - pk) * np.sqrt(A / (p + B))
    return 2 * a / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)


def sample(left, right, xi, g=1.4):
    """(rho, u, p) at the similarity coordinates xi = (x - x0)/t; left/right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps = brentq(lambda p: _f(p, rl, pl, g) + _f(p, rr, pr, g) + ur - ul, 1e-12, 100 * max(pl, pr), xtol=1e-15)
    us = 0.5 * (ul + ur) + 0.5 * (_f(ps, rr, pr, g) - _f(ps, rl, pl, g))
    xi = np.asarray(xi, dtype=np.float64)
    rho, u, p = np.zeros_like(xi), np.zeros_like(xi), np.zeros_like(xi)
    for k, s in enumerate(xi.flat):
        if s <= us:   # left of the contact
            if ps > pl:   # left shock
                rs = rl * (ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)
                S = ul - al * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                state = (rl, ul, pl) if s < S else (rs, us, ps)
            else:         # left rarefaction
                rs = rl * (ps / pl) ** (1 / g)
                head, tail = ul - al, us - al * (ps / pl) ** ((g - 1) / (2 * g))
                if s < head:
                    state = (rl, ul, pl)
                elif s > tail:
                    state = (rs, us, ps)
                else:
                    c = 2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)
                    state = (rl * c ** (2 / (g - 1)), 2 / (g + 1) * (al + (g - 1) / 2 * ul + s), pl * c ** (2 * g / (g - 1)))
        else:         # right of the contact
            if ps > pr:   # right shock
                rs = rr * (ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)
                S = ur + ar * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                state = (rr, ur, pr) if s > S else (rs, us, ps)
            else:         # right rarefaction
                rs = rr * (ps / pr) ** (1 / g)
                head, tail = ur + ar, us + ar * (ps / pr) ** ((g - 1) / (2 * g))
                if s > head:
                    state = (rr, ur, pr)
                elif s < tail:
                    state = (rs, us, ps)
                else:
                    c = 2 / (g + 1) - (g - 1) / ((g + 1) * ar) * (ur - s)
                    state = (rr * c ** (2 / (g - 1)), 2 / (g + 1) * (-ar + (g - 1) / 2 * ur + s), pr * c ** (2 * g / (g - 1)))
        rho.flat[k], u.flat[k], p.flat[k] = state
    return rho, u, p, ps, us
