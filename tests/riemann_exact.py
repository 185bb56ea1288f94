"""Exact solution of the 1-D Riemann problem of the Euler equations for an
ideal gas (E. F. Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", 3rd ed., chapter 4: the pressure function f(p) = f_L + f_R + du,
solved by Newton's method from the two-rarefaction guess, then sampled
along x/t).  Used as a closed-form check of the 1-D shock-tube runs."""
import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:   # shock
        A, B = 2.0 / ((g + 1.0) * rho), (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + B))
    # rarefaction
    r = (p / pk) ** ((g - 1.0) / (2.0 * g))
    return 2.0 * ck / (g - 1.0) * (r - 1.0), 1.0 / (rho * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))


def star_state(left, right, g=1.4):
    """(p*, u*) of left, right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    z = (g - 1.0) / (2.0 * g)
    p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(p - dp, 1e-14)
        if abs(pn - p) <= 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(left, right, s, g=1.4):
    """(rho, u, p) at the similarity coordinates s = (x - x0) / t."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps, us = star_state(left, right, g)
    gm = (g - 1.0) / (g + 1.0)
    out = np.zeros((3, len(s)))
    for i, si in enumerate(np.asarray(s, float)):
        if si <= us:   # left of the contact
            if ps > pl:
                sl = ul - cl * np.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                st = (rl * ((ps / pl + gm) / (gm * ps / pl + 1.0)), us, ps) if si >= sl else (rl, ul, pl)
            else:
                csl = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                if si <= ul - cl:
                    st = (rl, ul, pl)
                elif si >= us - csl:
                    st = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - si))
                    u = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + si)
                    st = (rl * (c / cl) ** (2.0 / (g - 1.0)), u, pl * (c / cl) ** (2.0 * g / (g - 1.0)))
        else:
            if ps > pr:
                sr = ur + cr * np.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                st = (rr * ((ps / pr + gm) / (gm * ps / pr + 1.0)), us, ps) if si <= sr else (rr, ur, pr)
            else:
                csr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                if si >= ur + cr:
                    st = (rr, ur, pr)
                elif si <= us + csr:
                    st = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - si))
                    u = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + si)
                    st = (rr * (c / cr) ** (2.0 / (g - 1.0)), u, pr * (c / cr) ** (2.0 * g / (g - 1.0)))
        out[:, i] = st
    return out
