"""Pins of the oracle's axisymmetric mode (AXISYMMETRIC build: the state is
r U with r = y; DGSEMOperator.cpp:161-424 recovery, r-weighted fluxes at
DGSEMIntegrator.cpp:190-205,290,337,1030,1110 and BdrFaceIntegrator.cpp:109,
the pressure source of DGSEMIntegrator.cpp:411-426, ZeroAxisRadialMom).

- Recovery: division by r off the axis; on the axis the ray fit through the
  origin, in closed form for a state linear in r, its minmod limiter (a
  sign change gives 0) and the density-floor fallback to the nearest
  off-axis node.
- Exact steady states of the axisymmetric Euler equations: uniform axial
  flow past the axis, and fluid at rest at uniform pressure with any smooth
  density, give a zero RHS (rectangles at every order; distorted elements
  from N = 2, where the products of r and the metric terms along a line are
  quadratic and the derivative matrix is exact).  A missing source, a
  missing or wrong r weighting, or a source of the wrong sign leaves O(1).
- A closed domain (slip walls and the axis) conserves mass and energy: the
  r-weighted two-point and face fluxes stay conservative.
- Away from the axis the RHS converges to the axisymmetric Euler operator
  -d(rF)/dz - d(rG)/dr + (0, 0, p, 0) of a smooth state.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
import quad_meshes as Q

G = 1.4


def cons(rho, uz, ur, p):
    return np.array([rho, rho * uz, rho * ur, p / (G - 1) + 0.5 * rho * (uz * uz + ur * ur)])


U0 = cons(1.0, 2.0, 0.0, 1.0 / 1.4)     # the ramp case's free stream


def pipe(nz, nr, L=2.0, R=1.0, perturb=0.0, seed=0, top="slip", right="supersonic-outflow"):
    """[0, L] x [0, R] with the axis at r = 0; attributes 1 inlet, 2 outlet,
    3 axis, 4 top."""
    m = Q.grid(nz, nr, (0.0, 0.0), (L, R), periodic=(False, False), perturb=perturb, seed=seed)
    st = np.zeros((4, 4))
    st[0] = U0
    return dict(m, attr_bc=["supersonic-inflow", right, "axis", top], attr_state=st)


def oracle(mesh, order, **kw):
    kw.setdefault("quirks", False)
    return O.Oracle.quad(mesh, order, axisymmetric=True, **kw)


def weighted(orc, U):
    """r U at the nodes, from U [4] or [4, NE, K]."""
    r = orc.coords()[1]
    U = np.asarray(U, float)
    if U.ndim == 1:
        U = U[:, None, None] * np.ones_like(r)
    return (r[None] * U).reshape(-1), r


def test_recover_off_axis_is_division():
    m = Q.grid(3, 2, (0.0, 0.5), (1.0, 1.5), periodic=(False, False), perturb=0.2, seed=3)
    m = dict(m, attr_bc=["slip"] * 4, attr_state=np.zeros((4, 4)))
    orc = oracle(m, 3)
    X = orc.coords()
    U = np.stack([1 + 0.2 * X[0], 0.3 * X[1], -0.1 + X[0] * X[1], 2.5 + X[1] ** 2])
    ru, r = weighted(orc, U)
    got = orc.recover(ru).reshape(U.shape)
    assert np.allclose(got, U, rtol=2e-16 * 4, atol=0)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_recover_constant_state_on_axis(order):
    orc = oracle(pipe(3, 2, perturb=0.0), order)
    ru, r = weighted(orc, U0)
    got = orc.recover(ru).reshape(4, orc.NE, orc.K)
    ax = r == 0.0
    assert ax.sum() == 3 * (order + 1)
    for v in range(4):
        assert np.allclose(got[v], U0[v], rtol=1e-14, atol=1e-14)
    assert np.all(got[2][ax] == 0.0)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_recover_linear_in_r_closed_form(order):
    """rho = a + b r: the fit through the origin over the two nearest nodes
    at the axis node's z gives a + b (r1^3 + r2^3) / (r1^2 + r2^2); a
    momentum changing sign between them is limited to 0; a density whose
    fit falls below the floor takes the nearest off-axis node's state."""
    h = 0.5
    orc = oracle(pipe(2, 2, L=1.0, R=2 * h), order)
    x, _ = O.lgl(order)
    r1, r2 = h * x[1], h * x[2]
    X = orc.coords()
    r = X[1]
    a, b = 1.3, 0.4
    rm = 0.5 * (r1 + r2)
    U = np.stack([a + b * r, 0.7 * (r - rm), 0 * r, 3.0 + r])
    got = orc.recover(weighted(orc, U)[0]).reshape(4, orc.NE, orc.K)
    ax = r == 0.0
    assert np.allclose(got[0][ax], a + b * (r1 ** 3 + r2 ** 3) / (r1 ** 2 + r2 ** 2), rtol=1e-14)
    assert np.all(got[1][ax] == 0.0)
    assert np.allclose(got[3][ax], 3.0 + (r1 ** 3 + r2 ** 3) / (r1 ** 2 + r2 ** 2), rtol=1e-14)
    # below the floor: the state of a node of the first off-axis row
    U2 = np.stack([-1.0 + 0 * r, 0.5 + 0.1 * r, 0.2 + 0 * r, 3.0 + r])
    got = orc.recover(weighted(orc, U2)[0]).reshape(4, orc.NE, orc.K)
    for v in (0, 1, 3):
        assert np.allclose(got[v][ax], U2[v][np.isclose(r, r1)][0], rtol=1e-13)
    assert np.all(got[2][ax] == 0.0)


def rhs_scale(orc, ru):
    return orc.Mult(ru).reshape(4, orc.NE, orc.K)


@pytest.mark.parametrize("order,perturb,blending", [(1, 0.0, True), (2, 0.0, False), (3, 0.0, True),
                                                    (2, 0.25, False), (3, 0.25, True), (4, 0.25, False)])
def test_uniform_axial_flow_is_steady(order, perturb, blending):
    orc = oracle(pipe(4, 3, perturb=perturb, seed=order), order, blending=blending)
    ru, r = weighted(orc, U0)
    d = rhs_scale(orc, ru)
    p = (G - 1) * (U0[3] - 0.5 * U0[1] ** 2 / U0[0])
    print("max |RHS| / p", np.abs(d).max() / p)
    assert np.abs(d).max() < 1e-12 * p * 10


@pytest.mark.parametrize("order,perturb", [(1, 0.0), (2, 0.0), (2, 0.25), (3, 0.25)])
def test_fluid_at_rest_with_uniform_pressure_is_steady(order, perturb):
    """Zero velocity, p = 1, any smooth density: the radial pressure flux
    d(r p)/dr balances the source p."""
    m = pipe(3, 3, perturb=perturb, seed=5, top="slip", right="slip")
    m["attr_bc"][0] = "slip"
    orc = oracle(m, order)
    X = orc.coords()
    rho = 1.0 + 0.3 * np.sin(2.0 * X[0] + 1.0) * np.cos(1.5 * X[1])
    U = np.stack([rho, 0 * rho, 0 * rho, np.full_like(rho, 1.0 / (G - 1))])
    d = rhs_scale(orc, weighted(orc, U)[0])
    print("max |RHS|", np.abs(d).max())
    assert np.abs(d).max() < 1e-12


@pytest.mark.parametrize("blending", [False, True])
def test_closed_domain_conserves_mass_and_energy(blending):
    m = pipe(4, 3, perturb=0.2, seed=7, top="slip", right="slip")
    m["attr_bc"][0] = "slip"
    orc = oracle(m, 3, blending=blending)
    X = orc.coords()
    z, r = X
    U = np.stack([1.0 + 0.2 * np.sin(3 * z) * np.cos(2 * r), 0.4 * np.cos(z + r), 0.3 * r * np.sin(2 * z),
                  2.5 + 0.3 * np.cos(2 * z - r)])
    d = rhs_scale(orc, weighted(orc, U)[0])
    WJ = orc.weights()
    for v in (0, 3):
        tot, mag = np.sum(WJ * d[v]), np.sum(WJ * np.abs(d[v]))
        print(v, tot, mag)
        assert abs(tot) < 1e-13 * mag
    if blending:
        assert orc.alpha().max() > 0.0


def exact_rhs(z, r):
    """-d(rF)/dz - d(rG)/dr + (0, 0, p, 0) of the smooth state below, by
    fourth-order central differences of the analytic fluxes."""
    def fluxes(z, r):
        rho = 1.0 + 0.2 * np.sin(2 * math.pi * z) * np.cos(math.pi * (r - 0.5))
        uz = 0.5 + 0.2 * np.cos(2 * math.pi * z)
        ur = 0.15 * np.sin(math.pi * (r - 0.5)) * np.sin(2 * math.pi * z)
        p = 1.0 + 0.1 * np.cos(2 * math.pi * z + r)
        E = p / (G - 1) + 0.5 * rho * (uz * uz + ur * ur)
        F = np.stack([rho * uz, rho * uz * uz + p, rho * uz * ur, uz * (E + p)])
        Gf = np.stack([rho * ur, rho * uz * ur, rho * ur * ur + p, ur * (E + p)])
        return np.stack([rho, rho * uz, rho * ur, E]), F, Gf, p
    h = 1e-3
    d = lambda f, a: (-f(a + 2 * h) + 8 * f(a + h) - 8 * f(a - h) + f(a - 2 * h)) / (12 * h)
    dF = d(lambda zz: r * fluxes(zz, r)[1], z)
    dG = d(lambda rr: rr * fluxes(z, rr)[2], r)
    U, _, _, p = fluxes(z, r)
    out = -dF - dG
    out[2] += p
    return U, out


@pytest.mark.parametrize("order", [2, 3])
def test_smooth_rhs_converges_to_axisymmetric_euler(order):
    """z periodic on [0, 1], slip walls at r = 0.5 and 1.5 (u_r = 0 there):
    the max-norm error of the RHS falls at least at rate N - 1.2."""
    errs = []
    for n in (4, 8, 16):
        m = Q.grid(n, n, (0.0, 0.5), (1.0, 1.5), periodic=(True, False), perturb=0.15, seed=2)
        m = dict(m, attr_bc=["slip"] * 4, attr_state=np.zeros((4, 4)))
        orc = oracle(m, order)
        z, r = orc.coords()
        U, ex = exact_rhs(z, r)
        d = rhs_scale(orc, (r[None] * U).reshape(-1))
        errs.append(np.abs(d - ex).max() / np.abs(ex).max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("errors", errs, "rates", rates)
    assert min(rates) > order - 1.2
    assert errs[-1] < 5e-2


def test_axis_radial_rhs_zeroed_and_rejections():
    orc = oracle(pipe(3, 2), 3)
    X = orc.coords()
    z, r = X
    U = np.stack([1.0 + 0.1 * np.sin(3 * z + r), 2.0 + 0.1 * r, 0.2 * r * np.cos(z), 4.0 + 0.2 * np.cos(z)])
    d = rhs_scale(orc, weighted(orc, U)[0])
    ax = r == 0.0
    assert np.all(d[2][ax] == 0.0) and np.abs(d[2][~ax]).min() >= 0.0 and np.abs(d[2][~ax]).max() > 1e-3
    with pytest.raises(ValueError):      # nodes on the axis without an axis boundary
        m = pipe(2, 2)
        m["attr_bc"][2] = "slip"
        oracle(m, 2)
    with pytest.raises(ValueError):      # nodes below the axis
        m = Q.grid(2, 2, (0.0, -0.5), (1.0, 0.5), periodic=(False, False))
        oracle(dict(m, attr_bc=["slip"] * 4, attr_state=np.zeros((4, 4))), 2)
    with pytest.raises(ValueError):      # Euler only
        oracle(pipe(2, 2), 2, parabolic=True, mu=0.01)
