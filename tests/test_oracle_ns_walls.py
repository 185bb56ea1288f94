"""Pins of the oracle's Navier-Stokes boundaries: no-slip adiabatic walls
(NoSlipAdiabWallBdrFaceIntegrator.cpp), supersonic inflow / outflow and
specified state with their BR1 boundary lifting (LiftingBR1.cpp:311-407) and
viscous boundary fluxes (BdrFaceIntegrator.cpp:141-175).

- Lifting fluxes in closed form: the wall's entropy variables (w0, beta V,
  -beta) minus w; outflow 0.5 (w(u(w)) - w), zero in corrected mode (the
  Entropy2Conserv / Conserv2Entropy round trip), and in faithful mode the
  velocities Entropy2Conserv leaves in w.
- A shear layer between walls, u = U sin(pi y / H) + U_w y / H (top wall
  moving at U_w), uniform rho and p: the x-momentum RHS converges to
  mu u_yy as the mesh is refined (a missing or mis-signed wall lifting or
  wall viscous flux leaves an O(1) error at the walls).
- Walls at rest with qn = 0 conserve mass and energy exactly; a wall heat
  flux qn adds qn times the wall length to the energy.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import cases

G = 1.4


def cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (G - 1) + 0.5 * rho * np.dot(v, v)])


@pytest.mark.parametrize("dim", [2, 3])
def test_lifting_flux_closed_forms(dim):
    s = cons(1.2, [0.3, -0.4, 0.2][:dim], 0.9)
    w = O.cons2entropy(dim, G, s)
    beta = -w[-1]
    V = np.array([0.7, -0.1, 0.05, 0.0, 0.0])
    f = O.bc_lift_flux(dim, G, 1, "no-slip-adiabatic", w, V)
    assert f[0] == 0.0 and f[-1] == 0.0
    assert np.allclose(f[1:1 + dim], beta * V[:dim] - w[1:1 + dim], rtol=1e-15, atol=1e-15)
    # outflow: the round trip, corrected; the velocities left in w, faithful
    fo = O.bc_lift_flux(dim, G, 0, "supersonic-outflow", w)
    assert np.abs(fo).max() < 1e-14
    fq = O.bc_lift_flux(dim, G, 1, "supersonic-outflow", w)
    vel = w[1:1 + dim] / beta
    assert np.allclose(fq[1:1 + dim], 0.5 * (w[1:1 + dim] - vel), rtol=1e-13, atol=1e-15)
    assert abs(fq[0]) < 1e-14 and abs(fq[-1]) < 1e-14
    # inflow: half the jump to the outer state's entropy variables
    g = cons(1.4, [2.0, 0.1, 0.0][:dim], 1.0)
    fi = O.bc_lift_flux(dim, G, 0, "supersonic-inflow", w, g)
    assert np.allclose(fi, 0.5 * (O.cons2entropy(dim, G, g) - w), rtol=1e-13, atol=1e-15)


def shear_setup(ny, order, mu=0.05, U=0.2, Uw=0.1, H=1.0, dim=2):
    Lx = 0.5
    nelem = (2, ny, 2)[:dim]
    hi = (Lx, H, 0.5)[:dim]
    bc = [("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic"), ("periodic", "periodic")][:dim]
    st = [(None, None), ((0.0, 0.0, 0.0, 0.0, 0.0), (Uw, 0.0, 0.0, 0.0, 0.0)), (None, None)][:dim]
    orc = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, parabolic=True, mu=mu, Pr=0.71, quirks=False,
                   bc=bc, bc_state=st)
    X = orc.coords()
    y = X[1]
    u = U * np.sin(math.pi * y / H) + Uw * y / H
    vel = [u, np.zeros_like(u)] + ([np.zeros_like(u)] if dim == 3 else [])
    state = cases.prim_to_cons(np.ones_like(u), vel, np.ones_like(u), G)
    exact = -mu * (math.pi / H) ** 2 * U * np.sin(math.pi * y / H)
    return orc, state, exact


@pytest.mark.parametrize("order", [2, 3, 4])
def test_shear_layer_between_walls_converges(order):
    """The x-momentum RHS error (max-norm, relative) falls at least at rate
    N - 1 (BR1's second derivative: N - 1 for even N, N for odd N; the same
    rates as the periodic problem), and the wall nodes' error stays that of
    the interior nodes."""
    errs, wall_over_interior = [], []
    for ny in (4, 8, 16):
        orc, st, exact = shear_setup(ny, order)
        r = orc.Mult(st).reshape(4, orc.NE, orc.K)
        y = orc.coords()[1]
        err = np.abs(r[1] - exact) / np.abs(exact).max()
        wall = (np.abs(y) < 1e-12) | (np.abs(y - 1.0) < 1e-12)
        errs.append(err.max())
        wall_over_interior.append(err[wall].max() / err[~wall].max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("errors", errs, "rates", rates, "wall/interior", wall_over_interior)
    assert min(rates) > order - 1.2
    assert wall_over_interior[-1] < 1.1


def test_shear_layer_3d_equals_2d():
    """A 3-D box periodic in z reproduces the 2-D operator on z-uniform states."""
    o2, s2, _ = shear_setup(4, 3)
    o3, s3, _ = shear_setup(4, 3, dim=3)
    r2 = o2.Mult(s2).reshape(4, o2.NE, o2.K)
    r3 = o3.Mult(s3).reshape(5, o3.NE, o3.K)
    Np = 4
    # z-layer k of 3-D element (ex, ey, ez) is 2-D element (ex, ey)
    for e3 in range(o3.NE):
        e2 = e3 % o2.NE
        for k in range(Np):
            sl = slice(k * Np * Np, (k + 1) * Np * Np)
            for v2, v3 in ((0, 0), (1, 1), (2, 2), (3, 4)):
                assert np.allclose(r3[v3, e3, sl], r2[v2, e2], rtol=0, atol=1e-12)
    assert np.abs(r3[3]).max() < 1e-12


@pytest.mark.parametrize("quirks", [False, True])
def test_walls_at_rest_conserve_mass_and_energy(quirks):
    orc = O.Oracle(2, 3, (3, 4), (0.0, 0.0), (1.0, 1.0), parabolic=True, mu=0.02, Pr=0.71, quirks=quirks,
                   bc=[("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic")],
                   bc_state=[(None, None), ((0,) * 5, (0,) * 5)])
    X = orc.coords()
    st = cases.random_smooth_state(X, 2, 4, L=[1.0, 1.0])
    r = orc.Mult(st).reshape(4, orc.NE, orc.K)
    WJ = orc.weights()
    for v in (0, 3):
        assert abs(np.sum(WJ * r[v])) < 1e-13 * np.sum(WJ * np.abs(r[v]))


def test_wall_heat_flux_enters_the_energy():
    qn = 0.37
    Lx = 1.3
    orc = O.Oracle(2, 3, (3, 2), (0.0, 0.0), (Lx, 1.0), parabolic=True, mu=0.02, Pr=0.71, quirks=False,
                   bc=[("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic")],
                   bc_state=[(None, None), ((0, 0, 0, qn, 0), (0,) * 5)])
    st = np.repeat(cons(1.1, [0.0, 0.0], 0.8)[:, None], orc.NE * orc.K, axis=1)
    r = orc.Mult(st).reshape(4, orc.NE, orc.K)
    WJ = orc.weights()
    assert abs(np.sum(WJ * r[3]) - qn * Lx) < 1e-13
    assert np.abs(r[:3]).max() < 1e-13


def test_rejected_combinations():
    kw = dict(parabolic=True, mu=0.01)
    with pytest.raises(ValueError):      # slip walls have no viscous flux in the reference
        O.Oracle(2, 3, (2, 2), (0, 0), (1, 1), bc=[("slip", "slip"), ("periodic", "periodic")], **kw)
    with pytest.raises(ValueError):      # no-slip walls are for Navier-Stokes
        O.Oracle(2, 3, (2, 2), (0, 0), (1, 1), bc=[("no-slip-adiabatic",) * 2, ("periodic",) * 2])
    with pytest.raises(ValueError):      # faithful mode reproduces qn = 0 only
        O.Oracle(2, 3, (2, 2), (0, 0), (1, 1), bc=[("no-slip-adiabatic",) * 2, ("periodic",) * 2],
                 bc_state=[((0, 0, 0, 0.5, 0), (0,) * 5), (None, None)], quirks=True, **kw)
