"""1-D Euler on a segment (TestCases/Euler/1D): pins of the oracle's 1-D
operator.

- A 2-D box whose state does not vary in y (periodic in y, no y-velocity)
  carries the 1-D problem: its RHS for rho, rho u and E equals the 1-D RHS
  node by node and its y-momentum RHS is zero, with blending, boundaries and
  every order (the 2-D operator is pinned against the reference's own
  vectors elsewhere).  In 1-D MFEM sets the face normal to +-1 from the side
  (DGSEMIntegrator.cpp:180-183) and the adjugate of the 1 x 1 Jacobian is 1,
  so a mistake there shows as a factor h.
- Free stream: a constant state gives a zero RHS, with specified-state ends
  equal to it and with slip walls at rest.
- Conservation: on a periodic segment the quadrature sums of the RHS vanish.
- The slip wall in 1-D flips the momentum by the normal's sign
  (SlipWallBdrFaceIntegrator.cpp:22-30): at rest its flux is (0, p n, 0).
- Sod's shock tube run to t = 0.25 (the reference case's settings: order 3,
  blending, 128 elements, cfl 1) stays close to the exact Riemann solution
  (Toro's solver), and Lax's to t = 0.14.
- The AcousticWave case's density wave (constant u and p) converges to its
  exact solution rho(x - u t) at order N + 1.
- The reference's 1-D mesh files (inline segment, periodic L2 nodes) read
  as the boxes of their cases.
"""
import numpy as np
import pytest

import oracle_lib as O
import riemann_exact as R
from prandtl_amd import cases

G = 1.4


def sod_like(X):
    rho = np.where(X < 0.5, 1.0, 0.125)
    p = np.where(X < 0.5, 1.0, 0.1)
    u = 0.3 * np.sin(2 * np.pi * X)
    return rho, rho * u, p / (G - 1) + 0.5 * rho * u * u


BCS = {
    "periodic": None,
    "specified": ("specified-state", "specified-state"),
    "outflow_slip": ("supersonic-outflow", "slip"),
    "inflow_outflow": ("supersonic-inflow", "supersonic-outflow"),
}
LEFT = (1.0, 0.2, 1.0 / (G - 1) + 0.02)
RIGHT = (0.125, 0.0, 0.1 / (G - 1))


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("order", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("blending", [False, True], ids=["dg", "blend"])
@pytest.mark.parametrize("bc", list(BCS))
def test_1d_equals_y_uniform_2d(order, blending, bc):
    nx, ny, L = 8, 2, (1.0, 0.37)
    b = BCS[bc]
    o1 = O.Oracle(1, order, [nx], [0.0], [L[0]], blending=blending,
                  bc=None if b is None else [b], bc_state=None if b is None else [(LEFT, RIGHT)])
    X1 = o1.coords()[0]
    r, m, E = sod_like(X1)
    d1 = o1.Mult(np.stack([r, m, E])).reshape(3, nx, order + 1)
    o2 = O.Oracle(2, order, [nx, ny], [0.0, 0.0], L, blending=blending,
                  bc=None if b is None else [b, ("periodic", "periodic")],
                  bc_state=None if b is None else [((LEFT[0], LEFT[1], 0.0, LEFT[2]),
                                                    (RIGHT[0], RIGHT[1], 0.0, RIGHT[2])), (None, None)])
    Np = order + 1
    u2 = np.zeros((4, ny, nx, Np, Np))
    u2[0], u2[1], u2[3] = (f.reshape(nx, Np)[None, :, None, :] for f in (r, m, E))
    d2 = o2.Mult(u2).reshape(4, ny, nx, Np, Np)
    sc = np.abs(d1).max()
    for a, c in ((0, 0), (1, 1), (2, 3)):
        err = np.abs(d2[c] - d1[a][None, :, None, :]).max() / sc
        assert err <= 1e-13, (a, err)
    assert np.abs(d2[2]).max() <= 1e-13 * sc
    if blending:
        assert np.allclose(o1.alpha(), o2.alpha().reshape(ny, nx)[0], rtol=0, atol=1e-12)
        assert (o1.alpha() > 0).any()


@pytest.mark.parametrize("bc", ["periodic", "specified", "slip"])
def test_1d_free_stream(bc):
    s = (1.3, 1.3 * 0.0, 2.1 / (G - 1)) if bc == "slip" else (1.3, 1.3 * 0.4, 2.1 / (G - 1) + 0.5 * 1.3 * 0.16)
    b = {"periodic": None, "specified": ("specified-state", "specified-state"), "slip": ("slip", "slip")}[bc]
    o = O.Oracle(1, 4, [6], [0.0], [2.0], blending=True, bc=None if b is None else [b],
                 bc_state=None if b is None else [(s, s)])
    u = np.repeat(np.asarray(s, float)[:, None], 6 * 5, axis=1)
    d = o.Mult(u)
    assert np.abs(d).max() <= 1e-12 * np.abs(u).max()


def test_1d_periodic_conservation():
    o = O.Oracle(1, 3, [16], [0.0], [1.0], blending=True)
    r, m, E = sod_like(o.coords()[0])
    d = o.Mult(np.stack([r, m, E])).reshape(3, 16, 4)
    W = o.weights()
    for v in range(3):
        assert abs(np.sum(W * d[v])) <= 1e-13 * np.sum(W * np.abs(d[v]))


@pytest.mark.parametrize("n", [-1.0, 1.0])
def test_1d_slip_wall_at_rest(n):
    s = np.array([0.8, 0.0, 1.7 / (G - 1)])
    F, lam = O.bc_flux(1, G, 1, "slip", s, np.array([n]))
    assert np.allclose(F, [0.0, 1.7 * n, 0.0], rtol=1e-15, atol=0)
    assert lam == pytest.approx(np.sqrt(G * 1.7 / 0.8), rel=1e-15)
    # moving into the wall: the shock branch raises p* above p
    F2, _ = O.bc_flux(1, G, 1, "slip", np.array([0.8, 0.8 * 0.5 * n, 1.7 / (G - 1) + 0.1]), np.array([n]))
    assert F2[0] == 0.0 and F2[2] == 0.0 and F2[1] * n > 1.7


def run_tube(name):
    kw, init = cases.shock_tube(name)
    o = O.Oracle(threads=4, **kw)
    X = o.coords()
    u = init(X).reshape(-1)
    t, T = 0.0, cases.SHOCK_TUBES[name]["t_final"]
    o.Mult(u)   # the reference's extra initial Mult sets the running wave speed
    while t < T - 1e-14:
        dt = min(1.0 * o.hmin() / (o.GetMaxCharSpeed() * (kw["order"] + 1) ** 2), T - t)
        u = o.Step(u, dt)
        t += dt
    return o, X, u.reshape(3, -1)


@pytest.mark.parametrize("name,l1max", [("sod", 4e-3), ("lax", 1.2e-2)])
def test_shock_tube_close_to_exact_riemann(name, l1max):
    o, X, U = run_tube(name)
    c = cases.SHOCK_TUBES[name]
    (x0, rl, ul, pl), (_, rr, ur, pr) = c["states"]
    x = X[0].reshape(-1)
    ex = R.sample((rl, ul, pl), (rr, ur, pr), (x - x0) / c["t_final"])
    W = o.weights().reshape(-1)
    l1 = float(np.sum(W * np.abs(U[0] - ex[0])) / np.sum(W))
    print(name, "L1(rho) =", l1, flush=True)
    assert np.isfinite(U).all() and (U[0] > 0).all()
    assert l1 <= l1max


def tube_ends(name):
    """The two primitive states (rho, u, p) of a two-state tube."""
    c = cases.SHOCK_TUBES[name]
    g = c.get("gamma", 1.4)
    (x0, a, b, q), (_, a2, b2, q2) = c["states"]
    if c.get("cons"):
        return x0, g, (a, b / a, (g - 1) * (q - 0.5 * b * b / a)), (a2, b2 / a2, (g - 1) * (q2 - 0.5 * b2 * b2 / a2))
    return x0, g, (a, b, q), (a2, b2, q2)


# relative L1 error of rho against the exact Riemann solution at t_final;
# the bounds are 1.3-1.6 times the oracle's own errors (order 3, blending,
# cfl 1), which a wrong state, gamma, alpha_max or end condition exceeds
TUBE_L1 = {"modified_sod": 8e-3, "problem123": 2.5e-2, "blast_left": 8e-2, "blast_right": 8e-2,
           "blast_collision": 3e-2, "leblanc": 1e-2}


@pytest.mark.parametrize("name", list(TUBE_L1))
def test_reference_tubes_close_to_exact_riemann(name):
    """The remaining 1-D cases of the reference (ModifiedSodShockTube,
    Problem123, the Woodward-Colella left / right halves and collision,
    LeBlancShockTube) run to their final times with the reference's dt
    policy, stay finite with positive density, and approach the exact
    Riemann solution."""
    o, X, U = run_tube(name)
    x0, g, left, right = tube_ends(name)
    ex = R.sample(left, right, (X[0].reshape(-1) - x0) / cases.SHOCK_TUBES[name]["t_final"], g=g)
    W = o.weights().reshape(-1)
    l1 = float(np.sum(W * np.abs(U[0] - ex[0])) / np.sum(W * np.abs(ex[0])))
    print(name, "relative L1(rho) =", l1, flush=True)
    assert np.isfinite(U).all() and (U[0] > 0).all()
    assert l1 <= TUBE_L1[name]


def acoustic_error(nelem, order, T):
    o = O.Oracle(1, order, [nelem], [0.0], [1.0], blending=False)
    X = o.coords()
    u = cases.acoustic_wave(X).reshape(-1)
    n = int(np.ceil(T / (0.2 / nelem / (order + 1) ** 2 / 2.0)))
    dt = T / n
    for _ in range(n):
        u = o.Step(u, dt)
    ex = cases.acoustic_wave(X, T)
    W = o.weights()
    return float(np.sqrt(np.sum(W * (u.reshape(3, nelem, -1)[0] - ex[0]) ** 2)))


def test_acoustic_wave_converges_at_order_n_plus_1():
    """The density wave of AcousticWave.hpp carried at constant u and p has
    the exact solution rho(x - u t): the L2 error falls as h^(N+1)."""
    errs = [acoustic_error(n, 3, 0.2) for n in (4, 8, 16)]
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("acoustic wave L2 errors", errs, "rates", rates, flush=True)
    assert rates[-1] > 3.6 and errs[-1] < 1e-5


def test_reads_reference_segment_meshes():
    """The AcousticWave mesh (periodic, L2 nodes) and Sod's inline mesh give
    the boxes of their cases after the configs' ser_ref_levels."""
    import os
    from prandtl_amd.operator import read_segment_mesh
    d = os.path.join(os.path.dirname(__file__), "golden", "meshes")
    a = read_segment_mesh(os.path.join(d, "AcousticWave.mesh"), refine=6)
    assert a == dict(dim=1, nelem=(256,), lo=(0.0,), hi=(1.0,), periodic=True)
    s = read_segment_mesh(os.path.join(d, "SodShockTube.mesh"), refine=5)
    assert s == dict(dim=1, nelem=(128,), lo=(0.0,), hi=(1.0,), periodic=False)
    assert s["nelem"] == cases.shock_tube("sod")[0]["nelem"]
    with pytest.raises(ValueError):
        read_segment_mesh(os.path.join(d, "IsentropicVortex.mesh"))


def _segment_file(tmp_path, bdr):
    """An MFEM v1.0 segment mesh of 4 elements on [0, 2] with the given
    boundary lines ("attribute 0 vertex")."""
    text = ("MFEM mesh v1.0\n\ndimension\n1\n\nelements\n4\n1 1 0 1\n1 1 1 2\n1 1 2 3\n1 1 3 4\n\n"
            "boundary\n%d\n%s\n\nvertices\n5\n1\n0\n0.5\n1\n1.5\n2\n" % (len(bdr), "\n".join(bdr)))
    p = tmp_path / "seg.mesh"
    p.write_text(text)
    return str(p)


def test_segment_mesh_boundary_attributes(tmp_path):
    """Attribute 1 ('left') must be the lo end and 2 ('right') the hi end
    (Simulation.cpp:283-292): the box's sides are by position, so a file
    with them the other way round is refused, not run with its two end
    conditions swapped."""
    from prandtl_amd.operator import read_segment_mesh
    ok = read_segment_mesh(_segment_file(tmp_path, ["1 0 0", "2 0 4"]), refine=1)
    assert ok == dict(dim=1, nelem=(8,), lo=(0.0,), hi=(2.0,), periodic=False)
    # the same points listed in the other order are the same mesh
    assert read_segment_mesh(_segment_file(tmp_path, ["2 0 4", "1 0 0"])) == dict(
        dim=1, nelem=(4,), lo=(0.0,), hi=(2.0,), periodic=False)
    for bad in (["1 0 4", "2 0 0"], ["1 0 0", "1 0 4"], ["1 0 0"], ["1 0 0", "2 0 7"], ["1 0 0", "2 0 2"]):
        with pytest.raises(ValueError):
            read_segment_mesh(_segment_file(tmp_path, bad))
