"""Navier-Stokes boundaries on the GPU against the oracle: no-slip adiabatic
walls (moving, with heat flux), supersonic inflow / outflow and specified
state, with the BR1 boundary lifting in k_visc and the viscous boundary
flux in k_rhs.

Bar: BR1 gradients and the plain FP64 RHS within 1e-12 relative max-norm of
the oracle's, one RK3 step within 1e-12; graph steps equal eager steps and
the device loop equals the host loop bitwise; the mixed form within 1e-6
(relative L2) of the FP64 RHS; over a run between walls at rest the mass
and energy totals stay constant to rounding.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import TimeLoop

pytestmark = pytest.mark.gpu

TOL = 1e-12
G = 1.4


def cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (G - 1) + 0.5 * rho * np.dot(v, v)])


IN2 = cons(1.2, [1.9, 0.1], 1.0)
IN3 = cons(1.2, [1.9, 0.1, -0.05], 1.0)
WALL_MOVING = (0.3, 0.0, 0.0, 0.0, 0.0)
WALL_HEAT = (0.0, 0.0, 0.0, 0.2, 0.0)
REST = (0.0,) * 5
NS = dict(parabolic=True, mu=0.02, Pr=0.71)

CASES = {
    "channel2d_n3": dict(dim=2, order=3, nelem=(4, 5), lo=(0.0, 0.0), hi=(1.0, 1.0), quirks=False,
                         bc=[("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic")],
                         bc_state=[(None, None), (WALL_HEAT, WALL_MOVING)], **NS),
    "channel2d_n4_faithful": dict(dim=2, order=4, nelem=(3, 4), lo=(0.0, 0.0), hi=(1.0, 0.8), quirks=True,
                                  bc=[("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic")],
                                  bc_state=[(None, None), (REST, WALL_MOVING)], **NS),
    "inout2d_n3_faithful": dict(dim=2, order=3, nelem=(5, 4), lo=(0.0, 0.0), hi=(1.2, 1.0), quirks=True,
                                bc=[("supersonic-inflow", "supersonic-outflow"),
                                    ("no-slip-adiabatic", "specified-state")],
                                bc_state=[(IN2, None), (REST, IN2)], **NS),
    "inout2d_n5_sutherland": dict(dim=2, order=5, nelem=(3, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), quirks=False,
                                  sutherland=True, R_gas=1.0,
                                  bc=[("supersonic-inflow", "supersonic-outflow"),
                                      ("no-slip-adiabatic", "no-slip-adiabatic")],
                                  bc_state=[(IN2, None), (WALL_HEAT, WALL_MOVING)], **NS),
    "duct3d_n3_faithful": dict(dim=3, order=3, nelem=(3, 3, 2), lo=(0.0,) * 3, hi=(1.0, 1.0, 0.8), quirks=True,
                               bc=[("supersonic-inflow", "supersonic-outflow"),
                                   ("no-slip-adiabatic", "no-slip-adiabatic"), ("periodic", "periodic")],
                               bc_state=[(IN3, None), (REST, WALL_MOVING), (None, None)], **NS),
    "duct3d_n3": dict(dim=3, order=3, nelem=(3, 3, 2), lo=(0.0,) * 3, hi=(1.0, 1.0, 0.8), quirks=False,
                      bc=[("supersonic-inflow", "supersonic-outflow"), ("no-slip-adiabatic", "no-slip-adiabatic"),
                          ("periodic", "periodic")],
                      bc_state=[(IN3, None), (WALL_HEAT, WALL_MOVING), (None, None)], **NS),
}


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def state(cfg, seed=6):
    X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    L = [cfg["hi"][r] - cfg["lo"][r] for r in range(cfg["dim"])]
    return cases.random_smooth_state(X, cfg["dim"], seed, L=L)


@pytest.mark.parametrize("name", list(CASES))
def test_ns_boundary_rhs_and_gradient_parity(cuda, name):
    import torch
    cfg = CASES[name]
    op = DGSEMOperator(**cfg)
    orc = O.Oracle(threads=8, **cfg)
    u = state(cfg)
    du = to_dev(u, cuda)
    out = torch.empty_like(du)
    op.Mult(du, out)
    g = op.gradients(du)
    torch.cuda.synchronize()
    ref = orc.Mult(u)
    r = rel_max(out.cpu().numpy(), ref)
    rg = rel_max(g.cpu().numpy(), orc.gradients())
    print(name, "rhs rel", r, "grad rel", rg, flush=True)
    assert rg <= TOL, (name, rg)
    assert r <= TOL, (name, r)
    assert abs(op.GetMaxCharSpeed() - orc.GetMaxCharSpeed()) <= 1e-13 * orc.GetMaxCharSpeed()


@pytest.mark.parametrize("name", ["channel2d_n3", "duct3d_n3", "duct3d_n3_faithful"])
def test_ns_boundary_rk3_step_and_graph(cuda, name):
    import torch
    cfg = CASES[name]
    orc = O.Oracle(threads=8, **cfg)
    u = state(cfg, seed=2)
    dt = 1e-3
    a, b = DGSEMOperator(**cfg), DGSEMOperator(**cfg)
    ua, ub = to_dev(u, cuda), to_dev(u, cuda)
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        a.Step(ua, dt)
    b.Step(ub, dt)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)
    r = rel_max(ua.cpu().numpy(), orc.Step(u, dt))
    print(name, "step rel", r, flush=True)
    assert r <= TOL


def test_ns_boundary_mixed_form(cuda):
    import torch
    cfg = dict(CASES["duct3d_n3"])
    fp = DGSEMOperator(**cfg)
    mx = DGSEMOperator(**dict(cfg, precision="mixed"))
    u = to_dev(state(cfg), cuda)
    a, b = torch.empty_like(u), torch.empty_like(u)
    fp.Mult(u, a)
    mx.Mult(u, b)
    torch.cuda.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    rel = np.linalg.norm(b - a) / np.linalg.norm(a)
    print("mixed vs fp64 rel L2", rel, flush=True)
    assert rel < 1e-6


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_ns_walls_device_loop_and_conservation(cuda, graph):
    """Walls at rest, zero heat flux, periodic in x: mass and energy totals
    constant over the run; device loop equal to the host loop."""
    import torch
    cfg = dict(dim=2, order=3, nelem=(6, 6), lo=(0.0, 0.0), hi=(1.0, 1.0), quirks=False,
               bc=[("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic")],
               bc_state=[(None, None), (REST, REST)], **NS)
    X = cases.node_coords(2, 3, cfg["nelem"], cfg["lo"], cfg["hi"])
    y = X[1]
    u0 = cases.prim_to_cons(1.0 + 0.1 * np.cos(2 * math.pi * X[0]), [0.3 * np.sin(math.pi * y), 0 * y],
                            1.0 + 0.05 * np.sin(2 * math.pi * X[0]), G)
    s = torch.cuda.Stream(device=cuda) if graph else torch.cuda.default_stream(cuda)
    with torch.cuda.stream(s):
        a, b = DGSEMOperator(**cfg), DGSEMOperator(**cfg)
        ua, ub = to_dev(u0, cuda), to_dev(u0, cuda)
        t_host, n_host = TimeLoop(a, 0.2, cfl=0.5, variable_dt=True).run(ua)
        t_dev, n_dev = b.Run(ub, 0.0, 0.2, cfl=0.5)
    torch.cuda.synchronize()
    print("host %d steps, device %d steps" % (n_host, n_dev), flush=True)
    assert n_dev == n_host and t_dev == t_host
    assert torch.equal(ua, ub)
    WJ = cases.quadrature_weights(2, 3, cfg["nelem"], cfg["lo"], cfg["hi"])
    U0 = u0.reshape(4, a.NE, a.K)
    U = ub.cpu().numpy().reshape(4, a.NE, a.K)
    for v in (0, 3):
        m0, m1 = np.sum(WJ * U0[v]), np.sum(WJ * U[v])
        assert abs(m1 - m0) <= 1e-13 * abs(m0), (v, m0, m1)
    # the walls slow the shear flow down
    assert np.sum(WJ * U[1]) < np.sum(WJ * U0[1])
