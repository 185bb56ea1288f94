"""Boundary conditions on the GPU: libprandtl_hip.so against the oracle.

Slip walls, symmetry, supersonic inflow / outflow and specified-state sides
in 2-D and 3-D, with and without blending, faithful and corrected modes.
Bar: the plain FP64 RHS within 1e-12 relative max-norm of the oracle's on
the same state, the wave speed (boundary points included) within 1e-13, one
RK3 step within 1e-12.  Graph-captured steps equal eager steps bitwise, and
the device time loop equals the host loop bitwise.  At the triple-point
problem's full size the check is size-independent: with slip walls on every
side total mass and energy stay constant to rounding over the run.
"""
import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import DGSEMOperator, PrandtlError, _lib, cases
from prandtl_amd.operator import TimeLoop

pytestmark = pytest.mark.gpu

TOL = 1e-12
G = 1.4


def cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (G - 1) + 0.5 * rho * np.dot(v, v)])


IN2 = cons(1.4, [2.2, 0.3], 1.0)          # supersonic inflow (|v| > c)
IN3 = cons(1.4, [2.2, 0.3, -0.2], 1.0)
SPEC3 = cons(0.9, [0.1, -0.2, 0.3], 1.2)

CASES = {
    "slip2d_n3": dict(dim=2, order=3, nelem=(6, 5), lo=(0.0, 0.0), hi=(1.0, 0.8),
                      bc=[("slip", "slip"), ("slip", "slip")]),
    "io2d_n4_blend": dict(dim=2, order=4, nelem=(7, 6), lo=(0.0, 0.0), hi=(1.2, 1.0), blending=True,
                          bc=[("supersonic-inflow", "supersonic-outflow"), ("slip", "symmetry")],
                          bc_state=[(IN2, None), (None, None)]),
    "sym2d_n5": dict(dim=2, order=5, nelem=(4, 5), lo=(0.0, 0.0), hi=(1.0, 1.0),
                     bc=[("symmetry", "symmetry"), ("periodic", "periodic")]),
    "spec2d_n2_corr": dict(dim=2, order=2, nelem=(8, 7), lo=(-1.0, 0.0), hi=(1.0, 1.5), quirks=False,
                           bc=[("specified-state", "specified-state"), ("slip", "supersonic-outflow")],
                           bc_state=[(IN2, IN2), (None, None)]),
    "slip3d_n3_corr": dict(dim=3, order=3, nelem=(4, 3, 3), lo=(0.0,) * 3, hi=(1.0, 0.9, 0.8), quirks=False,
                           bc=[("slip", "slip")] * 3),
    "mix3d_n4_blend": dict(dim=3, order=4, nelem=(3, 3, 4), lo=(0.0,) * 3, hi=(1.0, 1.0, 1.2), blending=True,
                           bc=[("supersonic-inflow", "supersonic-outflow"), ("slip", "slip"),
                               ("symmetry", "specified-state")],
                           bc_state=[(IN3, None), (None, None), (None, SPEC3)]),
    "zwall3d_n7_blend": dict(dim=3, order=7, nelem=(2, 2, 3), lo=(0.0,) * 3, hi=(1.0, 1.0, 1.5), blending=True,
                             bc=[("periodic", "periodic"), ("periodic", "periodic"), ("slip", "slip")]),
    "out3d_n2": dict(dim=3, order=2, nelem=(4, 4, 4), lo=(0.0,) * 3, hi=(1.0,) * 3,
                     bc=[("supersonic-outflow", "supersonic-outflow"), ("symmetry", "slip"),
                         ("periodic", "periodic")]),
}


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def state(cfg, seed=3):
    X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    L = [cfg["hi"][r] - cfg["lo"][r] for r in range(cfg["dim"])]
    return cases.random_smooth_state(X, cfg["dim"], seed, L=L)


@pytest.mark.parametrize("name", list(CASES))
def test_bc_rhs_parity(cuda, name):
    import torch
    cfg = CASES[name]
    op = DGSEMOperator(**cfg)
    orc = O.Oracle(threads=8, **cfg)
    u = state(cfg)
    du = to_dev(u, cuda)
    out = torch.empty_like(du)
    op.Mult(du, out)
    torch.cuda.synchronize()
    ref = orc.Mult(u)
    r = rel_max(out.cpu().numpy(), ref)
    lam_g, lam_o = op.GetMaxCharSpeed(), orc.GetMaxCharSpeed()
    print(name, "rhs rel", r, "lambda", lam_g, lam_o, flush=True)
    assert r <= TOL, (name, r)
    assert abs(lam_g - lam_o) <= 1e-13 * lam_o


@pytest.mark.parametrize("name", ["io2d_n4_blend", "mix3d_n4_blend", "slip2d_n3"])
def test_bc_rk3_step_parity(cuda, name):
    import torch
    cfg = CASES[name]
    op = DGSEMOperator(**cfg)
    orc = O.Oracle(threads=8, **cfg)
    u = state(cfg, seed=9)
    dt = 0.2 * min((cfg["hi"][r] - cfg["lo"][r]) / cfg["nelem"][r] for r in range(cfg["dim"])) \
        / (4.0 * (cfg["order"] + 1) ** 2)
    du = to_dev(u, cuda)
    op.Step(du, dt)
    torch.cuda.synchronize()
    ref = orc.Step(u, dt)
    r = rel_max(du.cpu().numpy(), ref)
    print(name, "step rel", r, flush=True)
    assert r <= TOL, (name, r)


def test_bc_graph_steps_equal_eager(cuda):
    import torch
    cfg = CASES["mix3d_n4_blend"]
    a, b = DGSEMOperator(**cfg), DGSEMOperator(**cfg)
    ua, ub = to_dev(state(cfg), cuda), to_dev(state(cfg), cuda)
    s = torch.cuda.Stream(device=cuda)
    for dt in (1e-4, 1e-4, 5e-5):
        with torch.cuda.stream(s):
            a.Step(ua, dt)
        b.Step(ub, dt)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)


def test_bc_shapes_natural_only(cuda):
    """Open boxes run k_rhs in its natural shape; the autotuner keeps it and
    forcing another shape is refused."""
    import torch
    cfg = CASES["slip2d_n3"]
    op = DGSEMOperator(**cfg)
    u = to_dev(state(cfg), cuda)
    shapes, times = op.autotune(u)
    torch.cuda.synchronize()
    assert shapes["rhs"] == 0
    assert times["rhs"][0] is not None and all(t is None for t in times["rhs"][1:])
    with pytest.raises(PrandtlError) as ei:
        op.set_shape("rhs", 2)
    assert ei.value.code == _lib.PRANDTL_ERR_UNSUPPORTED


def tp_config(nelem):
    cfg = dict(cases.TRIPLE_POINT)
    cfg["nelem"] = nelem
    return cfg


def tp_state(cfg, dev):
    X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    return to_dev(cases.triple_point(X), dev)


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_triple_point_device_loop_matches_host_loop_and_oracle(cuda, graph):
    import torch
    cfg = tp_config((28, 12))
    s = torch.cuda.Stream(device=cuda) if graph else torch.cuda.default_stream(cuda)
    with torch.cuda.stream(s):
        a, b = DGSEMOperator(**cfg), DGSEMOperator(**cfg)
        ua, ub = tp_state(cfg, cuda), tp_state(cfg, cuda)
        t_host, n_host = TimeLoop(a, 0.3, cfl=1.0, variable_dt=True).run(ua)
        t_dev, n_dev = b.Run(ub, 0.0, 0.3, cfl=1.0)
    torch.cuda.synchronize()
    print("host %d steps t=%.17g, device %d steps t=%.17g" % (n_host, t_host, n_dev, t_dev), flush=True)
    assert n_dev == n_host and t_dev == t_host
    assert torch.equal(ua, ub)
    # the evolved state has flow off and along the walls: RHS parity there
    u = ub.cpu().numpy()
    orc = O.Oracle(threads=8, **cfg)
    out = torch.empty_like(ub)
    b.Mult(ub, out)
    torch.cuda.synchronize()
    r = rel_max(out.cpu().numpy(), orc.Mult(u))
    print("triple point rhs rel", r, flush=True)
    assert r <= TOL


def test_triple_point_full_size_conserves_mass_and_energy(cuda):
    """112 x 48 elements of order 3 (the h = 1/16 mesh), 200 steps at cfl 1:
    slip walls pass no mass or energy, so both totals stay at their initial
    values to rounding (1e-13 relative)."""
    import torch
    cfg = tp_config(cases.TRIPLE_POINT["nelem"])
    op = DGSEMOperator(**cfg)
    u = tp_state(cfg, cuda)
    WJ = cases.quadrature_weights(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    U0 = u.cpu().numpy().reshape(4, op.NE, op.K)
    t, n = op.Run(u, 0.0, 4.0, cfl=1.0, max_steps=200)
    torch.cuda.synchronize()
    U = u.cpu().numpy().reshape(4, op.NE, op.K)
    assert n == 200 and np.isfinite(U).all() and (U[0] > 0).all()
    for v in (0, 3):
        m0, m1 = np.sum(WJ * U0[v]), np.sum(WJ * U[v])
        print("var", v, "total", m0, "->", m1, "rel", abs(m1 - m0) / abs(m0), flush=True)
        assert abs(m1 - m0) <= 1e-13 * abs(m0)
    # the walls push back: x-momentum starts at zero and leaves it
    assert abs(np.sum(WJ * U[1])) > 1e-3
    assert (op.alpha()[0] > 0).any()
