"""Quadrilateral meshes on the GPU: k_rhs's general-geometry variant against
the oracle.

Bar: the plain FP64 RHS within 1e-12 relative max-norm of the oracle's on the
same mesh and state, the wave speed within 1e-13, one RK3 step within 1e-12;
a uniform grid given as a quad mesh equals the Cartesian kernel to 1e-13;
graph-captured steps equal eager ones and the device loop equals the host
loop bitwise.  The reference's forward-facing-step mesh (refined once, as its
configuration does) runs 150 steps at CFL 1 with its boundary conditions,
stays physical, forms a blended shock, and its evolved state still matches
the oracle's RHS.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
import quad_meshes as Q
from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import QuadMesh, TimeLoop

pytestmark = pytest.mark.gpu

TOL = 1e-12
G = 1.4
HERE = os.path.dirname(os.path.abspath(__file__))


def cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (G - 1) + 0.5 * rho * np.dot(v, v)])


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def qmesh(d):
    return QuadMesh(d["vert"], d["elem"], d.get("bdr"), d.get("bdr_attr"), d.get("elem_coords"))


def bc_dict(d):
    return {i + 1: (k, st) for i, (k, st) in enumerate(zip(d["attr_bc"], d["attr_state"]))} \
        if "attr_bc" in d else None


IN = cons(1.4, [2.4, 0.2], 1.0)
CASES = {
    "periodic_n2": (dict(periodic=(True, True), perturb=0.25, seed=1), 2, False),
    "periodic_n3_blend": (dict(periodic=(True, True), perturb=0.25, seed=2), 3, True),
    "periodic_n4": (dict(periodic=(True, True), perturb=0.3, seed=3), 4, False),
    "periodic_n5_blend": (dict(periodic=(True, True), perturb=0.2, seed=4), 5, True),
    "walls_n3": (dict(periodic=(False, False), perturb=0.25, seed=5,
                      kinds=["supersonic-inflow", "supersonic-outflow", "slip", "symmetry"]), 3, False),
    "walls_n4_blend": (dict(periodic=(False, True), perturb=0.2, seed=6,
                            kinds=["specified-state", "slip", "slip", "slip"]), 4, True),
    "rotated_n3": (dict(periodic=(True, True), perturb=0.2, seed=7, rotate=0.6), 3, True),
}


def build(name):
    kw, order, blend = CASES[name]
    kw = dict(kw)
    kinds = kw.pop("kinds", None)
    d = Q.grid(6, 5, **kw)
    if kinds:
        d.update(attr_bc=kinds, attr_state=[IN] * 4)
    return d, order, blend


def state(orc):
    X = orc.coords()
    return cases.random_smooth_state(X, 2, 3, L=[1.0, 1.0])


@pytest.mark.parametrize("name", list(CASES))
def test_quad_rhs_parity(cuda, name):
    import torch
    d, order, blend = build(name)
    orc = O.Oracle.quad(d, order, blending=blend, threads=8)
    op = DGSEMOperator.quad(qmesh(d), order, bc=bc_dict(d), blending=blend)
    u = state(orc)
    du = to_dev(u, cuda)
    out = torch.empty_like(du)
    op.Mult(du, out)
    torch.cuda.synchronize()
    ref = orc.Mult(u)
    r = rel_max(out.cpu().numpy(), ref)
    lg, lo = op.GetMaxCharSpeed(), orc.GetMaxCharSpeed()
    print(name, "rhs rel", r, "lambda", lg, lo, flush=True)
    assert r <= TOL, (name, r)
    assert abs(lg - lo) <= 1e-13 * lo
    if blend:
        a, _ = op.alpha()
        # alpha amplifies eta rounding by 9.21 alpha (1 - alpha) / threshold (~4e3 at N = 3)
        assert np.allclose(a, orc.alpha(), rtol=0, atol=1e-11)


def test_quad_uniform_grid_equals_cartesian_kernel(cuda):
    import torch
    order = 3
    box = DGSEMOperator(2, order, (6, 5), (0.0, 0.0), (1.0, 1.0), blending=True)
    quad = DGSEMOperator.quad(qmesh(Q.grid(6, 5)), order, blending=True)
    X = cases.node_coords(2, order, (6, 5), (0.0, 0.0), (1.0, 1.0))
    u = to_dev(cases.random_smooth_state(X, 2, 5, L=[1.0, 1.0]), cuda)
    a, b = torch.empty_like(u), torch.empty_like(u)
    box.Mult(u, a)
    quad.Mult(u, b)
    torch.cuda.synchronize()
    assert rel_max(b.cpu().numpy(), a.cpu().numpy()) <= 1e-13
    assert abs(quad.hmin - box.hmin) <= 1e-15


@pytest.mark.parametrize("name", ["walls_n3", "periodic_n5_blend"])
def test_quad_rk3_step_parity(cuda, name):
    import torch
    d, order, blend = build(name)
    orc = O.Oracle.quad(d, order, blending=blend, threads=8)
    op = DGSEMOperator.quad(qmesh(d), order, bc=bc_dict(d), blending=blend)
    u = state(orc)
    dt = 0.2 * op.hmin / (4.0 * (order + 1) ** 2)
    du = to_dev(u, cuda)
    op.Step(du, dt)
    torch.cuda.synchronize()
    r = rel_max(du.cpu().numpy(), orc.Step(u, dt))
    print(name, "step rel", r, flush=True)
    assert r <= TOL


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_quad_device_loop_matches_host_loop(cuda, graph):
    import torch
    d, order, blend = build("walls_n4_blend")
    orc = O.Oracle.quad(d, order, blending=blend)
    u0 = state(orc)
    s = torch.cuda.Stream(device=cuda) if graph else torch.cuda.default_stream(cuda)
    with torch.cuda.stream(s):
        a = DGSEMOperator.quad(qmesh(d), order, bc=bc_dict(d), blending=blend)
        b = DGSEMOperator.quad(qmesh(d), order, bc=bc_dict(d), blending=blend)
        ua, ub = to_dev(u0, cuda), to_dev(u0, cuda)
        t_host, n_host = TimeLoop(a, 0.05, cfl=0.8, variable_dt=True).run(ua)
        t_dev, n_dev = b.Run(ub, 0.0, 0.05, cfl=0.8)
    torch.cuda.synchronize()
    print("host %d steps, device %d steps" % (n_host, n_dev), flush=True)
    assert n_dev == n_host and t_dev == t_host
    assert torch.equal(ua, ub)


def test_forward_facing_step_reference_mesh(cuda):
    """The reference's ForwardFacingStep case: its mesh refined once
    (ser_ref_levels 1: 13600 elements of order 3), Mach-3 inflow on the
    left, supersonic outflow on the right, slip walls on top, bottom and
    the step, uniform initial state, blending."""
    import torch
    mesh = QuadMesh.read(os.path.join(HERE, "golden", "meshes", "ForwardFacingStep.msh"), refine=1)
    u_in = np.array([1.4, 1.4 * 3.0, 0.0, 1.0 / (1.4 - 1.0) + 0.5 * 1.4 * 3.0 * 3.0])
    bc = {"left": ("supersonic-inflow", u_in), "right": "supersonic-outflow", "top": "slip", "bottom": "slip"}
    op = DGSEMOperator.quad(mesh, 3, bc=bc, blending=True, alpha_max=0.5)
    u = to_dev(np.repeat(u_in[:, None], op.NE * op.K, axis=1), cuda)
    t, n = op.Run(u, 0.0, 3.0, cfl=1.0, max_steps=150)
    torch.cuda.synchronize()
    U = u.cpu().numpy().reshape(4, op.NE, op.K)
    alpha, _ = op.alpha()
    print("ffs: %d steps to t = %.6f, alpha max %.3f, rho range %.4f..%.4f"
          % (n, t, alpha.max(), U[0].min(), U[0].max()), flush=True)
    assert n == 150 and np.isfinite(U).all() and (U[0] > 0).all()
    assert alpha.max() > 0.1 and U[0].max() > 2.0      # a shock stands in front of the step
    # parity of the RHS on the evolved state
    bc_attr = {mesh.names[k]: v for k, v in bc.items()}
    d = dict(vert=mesh.vert, elem=mesh.elem, bdr=mesh.bdr, bdr_attr=mesh.bdr_attr)
    nattr = int(mesh.bdr_attr.max())
    kinds = ["slip"] * nattr
    states = [u_in] * nattr
    for a, v in bc_attr.items():
        kinds[a - 1] = v if isinstance(v, str) else v[0]
    d.update(attr_bc=kinds, attr_state=states)
    orc = O.Oracle.quad(d, 3, blending=True, threads=8)
    out = torch.empty_like(u)
    op.Mult(u, out)
    torch.cuda.synchronize()
    r = rel_max(out.cpu().numpy(), orc.Mult(u.cpu().numpy()))
    print("ffs rhs rel", r, flush=True)
    assert r <= TOL
