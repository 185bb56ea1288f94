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


# ---------------------------------------------------------------- Navier-Stokes
NS = dict(parabolic=True, mu=0.02, Pr=0.71)
WALL = (0.0, 0.0, 0.0, 0.0)
LID = (0.4, 0.0, 0.0, 0.0)

NS_CASES = {
    "ns_periodic_n3_blend": (dict(periodic=(True, True), perturb=0.25, seed=31), 3, True, False, None),
    "ns_periodic_n4_faithful": (dict(periodic=(True, True), perturb=0.2, seed=32), 4, False, True, None),
    "ns_cavity_n3": (dict(periodic=(False, False), perturb=0.2, seed=33), 3, False, False,
                     (["no-slip-adiabatic"] * 4, [WALL, WALL, WALL, LID])),
    "ns_channel_n5_faithful": (dict(periodic=(False, True), perturb=0.15, seed=34), 5, True, True,
                               (["supersonic-inflow", "supersonic-outflow", "no-slip-adiabatic", "no-slip-adiabatic"],
                                [IN, IN, WALL, WALL])),
}


def ns_build(name):
    kw, order, blend, quirks, bcs = NS_CASES[name]
    d = Q.grid(5, 4, **kw)
    if bcs:
        d.update(attr_bc=bcs[0], attr_state=bcs[1])
    return d, order, blend, quirks


@pytest.mark.parametrize("name", list(NS_CASES))
def test_quad_ns_parity(cuda, name):
    import torch
    d, order, blend, quirks = ns_build(name)
    orc = O.Oracle.quad(d, order, blending=blend, quirks=quirks, threads=8, **NS)
    op = DGSEMOperator.quad(qmesh(d), order, bc=bc_dict(d), blending=blend, quirks=quirks, **NS)
    u = state(orc)
    du = to_dev(u, cuda)
    out = torch.empty_like(du)
    op.Mult(du, out)
    g = op.gradients(du)
    torch.cuda.synchronize()
    r = rel_max(out.cpu().numpy(), orc.Mult(u))
    rg = rel_max(g.cpu().numpy(), orc.gradients())
    print(name, "rhs rel", r, "grad rel", rg, flush=True)
    assert rg <= TOL and r <= TOL


def test_quad_ns_mixed_and_graph(cuda):
    import torch
    d, order, blend, quirks = ns_build("ns_cavity_n3")
    bc = bc_dict(d)
    fp = DGSEMOperator.quad(qmesh(d), order, bc=bc, quirks=False, **NS)
    mx = DGSEMOperator.quad(qmesh(d), order, bc=bc, quirks=False, precision="mixed", **NS)
    orc = O.Oracle.quad(d, order, quirks=False, **NS)
    u = to_dev(state(orc), cuda)
    a, b = torch.empty_like(u), torch.empty_like(u)
    fp.Mult(u, a)
    mx.Mult(u, b)
    torch.cuda.synchronize()
    rel = float(torch.linalg.norm(b - a) / torch.linalg.norm(a))
    print("mixed vs fp64", rel, flush=True)
    assert rel < 1e-6
    ua, ub = u.clone(), u.clone()
    s = torch.cuda.Stream(device=cuda)
    for _ in range(3):
        with torch.cuda.stream(s):
            fp.Step(ua, 1e-4)
        mx.Step(ub, 1e-4)
    torch.cuda.synchronize()
    ue = u.clone()
    fe = DGSEMOperator.quad(qmesh(d), order, bc=bc, quirks=False, **NS)
    for _ in range(3):
        fe.Step(ue, 1e-4)
    torch.cuda.synchronize()
    assert torch.equal(ua, ue)


def test_lid_driven_cavity_reference_mesh(cuda):
    """The reference's LidDrivenCavity case on its own mesh (256 quadrilaterals,
    N = 3, blending, mu = 0.02, Pr = 0.72, Ma = 0.1 initial state at rest,
    fixed dt = 1e-4, no-slip adiabatic walls, lid velocity (1, 0)): 300
    steps stay physical, the walls keep the mass constant, the lid drives
    the flow, and the RHS and gradients on the evolved state match the
    oracle."""
    import torch
    mesh = QuadMesh.read(os.path.join(HERE, "golden", "meshes", "LidDrivenCavity.msh"))
    wall, lid = (0.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0)
    bc = {"bottom": ("no-slip-adiabatic", wall), "right": ("no-slip-adiabatic", wall),
          "left": ("no-slip-adiabatic", wall), "lid": ("no-slip-adiabatic", lid)}
    kw = dict(blending=True, alpha_max=0.5, parabolic=True, mu=0.02, Pr=0.72, quirks=False)
    op = DGSEMOperator.quad(mesh, 3, bc=bc, **kw)
    p0 = 1.0 / (0.1 * 0.1 * 1.4)
    u0 = np.repeat(np.array([1.0, 0.0, 0.0, p0 / 0.4])[:, None], op.NE * op.K, axis=1)
    u = to_dev(u0, cuda)
    t, n = op.Run(u, 0.0, 3.0, dt=1e-4, max_steps=300)
    torch.cuda.synchronize()
    U = u.cpu().numpy().reshape(4, op.NE, op.K)
    x, w = cases.lgl_nodes(3)
    i = np.arange(16)
    c = mesh.corners()
    # w_i w_j J with J of each (affine) element
    area = 0.5 * np.abs((c[:, 2, 0] - c[:, 0, 0]) * (c[:, 3, 1] - c[:, 1, 1]) -
                        (c[:, 3, 0] - c[:, 1, 0]) * (c[:, 2, 1] - c[:, 0, 1]))
    WJ = (w[i % 4] * w[i // 4])[None, :] * area[:, None]
    m0, m1 = np.sum(WJ * u0.reshape(4, op.NE, op.K)[0]), np.sum(WJ * U[0])
    print("ldc: %d steps, mass %.15g -> %.15g, max |u| %.4f" % (n, m0, m1, np.abs(U[1] / U[0]).max()), flush=True)
    assert n == 300 and np.isfinite(U).all() and (U[0] > 0).all()
    assert abs(m1 - m0) <= 1e-13 * m0
    assert np.abs(U[1] / U[0]).max() > 0.1
    kinds = ["no-slip-adiabatic"] * 4
    d = dict(vert=mesh.vert, elem=mesh.elem, bdr=mesh.bdr, bdr_attr=mesh.bdr_attr, attr_bc=kinds,
             attr_state=[wall, wall, lid, wall])
    assert mesh.names["lid"] == 3
    orc = O.Oracle.quad(d, 3, threads=8, **kw)
    out = torch.empty_like(u)
    op.Mult(u, out)
    g = op.gradients(u)
    torch.cuda.synchronize()
    r = rel_max(out.cpu().numpy(), orc.Mult(u.cpu().numpy()))
    rg = rel_max(g.cpu().numpy(), orc.gradients())
    print("ldc rhs rel", r, "grad rel", rg, flush=True)
    assert r <= TOL and rg <= TOL


def rk4_step(mult, x, dt):
    """MFEM's RK4Solver::Step on a Mult: k1..k4 with y and the running sum z
    formed in the solver's order."""
    k = mult(x)
    y = x + (dt / 2) * k
    z = x + (dt / 6) * k
    k = mult(y)
    y = x + (dt / 2) * k
    z = z + (dt / 3) * k
    k = mult(y)
    y = x + dt * k
    z = z + (dt / 3) * k
    k = mult(y)
    return z + (dt / 6) * k


def test_backward_facing_step_reference_mesh_rk4(cuda):
    """The reference's BackwardFacingStep case on its own mesh (35000
    quadrilaterals, order 3, blending): the post-shock inflow, slip walls and
    supersonic outflow, with the shock at x = 0.5 initially.  Its config
    steps with RK4, which a caller composes from the operator's Mult as
    MFEM's RK4Solver does: 3 RK4 steps through the operator against the same
    steps through the oracle, the RHS of the initial state within 1e-12 and
    the stepped state within 1e-12."""
    import torch
    c = cases.BFS
    mesh = QuadMesh.read(os.path.join(HERE, "golden", "meshes", c["mesh"]))
    op = DGSEMOperator.quad(mesh, c["order"], bc=c["bc"], blending=c["blending"], alpha_max=c["alpha_max"])
    nattr = int(mesh.bdr_attr.max())
    kinds, states = ["slip"] * nattr, [c["bc"]["left"][1]] * nattr
    for key, v in c["bc"].items():
        kinds[mesh.names[key] - 1] = v if isinstance(v, str) else v[0]
    d = dict(vert=mesh.vert, elem=mesh.elem, bdr=mesh.bdr, bdr_attr=mesh.bdr_attr, attr_bc=kinds,
             attr_state=states)
    orc = O.Oracle.quad(d, c["order"], blending=c["blending"], alpha_max=c["alpha_max"], threads=16)
    u0 = cases.bfs_state(orc.coords()).reshape(-1)
    r0 = orc.Mult(u0)
    dt = 0.5 * orc.hmin() / (orc.GetMaxCharSpeed() * (c["order"] + 1) ** 2)
    du = to_dev(u0, cuda)
    out = torch.empty_like(du)

    def gpu_mult(x):
        y = torch.empty_like(x)
        op.Mult(x, y)
        return y

    op.Mult(du, out)
    torch.cuda.synchronize()
    r = rel_max(out.cpu().numpy(), r0)
    alpha, _ = op.alpha()
    print("bfs: RHS rel %.3e, alpha > 0 on %d elements, dt %.3g" % (r, int((alpha > 0).sum()), dt), flush=True)
    assert r <= TOL
    x, ref = du, u0.copy()
    for _ in range(3):
        x = rk4_step(gpu_mult, x, dt)
        ref = rk4_step(orc.Mult, ref, dt)
    torch.cuda.synchronize()
    r = rel_max(x.cpu().numpy(), ref)
    print("bfs: 3 RK4 steps rel %.3e" % r, flush=True)
    assert np.isfinite(ref).all() and r <= TOL
