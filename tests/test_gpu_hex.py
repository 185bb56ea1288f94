"""Hexahedral meshes on the GPU: the general-geometry k_rhs and k_visc in
3-D against the oracle.

Bar: the plain FP64 RHS (and BR1 gradients) within 1e-12 relative max-norm
of the oracle's on distorted periodic and bounded meshes, Euler and
Navier-Stokes (corrected and faithful), with and without blending; the wave speed within 1e-13; one
RK3 step within 1e-12; the mixed form within 1e-6 (relative L2) of FP64; a
uniform grid given as a hexahedral mesh equals the Cartesian kernels to
1e-13; renumbered elements with turned frames give the same RHS;
graph-captured steps equal eager ones and the device loop equals the host
steps bitwise.
"""
import numpy as np
import pytest

import oracle_lib as O
import quad_meshes as Q
from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import HexMesh

pytestmark = pytest.mark.gpu

TOL = 1e-12
G = 1.4
LO, HI = (0.0, 0.0, 0.0), (1.0, 0.8, 0.6)


def cons(rho, v, p):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (G - 1) + 0.5 * rho * np.dot(v, v)])


IN = cons(1.3, [1.9, 0.2, -0.1], 1.0)
WALL = (0.0, 0.0, 0.0, 0.0, 0.0)
LID = (0.3, 0.0, 0.1, 0.0, 0.0)
NS = dict(parabolic=True, mu=0.02, Pr=0.71)


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def hmesh(d):
    return HexMesh(d["vert"], d["elem"], d.get("bdr"), d.get("bdr_attr"), d.get("elem_coords"))


CASES = {
    "periodic_n2": (dict(periodic=(True,) * 3, perturb=0.25, seed=1), 2, dict(), None),
    "periodic_n3_blend": (dict(periodic=(True,) * 3, perturb=0.25, seed=2), 3, dict(blending=True), None),
    "periodic_n4": (dict(periodic=(True,) * 3, perturb=0.2, seed=3), 4, dict(), None),
    "walls_n3_blend": (dict(periodic=(False, False, False), perturb=0.2, seed=4), 3, dict(blending=True),
                       (["supersonic-inflow", "supersonic-outflow", "slip", "symmetry", "slip", "specified-state"],
                        [IN] * 6)),
    "walls_n2_faithful": (dict(periodic=(False, True, False), perturb=0.2, seed=5), 2, dict(quirks=True),
                          (["supersonic-inflow", "supersonic-outflow", "slip", "slip", "symmetry", "slip"],
                           [IN] * 6)),
    "ns_periodic_n3": (dict(periodic=(True,) * 3, perturb=0.2, seed=6), 3, dict(NS), None),
    "ns_periodic_n3_faithful": (dict(periodic=(True,) * 3, perturb=0.2, seed=8), 3, dict(NS, quirks=True), None),
    "ns_walls_n2_faithful": (dict(periodic=(False, True, False), perturb=0.2, seed=9), 2,
                             dict(NS, quirks=True, blending=True),
                             (["supersonic-inflow", "supersonic-outflow", "no-slip-adiabatic", "no-slip-adiabatic",
                               "no-slip-adiabatic", "no-slip-adiabatic"], [IN, IN, WALL, WALL, WALL, LID])),
    "ns_walls_n3_blend": (dict(periodic=(False, True, False), perturb=0.2, seed=7), 3, dict(NS, blending=True),
                          (["supersonic-inflow", "supersonic-outflow", "no-slip-adiabatic", "no-slip-adiabatic",
                            "no-slip-adiabatic", "no-slip-adiabatic"], [IN, IN, WALL, WALL, WALL, LID])),
}


def build(name):
    kw, order, extra, bcs = CASES[name]
    d = Q.hex_grid(3, 3, 4, LO, HI, **kw)
    bc = None
    if bcs:
        d.update(attr_bc=bcs[0], attr_state=bcs[1])
        bc = {i + 1: (k, st) for i, (k, st) in enumerate(zip(*bcs))}
    extra = dict(extra)
    extra.setdefault("quirks", False)
    op = DGSEMOperator.hex(hmesh(d), order, bc=bc, **extra)
    orc = O.Oracle.hex(d, order, threads=8, **extra)
    return op, orc, d, order, extra, bc


def state(orc, seed=3):
    return cases.random_smooth_state(orc.coords(), 3, seed, L=[1.0, 0.8, 0.6])


@pytest.mark.parametrize("name", list(CASES))
def test_hex_rhs_parity(cuda, name):
    import torch
    op, orc, *_ = build(name)
    u = state(orc)
    du = to_dev(u, cuda)
    out = torch.empty_like(du)
    op.Mult(du, out)
    torch.cuda.synchronize()
    ref = orc.Mult(u)
    r = rel_max(out.cpu().numpy(), ref)
    msg = [name, "rhs rel", r]
    if orc.params.parabolic:
        g = op.gradients(du)
        torch.cuda.synchronize()
        rg = rel_max(g.cpu().numpy(), orc.gradients())
        msg += ["grad rel", rg]
        assert rg <= TOL
    if orc.params.blending:
        a, _ = op.alpha()
        msg += ["alpha", float(np.abs(a - orc.alpha()).max()), a.max()]
        assert np.abs(a - orc.alpha()).max() <= 1e-11
    print(*msg, flush=True)
    assert r <= TOL
    assert abs(op.GetMaxCharSpeed() - orc.GetMaxCharSpeed()) <= 1e-13 * orc.GetMaxCharSpeed()
    assert abs(op.hmin - orc.hmin()) <= 1e-13 * orc.hmin()


@pytest.mark.parametrize("name", ["periodic_n3_blend", "ns_walls_n3_blend"])
def test_hex_rk3_step_graph_and_loop(cuda, name):
    import torch
    op, orc, d, order, extra, bc = build(name)
    op2 = DGSEMOperator.hex(hmesh(d), order, bc=bc, **extra)
    u = state(orc, seed=4)
    dt = 1e-3
    ua, ub = to_dev(u, cuda), to_dev(u, cuda)
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        op.Step(ua, dt)
    op2.Step(ub, dt)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)
    r = rel_max(ua.cpu().numpy(), orc.Step(u, dt))
    print(name, "step rel", r, flush=True)
    assert r <= TOL
    uc, ud = to_dev(u, cuda), to_dev(u, cuda)
    with torch.cuda.stream(s):
        t, n = op.Run(uc, 0.0, 1.0, dt=dt, max_steps=8)
    for _ in range(n):
        op2.Step(ud, dt)
    torch.cuda.synchronize()
    assert n == 8 and torch.equal(uc, ud)


def test_hex_ns_mixed_form(cuda):
    import torch
    _, orc, d, order, extra, bc = build("ns_walls_n3_blend")
    fp = DGSEMOperator.hex(hmesh(d), order, bc=bc, **extra)
    mx = DGSEMOperator.hex(hmesh(d), order, bc=bc, precision="mixed", **extra)
    u = to_dev(state(orc), cuda)
    a, b = torch.empty_like(u), torch.empty_like(u)
    fp.Mult(u, a)
    mx.Mult(u, b)
    torch.cuda.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    rel = np.linalg.norm(b - a) / np.linalg.norm(a)
    print("mixed vs fp64 rel L2", rel, flush=True)
    assert rel < 1e-6


@pytest.mark.parametrize("parabolic", [False, True])
def test_uniform_hex_grid_equals_box_kernels(cuda, parabolic):
    import torch
    kw = dict(quirks=False, blending=not parabolic, **(NS if parabolic else {}))
    d = Q.hex_grid(4, 3, 3, LO, HI)
    a = DGSEMOperator.hex(hmesh(d), 3, **kw)
    b = DGSEMOperator(3, 3, (4, 3, 3), LO, HI, **kw)
    X = cases.node_coords(3, 3, (4, 3, 3), LO, HI)
    u = to_dev(cases.random_smooth_state(X, 3, 9, L=[1.0, 0.8, 0.6]), cuda)
    oa, ob = torch.empty_like(u), torch.empty_like(u)
    a.Mult(u, oa)
    b.Mult(u, ob)
    torch.cuda.synchronize()
    r = rel_max(oa.cpu().numpy(), ob.cpu().numpy())
    print("hex vs box", parabolic, r, flush=True)
    assert r <= 1e-13


def test_relabelled_hex_mesh_same_rhs(cuda):
    import torch
    d = Q.hex_grid(3, 3, 4, LO, HI, perturb=0.2, seed=11)
    dr, perm, rots = Q.hex_relabel(d, seed=5)
    kw = dict(quirks=False, **NS)
    a, b = DGSEMOperator.hex(hmesh(d), 3, **kw), DGSEMOperator.hex(hmesh(dr), 3, **kw)
    Xa = hmesh(d).node_coords(3)
    Xb = hmesh(dr).node_coords(3)
    u = cases.random_smooth_state(Xa, 3, 8, L=[1.0, 0.8, 0.6])
    ub = np.zeros_like(u)
    match = []
    for k, p in enumerate(perm):
        dist = np.linalg.norm(Xb[:, k, :, None] - Xa[:, p, None, :], axis=0)
        idx = dist.argmin(axis=1)
        match.append(idx)
        ub[:, k, :] = u[:, p, idx]
    ua, ubb = to_dev(u, cuda), to_dev(ub, cuda)
    oa, ob = torch.empty_like(ua), torch.empty_like(ubb)
    a.Mult(ua, oa)
    b.Mult(ubb, ob)
    torch.cuda.synchronize()
    da = oa.cpu().numpy().reshape(5, a.NE, a.K)
    db = ob.cpu().numpy().reshape(5, b.NE, b.K)
    exp = np.stack([da[:, p, match[k]] for k, p in enumerate(perm)], axis=1)
    r = rel_max(db, exp)
    print("relabel", r, flush=True)
    assert r <= 1e-12
