"""The line kernel (k_rhs_ln, one thread per tensor line) of periodic 3-D
boxes against the FP64 oracle and against the tile kernel (k_rhs).

The two kernels restate the same operator (DGSEMIntegrator.cpp:151-1196)
with different summation orders and a reciprocal-based two-point flux, so
they agree to rounding.  Bars: the line kernel's RHS within 1e-12 relative
max-norm of the oracle (or, where FP64 rounding of the oracle itself exceeds
that, at most twice the oracle's distance from an extended-precision RHS);
line and tile within 1e-13 of each other on states whose RHS is not
cancellation-dominated; both blended (alpha, eta) outputs within 1e-12.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import DGSEMOperator, PrandtlError, cases

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi
NS = dict(parabolic=True, mu=1.0 / 1600.0, Pr=0.71)


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def rhs(op, u, dev):
    import torch
    du = to_dev(u, dev)
    out = torch.empty_like(du)
    op.Mult(du, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def blended_state(X, L):
    u = cases.random_smooth_state(X, 3, 5, amp=0.3, L=L)
    # element-wise density/pressure jumps so some elements blend
    return u * (1.0 + 0.2 * (np.arange(u.shape[1]) % 3 == 0))[None, :, None]


CASES = {
    "n%d_euler" % N: dict(order=N) for N in (2, 3, 4)
}
CASES.update({"n%d_ns%s" % (N, "_faithful" if q else ""): dict(order=N, quirks=q, **NS)
              for N in (2, 3, 4) for q in (False, True)})
CASES.update({"n%d_blend" % N: dict(order=N, blending=True, alpha_max=0.5) for N in (2, 3, 4)})
CASES["n3_ns_blend_faithful"] = dict(order=3, quirks=True, blending=True, alpha_max=0.1, **NS)
CASES["n3_ns_blend"] = dict(order=3, quirks=False, blending=True, alpha_max=0.1, **NS)
CASES["n3_euler_ec"] = dict(order=3, ec_surface=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_line_kernel_parity(cuda, name):
    extra = dict(CASES[name])
    N = extra.pop("order")
    nelem = (4, 3, 3)
    lo, hi = (0.0,) * 3, (TWO_PI, 2.0, 3.0)
    cfg = dict(dim=3, order=N, nelem=nelem, lo=lo, hi=hi, **extra)
    op = DGSEMOperator(**cfg)
    # the default: line at N = 2, 3 and blended N = 4, tile at unblended N = 4
    assert op.algorithm == ("tile" if N == 4 and not cfg.get("blending") else "line")
    op.set_algorithm("line")
    tile = DGSEMOperator(**cfg)
    tile.set_algorithm("tile")
    orc = O.Oracle(threads=8, **cfg)
    X = cases.node_coords(3, N, nelem, lo, hi)
    L = [hi[r] - lo[r] for r in range(3)]
    states = {"random": cases.random_smooth_state(X, 3, 17, L=L)}
    if cfg.get("blending"):
        states["jumps"] = blended_state(X, L)
    for sname, u in states.items():
        a, b = rhs(op, u, cuda), rhs(tile, u, cuda)
        ref = orc.Mult(u)
        r_or, r_lt = rel_max(a, ref), rel_max(a, b)
        print("line %s %s: vs oracle %.3e, vs tile %.3e" % (name, sname, r_or, r_lt), flush=True)
        if r_or > 1e-12:
            ext = orc.Mult_extended(u)
            e_gpu, e_orc = rel_max(a, ext), rel_max(ref, ext)
            assert e_orc > 1e-12 and e_gpu <= 2.0 * e_orc, (sname, r_or, e_gpu, e_orc)
        assert r_lt <= 1e-13, (sname, r_lt)
        if cfg.get("blending"):
            al, el = op.alpha()
            at, et = tile.alpha()
            assert np.abs(el - orc.eta()).max() <= 1e-12
            assert np.abs(al - orc.alpha()).max() <= 1e-10
            assert np.array_equal(al > 0, at > 0)
            if sname == "jumps":
                assert (al > 0).sum() > 0
    assert abs(op.GetMaxCharSpeed() - orc.GetMaxCharSpeed()) <= 1e-14 * orc.GetMaxCharSpeed()
    # one RK3 step (graph-captured) against the oracle's
    import torch
    u = states["random"]
    du = to_dev(u, cuda)
    op.Step(du, 1e-3)
    torch.cuda.synchronize()
    ref = orc.Step(u.reshape(-1).copy(), 1e-3)
    assert rel_max(du.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("N,nelem", [(3, (5, 3, 3)), (4, (5, 2, 2)), (2, (5, 2, 2))])
@pytest.mark.parametrize("ns", [True, False], ids=["ns", "euler"])
def test_line_kernel_ragged_last_workgroup(cuda, N, nelem, ns):
    """Element counts the line kernel's elements per workgroup (N = 2, 3, 4:
    7, 4, 3) do not divide, so the last workgroup is partly empty and its
    node passes, line ends and assembly loads are guarded: RHS and one RK3
    stage sequence against the oracle."""
    import torch
    lo, hi = (0.0,) * 3, (TWO_PI, 2.0, 3.0)
    cfg = dict(dim=3, order=N, nelem=nelem, lo=lo, hi=hi, quirks=False, **(NS if ns else {}))
    op = DGSEMOperator(**cfg)
    op.set_algorithm("line")
    orc = O.Oracle(threads=8, **cfg)
    X = cases.node_coords(3, N, nelem, lo, hi)
    u = cases.random_smooth_state(X, 3, 23, L=[hi[r] - lo[r] for r in range(3)])
    a = rhs(op, u, cuda)
    ref = orc.Mult(u)
    r = rel_max(a, ref)
    du = to_dev(u, cuda)
    op.Step(du, 1e-3)
    torch.cuda.synchronize()
    r_step = rel_max(du.cpu().numpy(), orc.Step(u.reshape(-1).copy(), 1e-3))
    print("line ragged N=%d %s %s: RHS %.3e, step %.3e" % (N, nelem, "ns" if ns else "euler", r, r_step), flush=True)
    assert r <= 1e-12 and r_step <= 1e-12


def test_line_kernel_not_for_other_meshes(cuda):
    # open boxes and 2-D boxes keep the tile kernel and refuse the line one
    op2 = DGSEMOperator(dim=2, order=3, nelem=(4, 4), lo=(0.0, 0.0), hi=(1.0, 1.0))
    assert op2.algorithm == "tile"
    with pytest.raises(PrandtlError):
        op2.set_algorithm("line")
    op3 = DGSEMOperator(dim=3, order=7, nelem=(2, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3)
    assert op3.algorithm == "tile"


def test_line_kernel_graph_and_loop(cuda):
    """Graph-captured steps equal eager ones bitwise, and the device loop equals
    host steps bitwise, with the line kernel."""
    import torch
    cfg = dict(dim=3, order=3, nelem=(4, 4, 4), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, quirks=False, **NS)
    X = cases.node_coords(3, 3, cfg["nelem"], cfg["lo"], cfg["hi"])
    u0 = to_dev(cases.taylor_green(X), cuda)
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        a = DGSEMOperator(**cfg)
        b = DGSEMOperator(**cfg)
        b.set_graph(False)
        ua, ub = u0.clone(), u0.clone()
        for _ in range(4):
            a.Step(ua, 2e-3)
            b.Step(ub, 2e-3)
        c = DGSEMOperator(**cfg)
        uc = u0.clone()
        _, n = c.Run(uc, 0.0, 1e30, dt=2e-3, max_steps=4)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)
    assert n == 4 and torch.equal(uc, ua)


def test_mixed_mfma_viscous_kernel(cuda):
    """The mixed form's viscous kernel on the FP32 matrix cores (k_visc_mx, line
    algorithm, 3-D N = 3) against the FP32 VALU kernel of the tile algorithm
    and against the FP64 oracle: gradients within FP32 rounding of the
    oracle's, RHS within 1e-6 relative L2 of the FP64 GPU RHS."""
    import torch
    cfg = dict(dim=3, order=3, nelem=(4, 3, 5), lo=(0.0,) * 3, hi=(TWO_PI, 2.0, 3.0), quirks=False, **NS)
    X = cases.node_coords(3, 3, cfg["nelem"], cfg["lo"], cfg["hi"])
    u = cases.random_smooth_state(X, 3, 41, L=[TWO_PI, 2.0, 3.0])
    line = DGSEMOperator(precision="mixed", **cfg)
    tile = DGSEMOperator(precision="mixed", **cfg)
    tile.set_algorithm("tile")
    fp64 = DGSEMOperator(precision="fp64", **cfg)
    assert line.algorithm == "line"
    a, b, ref = rhs(line, u, cuda), rhs(tile, u, cuda), rhs(fp64, u, cuda)
    l2 = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    print("mixed line vs fp64 %.3e, mixed tile vs fp64 %.3e, line vs tile %.3e" % (
        l2(a, ref), l2(b, ref), l2(a, b)), flush=True)
    assert l2(a, ref) <= 1e-6 and l2(b, ref) <= 1e-6
    orc = O.Oracle(threads=8, **cfg)
    orc.Mult(u)
    g_ref = orc.gradients()
    g = line.gradients(to_dev(u, cuda))
    torch.cuda.synchronize()
    rg = rel_max(g.cpu().numpy(), g_ref)
    print("mixed line gradients vs oracle %.3e" % rg, flush=True)
    assert rg <= 1e-5
    # graph-captured steps of the mixed line form equal eager ones bitwise
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        p = DGSEMOperator(precision="mixed", **cfg)
        q = DGSEMOperator(precision="mixed", **cfg)
        q.set_graph(False)
        up, uq = to_dev(u, cuda), to_dev(u, cuda)
        for _ in range(3):
            p.Step(up, 1e-3)
            q.Step(uq, 1e-3)
    torch.cuda.synchronize()
    assert torch.equal(up, uq)
