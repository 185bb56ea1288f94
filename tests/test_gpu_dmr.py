"""Double Mach reflection on the reference's own mesh (TestCases/Euler/2D/
DoubleMachReflection: DoubleMachReflection.msh, 36864 quadrilaterals, and
refined once as config.json's par_ref_levels 1 asks, 147456; N = 3,
blending), with its time-dependent top boundary (the post-shock state where
x < 1/6 + (y + 20 t)/sqrt(3), DoubleMachReflection.hpp:33-57, config.json
"time_dependent": true) evaluated on the device at each RK stage's time
(t, t + dt, t + dt/2), against the oracle evaluating the same (x, t)
function: RHS at two times, graph-captured RK3 steps and the device time
loop, each within 1e-12 relative max-norm, and the operator's time.  The
steps start from the discontinuous initial state at dt 2e-5: the configured
1e-4 loses positivity next to the shock foot in the first stage."""
import os

import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import QuadMesh

pytestmark = pytest.mark.gpu

MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes", "DoubleMachReflection.msh")
TOL = 1e-12


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


@pytest.fixture(scope="module", params=[0, 1], ids=["as_read", "refined_once"])
def dmr(request):
    c = cases.DMR
    mesh = QuadMesh.read(MESH, refine=request.param)
    op = DGSEMOperator.quad(mesh, c["order"], bc=c["bc"], blending=c["blending"], alpha_max=c["alpha_max"])
    nattr = int(mesh.bdr_attr.max())
    kinds, states = ["slip"] * nattr, [[1.0, 0.0, 0.0, 2.5]] * nattr
    for key, v in c["bc"].items():
        a = mesh.names[key]
        kinds[a - 1] = v if isinstance(v, str) else v[0]
        if not isinstance(v, str):
            states[a - 1] = list(v[1])
    d = dict(vert=mesh.vert, elem=mesh.elem, elem_coords=mesh.elem_coords, bdr=mesh.bdr, bdr_attr=mesh.bdr_attr,
             attr_bc=kinds, attr_state=states)
    orc = O.Oracle.quad(d, c["order"], blending=c["blending"], alpha_max=c["alpha_max"], threads=16)
    for name, sp in c["split"].items():
        op.set_boundary_split(name, sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
        orc.set_boundary_split(mesh.names[name], sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
    u0 = cases.dmr_state(mesh.node_coords(c["order"]))
    return mesh, op, orc, u0


def test_dmr_rhs_at_two_times(cuda, dmr):
    import torch
    mesh, op, orc, u0 = dmr
    du = to_dev(u0, cuda)
    out = torch.empty_like(du)
    for t in (0.0, 0.01):
        op.set_time(t)
        orc.set_time(t)
        op.Mult(du, out)
        torch.cuda.synchronize()
        ref = orc.Mult(u0)
        r = rel_max(out.cpu().numpy(), ref)
        a_gpu, _ = op.alpha()
        print("DMR RHS at t=%g: %.3e, alpha>0 on %d elements" % (t, r, int((orc.alpha() > 0).sum())), flush=True)
        assert r <= TOL
        assert np.abs(a_gpu - orc.alpha()).max() <= 1e-10
    # the top boundary state moved between the two times
    op.set_time(0.0)
    op.Mult(du, out)
    r0 = out.clone()
    op.set_time(0.01)
    op.Mult(du, out)
    torch.cuda.synchronize()
    assert not torch.equal(r0, out)


def test_dmr_steps_graph_and_device_loop(cuda, dmr):
    import torch
    mesh, op, orc, u0 = dmr
    dt = cases.DMR["dt_start"]
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        op.set_time(0.0)
        du = to_dev(u0, cuda)
        for _ in range(3):
            op.Step(du, dt)
    torch.cuda.synchronize()
    orc.set_time(0.0)
    ref = u0.reshape(-1).copy()
    for _ in range(3):
        ref = orc.Step(ref, dt)
    assert np.isfinite(ref).all()
    r = rel_max(du.cpu().numpy(), ref)
    print("DMR 3 RK3 steps (graph): %.3e, time %.17g / %.17g" % (r, op.time, orc.time), flush=True)
    assert r <= TOL
    assert op.time == orc.time
    # the device loop continues from there, reading the loop's time
    with torch.cuda.stream(s):
        t, n = op.Run(du, op.time, 1e30, dt=dt, max_steps=4)
    torch.cuda.synchronize()
    for _ in range(4):
        ref = orc.Step(ref, dt)
    r = rel_max(du.cpu().numpy(), ref)
    print("DMR device loop 4 more steps: %.3e, t %.17g / %.17g" % (r, t, orc.time), flush=True)
    assert n == 4 and r <= TOL
    assert abs(t - orc.time) <= 1e-15 and op.time == t
    assert op.check_state(du)


def test_split_set_after_steps_uses_the_advanced_time(cuda, dmr):
    """Steps taken before any time-dependent state exists advance only the
    host copy of the time; setting the split state then brings the device
    time to it, so the next RHS is evaluated at the advanced time.  25 steps
    move the split line by 5.8e-3 along the top boundary, more than the node
    spacing there, so the RHS at the old time differs (checked)."""
    import torch
    mesh, _, orc, u0 = dmr
    c = cases.DMR
    op = DGSEMOperator.quad(mesh, c["order"], bc=c["bc"], blending=c["blending"], alpha_max=c["alpha_max"])
    dt = c["dt_start"]
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        op.set_time(0.003)
        du = to_dev(u0, cuda)
        for _ in range(25):
            op.Step(du, dt)
    torch.cuda.synchronize()
    t = op.time
    assert abs(t - (0.003 + 25 * dt)) <= 1e-15
    for name, sp in c["split"].items():
        op.set_boundary_split(name, sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
    out = torch.empty_like(du)
    with torch.cuda.stream(s):
        op.Mult(du, out)
    torch.cuda.synchronize()
    state = du.cpu().numpy()
    orc.set_time(t)
    ref = orc.Mult(state)
    r = rel_max(out.cpu().numpy(), ref)
    orc.set_time(0.003)
    moved = rel_max(orc.Mult(state), ref)
    print("DMR RHS after split set at t=%.17g: %.3e (RHS at the start time differs by %.3e)" % (t, r, moved),
          flush=True)
    assert r <= TOL
    assert moved > 1e-3


def test_split_state_refusals(cuda):
    """prandtl_set_boundary_split refuses what it does not serve: a
    Cartesian box, a Navier-Stokes mesh, an attribute out of range and a
    state with non-positive pressure; none of them changes the operator."""
    import quad_meshes as Q
    from prandtl_amd import PrandtlError, _lib
    sp = cases.DMR_TOP_SPLIT
    box = DGSEMOperator(dim=2, order=3, nelem=(4, 4), lo=(0.0, 0.0), hi=(1.0, 1.0))
    with pytest.raises(PrandtlError) as e:
        box.set_boundary_split(1, sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
    assert e.value.code == _lib.PRANDTL_ERR_UNSUPPORTED
    grid = Q.grid(4, 4, periodic=(False, False))
    bc = {1: ("specified-state", cases.DMR_PRE), 2: "supersonic-outflow", 3: "slip", 4: "slip"}
    m = QuadMesh(**grid)
    ns = DGSEMOperator.quad(m, 3, bc={1: "no-slip-adiabatic", 2: "no-slip-adiabatic", 3: "no-slip-adiabatic",
                                      4: "no-slip-adiabatic"}, parabolic=True, mu=0.01, Pr=0.72)
    with pytest.raises(PrandtlError) as e:
        ns.set_boundary_split(1, sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
    assert e.value.code == _lib.PRANDTL_ERR_UNSUPPORTED
    op = DGSEMOperator.quad(m, 3, bc=bc)
    for attr in (0, 5):
        with pytest.raises(PrandtlError) as e:
            op.set_boundary_split(attr, sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
        assert e.value.code == _lib.PRANDTL_ERR_ARG
    cold = np.array([1.0, 0.0, 0.0, 0.0])   # zero pressure
    with pytest.raises(PrandtlError) as e:
        op.set_boundary_split(1, sp["normal"], sp["offset"], sp["speed"], sp["below"], cold)
    assert e.value.code == _lib.PRANDTL_ERR_ARG
