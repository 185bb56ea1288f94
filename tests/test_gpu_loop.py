"""The time loop on the device (prandtl_run) against the host loop of
single steps (TimeLoop over prandtl_rk3_step and prandtl_max_wave_speed).

Both form dt from the same running maximum wave speed with the same
arithmetic, so the bar is the same step count and time, and bitwise equal
states."""
import math

import numpy as np
import pytest

from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import TimeLoop

pytestmark = pytest.mark.gpu

IV = dict(dim=2, order=3, nelem=(16, 16), lo=(-5.0, -5.0), hi=(5.0, 5.0))
TGV = dict(dim=3, order=3, nelem=(4, 4, 4), lo=(0.0,) * 3, hi=(2 * math.pi,) * 3,
           parabolic=True, mu=1.0 / 1600.0, Pr=0.71, quirks=False)


def state(cfg, dev):
    import torch
    X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    u = cases.isentropic_vortex(X) if cfg["dim"] == 2 else cases.taylor_green(X)
    return torch.tensor(u.reshape(-1), dtype=torch.float64, device=dev)


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_device_loop_variable_dt_matches_host_loop(cuda, graph):
    import torch
    s = torch.cuda.Stream(device=cuda) if graph else torch.cuda.default_stream(cuda)
    with torch.cuda.stream(s):
        a, b = DGSEMOperator(**IV), DGSEMOperator(**IV)
        ua, ub = state(IV, cuda), state(IV, cuda)
        t_host, n_host = TimeLoop(a, 1.0, cfl=0.5, variable_dt=True).run(ua)
        t_dev, n_dev = b.Run(ub, 0.0, 1.0, cfl=0.5)
    torch.cuda.synchronize()
    print("host %d steps t=%.17g, device %d steps t=%.17g" % (n_host, t_host, n_dev, t_dev), flush=True)
    assert n_dev == n_host and t_dev == t_host
    assert torch.equal(ua, ub)
    assert b.GetMaxCharSpeed() == a.GetMaxCharSpeed()


def test_device_loop_fixed_dt_and_max_steps(cuda):
    import torch
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        a, b = DGSEMOperator(**TGV), DGSEMOperator(**TGV)
        ua, ub = state(TGV, cuda), state(TGV, cuda)
        dt = 1e-3
        t_host, n_host = TimeLoop(a, 0.05, dt=dt).run(ua)
        t_dev, n_dev = b.Run(ub, 0.0, 0.05, dt=dt)
        uc = state(TGV, cuda)
        t5, n5 = DGSEMOperator(**TGV).Run(uc, 0.0, 1.0, dt=dt, max_steps=5)
    torch.cuda.synchronize()
    assert n_dev == n_host == 50 and t_dev == t_host
    assert torch.equal(ua, ub)
    assert n5 == 5 and t5 == sum([dt] * 5)


@pytest.mark.parametrize("every", [1, 5])
def test_device_loop_state_check_stops(cuda, every):
    """A state with one non-physical node (negative density, set on the
    host) stops the loop at the first check: step `every`."""
    import torch
    from prandtl_amd import PrandtlError, _lib
    op = DGSEMOperator(**IV)
    u = state(IV, cuda)
    u[7] = -1.0                       # density of one node
    with pytest.raises(PrandtlError) as ei:
        op.Run(u, 0.0, 1.0, dt=1e-3, check_every=every)
    torch.cuda.synchronize()
    assert ei.value.code == _lib.PRANDTL_ERR_STATE
    assert ei.value.steps == every and ei.value.t == sum([1e-3] * every)
    good = DGSEMOperator(**IV)
    t, n = good.Run(state(IV, cuda), 0.0, 0.02, dt=1e-3, check_every=1)
    assert n == 20


def test_checkpoint_resume_bitwise(cuda, tmp_path):
    """10 steps, checkpoint, 10 more; a new operator that loads the
    checkpoint and takes the same 10 steps ends bitwise equal (fixed dt:
    a restart starts a new running wave speed, as the reference's does)."""
    import torch
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        a = DGSEMOperator(**TGV)
        u = state(TGV, cuda)
        t, n = a.Run(u, 0.0, 1.0, dt=1e-3, max_steps=10)
        a.save_checkpoint(u, tmp_path, n, t)
        t_end, _ = a.Run(u, t, 1.0, dt=1e-3, max_steps=10)
        b = DGSEMOperator(**TGV)
        v = torch.empty_like(u)
        t_load = b.load_checkpoint(v, tmp_path, 10)
        t_end2, _ = b.Run(v, t_load, 1.0, dt=1e-3, max_steps=10)
    torch.cuda.synchronize()
    assert t_load == t and t_end2 == t_end
    assert torch.equal(u, v)


@pytest.mark.parametrize("cfg", [IV, TGV], ids=["2d", "3d"])
def test_primitives(cuda, cfg):
    import torch
    op = DGSEMOperator(**cfg)
    u = state(cfg, cuda)
    prim = op.primitives(u)
    torch.cuda.synchronize()
    U = u.cpu().numpy().reshape(op.nv, op.NE, op.K)
    P = prim.cpu().numpy()
    dim = cfg["dim"]
    vel = [U[1 + d] / U[0] for d in range(dim)]
    vsq = vel[0] * vel[0]
    for d in range(1, dim):
        vsq = vsq + vel[d] * vel[d]
    p = 0.4 * (U[dim + 1] - 0.5 * U[0] * vsq)
    assert np.array_equal(P[0], U[0])
    for d in range(dim):
        assert np.array_equal(P[1 + d], vel[d])
    assert np.abs(P[dim + 1] - p).max() <= 1e-15 * np.abs(p).max()


def test_checkpoint_refused_on_another_mesh_of_the_same_size(cuda, tmp_path):
    """Two periodic quadrilateral meshes with the same element count and
    order (a uniform grid and a perturbed one): a checkpoint written on one
    does not load into the other (the header's geometry fingerprint), and
    still loads into an operator on its own mesh."""
    import torch
    import quad_meshes as Q
    from prandtl_amd import PrandtlError, _lib
    from prandtl_amd.operator import QuadMesh
    m1 = QuadMesh(**Q.grid(6, 5, lo=(-5.0, -5.0), hi=(5.0, 5.0)))
    m2 = QuadMesh(**Q.grid(6, 5, lo=(-5.0, -5.0), hi=(5.0, 5.0), perturb=0.2, seed=4))
    a, b, c = DGSEMOperator.quad(m1, 3), DGSEMOperator.quad(m2, 3), DGSEMOperator.quad(m1, 3)
    u = torch.tensor(cases.isentropic_vortex(m1.node_coords(3)).reshape(-1), dtype=torch.float64, device=cuda)
    a.save_checkpoint(u, tmp_path, 4, 0.25)
    v = torch.empty_like(u)
    with pytest.raises(PrandtlError) as e:
        b.load_checkpoint(v, tmp_path, 4)
    assert e.value.code == _lib.PRANDTL_ERR_IO and "another mesh" in str(e.value)
    assert c.load_checkpoint(v, tmp_path, 4) == 0.25
    torch.cuda.synchronize()
    assert torch.equal(u, v)
