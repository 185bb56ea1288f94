"""System-level checks on the GPU against the oracle.

 - C1: 2-D isentropic vortex, 16x16 N=3, CFL 0.5 with the reference's
   variable-dt policy to t = 1; L2 error against the exact solution from the
   GPU run and from the oracle run agree to 3 significant digits (L2 by the
   LGL nodal quadrature sum w_i J_i |u - u_exact|^2 over all variables).
 - C2 mixed form: kinetic-energy dissipation rate -dE_k/dt of the mixed
   (FP32 viscous) RHS within 1e-4 relative of the FP64 oracle, at t = 0 and
   after 40 RK3 steps of each.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import TimeLoop

pytestmark = pytest.mark.gpu


def to_dev(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def l2_error(U, X, t, WJ, lo, hi):
    Ue = cases.isentropic_vortex_exact(X, t, lo, hi)
    return math.sqrt(float(np.sum(WJ * np.sum((U - Ue) ** 2, axis=0))))


def test_isentropic_vortex_l2_matches_oracle(cuda):
    import torch
    cfg = dict(dim=2, order=3, nelem=(16, 16), lo=(-5.0, -5.0), hi=(5.0, 5.0))
    op = DGSEMOperator(**cfg)
    orc = O.Oracle(threads=8, **cfg)
    X = cases.node_coords(2, 3, cfg["nelem"], cfg["lo"], cfg["hi"])
    u0 = cases.isentropic_vortex(X)
    WJ = cases.quadrature_weights(2, 3, cfg["nelem"], cfg["lo"], cfg["hi"])
    du = to_dev(u0, cuda)
    t_gpu, n_gpu = TimeLoop(op, 1.0, cfl=0.5, variable_dt=True).run(du)
    torch.cuda.synchronize()
    ug = du.cpu().numpy().reshape(u0.shape)
    # the same loop on the oracle (Simulation.cpp:893-907, 1007-1022)
    u = u0.reshape(-1).copy()
    orc.Mult(u)
    hmin = orc.hmin()
    dt = 0.5 * hmin / (orc.GetMaxCharSpeed() * 16)
    t, n = 0.0, 0
    while True:
        dtr = min(dt, 1.0 - t)
        u = orc.Step(u, dtr)
        t += dtr
        dt = 0.5 * hmin / (orc.GetMaxCharSpeed() * 16)
        n += 1
        if t >= 1.0 - 1e-8 * dt:
            break
    uo = u.reshape(u0.shape)
    e_gpu = l2_error(ug, X, t_gpu, WJ, cfg["lo"], cfg["hi"])
    e_orc = l2_error(uo, X, t, WJ, cfg["lo"], cfg["hi"])
    print("IV L2 error gpu %.6e oracle %.6e steps %d/%d" % (e_gpu, e_orc, n_gpu, n), flush=True)
    assert n_gpu == n and abs(t_gpu - t) < 1e-14
    assert float("%.3g" % e_gpu) == float("%.3g" % e_orc)
    assert abs(e_gpu - e_orc) <= 1e-9 * e_orc
    assert e_orc < 1e-2


def test_tgv_mixed_dissipation_rate(cuda):
    import torch
    L = 2 * math.pi
    base = dict(dim=3, order=3, nelem=(16, 16, 16), lo=(0.0,) * 3, hi=(L,) * 3,
                parabolic=True, mu=1.0 / 1600.0, Pr=0.71, quirks=False)
    opm = DGSEMOperator(precision="mixed", **base)
    opd = DGSEMOperator(**base)
    orc = O.Oracle(threads=8, **base)
    X = cases.node_coords(3, 3, base["nelem"], base["lo"], base["hi"])
    u0 = cases.taylor_green(X)
    WJ = cases.quadrature_weights(3, 3, base["nelem"], base["lo"], base["hi"])
    vol = L ** 3

    def rate(op, u):
        du = to_dev(u, cuda)
        out = torch.empty_like(du)
        op.Mult(du, out)
        torch.cuda.synchronize()
        return cases.kinetic_energy_rate(u, out.cpu().numpy().reshape(u.shape), WJ, vol)

    r_ref = cases.kinetic_energy_rate(u0, orc.Mult(u0).reshape(u0.shape), WJ, vol)
    r_mix, r_dbl = rate(opm, u0), rate(opd, u0)
    print("t=0 -dEk/dt oracle %.10e fp64 %.10e mixed %.10e rel %.2e" % (
        -r_ref, -r_dbl, -r_mix, abs(r_mix - r_ref) / abs(r_ref)), flush=True)
    assert abs(r_mix - r_ref) <= 1e-4 * abs(r_ref)
    assert abs(r_dbl - r_ref) <= 1e-9 * abs(r_ref)
    # 40 steps of each form, then compare the rates of the evolved states
    dt = 2e-3
    dm, dd = to_dev(u0, cuda), to_dev(u0, cuda)
    uo = u0.reshape(-1).copy()
    for _ in range(40):
        opm.Step(dm, dt)
        opd.Step(dd, dt)
        uo = orc.Step(uo, dt)
    torch.cuda.synchronize()
    um = dm.cpu().numpy().reshape(u0.shape)
    ud = dd.cpu().numpy().reshape(u0.shape)
    uo = uo.reshape(u0.shape)
    r_ref = cases.kinetic_energy_rate(uo, orc.Mult(uo).reshape(u0.shape), WJ, vol)
    r_mix = rate(opm, um)
    ek = [cases.kinetic_energy(x, WJ, vol) for x in (uo, ud, um)]
    print("t=0.08 -dEk/dt oracle %.10e mixed %.10e rel %.2e; Ek %s" % (
        -r_ref, -r_mix, abs(r_mix - r_ref) / abs(r_ref), ek), flush=True)
    assert abs(r_mix - r_ref) <= 1e-4 * abs(r_ref)
    assert abs(ek[2] - ek[0]) <= 1e-8 * ek[0]
    assert np.abs(ud - uo).max() <= 1e-12 * np.abs(uo).max()
