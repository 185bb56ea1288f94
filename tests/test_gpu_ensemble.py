"""Batched ensemble: prandtl_ensemble_rk3_step against each case stepped alone.

The ensemble runs the same per-element arithmetic as prandtl_rk3_step, only
grouped into one launch per RK stage, so the bar is bitwise equality with
per-case steps on separately created operators; one case per kind is also
checked against the oracle's RK3 step to the 1e-12 parity bar.
"""
import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import DGSEMOperator, cases
from prandtl_amd.operator import Ensemble

pytestmark = pytest.mark.gpu


def small_cases():
    """All eight kernel variants (N 2..5 x blending), meshes whose element
    counts leave partial tiles."""
    out = []
    for i, N in enumerate((2, 3, 4, 5)):
        n = (3, 5, 6, 7)[i]
        out.append(dict(dim=2, order=N, nelem=(n, n + 1), lo=(-5.0, -5.0), hi=(5.0, 5.0), gamma=1.4,
                        precision="mixed", kind="iv", param=3.0 + 0.7 * i, seed=0))
        out.append(dict(dim=2, order=N, nelem=(n + 2, n + 1), lo=(-1.0, -1.0), hi=(1.0, 1.0), gamma=1.4,
                        precision="mixed", kind="kh", param=10.0 + 2.5 * i, seed=77 + i,
                        blending=True, alpha_max=0.5))
    return out


def states(case_list, dev):
    import torch
    return [torch.tensor(cases.ensemble_state(c).reshape(-1), dtype=torch.float64, device=dev)
            for c in case_list]


def step_sizes(case_list, scale=1.0):
    # CFL-like fixed steps (|v| + c <= 3 on these states)
    return [scale * 0.3 * min((c["hi"][r] - c["lo"][r]) / c["nelem"][r] for r in range(2))
            / (3.0 * (c["order"] + 1) ** 2) for c in case_list]


def run_pair(case_list, dev, dt_sets, graph_stream=True, mode=None):
    """Step an ensemble (launch form `mode`, default: as created) and separate
    per-case operators through the same sequence of step sizes; return both
    final state lists."""
    import torch
    ens = Ensemble(case_list)
    if mode is not None:
        ens.set_mode(mode)
    ops = [DGSEMOperator(**cases.operator_kwargs(c)) for c in case_list]
    ue = states(case_list, dev)
    us = [u.clone() for u in ue]
    s = torch.cuda.Stream(device=dev) if graph_stream else torch.cuda.default_stream(dev)
    with torch.cuda.stream(s):
        for dts in dt_sets:
            ens.Step(ue, dts)
            for op, u, dt in zip(ops, us, dts):
                op.Step(u, dt)
    torch.cuda.synchronize()
    ok = [op.check_state(u) for op, u in zip(ens.ops, ue)]
    ens.close()
    return ue, us, ok


def assert_same(ue, us, case_list):
    for i, (a, b) in enumerate(zip(ue, us)):
        c = case_list[i]
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()), \
            "case %d (%s N=%d %s): ensemble differs from the per-case step" % (
                i, c["kind"], c["order"], c["nelem"])


@pytest.mark.parametrize("graph_stream", [True, False], ids=["graph", "eager"])
def test_ensemble_equals_per_case_steps(cuda, graph_stream):
    cl = small_cases()
    d0, d1 = step_sizes(cl), step_sizes(cl, 0.5)
    # captured once, replayed, then replayed with new step sizes
    ue, us, ok = run_pair(cl, cuda, [d0, d0, d1, d1], graph_stream)
    assert all(ok)
    assert_same(ue, us, cl)


def test_ensemble_rebinds_new_state_pointers(cuda):
    import torch
    cl = small_cases()
    dts = step_sizes(cl)
    ens = Ensemble(cl)
    ops = [DGSEMOperator(**cases.operator_kwargs(c)) for c in cl]
    ua = states(cl, cuda)
    ref = [u.clone() for u in ua]
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        ens.Step(ua, dts)
        ub = [u.clone() for u in ua]          # new buffers holding the same states
        ens.Step(ub, dts)
        ens.Step(ub, dts)
        for op, u, dt in zip(ops, ref, dts):
            for _ in range(3):
                op.Step(u, dt)
    torch.cuda.synchronize()
    assert_same(ub, ref, cl)
    ens.close()


def test_ensemble_step_matches_oracle(cuda):
    import torch
    cl = [c for c in small_cases() if c["order"] == 4]
    dts = step_sizes(cl)
    ens = Ensemble(cl)
    us = states(cl, cuda)
    u0 = [u.cpu().numpy().copy() for u in us]
    ens.Step(us, dts)
    torch.cuda.synchronize()
    for c, u, u_init, dt in zip(cl, us, u0, dts):
        kw = cases.operator_kwargs(c)
        kw.pop("precision")
        orc = O.Oracle(threads=8, **kw)
        ref = orc.Step(u_init, dt)
        got = u.cpu().numpy()
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        assert err <= 1e-12, (c["kind"], err)
    ens.close()


def test_ensemble64_equals_per_case_steps(cuda):
    """The 64-case benchmark ensemble (meshes 8^2..64^2, N 2..5)."""
    cl = cases.ensemble_cases(64, 20261202)
    dts = step_sizes(cl)
    ue, us, ok = run_pair(cl, cuda, [dts, dts])
    assert all(ok)
    assert_same(ue, us, cl)


@pytest.mark.parametrize("mode", [1, 2, 4])
def test_persistent_ensemble_equals_per_case_steps(cuda, mode):
    """The persistent launch form (mode workgroups per CU over a device tile
    counter) gives bitwise the per-case steps."""
    cl = small_cases()
    d0 = step_sizes(cl)
    ue, us, ok = run_pair(cl, cuda, [d0, d0], True, mode)
    assert all(ok)
    assert_same(ue, us, cl)


def test_ensemble_autotune(cuda):
    import torch
    cl = cases.ensemble_cases(64, 20261202)
    dts = step_sizes(cl)
    ens = Ensemble(cl)
    us = states(cl, cuda)
    before = [u.clone() for u in us]
    mode, times = ens.autotune(us, dts)
    torch.cuda.synchronize()
    print("ensemble launch forms (us per stage):", ["%.1f" % t for t in times], "chosen", mode, flush=True)
    assert 0 <= mode <= 4 and times[mode] == min(times)
    assert all(torch.equal(a, b) for a, b in zip(us, before))   # the states are not modified
    # a step in the chosen form equals per-case steps
    ops = [DGSEMOperator(**cases.operator_kwargs(c)) for c in cl]
    ref = [u.clone() for u in us]
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        ens.Step(us, dts)
        for op, u, dt in zip(ops, ref, dts):
            op.Step(u, dt)
    torch.cuda.synchronize()
    assert_same(us, ref, cl)
    ens.close()
