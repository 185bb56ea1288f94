"""Launch shapes of the RHS kernel and the autotuner that picks among them.

Every shape runs the same per-node arithmetic in the same order, so the bar
is bitwise equality with the default shape, for the RHS and for RK3 steps
replayed from a graph.
"""
import math

import numpy as np
import pytest

from prandtl_amd import DGSEMOperator, PrandtlError, cases
from prandtl_amd import _lib

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi

CONFIGS = {
    "2d_n%d%s" % (N, "_blend" if b else ""): dict(dim=2, order=N, nelem=(6, 5), lo=(-1.0, -1.0), hi=(1.0, 1.0),
                                                  blending=b, alpha_max=0.5)
    for N in (2, 3, 4, 5) for b in (False, True)
}
CONFIGS.update({
    "3d_n%d_ns" % N: dict(dim=3, order=N, nelem=(4, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3,
                          parabolic=True, mu=1.0 / 1600.0, Pr=0.71, quirks=False)
    for N in (2, 3, 4)
})
CONFIGS["3d_n3_blend"] = dict(dim=3, order=3, nelem=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, blending=True,
                              alpha_max=0.5)
CONFIGS["3d_n3_ns_mixed"] = dict(dim=3, order=3, nelem=(4, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3,
                                 parabolic=True, mu=1.0 / 1600.0, Pr=0.71, quirks=False, precision="mixed")
CONFIGS["3d_n4_ns_faithful"] = dict(dim=3, order=4, nelem=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3,
                                    parabolic=True, mu=1.0 / 1600.0, Pr=0.71, quirks=True)


def state(cfg, dev):
    import torch
    X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    L = [cfg["hi"][r] - cfg["lo"][r] for r in range(cfg["dim"])]
    u = cases.random_smooth_state(X, cfg["dim"], 11, amp=0.3, L=L)
    if cfg.get("blending"):
        # element-wise jumps so some elements blend
        u = u * (1.0 + 0.2 * (np.arange(u.shape[1]) % 3 == 0))[None, :, None]
    return torch.tensor(u.reshape(-1), dtype=torch.float64, device=dev)


def shapes_of(op, kernel="rhs"):
    out = []
    for s in range(_lib.PRANDTL_SHAPES):
        try:
            op.set_shape(kernel, s)
        except PrandtlError as e:
            assert e.code == _lib.PRANDTL_ERR_UNSUPPORTED
            continue
        out.append(s)
    if out:
        op.set_shape(kernel, 0)
    return out


def combos(op, cfg):
    """(rhs shape, visc shape) pairs: every rhs shape with the default visc
    shape, and every visc shape with the default rhs shape."""
    rs = shapes_of(op, "rhs")
    vs = shapes_of(op, "visc") if cfg.get("parabolic") else []
    return [(r, 0) for r in rs] + [(0, v) for v in vs[1:]]


def set_combo(op, cfg, combo):
    op.set_shape("rhs", combo[0])
    if cfg.get("parabolic"):
        op.set_shape("visc", combo[1])


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_shapes_bitwise_equal(cuda, name):
    import torch
    cfg = CONFIGS[name]
    op = DGSEMOperator(**cfg)
    u = state(cfg, cuda)
    ref = torch.empty_like(u)
    op.Mult(u, ref)
    if cfg.get("quirks") and cfg.get("parabolic"):
        assert shapes_of(op, "rhs") == [0]
    if not cfg.get("parabolic"):
        assert shapes_of(op, "visc") == []
    for combo in combos(op, cfg)[1:]:
        set_combo(op, cfg, combo)
        out = torch.empty_like(u)
        op.Mult(u, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), "%s: shapes %s differ from the default" % (name, combo)


@pytest.mark.parametrize("name", ["2d_n4_blend", "3d_n3_ns", "3d_n3_ns_mixed"])
def test_shapes_graph_steps_equal(cuda, name):
    import torch
    cfg = CONFIGS[name]
    op = DGSEMOperator(**cfg)
    u0 = state(cfg, cuda)
    cs = combos(op, cfg)
    results = []
    stream = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(stream):
        for combo in cs:
            set_combo(op, cfg, combo)
            u = u0.clone()
            for _ in range(3):
                op.Step(u, 1e-3)
            results.append(u)
    torch.cuda.synchronize()
    for combo, r in zip(cs[1:], results[1:]):
        assert torch.equal(r, results[0]), "%s: shapes %s steps differ" % (name, combo)


def test_autotune(cuda):
    import torch
    cfg = CONFIGS["3d_n3_ns"]
    op = DGSEMOperator(**cfg)
    u = state(cfg, cuda)
    ref = torch.empty_like(u)
    op.Mult(u, ref)
    shapes, times = op.autotune(u)
    for kernel in ("rhs", "visc"):
        valid = shapes_of(op, kernel)
        t = times[kernel]
        assert shapes[kernel] in valid
        assert len(t) == _lib.PRANDTL_SHAPES
        assert all((t[s] is not None) == (s in valid) for s in range(_lib.PRANDTL_SHAPES))
        assert t[shapes[kernel]] == min(v for v in t if v is not None)
    op.set_shape("rhs", shapes["rhs"])
    op.set_shape("visc", shapes["visc"])
    out = torch.empty_like(u)
    op.Mult(u, out)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
