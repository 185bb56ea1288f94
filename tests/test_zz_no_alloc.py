"""prandtl_rhs_noalloc: the RHS of a box with no device allocation (its
scratch is the library's fixed per-GPU array), called through the C ABI and
checked bitwise against prandtl_rhs of a context created from the same mesh
and params, on two calls in a row, the second with a smaller case (the
array keeps the first call's contents); and its refusal of a case whose
scratch exceeds the array."""
import ctypes
import math

import numpy as np
import pytest

from prandtl_amd import DGSEMOperator, cases
from prandtl_amd import _lib
from prandtl_amd.operator import _descs, lib

TWO_PI = 2.0 * math.pi
NS = dict(parabolic=True, mu=1.0 / 1600.0, Pr=0.71)

# the first call: 3-D Navier-Stokes at N = 3 on 8 x 6 x 5 elements (the
# line kernel and k_visc); the second: a smaller blended 2-D Euler case (the
# tile kernel), then a smaller 3-D Navier-Stokes case again
CALLS = [
    dict(dim=3, order=3, nelem=(8, 6, 5), lo=(0.0,) * 3, hi=(TWO_PI, 2.0, 3.0), quirks=False, **NS),
    dict(dim=2, order=4, nelem=(6, 5), lo=(-1.0, -1.0), hi=(1.0, 1.0), blending=True, alpha_max=0.5),
    dict(dim=3, order=3, nelem=(3, 3, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, quirks=True, **NS),
]


def state(cfg):
    X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    L = [cfg["hi"][r] - cfg["lo"][r] for r in range(cfg["dim"])]
    u = cases.random_smooth_state(X, cfg["dim"], 23, amp=0.3, L=L)
    if cfg.get("blending"):
        u = u * (1.0 + 0.2 * (np.arange(u.shape[1]) % 3 == 0))[None, :, None]
    return u


def noalloc(cfg, u, out):
    import torch
    m, p = _descs(**cfg)
    rc = lib().prandtl_rhs_noalloc(ctypes.byref(m), ctypes.byref(p), ctypes.c_void_p(u.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


@pytest.mark.gpu
def test_rhs_noalloc_matches_rhs(cuda):
    import torch
    for i, cfg in enumerate(CALLS):
        u = torch.tensor(state(cfg).reshape(-1), dtype=torch.float64, device=cuda)
        print("no-alloc call %d: dim %d order %d nelem %s" % (i, cfg["dim"], cfg["order"], cfg["nelem"]),
              flush=True)
        out = torch.full_like(u, float("nan"))
        rc = noalloc(cfg, u, out)
        assert rc == 0, lib().prandtl_last_error(None)
        op = DGSEMOperator(**cfg)
        ref = torch.empty_like(u)
        op.Mult(u, ref)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), "call %d differs from prandtl_rhs" % i
        op.close()


def test_rhs_noalloc_refuses_large_case():
    # 3-D Navier-Stokes, N = 3, 40^3 elements: ~80 B of scratch per node
    # (5 M nodes) is far beyond the 32 MiB array; refused before any GPU call
    cfg = dict(dim=3, order=3, nelem=(40, 40, 40), lo=(0.0,) * 3, hi=(1.0,) * 3, quirks=False, **NS)
    m, p = _descs(**cfg)
    rc = lib().prandtl_rhs_noalloc(ctypes.byref(m), ctypes.byref(p), ctypes.c_void_p(1), ctypes.c_void_p(2),
                                   None)
    assert rc == _lib.PRANDTL_ERR_ARG
    msg = lib().prandtl_last_error(None)
    msg = msg.decode() if isinstance(msg, bytes) else msg
    assert "scratch" in msg
