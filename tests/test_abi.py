"""The C-ABI library loads, exports every entry point include/prandtl_hip.h
declares, and rejects bad arguments with the documented codes.  No compute
calls: these run without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from prandtl_amd import _lib
from prandtl_amd._lib import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "prandtl_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(prandtl_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    names = declared()
    assert len(names) >= 11
    L = lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_lgl_matches_oracle():
    import oracle_lib as O
    import prandtl_amd
    for N in range(1, 9):
        x, w = prandtl_amd.lgl_nodes(N)
        xo, wo = O.lgl(N)
        assert np.array_equal(x, xo) and np.array_equal(w, wo)


def _create(mesh_kw, prm_kw):
    m = _lib.MeshDesc()
    m.dim, m.order = mesh_kw.get("dim", 2), mesh_kw.get("order", 3)
    for r in range(3):
        m.nelem[r] = mesh_kw.get("nelem", (4, 4, 4))[r]
        m.lo[r], m.hi[r] = 0.0, mesh_kw.get("hi", 1.0)
        for side in range(2):
            m.bc[r][side] = mesh_kw.get("bc", ((0, 0),) * 3)[r][side]
            st = mesh_kw.get("bc_state", (1.0, 0.0, 0.0, 2.5, 0.0))
            for v in range(5):
                m.bc_state[r][side][v] = st[v]
    p = _lib.Params()
    p.gamma, p.Pr, p.mu = prm_kw.get("gamma", 1.4), 0.72, 0.0
    p.alpha_max, p.alpha_min = 0.5, 0.001
    p.parabolic = prm_kw.get("parabolic", 0)
    p.quirks = prm_kw.get("quirks", 1)
    p.precision = prm_kw.get("precision", 0)
    h = ctypes.c_void_p()
    rc = lib().prandtl_create(ctypes.byref(m), ctypes.byref(p), ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize("mesh_kw,prm_kw,code", [
    (dict(dim=4), {}, _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(dim=0), {}, _lib.PRANDTL_ERR_UNSUPPORTED),
    # 1-D: Euler only, orders 1..7; a 1-D side needs an outer state with
    # positive density and pressure like any other
    (dict(dim=1), dict(parabolic=1), _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(dim=1, order=8), {}, _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(dim=1, bc=((5, 5), (0, 0), (0, 0)), bc_state=(1.0, 0.0, -2.5, 0.0, 0.0)), {}, _lib.PRANDTL_ERR_ARG),
    (dict(dim=2, order=6), {}, _lib.PRANDTL_ERR_UNSUPPORTED),
    ({}, dict(parabolic=1, precision=1), _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(dim=3, order=1), {}, _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(nelem=(1, 4, 4)), {}, _lib.PRANDTL_ERR_ARG),
    (dict(hi=-1.0), {}, _lib.PRANDTL_ERR_ARG),
    ({}, dict(gamma=1.0), _lib.PRANDTL_ERR_ARG),
    ({}, dict(precision=7), _lib.PRANDTL_ERR_UNSUPPORTED),
    # boundary tables: one periodic side, unknown type, Navier-Stokes walls,
    # an outer state with negative pressure or density
    (dict(bc=((1, 0), (0, 0), (0, 0))), {}, _lib.PRANDTL_ERR_ARG),
    (dict(bc=((7, 7), (0, 0), (0, 0))), {}, _lib.PRANDTL_ERR_ARG),
    (dict(bc=((1, 1), (0, 0), (0, 0))), dict(parabolic=1), _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(bc=((2, 3), (0, 0), (0, 0)), bc_state=(1.0, 3.0, 0.0, 1.0, 0.0)), {}, _lib.PRANDTL_ERR_ARG),
    (dict(bc=((1, 1), (5, 5), (0, 0)), bc_state=(-1.0, 0.0, 0.0, 2.5, 0.0)), {}, _lib.PRANDTL_ERR_ARG),
    # Navier-Stokes sides: no slip / symmetry (no viscous flux in the
    # reference), no-slip walls without viscosity, heat flux in faithful
    # mode
    (dict(bc=((4, 4), (0, 0), (0, 0))), dict(parabolic=1), _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(bc=((6, 6), (0, 0), (0, 0)), bc_state=(0.0,) * 5), {}, _lib.PRANDTL_ERR_UNSUPPORTED),
    (dict(bc=((6, 6), (0, 0), (0, 0)), bc_state=(0.0, 0.0, 0.0, 0.5, 0.0)), dict(parabolic=1),
     _lib.PRANDTL_ERR_UNSUPPORTED),
])
def test_create_rejects(mesh_kw, prm_kw, code):
    rc, h = _create(mesh_kw, prm_kw)
    assert rc == code
    assert not h.value
    assert lib().prandtl_last_error(None)


def test_null_arguments():
    L = lib()
    assert L.prandtl_rhs(None, None, None, None) == _lib.PRANDTL_ERR_ARG
    assert L.prandtl_rk3_step(None, None, 0.1, None) == _lib.PRANDTL_ERR_ARG
    assert L.prandtl_sizes(None, None, None, None) == _lib.PRANDTL_ERR_ARG
    assert L.prandtl_lgl(0, None, None) == _lib.PRANDTL_ERR_ARG
    L.prandtl_destroy(None)


def _ensemble_create(case_kw):
    n = len(case_kw)
    meshes = (_lib.MeshDesc * n)()
    params = (_lib.Params * n)()
    for i, kw in enumerate(case_kw):
        m, p = meshes[i], params[i]
        m.dim, m.order = kw.get("dim", 2), kw.get("order", 3)
        for r in range(3):
            m.nelem[r] = 4
            m.lo[r], m.hi[r] = 0.0, 1.0
            for side in range(2):
                m.bc[r][side] = kw.get("bc", 0) if r == 0 else 0
        p.gamma, p.Pr, p.mu = 1.4, 0.72, 0.0
        p.alpha_max, p.alpha_min = 0.5, 0.001
        p.parabolic = kw.get("parabolic", 0)
        p.ec_surface = kw.get("ec_surface", 0)
        p.quirks = 1
    h = ctypes.c_void_p()
    rc = lib().prandtl_ensemble_create(n, meshes, params, ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize("case_kw,code", [
    ([{}, dict(dim=3)], _lib.PRANDTL_ERR_UNSUPPORTED),
    ([dict(order=6)], _lib.PRANDTL_ERR_UNSUPPORTED),
    ([dict(order=1)], _lib.PRANDTL_ERR_UNSUPPORTED),
    ([{}, dict(parabolic=1)], _lib.PRANDTL_ERR_UNSUPPORTED),
    ([dict(ec_surface=1)], _lib.PRANDTL_ERR_UNSUPPORTED),
    ([{}, dict(bc=1)], _lib.PRANDTL_ERR_UNSUPPORTED),
    ([{}] * 257, _lib.PRANDTL_ERR_ARG),
])
def test_ensemble_create_rejects(case_kw, code):
    rc, h = _ensemble_create(case_kw)
    assert rc == code
    assert not h.value
    assert lib().prandtl_ensemble_last_error(None)


def test_ensemble_null_arguments():
    L = lib()
    h = ctypes.c_void_p()
    assert L.prandtl_ensemble_create(0, None, None, ctypes.byref(h)) == _lib.PRANDTL_ERR_ARG
    assert L.prandtl_ensemble_rk3_step(None, None, None, None) == _lib.PRANDTL_ERR_ARG
    assert L.prandtl_ensemble_case(None, 0, ctypes.byref(h)) == _lib.PRANDTL_ERR_ARG
    L.prandtl_ensemble_destroy(None)


def test_run_null_arguments():
    L = lib()
    t = ctypes.c_double(0.0)
    assert L.prandtl_run(None, None, ctypes.byref(t), 1.0, 0.5, 0.0, 0, 0, None, None) == _lib.PRANDTL_ERR_ARG
