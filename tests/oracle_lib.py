"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(_HERE), "oracle")
_lib = None

_D = ctypes.c_double
_P = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("order", ctypes.c_int), ("nelem", ctypes.c_int * 3),
        ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
        ("gamma", ctypes.c_double), ("Pr", ctypes.c_double), ("mu", ctypes.c_double),
        ("parabolic", ctypes.c_int), ("blending", ctypes.c_int),
        ("alpha_max", ctypes.c_double), ("alpha_min", ctypes.c_double),
        ("ec_surface", ctypes.c_int), ("quirks", ctypes.c_int), ("threads", ctypes.c_int),
        ("sutherland", ctypes.c_int), ("R_gas", ctypes.c_double),
        ("bc", (ctypes.c_int * 2) * 3), ("bc_state", ((ctypes.c_double * 5) * 2) * 3),
        ("axisymmetric", ctypes.c_int), ("axis_rho_floor", ctypes.c_double),
        ("plant", ctypes.c_int),
    ]


# planted misreadings (orc_params.plant; tests/test_oracle_faithful.py only)
PLANT_STALE_NODE0, PLANT_ZETA_F, PLANT_ZETA_XI = 1, 2, 4
PLANT_FACE_NO_ALIAS, PLANT_GRAD_NO_ALIAS, PLANT_THRESH_NP1 = 8, 16, 32


# boundary types by the names of the reference's configuration files
# (Simulation.cpp:518-735); the values are the oracle's ORC_BC_* codes
BC_CODES = {"periodic": 0, "slip": 1, "supersonic-inflow": 2, "supersonic-outflow": 3,
            "symmetry": 4, "specified-state": 5, "no-slip-adiabatic": 6, "axis": 7}


def bc_code(x):
    return BC_CODES[x] if isinstance(x, str) else int(x)


def set_bcs(p, dim, bc, bc_state):
    """bc: per direction a (low, high) pair of boundary types (None: all
    periodic); bc_state: per direction a (low, high) pair of outer states
    (None where unused)."""
    for r in range(3):
        for side in range(2):
            p.bc[r][side] = bc_code(bc[r][side]) if (bc is not None and r < dim) else 0
            st = bc_state[r][side] if (bc_state is not None and r < dim) else None
            for v in range(5):
                p.bc_state[r][side][v] = float(st[v]) if (st is not None and v < dim + 2) else 0.0


class OrcQuadMesh(ctypes.Structure):
    _fields_ = [("nvert", ctypes.c_int), ("vert", ctypes.POINTER(ctypes.c_double)),
                ("nelem", ctypes.c_int), ("elem", ctypes.POINTER(ctypes.c_int)),
                ("elem_coords", ctypes.POINTER(ctypes.c_double)),
                ("nbdr", ctypes.c_int), ("bdr", ctypes.POINTER(ctypes.c_int)),
                ("bdr_attr", ctypes.POINTER(ctypes.c_int)),
                ("nattr", ctypes.c_int), ("attr_bc", ctypes.POINTER(ctypes.c_int)),
                ("attr_state", ctypes.POINTER(ctypes.c_double))]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_logmean.restype = _D
        L.orc_logmean.argtypes = [_D, _D]
        L.orc_volume_flux.restype = _D
        L.orc_volume_flux.argtypes = [ctypes.c_int, _D, _P, _P, _P, _P, _P]
        L.orc_face_flux.restype = _D
        L.orc_face_flux.argtypes = [ctypes.c_int, _D, ctypes.c_int, _P, _P, _P, _P]
        L.orc_bc_flux.restype = _D
        L.orc_bc_flux.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]
        L.orc_bc_lift_flux.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.orc_cons2entropy.argtypes = [ctypes.c_int, _D, _P, _P]
        L.orc_entgrad2primgrad.argtypes = [ctypes.c_int, _D, ctypes.c_int, _P, _P]
        L.orc_viscous_flux.argtypes = [ctypes.c_int, _D, _D, _D, _D, _P, _P, _P, _P, _P]
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(ctypes.c_void_p)]
        L.orc_create_quad.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcQuadMesh),
                                      ctypes.POINTER(ctypes.c_void_p)]
        # orc_hex_mesh has orc_quad_mesh's layout
        L.orc_create_hex.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcQuadMesh),
                                     ctypes.POINTER(ctypes.c_void_p)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sizes.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 3
        L.orc_axis_recover.argtypes = [ctypes.c_void_p, _P, _P]
        for name in ("orc_coords", "orc_weights", "orc_alpha", "orc_eta", "orc_last_gradients"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _P]
        L.orc_rhs.argtypes = [ctypes.c_void_p, _P, _P]
        L.orc_rhs_ld.argtypes = [ctypes.c_void_p, _P, _P]
        L.orc_rk3_step.argtypes = [ctypes.c_void_p, _P, _D]
        L.orc_set_time.argtypes = [ctypes.c_void_p, _D]
        L.orc_get_time.restype = _D
        L.orc_get_time.argtypes = [ctypes.c_void_p]
        L.orc_set_boundary_split.argtypes = [ctypes.c_void_p, ctypes.c_int, _P, _D, _D, _P, _P]
        L.orc_max_wave_speed.restype = _D
        L.orc_max_wave_speed.argtypes = [ctypes.c_void_p]
        L.orc_reset_wave_speed.argtypes = [ctypes.c_void_p]
        L.orc_hmin.restype = _D
        L.orc_hmin.argtypes = [ctypes.c_void_p]
        L.orc_count_rhs_flops.argtypes = [ctypes.c_void_p, _P, _P]
        for name in ("orc_lgl", "orc_dmat", "orc_legendre_vdm"):
            getattr(L, name).argtypes = [ctypes.c_int] + [_P] * (2 if name != "orc_dmat" else 3)
        _lib = L
    return _lib


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def lgl(order):
    x = np.zeros(order + 1)
    w = np.zeros(order + 1)
    assert lib().orc_lgl(order, ptr(x), ptr(w)) == 0
    return x, w


def dmat(order):
    n = order + 1
    D, Dh, Dh2 = (np.zeros((n, n)) for _ in range(3))
    assert lib().orc_dmat(order, ptr(D), ptr(Dh), ptr(Dh2)) == 0
    return D, Dh, Dh2


def legendre_vdm(order):
    n = order + 1
    V, Vi = np.zeros((n, n)), np.zeros((n, n))
    assert lib().orc_legendre_vdm(order, ptr(V), ptr(Vi)) == 0
    return V, Vi


def logmean(x, y):
    return lib().orc_logmean(float(x), float(y))


def volume_flux(dim, gamma, s1, s2, m1, m2):
    F = np.zeros(dim + 2)
    lam = lib().orc_volume_flux(dim, gamma, ptr(np.ascontiguousarray(s1, float)),
                                ptr(np.ascontiguousarray(s2, float)),
                                ptr(np.ascontiguousarray(m1, float)),
                                ptr(np.ascontiguousarray(m2, float)), ptr(F))
    return F, lam


def face_flux(dim, gamma, quirks, s1, s2, nor):
    F = np.zeros(dim + 2)
    lam = lib().orc_face_flux(dim, gamma, quirks, ptr(np.ascontiguousarray(s1, float)),
                              ptr(np.ascontiguousarray(s2, float)),
                              ptr(np.ascontiguousarray(nor, float)), ptr(F))
    return F, lam


def bc_flux(dim, gamma, quirks, kind, s1, nor, outer=None):
    """Boundary flux of type `kind` (name or code) for the interior state s1
    and the outward face normal nor; returns (F, wave speed)."""
    F = np.zeros(dim + 2)
    o = np.zeros(dim + 2) if outer is None else np.ascontiguousarray(outer, float)
    lam = lib().orc_bc_flux(dim, gamma, quirks, bc_code(kind), ptr(np.ascontiguousarray(s1, float)),
                            ptr(o), ptr(np.ascontiguousarray(nor, float)), ptr(F))
    return F, lam


def bc_lift_flux(dim, gamma, quirks, kind, w, outer=None):
    """Boundary lifting flux of type `kind` for interior entropy variables w."""
    f = np.zeros(dim + 2)
    o = np.zeros(5) if outer is None else np.ascontiguousarray(outer, float)
    lib().orc_bc_lift_flux(dim, gamma, quirks, bc_code(kind), ptr(np.ascontiguousarray(w, float)), ptr(o), ptr(f))
    return f


def cons2entropy(dim, gamma, s):
    w = np.zeros(dim + 2)
    lib().orc_cons2entropy(dim, gamma, ptr(np.ascontiguousarray(s, float)), ptr(w))
    return w


def entgrad2primgrad(dim, gamma, quirks, s, g):
    g = np.array(g, dtype=float)
    lib().orc_entgrad2primgrad(dim, gamma, quirks, ptr(np.ascontiguousarray(s, float)), ptr(g))
    return g


def viscous_flux(dim, gamma, Pr, mu, s, dq, R_gas=0.0):
    """R_gas > 0: Sutherland viscosity."""
    fl = np.zeros((dim + 2) * dim)
    dqs = [np.ascontiguousarray(d, float) for d in dq] + [np.zeros(dim + 2)] * (3 - dim)
    lib().orc_viscous_flux(dim, gamma, Pr, mu, R_gas, ptr(np.ascontiguousarray(s, float)),
                           ptr(dqs[0]), ptr(dqs[1]), ptr(dqs[2]), ptr(fl))
    return fl.reshape(dim + 2, dim)


class Oracle:
    """Operator-shaped handle on the CPU restatement (Mult / Step / GetMaxCharSpeed)."""

    def __init__(self, dim, order, nelem, lo, hi, gamma=1.4, Pr=0.72, mu=0.0,
                 parabolic=False, blending=False, alpha_max=0.5, alpha_min=0.001,
                 ec_surface=False, quirks=True, threads=1, sutherland=False, R_gas=287.05,
                 bc=None, bc_state=None, plant=0):
        p = OrcParams()
        p.dim, p.order = dim, order
        for r in range(3):
            p.nelem[r] = int(nelem[r]) if r < dim else 1
            p.lo[r] = float(lo[r]) if r < dim else 0.0
            p.hi[r] = float(hi[r]) if r < dim else 1.0
        p.gamma, p.Pr, p.mu = gamma, Pr, mu
        p.parabolic, p.blending = int(parabolic), int(blending)
        p.alpha_max, p.alpha_min = alpha_max, alpha_min
        p.ec_surface, p.quirks, p.threads = int(ec_surface), int(quirks), int(threads)
        p.sutherland, p.R_gas = int(sutherland), float(R_gas)
        set_bcs(p, dim, bc, bc_state)
        p.plant = int(plant)
        self.params = p
        h = ctypes.c_void_p()
        rc = lib().orc_create(ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise ValueError("orc_create failed: %d" % rc)
        self.h = h
        NE, K, nv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().orc_sizes(h, ctypes.byref(NE), ctypes.byref(K), ctypes.byref(nv))
        self.NE, self.K, self.nv, self.dim = NE.value, K.value, nv.value, dim

    @classmethod
    def hex(cls, mesh, order, **kw):
        """The oracle on a hexahedral mesh: `mesh` a dict with vert [nv,3],
        elem [ne,8], optional elem_coords [ne,8,3], bdr [nb,4], bdr_attr [nb],
        attr_bc [na], attr_state [na,5]; keywords as quad()."""
        return cls.quad(mesh, order, _dim=3, **kw)

    @classmethod
    def quad(cls, mesh, order, gamma=1.4, blending=False, alpha_max=0.5, alpha_min=0.001,
             quirks=True, threads=1, ec_surface=False, parabolic=False, mu=0.0, Pr=0.72,
             sutherland=False, R_gas=287.05, axisymmetric=False, axis_rho_floor=1e-12, _dim=2):
        """The oracle on a quadrilateral mesh: `mesh` a dict with vert [nv,2],
        elem [ne,4], optional elem_coords [ne,4,2], bdr [nb,2], bdr_attr [nb],
        attr_bc [na] (names or codes), attr_state [na,4].  axisymmetric: the
        state is r U (r = y)."""
        self = cls.__new__(cls)
        p = OrcParams()
        p.dim, p.order = _dim, order
        for r in range(3):
            p.nelem[r], p.lo[r], p.hi[r] = 1, 0.0, 1.0
        p.gamma, p.Pr, p.mu = gamma, Pr, mu
        p.parabolic, p.blending = int(parabolic), int(blending)
        p.alpha_max, p.alpha_min = alpha_max, alpha_min
        p.ec_surface, p.quirks, p.threads = int(ec_surface), int(quirks), int(threads)
        p.sutherland, p.R_gas = int(sutherland), float(R_gas)
        p.axisymmetric, p.axis_rho_floor = int(axisymmetric), float(axis_rho_floor)
        set_bcs(p, 2, None, None)
        keep = {}

        def arr(name, dtype, ctype):
            if mesh.get(name) is None:
                return ctypes.POINTER(ctype)()
            a = np.ascontiguousarray(mesh[name], dtype=dtype).reshape(-1)
            keep[name] = a
            return a.ctypes.data_as(ctypes.POINTER(ctype))

        q = OrcQuadMesh()
        q.nvert = len(mesh["vert"]) if mesh.get("vert") is not None else 0
        q.vert = arr("vert", np.float64, ctypes.c_double)
        q.nelem = len(mesh["elem"])
        q.elem = arr("elem", np.int32, ctypes.c_int)
        q.elem_coords = arr("elem_coords", np.float64, ctypes.c_double)
        q.nbdr = len(mesh["bdr_attr"]) if mesh.get("bdr_attr") is not None else 0
        q.bdr = arr("bdr", np.int32, ctypes.c_int)
        q.bdr_attr = arr("bdr_attr", np.int32, ctypes.c_int)
        codes = [bc_code(t) for t in mesh.get("attr_bc", [])]
        q.nattr = len(codes)
        mesh = dict(mesh, attr_bc=codes if codes else None)
        q.attr_bc = arr("attr_bc", np.int32, ctypes.c_int)
        q.attr_state = arr("attr_state", np.float64, ctypes.c_double)
        h = ctypes.c_void_p()
        create = lib().orc_create_hex if _dim == 3 else lib().orc_create_quad
        rc = create(ctypes.byref(p), ctypes.byref(q), ctypes.byref(h))
        if rc != 0:
            raise ValueError("orc_create_%s failed: %d" % ("hex" if _dim == 3 else "quad", rc))
        self.params, self.h = p, h
        NE, K, nv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().orc_sizes(h, ctypes.byref(NE), ctypes.byref(K), ctypes.byref(nv))
        self.NE, self.K, self.nv, self.dim = NE.value, K.value, nv.value, _dim
        return self

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def size(self):
        return self.nv * self.NE * self.K

    def coords(self):
        X = np.zeros((self.dim, self.NE, self.K))
        lib().orc_coords(self.h, ptr(X))
        return X

    def weights(self):
        W = np.zeros((self.NE, self.K))
        lib().orc_weights(self.h, ptr(W))
        return W

    def Mult(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        out = np.zeros_like(u)
        assert lib().orc_rhs(self.h, ptr(u), ptr(out)) == 0
        return out

    def Mult_extended(self, u):
        """The RHS in long double (extended precision), rounded to double."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        out = np.zeros_like(u)
        assert lib().orc_rhs_ld(self.h, ptr(u), ptr(out)) == 0
        return out

    def set_time(self, t):
        assert lib().orc_set_time(self.h, float(t)) == 0

    @property
    def time(self):
        return lib().orc_get_time(self.h)

    def set_boundary_split(self, attr, normal, offset, speed, below, above):
        """below where normal . x - offset - speed t < 0 at a boundary point of
        attribute attr (1-based), above elsewhere."""
        nv = np.zeros(3)
        nv[:len(normal)] = normal
        b = np.zeros(5)
        b[:len(below)] = below
        a = np.zeros(5)
        a[:len(above)] = above
        assert lib().orc_set_boundary_split(self.h, int(attr), ptr(nv), float(offset), float(speed), ptr(b),
                                            ptr(a)) == 0

    def Step(self, u, dt):
        u = np.array(u, dtype=np.float64).reshape(-1)
        assert lib().orc_rk3_step(self.h, ptr(u), float(dt)) == 0
        return u

    def GetMaxCharSpeed(self):
        return lib().orc_max_wave_speed(self.h)

    def reset_wave_speed(self):
        lib().orc_reset_wave_speed(self.h)

    def hmin(self):
        return lib().orc_hmin(self.h)

    def alpha(self):
        a = np.zeros(self.NE)
        lib().orc_alpha(self.h, ptr(a))
        return a

    def eta(self):
        a = np.zeros(self.NE)
        lib().orc_eta(self.h, ptr(a))
        return a

    def recover(self, rU):
        """The state U of the weighted state r U (axisymmetric oracles)."""
        rU = np.ascontiguousarray(rU, dtype=np.float64).reshape(-1)
        U = np.zeros_like(rU)
        assert lib().orc_axis_recover(self.h, ptr(rU), ptr(U)) == 0
        return U

    def gradients(self):
        g = np.zeros((self.dim, self.nv, self.NE, self.K))
        assert lib().orc_last_gradients(self.h, ptr(g)) == 0
        return g

    def count_flops(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        out = np.zeros(8)
        lib().orc_count_rhs_flops(self.h, ptr(u), ptr(out))
        keys = ["total", "indicator", "lifting", "viscous", "volume", "surface", "fv", "assembly"]
        return dict(zip(keys, out.tolist()))
