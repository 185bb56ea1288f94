"""Host-side mirror of the reference's operator interface, over the HIP library.

`DGSEMOperator` mirrors Prandtl::DGSEMOperator (DGSEMOperator.hpp:15-141):
`Mult(u, dudt)` overwrites dudt with the semi-discrete RHS,
`GetMaxCharSpeed()` returns the running maximum wave speed.
`RK3SSPSolver` mirrors mfem::RK3SSPSolver as used at Simulation.cpp:186-189
and 1010 (`Init(op)`, `Step(x, t, dt)`), with the three stages fused into
the RHS kernels.  `TimeLoop` restates the dt policy of Simulation::Run
(Simulation.cpp:893-907, 1007-1022).

State vectors are torch float64 tensors on the GPU in byNODES layout
(var, element, node).  Every call goes to libprandtl_hip.so on the current
torch stream; there is no CPU path.
"""
import ctypes
import inspect

import numpy as np

from . import _lib
from ._lib import check, lib


def lgl_nodes(order):
    """LGL nodes and weights on [0,1] as the operator uses them (prandtl_lgl)."""
    x = np.zeros(order + 1)
    w = np.zeros(order + 1)
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().prandtl_lgl(order, x.ctypes.data_as(dp), w.ctypes.data_as(dp)))
    return x, w


def _bc_code(kind):
    if isinstance(kind, str):
        if kind not in _lib.BC_TYPES:
            raise ValueError("unknown boundary type %r (one of %s)" % (kind, ", ".join(_lib.BC_TYPES)))
        return _lib.BC_TYPES[kind]
    return int(kind)


def _descs(dim, order, nelem, lo, hi, gamma=1.4, Pr=0.72, mu=0.0, parabolic=False, blending=False,
           alpha_max=0.5, alpha_min=0.001, ec_surface=False, quirks=True, precision="fp64",
           sutherland=False, R_gas=287.05, bc=None, bc_state=None, axisymmetric=False, axis_rho_floor=1e-12):
    """prandtl_mesh_desc and prandtl_params for one case.  bc: per direction
    a (low, high) pair of boundary types by name ("slip", "symmetry",
    "supersonic-inflow", "supersonic-outflow", "specified-state",
    "periodic"), None for a periodic box; bc_state: per direction a
    (low, high) pair of conservative outer states (None where unused)."""
    m = _lib.MeshDesc()
    m.dim, m.order = dim, order
    for r in range(3):
        m.nelem[r] = int(nelem[r]) if r < dim else 1
        m.lo[r] = float(lo[r]) if r < dim else 0.0
        m.hi[r] = float(hi[r]) if r < dim else 1.0
        for side in range(2):
            m.bc[r][side] = _bc_code(bc[r][side]) if (bc is not None and r < dim) else 0
            st = bc_state[r][side] if (bc_state is not None and r < dim) else None
            for v in range(5):
                m.bc_state[r][side][v] = float(st[v]) if (st is not None and v < dim + 2) else 0.0
    p = _lib.Params()
    p.gamma, p.Pr, p.mu = float(gamma), float(Pr), float(mu)
    p.alpha_max, p.alpha_min = float(alpha_max), float(alpha_min)
    p.parabolic, p.blending = int(parabolic), int(blending)
    p.ec_surface, p.quirks = int(ec_surface), int(quirks)
    p.precision = {"fp64": _lib.PRANDTL_FP64, "mixed": _lib.PRANDTL_MIXED}[precision]
    p.sutherland, p.R_gas = int(sutherland), float(R_gas)
    p.axisymmetric, p.axis_rho_floor = int(axisymmetric), float(axis_rho_floor)
    return m, p


_DESC_KEYS = frozenset(inspect.signature(_descs).parameters)


def checkpoint_write(directory, cycle, t, u, dim, order, nelem, lo, hi):
    """Write a host state (numpy float64, byNODES) as checkpoint `cycle`
    under `directory` (prandtl_checkpoint_write; no GPU needed)."""
    m, _ = _descs(dim, order, nelem, lo, hi)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    dp = ctypes.POINTER(ctypes.c_double)
    nv = dim + 2
    check(lib().prandtl_checkpoint_write(str(directory).encode(), int(cycle), float(t), ctypes.byref(m), nv,
                                         u.ctypes.data_as(dp)))


def checkpoint_read(directory, cycle, dim, order, nelem, lo, hi):
    """Read checkpoint `cycle` for this mesh into a new host array; returns
    (u, t).  Raises PrandtlError (PRANDTL_ERR_IO) for a missing, truncated,
    corrupt or mismatched checkpoint."""
    m, _ = _descs(dim, order, nelem, lo, hi)
    nv = dim + 2
    NE = int(np.prod([int(nelem[r]) for r in range(dim)]))
    u = np.empty(nv * NE * (order + 1) ** dim)
    t = ctypes.c_double()
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().prandtl_checkpoint_read(str(directory).encode(), int(cycle), ctypes.byref(m), nv,
                                        u.ctypes.data_as(dp), ctypes.byref(t)))
    return u, t.value


def _stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def read_segment_mesh(path, refine=0):
    """The box of a 1-D MFEM mesh file (the reference's TestCases/Euler/1D
    meshes) after `refine` uniform refinements: dict(dim=1, nelem=(n,),
    lo=(a,), hi=(b,), periodic=bool).  Reads "MFEM INLINE mesh v1.0" with
    type = segment, nx, sx (the segment [0, sx]) and "MFEM mesh v1.0" of
    dimension 1 with uniform elements, periodic when it has no boundary
    points (its elements then close a cycle and L2_T1_1D_P1 nodes give their
    coordinates).  A bounded file has two boundary points; the reference
    names attribute 1 "left" and 2 "right" (Simulation.cpp:283-292), and the
    box's side 0 is lo, so attribute 1 must sit at lo and 2 at hi (an inline
    segment has them there by construction).  Raises ValueError for anything
    else."""
    with open(path) as f:
        lines = [ln.split("#")[0].strip() for ln in f]
    lines = [ln for ln in lines if ln]
    if not lines:
        raise ValueError("%s: empty mesh file" % path)
    if lines[0].startswith("MFEM INLINE mesh"):
        kv = {}
        for ln in lines[1:]:
            if "=" in ln:
                k, v = (s.strip() for s in ln.split("=", 1))
                kv[k] = v
        if kv.get("type") != "segment":
            raise ValueError("%s: inline mesh type %r is not a segment" % (path, kv.get("type")))
        n, sx = int(kv["nx"]), float(kv.get("sx", 1.0))
        return dict(dim=1, nelem=(n << refine,), lo=(0.0,), hi=(sx,), periodic=False)
    if not lines[0].startswith("MFEM mesh v1.0"):
        raise ValueError("%s: not an MFEM v1.0 mesh" % path)

    def section(name):
        if name not in lines:
            raise ValueError("%s: no %s section" % (path, name))
        return lines.index(name) + 1

    i = section("dimension")
    if int(lines[i]) != 1:
        raise ValueError("%s: not a 1-D mesh" % path)
    i = section("elements")
    ne = int(lines[i])
    elems = [tuple(int(t) for t in lines[i + 1 + k].split()[2:4]) for k in range(ne)]
    ib = section("boundary")
    nb = int(lines[ib])
    # boundary points: "attribute geometry vertex" (geometry 0, a point)
    bpts = [tuple(int(t) for t in lines[ib + 1 + k].split()[:3]) for k in range(nb)]
    if "nodes" in lines:
        # header lines (FiniteElementSpace, FiniteElementCollection, VDim,
        # Ordering) up to the first coordinate line
        j = lines.index("nodes") + 1
        while not (lines[j][0].isdigit() or lines[j][0] in "-."):
            if lines[j].startswith("FiniteElementCollection") and "L2_T1_1D_P1" not in lines[j]:
                raise ValueError("%s: nodes of %s are not read" % (path, lines[j]))
            j += 1
        x = np.array([[float(t) for t in lines[j + k].split()] for k in range(ne)])
    else:
        j = section("vertices")
        nv, sd = int(lines[j]), int(lines[j + 1])
        if sd != 1:
            raise ValueError("%s: vertices are not 1-D" % path)
        v = np.array([float(lines[j + 2 + k]) for k in range(nv)])
        x = np.array([[v[a], v[b]] for a, b in elems])
        if nb:
            xb = {}
            for attr, geom, vert in bpts:
                if geom != 0 or not 0 <= vert < nv:
                    raise ValueError("%s: boundary element %r is not a vertex point" % (path, (attr, geom, vert)))
                xb[attr] = v[vert]
    lo, hi = float(x.min()), float(x.max())
    h = (hi - lo) / ne
    if not np.allclose(np.sort(x[:, 0]), lo + h * np.arange(ne), rtol=0, atol=1e-12 * (hi - lo)) or \
            not np.allclose(x[:, 1] - x[:, 0], h, rtol=1e-12, atol=0):
        raise ValueError("%s: elements are not a uniform left-to-right segment" % path)
    if nb:
        if "vertices" not in lines or nb != 2 or sorted(xb) != [1, 2]:
            raise ValueError("%s: a bounded segment has boundary attributes 1 and 2, one point each" % path)
        if not (xb[1] == lo and xb[2] == hi):
            raise ValueError("%s: boundary attribute 1 ('left') is at x = %g and 2 ('right') at x = %g; the box's "
                             "side 0 is x = %g" % (path, xb[1], xb[2], lo))
    return dict(dim=1, nelem=(ne << refine,), lo=(lo,), hi=(hi,), periodic=nb == 0)


class QuadMesh:
    """A 2-D quadrilateral mesh: vert [nv,2], elem [ne,4] (counterclockwise),
    optional elem_coords [ne,4,2] (per-element corners of periodic meshes),
    bdr [nb,2] boundary segments with bdr_attr [nb], names {name: attribute}."""

    def __init__(self, vert, elem, bdr=None, bdr_attr=None, elem_coords=None, names=None):
        self.vert = np.ascontiguousarray(vert, dtype=np.float64).reshape(-1, 2)
        self.elem = np.ascontiguousarray(elem, dtype=np.int32).reshape(-1, 4)
        self.elem_coords = None if elem_coords is None else \
            np.ascontiguousarray(elem_coords, dtype=np.float64).reshape(-1, 4, 2)
        self.bdr = np.zeros((0, 2), np.int32) if bdr is None else np.ascontiguousarray(bdr, dtype=np.int32).reshape(-1, 2)
        self.bdr_attr = np.zeros(0, np.int32) if bdr_attr is None else np.ascontiguousarray(bdr_attr, dtype=np.int32)
        self.names = dict(names or {})

    @property
    def nelem(self):
        return len(self.elem)

    @classmethod
    def read(cls, path, refine=0):
        """Read an MFEM v1.0 or Gmsh 2.2 file (prandtl_mesh_read) and refine it
        `refine` times (ser_ref_levels)."""
        h = ctypes.c_void_p()
        rc = lib().prandtl_mesh_read(str(path).encode(), ctypes.byref(h))
        if rc != _lib.PRANDTL_OK:
            raise _lib.PrandtlError(rc, lib().prandtl_mesh_last_error().decode())
        try:
            if lib().prandtl_mesh_dim(h) != 2:
                raise ValueError("%s is not a quadrilateral mesh (use HexMesh.read)" % path)
            for _ in range(refine):
                check(lib().prandtl_mesh_refine(h))
            v = _lib.QuadMesh()
            check(lib().prandtl_mesh_view(h, ctypes.byref(v)))
            vert = np.ctypeslib.as_array(v.vert, shape=(v.nvert * 2,)).copy() if v.nvert else np.zeros(0)
            elem = np.ctypeslib.as_array(v.elem, shape=(v.nelem * 4,)).copy()
            ec = np.ctypeslib.as_array(v.elem_coords, shape=(v.nelem * 8,)).copy() if v.elem_coords else None
            bdr = np.ctypeslib.as_array(v.bdr, shape=(v.nbdr * 2,)).copy() if v.nbdr else None
            battr = np.ctypeslib.as_array(v.bdr_attr, shape=(v.nbdr,)).copy() if v.nbdr else None
            names, i = {}, 0
            while True:
                a = ctypes.c_int()
                nm = lib().prandtl_mesh_name(h, i, ctypes.byref(a))
                if nm is None:
                    break
                names[nm.decode()] = a.value
                i += 1
        finally:
            lib().prandtl_mesh_free(h)
        return cls(vert, elem, bdr, battr, ec, names)

    def corners(self):
        """[ne,4,2] corner coordinates."""
        return self.elem_coords if self.elem_coords is not None else self.vert[self.elem]

    def node_coords(self, order):
        """Node coordinates X[2][NE][K] of the bilinear element maps."""
        x, _ = lgl_nodes(order)
        Np = order + 1
        i = np.arange(Np * Np)
        xi, eta = x[i % Np], x[i // Np]
        N = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta])   # [4][K]
        c = self.corners()
        return np.einsum("ek d,kn->den".replace(" ", ""), c, N)

    def attribute_tables(self, bc):
        """bc: {name or attribute: type or (type, state)} -> (attr_bc, attr_state)."""
        nattr = int(self.bdr_attr.max()) if len(self.bdr_attr) else 0
        for key in bc or {}:
            a = self.names[key] if isinstance(key, str) else int(key)
            nattr = max(nattr, a)
        attr_bc = np.zeros(max(nattr, 1), np.int32)
        attr_state = np.zeros((max(nattr, 1), 4))
        for key, val in (bc or {}).items():
            a = self.names[key] if isinstance(key, str) else int(key)
            kind, st = (val, None) if isinstance(val, (str, int)) else (val[0], val[1])
            attr_bc[a - 1] = _bc_code(kind)
            if st is not None:
                attr_state[a - 1] = np.asarray(st, float)[:4]
        return nattr, attr_bc, attr_state


class HexMesh:
    """A 3-D hexahedral mesh: vert [nv,3], elem [ne,8] (MFEM's vertex order),
    optional elem_coords [ne,8,3] (periodic meshes), bdr [nb,4] boundary
    quadrilaterals with bdr_attr [nb], names {name: attribute}."""

    def __init__(self, vert, elem, bdr=None, bdr_attr=None, elem_coords=None, names=None):
        self.vert = np.ascontiguousarray(vert, dtype=np.float64).reshape(-1, 3)
        self.elem = np.ascontiguousarray(elem, dtype=np.int32).reshape(-1, 8)
        self.elem_coords = None if elem_coords is None else \
            np.ascontiguousarray(elem_coords, dtype=np.float64).reshape(-1, 8, 3)
        self.bdr = np.zeros((0, 4), np.int32) if bdr is None else np.ascontiguousarray(bdr, dtype=np.int32).reshape(-1, 4)
        self.bdr_attr = np.zeros(0, np.int32) if bdr_attr is None else np.ascontiguousarray(bdr_attr, dtype=np.int32)
        self.names = dict(names or {})

    @property
    def nelem(self):
        return len(self.elem)

    @classmethod
    def read(cls, path, refine=0):
        """Read a Gmsh 2.2 or MFEM v1.0 hexahedral mesh (prandtl_mesh_read) and
        refine it `refine` times."""
        h = ctypes.c_void_p()
        rc = lib().prandtl_mesh_read(str(path).encode(), ctypes.byref(h))
        if rc != _lib.PRANDTL_OK:
            raise _lib.PrandtlError(rc, lib().prandtl_mesh_last_error().decode())
        try:
            if lib().prandtl_mesh_dim(h) != 3:
                raise ValueError("%s is not a hexahedral mesh (use QuadMesh.read)" % path)
            for _ in range(refine):
                check(lib().prandtl_mesh_refine(h))
            v = _lib.HexMesh()
            check(lib().prandtl_mesh_view_hex(h, ctypes.byref(v)))
            vert = np.ctypeslib.as_array(v.vert, shape=(v.nvert * 3,)).copy()
            elem = np.ctypeslib.as_array(v.elem, shape=(v.nelem * 8,)).copy()
            ec = np.ctypeslib.as_array(v.elem_coords, shape=(v.nelem * 24,)).copy() if v.elem_coords else None
            bdr = np.ctypeslib.as_array(v.bdr, shape=(v.nbdr * 4,)).copy() if v.nbdr else None
            battr = np.ctypeslib.as_array(v.bdr_attr, shape=(v.nbdr,)).copy() if v.nbdr else None
            names, i = {}, 0
            while True:
                a = ctypes.c_int()
                nm = lib().prandtl_mesh_name(h, i, ctypes.byref(a))
                if nm is None:
                    break
                names[nm.decode()] = a.value
                i += 1
        finally:
            lib().prandtl_mesh_free(h)
        return cls(vert, elem, bdr, battr, ec, names)

    def corners(self):
        """[ne,8,3] corner coordinates."""
        return self.elem_coords if self.elem_coords is not None else self.vert[self.elem]

    def node_coords(self, order):
        """Node coordinates X[3][NE][K] of the trilinear element maps."""
        x, _ = lgl_nodes(order)
        Np = order + 1
        i = np.arange(Np ** 3)
        xi, eta, zeta = x[i % Np], x[(i // Np) % Np], x[i // (Np * Np)]
        cv = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
        N = np.stack([(xi if a else 1 - xi) * (eta if b else 1 - eta) * (zeta if c else 1 - zeta) for a, b, c in cv])
        return np.einsum("ekd,kn->den", self.corners(), N)

    def attribute_tables(self, bc):
        """bc: {name or attribute: type or (type, state)} -> (nattr, attr_bc, attr_state [nattr][5])."""
        nattr = int(self.bdr_attr.max()) if len(self.bdr_attr) else 0
        for key in bc or {}:
            nattr = max(nattr, self.names[key] if isinstance(key, str) else int(key))
        attr_bc = np.zeros(max(nattr, 1), np.int32)
        attr_state = np.zeros((max(nattr, 1), 5))
        for key, val in (bc or {}).items():
            a = self.names[key] if isinstance(key, str) else int(key)
            kind, st = (val, None) if isinstance(val, (str, int)) else (val[0], val[1])
            attr_bc[a - 1] = _bc_code(kind)
            if st is not None:
                attr_state[a - 1] = np.asarray(st, float)[:5]
        return nattr, attr_bc, attr_state


class DGSEMOperator:
    """One GPU's DGSEM semi-discrete operator on a Cartesian box, periodic or
    with boundary conditions on its sides (bc, bc_state: see _descs), or on a
    2-D quadrilateral mesh (DGSEMOperator.quad)."""

    def __init__(self, dim, order, nelem, lo, hi, gamma=1.4, Pr=0.72, mu=0.0,
                 parabolic=False, blending=False, alpha_max=0.5, alpha_min=0.001,
                 ec_surface=False, quirks=True, precision="fp64", sutherland=False, R_gas=287.05,
                 bc=None, bc_state=None):
        m, p = _descs(dim, order, nelem, lo, hi, gamma, Pr, mu, parabolic, blending, alpha_max,
                      alpha_min, ec_surface, quirks, precision, sutherland, R_gas, bc, bc_state)
        h = ctypes.c_void_p()
        rc = lib().prandtl_create(ctypes.byref(m), ctypes.byref(p), ctypes.byref(h))
        check(rc, None)
        self._h = h
        self._owner = True
        self._describe(dim, order, nelem, lo, hi)

    @classmethod
    def quad(cls, mesh, order, bc=None, gamma=1.4, blending=False, alpha_max=0.5, alpha_min=0.001,
             quirks=True, precision="fp64", parabolic=False, mu=0.0, Pr=0.72, sutherland=False,
             R_gas=287.05, axisymmetric=False, axis_rho_floor=1e-12):
        """The operator on QuadMesh `mesh` (prandtl_create_quad); bc maps
        boundary names or attributes to a type or (type, state): the outer
        conservative state of inflow / specified-state boundaries, or
        (V0, V1, 0, qn) of a no-slip wall.  axisymmetric: the state is r U
        with r = y (an "axis" boundary marks the axis)."""
        _, p = _descs(2, order, (1, 1), (0.0, 0.0), (1.0, 1.0), gamma=gamma, blending=blending,
                      alpha_max=alpha_max, alpha_min=alpha_min, quirks=quirks, precision=precision,
                      parabolic=parabolic, mu=mu, Pr=Pr, sutherland=sutherland, R_gas=R_gas,
                      axisymmetric=axisymmetric, axis_rho_floor=axis_rho_floor)
        nattr, attr_bc, attr_state = mesh.attribute_tables(bc)
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        q = _lib.QuadMesh()
        q.nvert = len(mesh.vert)
        q.vert = mesh.vert.ctypes.data_as(dp)
        q.nelem = len(mesh.elem)
        q.elem = mesh.elem.ctypes.data_as(ip)
        q.elem_coords = mesh.elem_coords.ctypes.data_as(dp) if mesh.elem_coords is not None else dp()
        q.nbdr = len(mesh.bdr_attr)
        q.bdr = mesh.bdr.ctypes.data_as(ip)
        q.bdr_attr = mesh.bdr_attr.ctypes.data_as(ip)
        q.nattr = nattr
        q.attr_bc = attr_bc.ctypes.data_as(ip)
        q.attr_state = np.ascontiguousarray(attr_state).ctypes.data_as(dp)
        h = ctypes.c_void_p()
        check(lib().prandtl_create_quad(ctypes.byref(q), int(order), ctypes.byref(p), ctypes.byref(h)), None)
        op = cls.__new__(cls)
        op._h = h
        op._owner = True
        op._describe(2, order, (len(mesh.elem), 1), (0.0, 0.0), (1.0, 1.0))
        op.mesh = mesh
        return op

    @classmethod
    def hex(cls, mesh, order, bc=None, gamma=1.4, blending=False, alpha_max=0.5, alpha_min=0.001,
            quirks=True, precision="fp64", parabolic=False, mu=0.0, Pr=0.72, sutherland=False,
            R_gas=287.05):
        """The operator on HexMesh `mesh` (prandtl_create_hex); bc as quad()
        with 5-value states (no-slip walls: (V0, V1, V2, qn, 0))."""
        _, p = _descs(3, order, (1, 1, 1), (0.0,) * 3, (1.0,) * 3, gamma=gamma, blending=blending,
                      alpha_max=alpha_max, alpha_min=alpha_min, quirks=quirks, precision=precision,
                      parabolic=parabolic, mu=mu, Pr=Pr, sutherland=sutherland, R_gas=R_gas)
        nattr, attr_bc, attr_state = mesh.attribute_tables(bc)
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        q = _lib.HexMesh()
        q.nvert = len(mesh.vert)
        q.vert = mesh.vert.ctypes.data_as(dp)
        q.nelem = len(mesh.elem)
        q.elem = mesh.elem.ctypes.data_as(ip)
        q.elem_coords = mesh.elem_coords.ctypes.data_as(dp) if mesh.elem_coords is not None else dp()
        q.nbdr = len(mesh.bdr_attr)
        q.bdr = mesh.bdr.ctypes.data_as(ip)
        q.bdr_attr = mesh.bdr_attr.ctypes.data_as(ip)
        q.nattr = nattr
        q.attr_bc = attr_bc.ctypes.data_as(ip)
        q.attr_state = np.ascontiguousarray(attr_state).ctypes.data_as(dp)
        h = ctypes.c_void_p()
        check(lib().prandtl_create_hex(ctypes.byref(q), int(order), ctypes.byref(p), ctypes.byref(h)), None)
        op = cls.__new__(cls)
        op._h = h
        op._owner = True
        op._describe(3, order, (len(mesh.elem), 1, 1), (0.0,) * 3, (1.0,) * 3)
        op.mesh = mesh
        return op

    @classmethod
    def _view(cls, handle, dim, order, nelem, lo, hi):
        """A non-owning operator over a context that another object frees."""
        op = cls.__new__(cls)
        op._h = handle
        op._owner = False
        op._describe(dim, order, nelem, lo, hi)
        return op

    def _describe(self, dim, order, nelem, lo, hi):
        h = self._h
        NE, K, nv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().prandtl_sizes(h, ctypes.byref(NE), ctypes.byref(K), ctypes.byref(nv)), h)
        self.NE, self.K, self.nv = NE.value, K.value, nv.value
        self.dim, self.order = dim, order
        self.nelem = tuple(int(nelem[r]) for r in range(dim))
        self.lo = tuple(float(v) for v in lo[:dim])
        self.hi = tuple(float(v) for v in hi[:dim])
        self.height = self.width = self.nv * self.NE * self.K
        hm = ctypes.c_double()
        check(lib().prandtl_hmin(h, ctypes.byref(hm)), h)
        self.hmin = hm.value

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            if self._owner:
                lib().prandtl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- checks -------------------------------------------------------------
    def _vec(self, x, name):
        import torch
        if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or not x.is_cuda:
            raise TypeError("%s must be a float64 CUDA tensor" % name)
        if not x.is_contiguous() or x.numel() != self.height:
            raise ValueError("%s must be contiguous with %d entries" % (name, self.height))
        return ctypes.c_void_p(x.data_ptr())

    # -- Operator surface ---------------------------------------------------
    def Mult(self, u, dudt):
        """dudt <- L(u) (DGSEMOperator::Mult, DGSEMOperator.cpp:537-591)."""
        check(lib().prandtl_rhs(self._h, self._vec(u, "u"), self._vec(dudt, "dudt"), _stream_ptr()), self._h)

    def Step(self, u, dt):
        """u <- one SSP-RK3 step of size dt, in place."""
        check(lib().prandtl_rk3_step(self._h, self._vec(u, "u"), float(dt), _stream_ptr()), self._h)

    def Run(self, u, t, t_final, cfl=None, dt=None, max_steps=0, check_every=0):
        """The time loop on the device (prandtl_run): steps u from t to
        t_final with variable dt from cfl, or fixed dt; returns (t, steps).
        check_every > 0: state check after every such number of steps; a bad
        state raises PrandtlError (code PRANDTL_ERR_STATE) whose .t and
        .steps say where the loop stopped."""
        tt = ctypes.c_double(float(t))
        steps = ctypes.c_int()
        rc = lib().prandtl_run(self._h, self._vec(u, "u"), ctypes.byref(tt), float(t_final),
                               float(cfl) if cfl else 0.0, float(dt) if dt else 0.0, int(max_steps),
                               int(check_every), ctypes.byref(steps), _stream_ptr())
        try:
            check(rc, self._h)
        except _lib.PrandtlError as e:
            e.t, e.steps = tt.value, steps.value
            raise
        return tt.value, steps.value

    def save_checkpoint(self, u, directory, cycle, t):
        """Checkpoint `cycle` of device state u at time t (prandtl_checkpoint_save)."""
        check(lib().prandtl_checkpoint_save(self._h, self._vec(u, "u"), str(directory).encode(), int(cycle),
                                            float(t), _stream_ptr()), self._h)

    def load_checkpoint(self, u, directory, cycle):
        """Load checkpoint `cycle` into device state u; returns its time."""
        t = ctypes.c_double()
        check(lib().prandtl_checkpoint_load(self._h, self._vec(u, "u"), str(directory).encode(), int(cycle),
                                            ctypes.byref(t), _stream_ptr()), self._h)
        return t.value

    def set_graph(self, enable):
        """Capture/replay the RK3 step as a hipGraph on non-default streams."""
        check(lib().prandtl_set_graph(self._h, int(enable)), self._h)

    def set_shape(self, kernel, shape):
        """Force a launch shape (0 = default) of kernel "rhs" or "visc"."""
        k = {"rhs": _lib.PRANDTL_KERNEL_RHS, "visc": _lib.PRANDTL_KERNEL_VISC}[kernel]
        check(lib().prandtl_set_shape(self._h, k, int(shape)), self._h)

    def set_time(self, t):
        """TimeDependentOperator::SetTime: the time of the next RHS (RK3 steps
        evaluate their stages at t, t + dt, t + dt/2 and advance it by dt)."""
        check(lib().prandtl_set_time(self._h, float(t), _stream_ptr()), self._h)

    @property
    def time(self):
        t = ctypes.c_double()
        check(lib().prandtl_get_time(self._h, ctypes.byref(t)), self._h)
        return t.value

    def set_boundary_split(self, boundary, normal, offset, speed, below, above):
        """A time-dependent outer state of a specified-state or supersonic-inflow
        boundary (name or attribute) of a quadrilateral mesh: `below` where
        normal . x - offset - speed t < 0 at the boundary point, `above`
        elsewhere (conservative states)."""
        a = self.mesh.names[boundary] if isinstance(boundary, str) else int(boundary)
        dp = ctypes.POINTER(ctypes.c_double)
        nv = np.zeros(3)
        nv[:len(normal)] = normal
        b = np.ascontiguousarray(below, dtype=np.float64)
        ab = np.ascontiguousarray(above, dtype=np.float64)
        check(lib().prandtl_set_boundary_split(self._h, a, nv.ctypes.data_as(dp), float(offset), float(speed),
                                               b.ctypes.data_as(dp), ab.ctypes.data_as(dp)), self._h)

    def set_algorithm(self, algo):
        """Kernel algorithm of the flux-differencing RHS: "tile" (k_rhs, every
        mesh kind) or "line" (k_rhs_ln, one thread per tensor line: periodic
        3-D boxes of order 2..4, the default at orders 2 and 3 and for
        blended order 4)."""
        a = {"tile": _lib.PRANDTL_ALGO_TILE, "line": _lib.PRANDTL_ALGO_LINE}[algo]
        check(lib().prandtl_set_algorithm(self._h, _lib.PRANDTL_KERNEL_RHS, a), self._h)

    @property
    def algorithm(self):
        a = ctypes.c_int()
        check(lib().prandtl_get_algorithm(self._h, _lib.PRANDTL_KERNEL_RHS, ctypes.byref(a)), self._h)
        return "line" if a.value == _lib.PRANDTL_ALGO_LINE else "tile"

    def autotune(self, u):
        """Time every launch shape of both kernels on state u; keep the fastest.

        Returns ({"rhs": shape, "visc": shape}, {"rhs": times, "visc": times}),
        times[s] the mean microseconds per launch of shape s, None for shapes
        this case does not have.
        """
        shapes = (ctypes.c_int * 2)()
        us = (ctypes.c_double * (2 * _lib.PRANDTL_SHAPES))()
        check(lib().prandtl_autotune(self._h, self._vec(u, "u"), shapes, us, _stream_ptr()), self._h)
        n = _lib.PRANDTL_SHAPES
        times = [v if v >= 0 else None for v in us]
        return ({"rhs": shapes[0], "visc": shapes[1]},
                {"rhs": times[:n], "visc": times[n:]})

    def GetMaxCharSpeed(self, reset=False):
        out = ctypes.c_double()
        check(lib().prandtl_max_wave_speed(self._h, ctypes.byref(out), int(reset), _stream_ptr()), self._h)
        return out.value

    def alpha(self):
        a = np.zeros(self.NE)
        e = np.zeros(self.NE)
        dp = ctypes.POINTER(ctypes.c_double)
        check(lib().prandtl_alpha(self._h, a.ctypes.data_as(dp), e.ctypes.data_as(dp), _stream_ptr()), self._h)
        return a, e

    def check_state(self, u):
        """True if every node has finite values, rho > 0 and p > 0."""
        rc = lib().prandtl_check_state(self._h, self._vec(u, "u"), _stream_ptr())
        if rc == _lib.PRANDTL_ERR_STATE:
            return False
        check(rc, self._h)
        return True

    def profile_begin(self):
        """Bracket every following kernel launch with HIP events."""
        check(lib().prandtl_profile_begin(self._h), self._h)

    def profile_end(self):
        """Synchronise; return {'rhs': (ms, launches), 'visc': (ms, launches)}."""
        ms = (ctypes.c_double * 2)()
        cnt = (ctypes.c_int * 2)()
        check(lib().prandtl_profile_end(self._h, ms, cnt), self._h)
        return {"rhs": (ms[0], cnt[0]), "visc": (ms[1], cnt[1])}

    def primitives(self, u):
        """rho, velocity components and p of state u, [nv][NE][K]
        (the reference's output fields, Simulation.cpp:1046-1061)."""
        import torch
        prim = torch.empty((self.nv, self.NE, self.K), dtype=torch.float64, device=u.device)
        check(lib().prandtl_primitives(self._h, self._vec(u, "u"), ctypes.c_void_p(prim.data_ptr()),
                                       _stream_ptr()), self._h)
        return prim

    def recover(self, ru):
        """The state U of the weighted state r U (axisymmetric operators,
        prandtl_axis_recover), [nv][NE][K]."""
        import torch
        U = torch.empty((self.nv, self.NE, self.K), dtype=torch.float64, device=ru.device)
        check(lib().prandtl_axis_recover(self._h, self._vec(ru, "u"), ctypes.c_void_p(U.data_ptr()),
                                         _stream_ptr()), self._h)
        return U

    def gradients(self, u):
        """Primitive-variable BR1 gradients [dim][nv][NE][K] (parabolic contexts)."""
        import torch
        g = torch.empty((self.dim, self.nv, self.NE, self.K), dtype=torch.float64, device=u.device)
        check(lib().prandtl_debug_gradients(self._h, self._vec(u, "u"), ctypes.c_void_p(g.data_ptr()),
                                            _stream_ptr()), self._h)
        return g


class Ensemble:
    """Independent 2-D Euler cases advanced together (prandtl_ensemble_*).

    `cases` is a sequence of dicts with the DGSEMOperator keyword arguments
    (dim, order, nelem, lo, hi, gamma, blending, alpha_max, ...); other keys
    are ignored.  `Step(us,
    dts)` advances case i's state us[i] by dts[i], one SSP-RK3 step each, in
    one launch per RK stage; `ops[i]` is a non-owning DGSEMOperator over case
    i's context (sizes, wave speed, alpha).
    """

    def __init__(self, cases):
        n = len(cases)
        meshes = (_lib.MeshDesc * n)()
        params = (_lib.Params * n)()
        for i, c in enumerate(cases):
            meshes[i], params[i] = _descs(**{k: v for k, v in c.items() if k in _DESC_KEYS})
        h = ctypes.c_void_p()
        rc = lib().prandtl_ensemble_create(n, meshes, params, ctypes.byref(h))
        _lib.check_ensemble(rc, None)
        self._h = h
        self.n = n
        self.ops = []
        for i, c in enumerate(cases):
            ch = ctypes.c_void_p()
            _lib.check_ensemble(lib().prandtl_ensemble_case(h, i, ctypes.byref(ch)), h)
            self.ops.append(DGSEMOperator._view(ch, c["dim"], c["order"], c["nelem"], c["lo"], c["hi"]))

    def close(self):
        if getattr(self, "_h", None):
            for op in self.ops:
                op.close()
            lib().prandtl_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Step(self, us, dts):
        if len(us) != self.n or len(dts) != self.n:
            raise ValueError("need %d states and %d step sizes" % (self.n, self.n))
        ptrs = (ctypes.c_void_p * self.n)(*[op._vec(u, "u").value for op, u in zip(self.ops, us)])
        dt = (ctypes.c_double * self.n)(*[float(v) for v in dts])
        _lib.check_ensemble(lib().prandtl_ensemble_rk3_step(self._h, ptrs, dt, _stream_ptr()), self._h)

    def set_mode(self, mode):
        """Launch form: 0 one workgroup per tile, g > 0 persistent with g
        workgroups per CU over a device tile counter (bitwise the same)."""
        _lib.check_ensemble(lib().prandtl_ensemble_set_mode(self._h, int(mode)), self._h)

    def autotune(self, us, dts):
        """Time the first RK stage in every launch form on states us (not
        modified) and keep the fastest; returns (mode, [us per form])."""
        ptrs = (ctypes.c_void_p * self.n)(*[op._vec(u, "u").value for op, u in zip(self.ops, us)])
        dt = (ctypes.c_double * self.n)(*[float(v) for v in dts])
        mode = ctypes.c_int()
        times = (ctypes.c_double * (_lib.PRANDTL_ENS_MAX_PER_CU + 1))()
        _lib.check_ensemble(lib().prandtl_ensemble_autotune(self._h, ptrs, dt, ctypes.byref(mode), times,
                                                            _stream_ptr()), self._h)
        return mode.value, list(times)


class RK3SSPSolver:
    """mfem::RK3SSPSolver surface: Init(op) then Step(x, t, dt) -> (t, dt)."""

    def Init(self, f):
        self.f = f

    def Step(self, x, t, dt):
        self.f.Step(x, dt)
        return t + dt, dt


class TimeLoop:
    """Simulation::Run's stepping and dt policy (Simulation.cpp:893-907, 1007-1022).

    variable_dt: dt = cfl*hmin/(lambda*(N+1)^2) from the running maximum wave
    speed, recomputed after every step; otherwise dt is fixed.  The last step
    is clipped to t_final.
    """

    def __init__(self, op, t_final, cfl=None, dt=None, variable_dt=False):
        self.op, self.t_final, self.cfl = op, float(t_final), cfl
        self.variable_dt = variable_dt
        self.dt = dt
        self.ode = RK3SSPSolver()
        self.ode.Init(op)

    def _cfl_dt(self):
        lam = self.op.GetMaxCharSpeed()
        return self.cfl * self.op.hmin / (lam * (self.op.order + 1) ** 2)

    def run(self, u, t=0.0, max_steps=None, callback=None):
        """Advance u to t_final; callback(t, steps, done), if given, runs after
        every step."""
        import torch
        if self.variable_dt and self.cfl:
            z = torch.empty_like(u)
            self.op.Mult(u, z)
            self.dt = self._cfl_dt()
        done, steps = False, 0
        while not done:
            dt_real = min(self.dt, self.t_final - t)
            t, _ = self.ode.Step(u, t, dt_real)
            if self.variable_dt and self.cfl:
                self.dt = self._cfl_dt()
            steps += 1
            done = t >= self.t_final - 1e-8 * self.dt
            if max_steps is not None and steps >= max_steps:
                done = True
            if callback is not None:
                callback(t, steps, done)
        return t, steps
