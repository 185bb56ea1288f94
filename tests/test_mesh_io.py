"""Mesh files and quadrilateral-mesh validation on the CPU.

The readers (prandtl_mesh_read / _refine / _view / _name) are host code; the
fixtures are the reference's own mesh files (tests/golden/meshes).  Checks:
element, vertex and boundary counts, physical names, total area (the domain's
exact area: the forward-facing step is 3 x 1 less the 2.4 x 0.2 step),
refinement (4x elements, 2x boundary segments, area kept), orientation of
every element, and the errors prandtl_create_quad reports before touching
the GPU for a mesh that does not close.
"""
import ctypes
import os

import numpy as np
import pytest

from prandtl_amd import PrandtlError, _lib
from prandtl_amd._lib import lib
from prandtl_amd.operator import DGSEMOperator, QuadMesh, _descs

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def area(m):
    c = m.corners()
    return 0.5 * np.sum(c[:, :, 0] * np.roll(c[:, :, 1], -1, axis=1) - np.roll(c[:, :, 0], -1, axis=1) * c[:, :, 1])


def signed_areas(m):
    c = m.corners()
    return 0.5 * (c[:, :, 0] * np.roll(c[:, :, 1], -1, axis=1) - np.roll(c[:, :, 0], -1, axis=1) * c[:, :, 1]).sum(axis=1)


def test_gmsh_forward_facing_step():
    m = QuadMesh.read(os.path.join(MESHES, "ForwardFacingStep.msh"))
    assert m.nelem == 3400 and len(m.vert) == 3531 and len(m.bdr_attr) == 260
    assert m.names == {"left": 2, "right": 3, "top": 4, "bottom": 5}
    assert abs(area(m) - (3.0 - 2.4 * 0.2)) < 1e-12
    assert (signed_areas(m) > 0).all()
    assert set(np.unique(m.bdr_attr)) == {2, 3, 4, 5}
    r = QuadMesh.read(os.path.join(MESHES, "ForwardFacingStep.msh"), refine=1)
    assert r.nelem == 4 * 3400 and len(r.bdr_attr) == 2 * 260
    assert abs(area(r) - area(m)) < 1e-12
    assert (signed_areas(r) > 0).all()
    # children of element i are 4i..4i+3 and tile it
    assert np.allclose(signed_areas(r).reshape(-1, 4).sum(axis=1), signed_areas(m), rtol=1e-13)


def test_gmsh_lid_driven_cavity():
    m = QuadMesh.read(os.path.join(MESHES, "LidDrivenCavity.msh"))
    assert m.nelem == 256 and len(m.bdr_attr) == 64
    assert m.names == {"bottom": 1, "right": 2, "lid": 3, "left": 4}
    assert abs(area(m) - 4.0) < 1e-13


def test_mfem_periodic_mesh():
    m = QuadMesh.read(os.path.join(MESHES, "IsentropicVortex.mesh"))
    assert m.nelem == 9 and m.elem_coords is not None and len(m.bdr_attr) == 0
    assert abs(area(m) - 4.0) < 1e-8            # the file's corners are 9-digit
    # element 0: (-1,-1), (-1/3,-1), (-1/3,-1/3), (-1,-1/3) in vertex order
    assert np.allclose(m.corners()[0], [[-1, -1], [-0.333333333, -1], [-0.333333333, -0.333333333],
                                        [-1, -0.333333333]])
    r = QuadMesh.read(os.path.join(MESHES, "IsentropicVortex.mesh"), refine=2)
    assert r.nelem == 9 * 16 and abs(area(r) - area(m)) < 1e-12


def test_mfem_vertex_mesh_and_orientation(tmp_path):
    # a 2 x 1 mesh whose second element is listed clockwise
    p = tmp_path / "two.mesh"
    p.write_text("MFEM mesh v1.0\n\ndimension\n2\n\nelements\n2\n1 3 0 1 4 3\n1 3 1 4 5 2\n\n"
                 "boundary\n6\n1 1 0 1\n1 1 1 2\n2 1 2 5\n1 1 5 4\n1 1 4 3\n3 1 3 0\n\n"
                 "vertices\n6\n2\n0 0\n1 0\n2 0\n0 1\n1 1\n2 1\n")
    m = QuadMesh.read(p)
    assert m.nelem == 2 and len(m.bdr_attr) == 6
    assert (signed_areas(m) > 0).all()
    assert list(m.elem[1]) == [1, 2, 5, 4]


@pytest.mark.parametrize("text,needle", [
    ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$EndNodes\n"
     "$Elements\n1\n1 2 2 1 1 1 2 3\n$EndElements\n", "element type 2"),
    ("MFEM mesh v1.0\ndimension\n1\n", "2-D and 3-D"),
    ("something else\n", "unknown mesh format"),
])
def test_reader_errors(tmp_path, text, needle):
    p = tmp_path / "bad.msh"
    p.write_text(text)
    with pytest.raises(PrandtlError) as ei:
        QuadMesh.read(p)
    assert ei.value.code == _lib.PRANDTL_ERR_IO and needle in str(ei.value)
    with pytest.raises(PrandtlError):
        QuadMesh.read(tmp_path / "missing.msh")


def test_create_quad_rejects_open_meshes():
    """Checked on the host before any device allocation."""
    m = QuadMesh.read(os.path.join(MESHES, "LidDrivenCavity.msh"))
    with pytest.raises(PrandtlError) as ei:            # the lid has no boundary condition
        DGSEMOperator.quad(m, 3, bc={"bottom": "slip", "right": "slip", "left": "slip"})
    assert ei.value.code == _lib.PRANDTL_ERR_ARG and "attribute 3" in str(ei.value)
    bad = QuadMesh(m.vert, m.elem)                     # no boundary segments at all
    with pytest.raises(PrandtlError) as ei:
        DGSEMOperator.quad(bad, 3)
    assert "no neighbour and no boundary segment" in str(ei.value)
    with pytest.raises(PrandtlError) as ei:            # negative pressure outer state
        DGSEMOperator.quad(m, 3, bc={"bottom": ("supersonic-inflow", [1.0, 3.0, 0.0, 1.0]),
                                     "right": "slip", "lid": "slip", "left": "slip"})
    assert ei.value.code == _lib.PRANDTL_ERR_ARG
    # Navier-Stokes with slip walls (no viscous flux in the reference), and
    # no-slip walls without viscosity, are refused
    with pytest.raises(PrandtlError) as ei:
        DGSEMOperator.quad(m, 3, parabolic=True, mu=0.01,
                           bc={"bottom": "slip", "right": "slip", "lid": "slip", "left": "slip"})
    assert ei.value.code == _lib.PRANDTL_ERR_UNSUPPORTED
    wall = ("no-slip-adiabatic", (0.0, 0.0, 0.0, 0.0))
    with pytest.raises(PrandtlError) as ei:
        DGSEMOperator.quad(m, 3, bc={"bottom": wall, "right": wall, "lid": wall, "left": wall})
    assert ei.value.code == _lib.PRANDTL_ERR_UNSUPPORTED
    # an empty mesh
    _, prm = _descs(2, 3, (1, 1), (0, 0), (1, 1))
    q = _lib.QuadMesh()
    h = ctypes.c_void_p()
    rc = lib().prandtl_create_quad(ctypes.byref(q), 3, ctypes.byref(prm), ctypes.byref(h))
    assert rc == _lib.PRANDTL_ERR_ARG and not h.value


def _hex_files(tmp_path):
    """A distorted 3 x 2 x 2 hexahedral box written as Gmsh 2.2 and MFEM v1.0
    files (attributes 1..6 per side, named in the Gmsh file)."""
    import quad_meshes as Q
    m = Q.hex_grid(3, 2, 2, (0.0, 0.0, 0.0), (1.5, 1.0, 0.8), periodic=(False, False, False), perturb=0.2, seed=3)
    names = ["xlo", "xhi", "ylo", "yhi", "zlo", "zhi"]
    g = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "7"]
    g += ['2 %d "%s"' % (i + 1, n) for i, n in enumerate(names)] + ['3 7 "fluid"', "$EndPhysicalNames"]
    g += ["$Nodes", str(len(m["vert"]))] + ["%d %.17g %.17g %.17g" % (i + 1, *v) for i, v in enumerate(m["vert"])]
    g += ["$EndNodes", "$Elements", str(len(m["bdr"]) + len(m["elem"]))]
    k = 1
    for q, a in zip(m["bdr"], m["bdr_attr"]):
        g.append("%d 3 2 %d %d %s" % (k, a, a, " ".join(str(v + 1) for v in q)))
        k += 1
    for e in m["elem"]:
        g.append("%d 5 2 7 7 %s" % (k, " ".join(str(v + 1) for v in e)))
        k += 1
    g += ["$EndElements"]
    (tmp_path / "box.msh").write_text("\n".join(g) + "\n")
    f = ["MFEM mesh v1.0", "", "dimension", "3", "", "elements", str(len(m["elem"]))]
    f += ["1 5 " + " ".join(str(v) for v in e) for e in m["elem"]]
    f += ["", "boundary", str(len(m["bdr"]))]
    f += ["%d 3 %s" % (a, " ".join(str(v) for v in q)) for q, a in zip(m["bdr"], m["bdr_attr"])]
    f += ["", "vertices", str(len(m["vert"])), "3"] + ["%.17g %.17g %.17g" % tuple(v) for v in m["vert"]]
    (tmp_path / "box.mesh").write_text("\n".join(f) + "\n")
    return m, names


@pytest.mark.parametrize("fname", ["box.msh", "box.mesh"])
def test_hexahedral_mesh_files(tmp_path, fname):
    """Gmsh 8-node hexahedra with 4-node boundary quadrilaterals and MFEM
    geometry-5 meshes read back to the same topology and coordinates; the
    oracle builds the same operator from either."""
    import oracle_lib as O
    from prandtl_amd import cases
    from prandtl_amd.operator import HexMesh
    m, names = _hex_files(tmp_path)
    h = HexMesh.read(tmp_path / fname)
    assert np.array_equal(h.elem, m["elem"]) and np.array_equal(h.bdr, m["bdr"])
    assert np.array_equal(h.bdr_attr, m["bdr_attr"]) and np.abs(h.vert - m["vert"]).max() < 1e-15
    if fname.endswith(".msh"):
        assert h.names == {n: i + 1 for i, n in enumerate(names)}
    kinds = ["supersonic-inflow", "supersonic-outflow", "slip", "slip", "symmetry", "slip"]
    st = [np.array([1.2, 2.0, 0.1, 0.0, 4.5])] * 6
    d = dict(vert=h.vert, elem=h.elem, bdr=h.bdr, bdr_attr=h.bdr_attr, attr_bc=kinds, attr_state=st)
    a = O.Oracle.hex(d, 3, quirks=False)
    b = O.Oracle.hex(dict(m, attr_bc=kinds, attr_state=st), 3, quirks=False)
    assert np.abs(h.node_coords(3) - a.coords()).max() < 1e-14
    u = cases.random_smooth_state(a.coords(), 3, 1, L=[1.5, 1.0, 0.8])
    assert np.array_equal(a.Mult(u), b.Mult(u))
    with pytest.raises(ValueError):
        QuadMesh.read(tmp_path / fname)


@pytest.mark.parametrize("dim", [2, 3])
def test_gmsh_periodic_section(tmp_path, dim):
    """A Gmsh 2.2 box whose high faces are slaves of the low ones ($Periodic,
    with an Affine line and a chained corner): the mesh identifies the
    slave nodes with their masters, keeps each element's own corner
    coordinates, and the oracle on it equals the periodic box operator."""
    import oracle_lib as O
    from prandtl_amd import cases
    from prandtl_amd.operator import HexMesh
    n, L = (3, 3, 3)[:dim], (1.0, 0.8, 0.6)[:dim]
    nv = [k + 1 for k in n]
    ids = {}
    nodes = []
    for idx in np.ndindex(*reversed(nv)):
        ijk = tuple(reversed(idx))
        ids[ijk] = len(nodes) + 1
        nodes.append([ijk[d] * L[d] / n[d] for d in range(dim)] + [0.0] * (3 - dim))
    g = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(nodes))]
    g += ["%d %.17g %.17g %.17g" % (i + 1, *p) for i, p in enumerate(nodes)] + ["$EndNodes"]
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)] if dim == 2 else \
        [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    els = []
    for idx in np.ndindex(*reversed(n)):
        base = tuple(reversed(idx))
        els.append([ids[tuple(base[d] + c[d] for d in range(dim))] for c in corners])
    g += ["$Elements", str(len(els))]
    g += ["%d %d 2 1 1 %s" % (k + 1, 3 if dim == 2 else 5, " ".join(map(str, e))) for k, e in enumerate(els)]
    g += ["$EndElements", "$Periodic", str(dim)]
    for d in range(dim):
        pairs = [(ids[ijk], ids[tuple(0 if q == d else ijk[q] for q in range(dim))]) for ijk in ids if ijk[d] == n[d]]
        g += ["%d %d %d" % (dim - 1, 2 * d + 2, 2 * d + 1), "Affine 1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1", str(len(pairs))]
        g += ["%d %d" % p for p in pairs]
    g += ["$EndPeriodic"]
    (tmp_path / "per.msh").write_text("\n".join(g) + "\n")
    if dim == 3:
        h = HexMesh.read(tmp_path / "per.msh")
        d = dict(vert=h.vert, elem=h.elem, elem_coords=h.elem_coords)
        orc = O.Oracle.hex(d, 3, quirks=False, blending=True)
    else:
        h = QuadMesh.read(tmp_path / "per.msh")
        d = dict(vert=h.vert, elem=h.elem, elem_coords=h.elem_coords)
        orc = O.Oracle.quad(d, 3, quirks=False, blending=True)
    assert len(np.unique(h.elem)) == int(np.prod(n))      # one vertex per periodic class
    box = O.Oracle(dim, 3, n, (0.0,) * dim, L, quirks=False, blending=True)
    X = box.coords()
    assert np.abs(orc.coords() - X).max() < 1e-15
    u = cases.random_smooth_state(X, dim, 2, L=list(L))
    assert np.abs(orc.Mult(u) - box.Mult(u)).max() < 1e-13 * np.abs(box.Mult(u)).max()


def test_mfem_periodic_hex_nodes(tmp_path):
    """An MFEM v1.0 periodic hexahedral mesh (L2_T1_3D_P1 nodes, one corner
    set per element in lexicographic order, byVDIM): read with per-element
    corners, the oracle on it equals the periodic box operator."""
    import oracle_lib as O
    from prandtl_amd import cases
    from prandtl_amd.operator import HexMesh
    n, L = (3, 3, 3), (1.0, 0.8, 0.6)
    h = [L[d] / n[d] for d in range(3)]
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    lex = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
    els, nodes = [], []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                els.append([(((k + c) % n[2]) * n[1] + (j + b) % n[1]) * n[0] + (i + a) % n[0] for a, b, c in order])
                nodes += [((i + a) * h[0], (j + b) * h[1], (k + c) * h[2]) for a, b, c in lex]
    f = ["MFEM mesh v1.0", "", "dimension", "3", "", "elements", str(len(els))]
    f += ["1 5 " + " ".join(map(str, e)) for e in els]
    f += ["", "boundary", "0", "", "vertices", str(int(np.prod(n))), "", "nodes", "FiniteElementSpace",
          "FiniteElementCollection: L2_T1_3D_P1", "VDim: 3", "Ordering: 1", ""]
    f += ["%.17g %.17g %.17g" % p for p in nodes]
    (tmp_path / "per.mesh").write_text("\n".join(f) + "\n")
    hm = HexMesh.read(tmp_path / "per.mesh")
    assert hm.elem_coords is not None
    orc = O.Oracle.hex(dict(vert=hm.vert, elem=hm.elem, elem_coords=hm.elem_coords), 3, quirks=False)
    box = O.Oracle(3, 3, n, (0.0,) * 3, L, quirks=False)
    X = box.coords()
    assert np.abs(orc.coords() - X).max() < 1e-15
    u = cases.random_smooth_state(X, 3, 4, L=list(L))
    assert np.abs(orc.Mult(u) - box.Mult(u)).max() < 1e-13 * np.abs(box.Mult(u)).max()


@pytest.mark.parametrize("periodic", [False, True])
def test_hexahedral_refinement(tmp_path, periodic):
    """One uniform refinement of a hexahedral box equals the box of twice the
    elements (elements matched by centre); boundary quadrilaterals split into
    four with their attributes; refining a distorted mesh keeps its volume."""
    import oracle_lib as O
    from prandtl_amd import cases
    from prandtl_amd.operator import HexMesh
    n, L = (3, 3, 3), (1.5, 1.0, 0.8)
    if periodic:
        f = tmp_path / "per.mesh"
        order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        lex = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
        h = [L[d] / n[d] for d in range(3)]
        els, nodes = [], []
        for k in range(n[2]):
            for j in range(n[1]):
                for i in range(n[0]):
                    els.append([(((k + c) % n[2]) * n[1] + (j + b) % n[1]) * n[0] + (i + a) % n[0]
                                for a, b, c in order])
                    nodes += [((i + a) * h[0], (j + b) * h[1], (k + c) * h[2]) for a, b, c in lex]
        txt = ["MFEM mesh v1.0", "dimension", "3", "elements", str(len(els))]
        txt += ["1 5 " + " ".join(map(str, e)) for e in els]
        txt += ["boundary", "0", "vertices", str(int(np.prod(n))), "nodes", "FiniteElementSpace",
                "FiniteElementCollection: L2_T1_3D_P1", "VDim: 3", "Ordering: 1", ""]
        txt += ["%.17g %.17g %.17g" % p for p in nodes]
        f.write_text("\n".join(txt) + "\n")
    else:
        _hex_files(tmp_path)
        f = tmp_path / "box.mesh"
    base = HexMesh.read(f)
    ref = HexMesh.read(f, refine=1)
    assert ref.nelem == 8 * base.nelem and len(ref.bdr_attr) == 4 * len(base.bdr_attr)
    assert len(np.unique(ref.elem)) == (int(np.prod([2 * v for v in n])) if periodic else
                                        int(np.prod([2 * v + 1 for v in (3, 2, 2)])))
    if not periodic:
        assert sorted(ref.bdr_attr.tolist()) == sorted(np.repeat(base.bdr_attr, 4).tolist())
        # the distorted mesh keeps its volume under refinement
        va = O.Oracle.hex(dict(vert=base.vert, elem=base.elem, bdr=base.bdr, bdr_attr=base.bdr_attr,
                               attr_bc=["slip"] * 6, attr_state=[np.array([1, 0, 0, 0, 2.5])] * 6), 2).weights().sum()
        vb = O.Oracle.hex(dict(vert=ref.vert, elem=ref.elem, bdr=ref.bdr, bdr_attr=ref.bdr_attr,
                               attr_bc=["slip"] * 6, attr_state=[np.array([1, 0, 0, 0, 2.5])] * 6), 2).weights().sum()
        assert abs(va - vb) < 1e-13 * va
        return
    d = dict(vert=ref.vert, elem=ref.elem, elem_coords=ref.elem_coords)
    orc = O.Oracle.hex(d, 2, quirks=False, blending=True)
    n2 = tuple(2 * v for v in n)
    box = O.Oracle(3, 2, n2, (0.0,) * 3, L, quirks=False, blending=True)
    Xr, Xb = orc.coords(), box.coords()
    cr, cb = Xr.mean(axis=2).T, Xb.mean(axis=2).T
    perm = [int(np.argmin(np.linalg.norm(cb - c, axis=1))) for c in cr]
    assert len(set(perm)) == len(perm)
    assert np.abs(Xr - Xb[:, perm, :]).max() < 1e-14
    u = cases.random_smooth_state(Xb, 3, 5, L=list(L))
    ur = u[:, perm, :]
    a = orc.Mult(ur).reshape(5, -1, orc.K)
    b = box.Mult(u).reshape(5, -1, box.K)[:, perm, :]
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
