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
    ("MFEM mesh v1.0\ndimension\n3\n", "2-D"),
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
