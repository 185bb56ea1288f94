"""Synthetic 2-D quadrilateral meshes for the tests (plain numpy).

grid(): an nx x ny grid of a box, optionally periodic per direction (the
wrapped vertices are identified in the topology; per-element corner
coordinates carry the geometry, as MFEM's L2 mesh nodes do), with interior
vertices displaced by a smooth or random perturbation, boundary segments
with attributes 1 = x-low, 2 = x-high, 3 = y-low, 4 = y-high.
"""
import numpy as np


def grid(nx, ny, lo=(0.0, 0.0), hi=(1.0, 1.0), periodic=(True, True), perturb=0.0, seed=0,
         rotate=0.0):
    rng = np.random.default_rng(seed)
    hx, hy = (hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny
    vx = nx if periodic[0] else nx + 1
    vy = ny if periodic[1] else ny + 1

    def vid(i, j):
        return (j % vy) * vx + (i % vx)

    # displacement per topological vertex (zero on non-periodic box sides)
    dsp = perturb * rng.uniform(-1, 1, size=(vy, vx, 2)) * np.array([hx, hy])
    if not periodic[0]:
        dsp[:, 0, 0] = dsp[:, -1, 0] = 0.0
    if not periodic[1]:
        dsp[0, :, 1] = dsp[-1, :, 1] = 0.0

    def pos(i, j):
        x = lo[0] + i * hx + dsp[j % vy, i % vx, 0]
        y = lo[1] + j * hy + dsp[j % vy, i % vx, 1]
        return x, y

    elem, ec = [], []
    for j in range(ny):
        for i in range(nx):
            cs = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            elem.append([vid(a, b) for a, b in cs])
            ec.append([pos(a, b) for a, b in cs])
    ec = np.array(ec)
    vert = np.zeros((vx * vy, 2))
    for j in range(vy):
        for i in range(vx):
            vert[vid(i, j)] = pos(i, j)
    bdr, battr = [], []
    if not periodic[0]:
        for j in range(ny):
            bdr.append([vid(0, j), vid(0, j + 1)]); battr.append(1)
            bdr.append([vid(nx, j), vid(nx, j + 1)]); battr.append(2)
    if not periodic[1]:
        for i in range(nx):
            bdr.append([vid(i, 0), vid(i + 1, 0)]); battr.append(3)
            bdr.append([vid(i, ny), vid(i + 1, ny)]); battr.append(4)
    if rotate:
        c, s = np.cos(rotate), np.sin(rotate)
        R = np.array([[c, -s], [s, c]])
        ec = ec @ R.T
        vert = vert @ R.T
    mesh = dict(vert=vert, elem=np.array(elem), bdr=np.array(bdr).reshape(-1, 2),
                bdr_attr=np.array(battr, dtype=np.int32))
    if periodic[0] or periodic[1]:
        mesh["elem_coords"] = ec
    return mesh


def relabel(mesh, seed=1):
    """The same mesh with elements in another order and each element's
    vertex list rotated (its local frame turned by 0, 90, 180 or 270
    degrees); returns (mesh, perm) with new element k = old element perm[k]."""
    rng = np.random.default_rng(seed)
    ne = len(mesh["elem"])
    perm = rng.permutation(ne)
    shift = rng.integers(0, 4, size=ne)
    elem = np.array([np.roll(mesh["elem"][p], -shift[k]) for k, p in enumerate(perm)])
    out = dict(mesh, elem=elem)
    if "elem_coords" in mesh:
        out["elem_coords"] = np.array([np.roll(mesh["elem_coords"][p], -shift[k], axis=0)
                                       for k, p in enumerate(perm)])
    return out, perm


def area(mesh):
    c = mesh["elem_coords"] if "elem_coords" in mesh else mesh["vert"][mesh["elem"]]
    return 0.5 * np.sum(c[:, :, 0] * np.roll(c[:, :, 1], -1, axis=1) - np.roll(c[:, :, 0], -1, axis=1) * c[:, :, 1])
