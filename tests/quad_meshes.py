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


# vertex order of a hexahedron (MFEM's): the z = 0 face counterclockwise,
# then z = 1
HEX_CORNERS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
# local face f = 2 r + side: its corners (the oracle's face-corner order)
HEX_FACES = ((0, 3, 4, 7), (1, 2, 5, 6), (0, 1, 4, 5), (3, 2, 7, 6), (0, 1, 3, 2), (4, 5, 7, 6))


def hex_grid(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), periodic=(True, True, True), perturb=0.0,
             seed=0, rotate=None):
    """An nx x ny x nz hexahedral grid of a box, periodic per direction,
    interior vertices displaced by perturb * h (uniform random), boundary
    quadrilaterals with attributes 1/2 x-low/high, 3/4 y, 5/6 z.  rotate: a
    3x3 rotation applied to all coordinates."""
    rng = np.random.default_rng(seed)
    n = (nx, ny, nz)
    h = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    nvx = [n[d] if periodic[d] else n[d] + 1 for d in range(3)]

    def vid(i, j, k):
        return ((k % nvx[2]) * nvx[1] + (j % nvx[1])) * nvx[0] + (i % nvx[0])

    dsp = perturb * rng.uniform(-1, 1, size=(nvx[2], nvx[1], nvx[0], 3)) * np.array(h)
    for d in range(3):
        if not periodic[d]:
            idx = [slice(None)] * 3
            for end in (0, -1):
                idx[2 - d] = end
                dsp[tuple(idx) + (d,)] = 0.0

    def pos(i, j, k):
        return np.array([lo[0] + i * h[0], lo[1] + j * h[1], lo[2] + k * h[2]]) + \
            dsp[k % nvx[2], j % nvx[1], i % nvx[0]]

    elem, ec = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                cs = [(i + a, j + b, k + c) for a, b, c in HEX_CORNERS]
                elem.append([vid(*c) for c in cs])
                ec.append([pos(*c) for c in cs])
    ec = np.array(ec)
    vert = np.zeros((nvx[0] * nvx[1] * nvx[2], 3))
    for k in range(nvx[2]):
        for j in range(nvx[1]):
            for i in range(nvx[0]):
                vert[vid(i, j, k)] = pos(i, j, k)
    bdr, battr = [], []
    for d in range(3):
        if periodic[d]:
            continue
        o = [x for x in range(3) if x != d]
        for side in (0, 1):
            for b in range(n[o[1]]):
                for a in range(n[o[0]]):
                    quad = []
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        c = [0, 0, 0]
                        c[d] = side * n[d]
                        c[o[0]] = a + da
                        c[o[1]] = b + db
                        quad.append(vid(*c))
                    bdr.append(quad)
                    battr.append(2 * d + 1 + side)
    if rotate is not None:
        R = np.asarray(rotate, float)
        ec = ec @ R.T
        vert = vert @ R.T
    mesh = dict(vert=vert, elem=np.array(elem), bdr=np.array(bdr, dtype=np.int64).reshape(-1, 4),
                bdr_attr=np.array(battr, dtype=np.int32))
    if any(periodic):
        mesh["elem_coords"] = ec
    return mesh


def hex_relabel(mesh, seed=1):
    """The same hexahedral mesh with elements in another order and each
    element's frame turned by a rotation of the cube (one of the 24, chosen
    at random; the vertex list permuted accordingly); returns (mesh, perm,
    rots) with new element k = old element perm[k] under rotation rots[k]."""
    rng = np.random.default_rng(seed)
    ne = len(mesh["elem"])
    perm = rng.permutation(ne)
    R = cube_rotations()
    rots = rng.integers(0, len(R), size=ne)
    elem = np.array([mesh["elem"][p][R[rots[k]]] for k, p in enumerate(perm)])
    out = dict(mesh, elem=elem)
    if "elem_coords" in mesh:
        out["elem_coords"] = np.array([mesh["elem_coords"][p][R[rots[k]]] for k, p in enumerate(perm)])
    return out, perm, rots


def cube_rotations():
    """The 24 proper rotations of the reference cube as vertex permutations:
    new vertex v is old vertex perm[v]."""
    import itertools
    pos = {c: i for i, c in enumerate(HEX_CORNERS)}
    out = []
    for axes in itertools.permutations(range(3)):
        for flips in itertools.product((0, 1), repeat=3):
            M = np.zeros((3, 3), int)
            for r in range(3):
                M[r, axes[r]] = -1 if flips[r] else 1
            if round(np.linalg.det(M)) != 1:
                continue
            perm = []
            for c in HEX_CORNERS:
                x = np.array(c) * 2 - 1
                y = M @ x
                perm.append(pos[tuple(((y + 1) // 2).astype(int))])
            out.append(perm)
    return out
