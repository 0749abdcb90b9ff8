"""polyMesh arrays of single-cell hex blocks (blockMesh with (1 1 1) divisions):
points, hex cells in OpenFOAM vertex order, named boundary patches given as
vertex quads; every other boundary face goes to "defaultFaces".  Faces are in
OpenFOAM order: internal faces upper-triangular (sorted by owner, then
neighbour) and pointing owner -> neighbour, then boundary faces patch by patch."""
import numpy as np

# OpenFOAM hex model faces, outward for a right-handed vertex order
HEX_FACES = [(0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2), (0, 3, 2, 1), (4, 5, 6, 7)]

# patchTest tutorial (tutorials/solids/linearElasticity/patchTest/constant/polyMesh/blockMeshDict)
PATCH_TEST_POINTS = [(0, 0, 0), (3.5, 0, 0), (2.5, 3.5, 0), (0, 3, 0), (1, 1.2, 0), (2, 1, 0), (2, 1.8, 0), (1, 2, 0),
                     (0, 0, 0.1), (3.5, 0, 0.1), (2.5, 3.5, 0.1), (0, 3, 0.1), (1, 1.2, 0.1), (2, 1, 0.1),
                     (2, 1.8, 0.1), (1, 2, 0.1)]
PATCH_TEST_HEXES = [(0, 1, 5, 4, 8, 9, 13, 12), (5, 1, 2, 6, 13, 9, 10, 14), (7, 6, 2, 3, 15, 14, 10, 11),
                    (0, 4, 7, 3, 8, 12, 15, 11), (4, 5, 6, 7, 12, 13, 14, 15)]
PATCH_TEST_PATCHES = [("boundary", [(0, 1, 9, 8), (1, 2, 10, 9), (2, 3, 11, 10), (3, 0, 8, 11)])]


def hex_blocks(points, hexes, patches):
    faces = {}
    for c, hx in enumerate(hexes):
        for fv in HEX_FACES:
            verts = tuple(hx[i] for i in fv)
            faces.setdefault(frozenset(verts), []).append((c, verts))
    internal, boundary = [], {}
    for key, lst in faces.items():
        if len(lst) == 2:
            (c0, v0), (c1, v1) = sorted(lst)
            internal.append((c0, c1, v0))  # owner's outward orientation points to the neighbour
        else:
            boundary[key] = lst[0]
    internal.sort(key=lambda t: (t[0], t[1]))
    names, groups = [], []
    used = set()
    for name, quads in patches:
        grp = []
        for q in quads:
            key = frozenset(q)
            c, v = boundary[key]
            grp.append((c, v))
            used.add(key)
        names.append(name)
        groups.append(grp)
    rest = sorted((boundary[k] for k in boundary if k not in used), key=lambda t: t[0])
    names.append("defaultFaces")
    groups.append(rest)
    fverts, owner, neighbour = [], [], []
    for c0, c1, v in internal:
        fverts.append(v)
        owner.append(c0)
        neighbour.append(c1)
    start, size = [], []
    for grp in groups:
        start.append(len(fverts))
        size.append(len(grp))
        for c, v in grp:
            fverts.append(v)
            owner.append(c)
    fOff = np.concatenate([[0], np.cumsum([len(v) for v in fverts])]).astype(np.int32)
    return dict(points=np.asarray(points, np.float64), fOff=fOff,
                fVerts=np.asarray([i for v in fverts for i in v], np.int32), owner=np.asarray(owner, np.int32),
                neighbour=np.asarray(neighbour, np.int32), nCells=len(hexes), patchStart=np.asarray(start, np.int32),
                patchSize=np.asarray(size, np.int32), names=names)


def linear_spatial(Cf, a, b):
    """linearSpatialDisplacement value a + (b & Cf)"""
    return np.asarray(a) + np.asarray(Cf) @ np.asarray(b).T


PATCH_TEST_A = (1e-6, 4e-6, 0.0)
PATCH_TEST_B = ((2e-6, 3e-6, 0.0), (5e-6, 6e-6, 0.0), (0.0, 0.0, 0.0))
