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


# hex block faces as blockMesh names them: (vertex indices of the block's
# 8 corners, the fixed parametric axis and its side)
BLOCK_FACES = {"x0": (0, 0), "x1": (0, 1), "y0": (1, 0), "y1": (1, 1), "z0": (2, 0), "z1": (2, 1)}


def single_block(corners, n, patches):
    """blockMesh of one hex block with straight edges and uniform grading:
    point (i, j, k) is the trilinear map of (i/nx, j/ny, k/nz) from the 8
    corner vertices (blockMesh vertex order), points and cells i-fastest;
    `patches` is a list of (name, [block faces "x0".."z1"]).  Returns the
    hex_blocks() arrays."""
    corners = np.asarray(corners, np.float64)
    nx, ny, nz = n
    u = np.arange(nx + 1) / nx
    v = np.arange(ny + 1) / ny
    w = np.arange(nz + 1) / nz
    W, V, U = np.meshgrid(w, v, u, indexing="ij")
    U, V, W = U[..., None], V[..., None], W[..., None]
    c = corners
    pts = ((1 - U) * (1 - V) * (1 - W) * c[0] + U * (1 - V) * (1 - W) * c[1] + U * V * (1 - W) * c[2]
           + (1 - U) * V * (1 - W) * c[3] + (1 - U) * (1 - V) * W * c[4] + U * (1 - V) * W * c[5]
           + U * V * W * c[6] + (1 - U) * V * W * c[7]).reshape(-1, 3)

    def pid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    hexes = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                hexes.append((pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i, j + 1, k),
                              pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1)))
    named = []
    for name, faces in patches:
        quads = []
        for f in faces:
            ax, side = BLOCK_FACES[f]
            m = (nx, ny, nz)[ax]
            fix = m if side else 0
            r = [range(nx + 1), range(ny + 1), range(nz + 1)]
            r[ax] = [fix]
            o = [a for a in range(3) if a != ax]
            for b in range((nx, ny, nz)[o[1]]):
                for a in range((nx, ny, nz)[o[0]]):
                    idx = [0, 0, 0]
                    idx[ax] = fix
                    quad = []
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        idx[o[0]], idx[o[1]] = a + da, b + db
                        quad.append(pid(*idx))
                    quads.append(tuple(quad))
        named.append((name, quads))
    return hex_blocks(pts, hexes, named)


# tutorials/solids/linearElasticity/cooksMembrane (constant/polyMesh/blockMeshDict,
# convertToMeters 0.001): trapezoid (0,0), (48,44), (48,60), (0,44) mm, 20x20x1
COOK_CORNERS = [(0.0, 0.0, -0.0005), (0.048, 0.044, -0.0005), (0.048, 0.060, -0.0005), (0.0, 0.044, -0.0005),
                (0.0, 0.0, 0.0005), (0.048, 0.044, 0.0005), (0.048, 0.060, 0.0005), (0.0, 0.044, 0.0005)]
COOK_PATCHES = [("topAndBottom", ["y1", "y0"]), ("left", ["x0"]), ("right", ["x1"]), ("frontAndBack", ["z1", "z0"])]
COOK_E, COOK_NU, COOK_TRACTION = 70e6, 0.3333, (0.0, 6.25e6, 0.0)
COOK_TOP_RIGHT = (0.048, 0.060, 0.0)
COOK_UY_README = 0.03212  # README.md: solids4foam (foam-extend-4.1), 20x20 cells, top-right corner uy
