"""Test helper: periodic box meshes in the reference's mesh-file layout
(src/mesh/mesh.cpp:79-175): NodeCoords [n][3], Elem2Nodes [E][8] (node
4k + 2j + i of each element's own frame) and IFaceData [F][6] = (elemL,
faceL, orientL, elemR, faceR, orientR).

The elements of the periodic box (nodes from the oracle's orc_box_node, so
the same warp) can be renumbered and each given its own local frame (one of
the 24 proper rotations of the cube), and the left orientations drawn at
random; the right orientation of every face is then found by matching the
physical face points (modulo the period) for the eight orientations of the
reference's face point orderings (src/numerics/basis/shape.cpp:296-428)."""
import ctypes as C
import itertools

import numpy as np

import oracle as O

# outward normal of each reference face in the element frame
# (faces 0: z-, 1: y-, 2: x+, 3: y+, 4: x-, 5: z+; shape.h:80-86)
FACE_NORMAL = {0: (0, 0, -1), 1: (0, -1, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (-1, 0, 0), 5: (0, 0, 1)}


def rotations():
    """The 24 proper rotations of the cube as signed permutation matrices."""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), dtype=int)
            for r in range(3):
                R[r, perm[r]] = signs[r]
            if round(np.linalg.det(R)) == 1:
                out.append(R)
    return out


def trilinear(X, xi):
    """Physical points of reference points xi [n][3] of an element X [8][3]."""
    xi = np.atleast_2d(xi)
    out = np.zeros((len(xi), 3))
    for n in range(8):
        i, j, k = n & 1, (n >> 1) & 1, n >> 2
        w = ((1 + (2 * i - 1) * xi[:, 0]) * (1 + (2 * j - 1) * xi[:, 1]) * (1 + (2 * k - 1) * xi[:, 2])) / 8.0
        out += w[:, None] * X[n]
    return out


def box_mesh(nelem, lo, hi, warp=0.0, seed=None, rotate=False, shuffle=False, random_left=False, m=3):
    nx, ny, nz = nelem
    prm = O.make_params(nelem=nelem, lo=lo, hi=hi, warp=warp)
    lib = O.lib()
    x = (C.c_double * 3)()
    coords = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                lib.orc_box_node(C.byref(prm), i, j, k, x)
                coords[nid(i, j, k)] = list(x)
    rng = np.random.default_rng(seed)
    rots = rotations()
    E = nx * ny * nz
    perm = rng.permutation(E) if shuffle else np.arange(E)
    R = [rots[rng.integers(len(rots))] if rotate else np.eye(3, dtype=int) for _ in range(E)]
    e2n = np.zeros((E, 8), dtype=np.int64)
    boxpos = {}
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                b = ix + nx * (iy + ny * iz)
                boxpos[b] = (ix, iy, iz)
                for n in range(8):
                    v = np.array([2 * (n & 1) - 1, 2 * ((n >> 1) & 1) - 1, 2 * (n >> 2) - 1])
                    g = (R[b] @ v + 1) // 2
                    e2n[perm[b], n] = nid(ix + g[0], iy + g[1], iz + g[2])
    period = np.array(hi, dtype=float) - np.array(lo, dtype=float)
    xg, _ = O.gl_rule(m)
    fp = np.array([(xg[i], xg[j]) for j in range(m) for i in range(m)])

    def local_face(b, d, side):
        want = np.zeros(3, dtype=int)
        want[d] = 1 if side else -1
        for f, nrm in FACE_NORMAL.items():
            if np.array_equal(R[b] @ np.array(nrm), want):
                return f
        raise AssertionError

    def face_xyz(e, f, o):
        return trilinear(coords[e2n[e]], O.points_on_face(f, o, fp))

    faces = []
    for b in range(E):
        ix, iy, iz = boxpos[b]
        for d in range(3):
            c = [ix, iy, iz]
            c[d] = (c[d] + 1) % nelem[d]
            bn = c[0] + nx * (c[1] + ny * c[2])
            eL, eR = perm[b], perm[bn]
            fL, fR = local_face(b, d, 1), local_face(bn, d, 0)
            oL = int(rng.integers(8)) if random_left else 0
            pL = face_xyz(eL, fL, oL)
            found = []
            for o in range(8):
                dd = pL - face_xyz(eR, fR, o)
                dd -= period * np.round(dd / period)
                if np.max(np.abs(dd)) < 1e-9 * np.max(period):
                    found.append(o)
            assert len(found) == 1, found
            faces.append((eL, fL, oL, eR, fR, found[0]))
    return coords, e2n, np.array(faces, dtype=np.int32)
