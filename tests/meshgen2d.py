"""Test helper: periodic boxes of quadrilaterals in the reference's mesh
layout (src/mesh/mesh.cpp:79-175): NodeCoords [n][2], Elem2Nodes [E][4]
(node 2j + i of each element's frame), IFaceData [F][6].  Elements can be
renumbered and turned (one of the four rotations of the square), left
orientations drawn at random; the right orientation of each face is found
by matching the physical face points modulo the period.  warp > 0 moves node
(i, j) by warp h_x sin(2 pi j / ny) in x and warp h_y sin(2 pi i / nx) in y
(bilinear, non-affine elements; opposite box sides stay translates)."""
import numpy as np

import dg2d
import oracle as O

# outward normals of the faces in the element frame (shape.h:32-40)
FACE_NORMAL = {0: (0, -1), 1: (1, 0), 2: (0, 1), 3: (-1, 0)}
ROT = [np.array(r) for r in ([[1, 0], [0, 1]], [[0, -1], [1, 0]], [[-1, 0], [0, -1]], [[0, 1], [-1, 0]])]


def box_mesh2d(nelem, lo, hi, warp=0.0, seed=None, rotate=False, shuffle=False, random_left=False, m=3):
    nx, ny = nelem
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    h = (hi - lo) / np.array(nelem)
    coords = np.zeros(((nx + 1) * (ny + 1), 2))

    def nid(i, j):
        return i + (nx + 1) * j

    for j in range(ny + 1):
        for i in range(nx + 1):
            x = lo + h * np.array([i, j])
            x[0] += warp * h[0] * np.sin(2 * np.pi * (j % ny) / ny)
            x[1] += warp * h[1] * np.sin(2 * np.pi * (i % nx) / nx)
            coords[nid(i, j)] = x
    rng = np.random.default_rng(seed)
    E = nx * ny
    perm = rng.permutation(E) if shuffle else np.arange(E)
    R = [ROT[rng.integers(4)] if rotate else ROT[0] for _ in range(E)]
    e2n = np.zeros((E, 4), dtype=np.int64)
    for iy in range(ny):
        for ix in range(nx):
            b = ix + nx * iy
            for n in range(4):
                v = np.array([2 * (n & 1) - 1, 2 * (n >> 1) - 1])
                g = (R[b] @ v + 1) // 2
                e2n[perm[b], n] = nid(ix + g[0], iy + g[1])
    period = hi - lo
    s, _ = O.gl_rule(m)

    def local_face(b, d, side):
        want = np.zeros(2, dtype=int)
        want[d] = 1 if side else -1
        for f, nrm in FACE_NORMAL.items():
            if np.array_equal(R[b] @ np.array(nrm), want):
                return f
        raise AssertionError

    def face_xy(e, f, o):
        x, _, _, _ = dg2d.bilinear(coords[e2n[e]], dg2d.points_on_face(f, o, s))
        return x

    faces = []
    for iy in range(ny):
        for ix in range(nx):
            b = ix + nx * iy
            for d in range(2):
                c = [ix, iy]
                c[d] = (c[d] + 1) % nelem[d]
                bn = c[0] + nx * c[1]
                eL, eR = perm[b], perm[bn]
                fL, fR = local_face(b, d, 1), local_face(bn, d, 0)
                oL = int(rng.integers(2)) if random_left else 0
                pL = face_xy(eL, fL, oL)
                found = []
                for o in range(2):
                    dd = pL - face_xy(eR, fR, o)
                    dd -= period * np.round(dd / period)
                    if np.max(np.abs(dd)) < 1e-9 * np.max(period):
                        found.append(o)
                assert len(found) == 1, found
                faces.append((eL, fL, oL, eR, fR, found[0]))
    return coords, e2n, np.array(faces, dtype=np.int32)
