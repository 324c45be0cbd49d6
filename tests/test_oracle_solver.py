"""Pins of the oracle's mesh conventions, mass matrices, residual
invariants and the reference's end-to-end 3D isentropic-vortex regression."""
import json
import os

import numpy as np
import pytest

import oracle as O


def _solver(**kw):
    prm = O.make_params(**kw)
    return O.Solver(prm)


def test_face_table_and_orientations():
    s = _solver(order=2, nelem=(3, 2, 4))
    t = s.face_table()
    assert t.shape == (3 * 24, 8)
    # [rankL, elemL, faceL, orientL, rankR, elemR, faceR, orientR]
    assert np.all(t[:, 0] == 0) and np.all(t[:, 4] == 0) and np.all(t[:, 3] == 0)
    d = np.arange(len(t)) % 3
    assert np.all(t[d == 0, 2] == 2) and np.all(t[d == 0, 6] == 4)   # x+ | x-
    assert np.all(t[d == 1, 2] == 3) and np.all(t[d == 1, 6] == 1)   # y+ | y-
    assert np.all(t[d == 2, 2] == 5) and np.all(t[d == 2, 6] == 0)   # z+ | z-
    # orientation codes derived by matching points (shape.cpp:326-428)
    assert np.all(t[d == 0, 7] == 5) and np.all(t[d == 1, 7] == 5) and np.all(t[d == 2, 7] == 4)
    # every element appears exactly 3 times as L and 3 times as R
    assert np.all(np.bincount(t[:, 1], minlength=24) == 3)
    assert np.all(np.bincount(t[:, 5], minlength=24) == 3)


def test_orientation_permutations():
    # orientation 5 reverses each row, 4 transposes (shape.cpp:372-396)
    m = 5
    o5 = O.face_pts_order(5, m * m).reshape(m, m)
    o4 = O.face_pts_order(4, m * m).reshape(m, m)
    idx = np.arange(m * m).reshape(m, m)
    assert np.array_equal(o5, idx[:, ::-1])
    assert np.array_equal(o4, idx.T)
    for o in range(8):
        assert sorted(O.face_pts_order(o, 16)) == list(range(16))


def test_face_points_map_to_faces():
    fp = np.array([[-0.3, 0.6], [0.2, -0.9]])
    fixed = {0: (2, -1), 1: (1, -1), 2: (0, 1), 3: (1, 1), 4: (0, -1), 5: (2, 1)}
    for f, (d, val) in fixed.items():
        for o in range(8):
            x = O.points_on_face(f, o, fp)
            assert np.all(x[:, d] == val)


def test_normals_are_outward_area_weighted():
    s = _solver(order=3, nelem=(2, 3, 4), lo=(-1, 0, 2), hi=(3, 1.5, 4))
    h = np.array([4 / 2, 1.5 / 3, 2 / 4])
    nqf = 25
    N = s.normals(nqf)
    d = np.arange(s.F) % 3
    for k in range(3):
        o = [i for i in range(3) if i != k]
        expect = np.zeros(3); expect[k] = h[o[0]] * h[o[1]] / 4
        assert np.allclose(N[d == k], expect, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("basis", ["LagrangeHex", "LegendreHex"])
def test_inverse_mass_matrix(basis):
    p = 3
    s = _solver(order=p, basis=basis, nelem=(2, 2, 2), lo=(0, 0, 0), hi=(1, 2, 3))
    iMM = s.imm(5)
    # closed form: M = |J| (M1 x M1 x M1), M1 from a GL rule exact for degree 2p
    x, w = O.gl_rule(p + 1)
    V = np.array([O.basis_1d(basis, p, xi)[0] for xi in x])
    M1 = V.T @ np.diag(w) @ V
    J = (0.5 * 1.0 / 2) * (0.5 * 2.0 / 2) * (0.5 * 3.0 / 2)
    M = J * np.kron(np.kron(M1, M1), M1)
    assert np.allclose(iMM @ M, np.eye(64), atol=1e-11)
    if basis == "LegendreHex":
        n = np.arange(p + 1)
        d1 = 2.0 / (2 * n + 1)
        dd = np.kron(np.kron(d1, d1), d1) * J
        assert np.allclose(np.diag(iMM), 1 / dd, rtol=1e-12)
        assert np.max(np.abs(iMM - np.diag(np.diag(iMM)))) < 1e-10 * np.max(np.abs(iMM))


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
@pytest.mark.parametrize("basis", ["LagrangeHex", "LegendreHex"])
def test_free_stream_preservation(flux, basis):
    s = _solver(order=2, basis=basis, flux=flux, nelem=(3, 2, 2), lo=(-1, -1, -1), hi=(1, 1, 1),
                ic="Uniform", ic_data=(1.2, 0.3, -0.2, 0.5, 3.1))
    s.init_state()
    U = s.get_state()
    R = s.residual(U)
    assert np.max(np.abs(R)) < 1e-13


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_discrete_conservation(flux):
    # sum over elements of the residual tested against phi = 1 vanishes on a
    # periodic mesh (Lagrange: partition of unity over all basis functions)
    s = _solver(order=2, basis="LagrangeHex", flux=flux, nelem=(4, 4, 4), lo=(-5, -5, -5), hi=(5, 5, 5),
                ic="ExtrudedIsentropicVortex", ic_data=(1, 1, 1, 1, 5))
    s.init_state()
    U = s.get_state()
    R = s.residual(U)
    tot = R.sum(axis=(0, 1))
    scale = np.abs(R).sum(axis=(0, 1))
    assert np.all(np.abs(tot) < 1e-13 * np.max(scale))
    assert np.max(np.abs(R)) > 1e-3


def test_l2_projection_reproduces_polynomials():
    # a uniform state is projected exactly; the L2 error against it is ~0
    s = _solver(order=3, basis="LegendreHex", nelem=(2, 2, 2), ic="Uniform", ic_data=(1.0, 0.1, 0.2, 0.3, 2.5))
    s.init_state()
    U = s.get_state()
    assert np.allclose(U[:, 0, :], [1.0, 0.1, 0.2, 0.3, 2.5] / np.array(1.0), rtol=1e-13)
    assert np.max(np.abs(U[:, 1:, :])) < 1e-13
    assert np.max(s.l2_error(0.0)) < 1e-14


def test_reference_3d_vortex_regression(golden_dir):
    """examples/euler/3D/isentropic_vortex/regression_data.txt:11-17 on a
    regenerated periodic 4^3 unit cube, p=3, LagrangeHex, HLLC, RK4,
    dt = 1e-3, end = 1.0 (int(1/1e-3) = 1000 steps)."""
    g = json.load(open(os.path.join(golden_dir, "vortex3d_regression.json")))
    s = _solver(order=3, basis="LagrangeHex", flux="HLLC", nelem=(4, 4, 4), lo=(0, 0, 0), hi=(1, 1, 1),
                ic="IsentropicVortex", ic_data=(1, 1, 1, 1, 5))
    s.init_state()
    nsteps = int(g["end"] / g["timestep"])
    assert nsteps == 1000
    s.step("RK4", g["timestep"], nsteps)
    e = s.l2_error(s.time)
    ref = np.array([float(v) for v in g["l2_errors"]])
    for v in (0, 1, 3, 4):
        assert abs(e[v] - ref[v]) < 1e-10 * ref[v], (v, e[v], ref[v])
    assert e[2] < 1e-14 and ref[2] < 1e-14     # rho v stays at round-off


@pytest.mark.parametrize("order,N", [(1, 2), (1, 4), (1, 8), (2, 2), (2, 4), (3, 2)])
def test_extruded_vortex_is_consistent_with_the_reference_2d_convergence_table(golden_dir, order, N):
    """The z-extruded vortex on an N x N x 1 periodic box [-5,5]^2 (HLLC,
    LagrangeHex, RK4, dt 1e-3, t = 1) against the reference's 2D
    quadrilateral convergence study (examples/isentropic_vortex/
    convergence.py:22-86, density L2 error).  The reference's 2D runs use
    their own quadrature and mesh files and its two HLLC series differ from
    each other by up to 0.2%, so this is a consistency check at 2%, not a
    parity pin."""
    series = json.load(open(os.path.join(golden_dir, "vortex2d_convergence.json")))["series"]
    ref = series[f"err_p{order}"][[2, 4, 8].index(N)]
    h = 10.0 / N
    prm = O.make_params(order=order, basis="LagrangeHex", flux="HLLC", nelem=(N, N, 1), lo=(-5, -5, 0),
                        hi=(5, 5, h), gamma=1.4, R=1.0, ic="ExtrudedIsentropicVortex", ic_data=(1, 1, 1, 1, 5))
    s = O.Solver(prm)
    s.init_state()
    s.step("RK4", 1e-3, 1000)
    err = s.l2_error(1.0)
    assert abs(err[0] - ref) / ref < 0.02
    assert err[3] < 1e-14            # no z momentum appears


# ---------------------------------------------------------------------------
# Warped (non-affine, trilinear) periodic box: orc_box_node
# ---------------------------------------------------------------------------

def _warped(order=2, basis="LagrangeHex", flux="Roe", nelem=(3, 4, 3), warp=0.12, **kw):
    kw.setdefault("lo", (-1.0, -1.5, 0.0))
    kw.setdefault("hi", (2.0, 1.5, 1.0))
    kw.setdefault("ic", "ExtrudedIsentropicVortex")
    kw.setdefault("ic_data", (1, 1, 1, 1, 5))
    return O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=nelem, warp=warp, **kw))


def test_warped_nodes_are_periodic_and_displaced():
    prm = O.make_params(nelem=(3, 4, 3), lo=(-1, -1.5, 0), hi=(2, 1.5, 1), warp=0.12)
    lib = O.lib()
    import ctypes as C
    x0, x1 = (C.c_double * 3)(), (C.c_double * 3)()
    moved = 0.0
    for j in range(5):
        for k in range(4):
            lib.orc_box_node(C.byref(prm), 0, j, k, x0)
            lib.orc_box_node(C.byref(prm), 3, j, k, x1)
            assert x1[0] - x0[0] == pytest.approx(3.0, abs=1e-15)
            assert x1[1] == x0[1] and x1[2] == x0[2]
    for i in range(4):
        for j in range(5):
            lib.orc_box_node(C.byref(prm), i, j, 1, x0)
            moved = max(moved, abs(x0[0] - (-1 + i)))
    assert moved > 0.05                     # the mesh is really non-affine


def test_warped_mesh_volume_and_closed_element_surfaces():
    """The displacement maps the periodic cell onto itself, so the elements
    tile a cell of the box's volume; each element's area-weighted normals
    (taken from the left element, helpers.cpp:438-456) close."""
    s = _warped(order=2)
    assert s.total_volume() == pytest.approx(3.0 * 3.0 * 1.0, rel=1e-13)
    nqf = 16
    _, wq = O.gl_rule(4)
    w2 = np.outer(wq, wq).ravel()
    N = s.normals(nqf)
    ft = s.face_table()
    acc = np.zeros((s.E, 3))
    for f in range(s.F):
        v = (w2[:, None] * N[f]).sum(axis=0)
        acc[ft[f, 1]] += v
        acc[ft[f, 5]] -= v
    assert np.max(np.abs(acc)) < 1e-14
    # non-affine: normals vary along a face
    assert np.max(np.abs(N[:, :, 0] - N[:, :1, 0])) > 1e-3


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_warped_free_stream_and_conservation(flux):
    """With normals at the flux points (exact_normals) a uniform state is
    an exact steady solution on the trilinear mesh: the cofactor metric
    terms satisfy the metric identities and the (p+2)-point rules integrate
    them exactly.  The reference places the normals at each face's swapped
    parameters (basis.cpp:332), which breaks this on non-affine faces (the
    residual is O(0.1) here); conservation holds either way."""
    for exact in (True, False):
        s = _warped(order=3, flux=flux, ic="Uniform", ic_data=(1.2, 0.3, -0.2, 0.5, 3.1), exact_normals=exact)
        s.init_state()
        U = s.get_state()
        assert np.max(np.abs(U - U[0, 0])) < 1e-12      # a constant is projected exactly
        r = np.max(np.abs(s.residual(U)))
        if exact:
            assert r < 1e-13
        else:
            assert r > 1e-3
        v = _warped(order=2, flux=flux, exact_normals=exact)
        v.init_state()
        R = v.residual(v.get_state())
        assert np.all(np.abs(R.sum(axis=(0, 1))) < 1e-13 * np.abs(R).sum(axis=(0, 1)).max())
        assert np.max(np.abs(R)) > 1e-3


def test_reference_normals_are_the_exact_normals_at_swapped_face_parameters():
    m = 4                                      # p = 2: (p+2)^2 face points
    ref = _warped(order=2, exact_normals=False).normals(m * m)
    ex = _warped(order=2, exact_normals=True).normals(m * m)
    perm = np.array([(q % m) * m + q // m for q in range(m * m)])
    assert np.allclose(ref, ex[:, perm, :], rtol=0, atol=1e-15)
    assert np.max(np.abs(ref - ex)) > 1e-3


def test_warped_projection_converges_at_order_p_plus_1():
    """L2 projection of the Gaussian bump on warped meshes: the error falls
    like h^(p+1)."""
    errs = []
    for n in (4, 8):
        s = O.Solver(O.make_params(order=2, basis="LegendreHex", flux="Roe", nelem=(n, n, n), lo=(0, 0, 0),
                                   hi=(2, 2, 2), warp=0.1, ic="GaussianBump",
                                   ic_data=(0.3, 0.4, 0.2, -0.1, 0.3, 1.0, 1.0, 1.0, 1.0)))
        s.init_state()
        errs.append(s.l2_error(0.0)[0])
    rate = np.log2(errs[0] / errs[1])
    assert 2.6 < rate < 3.6, (errs, rate)


# ---------------------------------------------------------------------------
# Meshes in the reference's mesh-file layout (orc_create_mesh)
# ---------------------------------------------------------------------------

def _box_vs_mesh(warp, exact, rotate, seed, order=2, basis="LegendreHex", flux="Roe"):
    import meshgen
    ne, lo, hi = (3, 4, 3), (-5.0, -5.0, -5.0), (5.0, 5.0, 5.0)
    kw = dict(order=order, basis=basis, flux=flux, nelem=ne, lo=lo, hi=hi, ic="ExtrudedIsentropicVortex",
              ic_data=(1, 1, 1, 1, 5), warp=warp, exact_normals=exact)
    box = O.Solver(O.make_params(**kw))
    mesh = meshgen.box_mesh(ne, lo, hi, warp=warp, seed=seed, rotate=rotate, shuffle=rotate, random_left=rotate)
    gen = O.Solver(O.make_params(**kw), mesh=mesh)
    return box, gen, mesh


def test_mesh_input_reproduces_the_box_bitwise():
    box, gen, mesh = _box_vs_mesh(warp=0.1, exact=False, rotate=False, seed=0)
    ft = box.face_table()
    assert np.array_equal(ft[:, [1, 2, 3, 5, 6, 7]].astype(np.int64), mesh[2].astype(np.int64))
    box.init_state()
    gen.init_state()
    U = box.get_state()
    assert np.array_equal(gen.get_state(), U)
    assert np.array_equal(gen.residual(U), box.residual(U))


@pytest.mark.parametrize("warp,exact", [(0.0, False), (0.1, True)])
def test_rotated_shuffled_mesh_gives_the_same_solution(warp, exact):
    """Renumbered elements, each with its own local frame (one of the 24
    rotations of the cube) and random left orientations: the discrete
    solution is the same function (Q_p is invariant under the cube's
    rotations), so the L2 errors after projection and after RK4 steps agree
    with the box.  With the reference's normal placement on non-affine faces
    this invariance does not hold (the placement depends on the frame)."""
    box, gen, mesh = _box_vs_mesh(warp=warp, exact=exact, rotate=True, seed=3)
    assert len(set(mesh[2][:, 2])) > 1 and len(set(mesh[2][:, 1])) > 2
    for s in (box, gen):
        s.init_state()
    assert np.allclose(gen.l2_error(0.0), box.l2_error(0.0), rtol=1e-12, atol=1e-16)
    for s in (box, gen):
        s.step("RK4", 0.02, 3)
    assert np.allclose(gen.l2_error(0.06), box.l2_error(0.06), rtol=1e-11, atol=1e-16)
    assert gen.total_volume() == pytest.approx(1000.0, rel=1e-13)
