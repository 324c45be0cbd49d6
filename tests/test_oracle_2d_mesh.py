"""Pins of the 2D mesh oracle (oracle/dg2d.py Mesh2D: bilinear
quadrilaterals given as NodeCoords / Elem2Nodes / IFaceData,
src/mesh/mesh.cpp:79-175) against the 2D box oracle, which the reference's
own convergence data pins, and against what the geometry fixes:

* the box as a mesh (identity numbering and frames) gives the box's
  projection, residual, inverse mass and errors;
* renumbered and turned elements (the four rotations of the square, random
  left orientations) give the same physical solution: the L2 errors and the
  integrals after steps equal the box's;
* on warped (non-affine) elements a uniform state is steady (discrete metric
  identities of the cofactor form and of the straight-edge normals), the
  conserved integrals stay constant, and the vortex error falls at about the
  p+1 rate under refinement;
* a wrong sign or a transposed cofactor (planted in a copy of the oracle)
  breaks the free stream on warped elements.
"""
import numpy as np
import pytest

import dg2d
from meshgen2d import box_mesh2d

VORTEX = dict(ic="IsentropicVortex", ic_data=(1.0, 1.0, 1.0, 1.0, 5.0))


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("p,basis,flux", [(1, "LagrangeQuad", "HLLC"), (2, "LegendreQuad", "Roe"),
                                          (3, "LagrangeQuad", "LaxFriedrichs")])
def test_box_as_a_mesh_is_the_box(p, basis, flux):
    ne, lo, hi = (3, 4), (-5, -5), (5, 5)
    b = dg2d.Box2D(p, basis, flux, ne, lo, hi, **VORTEX)
    m = dg2d.Mesh2D(p, basis, flux, box_mesh2d(ne, lo, hi), lo, hi, **VORTEX)
    b.init_state()
    m.init_state()
    assert relerr(m.U, b.U) < 1e-13
    assert relerr(m.residual(b.U), b.residual(b.U)) < 1e-13
    R = b.residual(b.U)
    assert relerr(m.inv_mass(0.3, R), b.inv_mass(0.3, R)) < 1e-13
    assert np.allclose(m.l2_error(0.2), b.l2_error(0.2), rtol=1e-12, atol=1e-16)


@pytest.mark.parametrize("seed", [1, 2])
def test_turned_and_renumbered_elements(seed):
    ne, lo, hi = (3, 3), (-5, -5), (5, 5)
    b = dg2d.Box2D(2, "LagrangeQuad", "HLLC", ne, lo, hi, **VORTEX)
    mesh = box_mesh2d(ne, lo, hi, seed=seed, rotate=True, shuffle=True, random_left=True)
    m = dg2d.Mesh2D(2, "LagrangeQuad", "HLLC", mesh, lo, hi, **VORTEX)
    for s in (b, m):
        s.init_state()
        s.step("RK4", 0.01, 4)
    assert np.allclose(m.l2_error(m.time), b.l2_error(b.time), rtol=1e-11, atol=1e-16)
    assert np.allclose(m.integrals(), b.integrals(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_warped_free_stream_and_conservation(flux):
    mesh = box_mesh2d((4, 3), (0, 0), (4, 3), warp=0.15, seed=5, rotate=True, random_left=True)
    m = dg2d.Mesh2D(3, "LegendreQuad", flux, mesh, (0, 0), (4, 3), ic="Uniform", ic_data=(1.1, 0.4, -0.3, 2.9))
    m.init_state()
    assert np.max(np.abs(m.residual(m.U))) < 1e-13
    U0 = m.U.copy()
    m.step("RK4", 0.01, 2)
    assert np.max(np.abs(m.U - U0)) < 1e-13
    v = dg2d.Mesh2D(2, "LagrangeQuad", flux, box_mesh2d((4, 4), (-5, -5), (5, 5), warp=0.2), (-5, -5), (5, 5),
                    **VORTEX)
    v.init_state()
    I0 = v.integrals()
    v.step("RK4", 0.01, 5)
    assert np.allclose(v.integrals(), I0, rtol=1e-13, atol=1e-12)


def test_warped_vortex_converges():
    errs = []
    for n in (4, 8):
        m = dg2d.Mesh2D(2, "LagrangeQuad", "HLLC", box_mesh2d((n, n), (-5, -5), (5, 5), warp=0.2), (-5, -5), (5, 5),
                        **VORTEX)
        m.init_state()
        m.step("RK4", 0.02, 10)
        errs.append(m.l2_error(m.time)[0])
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.3, (errs, rate)


@pytest.mark.parametrize("plant", ["sign", "transpose"])
def test_planted_metric_mistake_breaks_free_stream(plant):
    mesh = box_mesh2d((3, 3), (0, 0), (3, 3), warp=0.15)
    m = dg2d.Mesh2D(2, "LagrangeQuad", "Roe", mesh, (0, 0), (3, 3), ic="Uniform", ic_data=(1.1, 0.4, -0.3, 2.9))
    m.init_state()
    if plant == "sign":
        m.cof[:, :, 0, 1] *= -1
    else:
        m.cof = np.swapaxes(m.cof, 2, 3).copy()
    assert np.max(np.abs(m.residual(m.U))) > 1e-6
