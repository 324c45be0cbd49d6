"""Pins of the oracle's pointwise physics: closed forms from the reference's
test/test_physics.cpp and textbook properties of the numerical fluxes."""
import numpy as np
import pytest

import oracle as O

GAM = 1.4


def cons(rho, u, v, w, p, gam=GAM):
    return np.array([rho, rho * u, rho * v, rho * w, p / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def test_flux_interior_3d():
    # test/test_physics.cpp:58-110
    P, rho, u, v, w = 101325., 1.1, 2.5, 3.5, -4.5
    U = cons(rho, u, v, w, P)
    rhoE = U[4]
    F = O.flux_interior(U)
    Fref = np.array([
        [rho * u, rho * u * u + P, rho * u * v, rho * u * w, (rhoE + P) * u],
        [rho * v, rho * u * v, rho * v * v + P, rho * v * w, (rhoE + P) * v],
        [rho * w, rho * u * w, rho * v * w, rho * w * w + P, (rhoE + P) * w]])
    assert np.allclose(F, Fref, rtol=1e-13, atol=0)


def test_pressure_and_wavespeed():
    # test/test_physics.cpp:344-443
    U = cons(1.1, 2.5, 3.5, -4.5, 101325.)
    assert abs(O.pressure(U) - 101325.) < 1e-9 * 101325.
    assert abs(O.maxwavespeed(cons(1., 0., 0., 0., 1.)) - np.sqrt(1.4)) < 1e-13
    assert abs(O.maxwavespeed(cons(1., 1., 2., 2., 1.)) - (3.0 + np.sqrt(1.4))) < 1e-13


STATES = [
    (cons(1.1, 2.5, -3.5, 0.7, 101325.), cons(0.7, -3.5, -6.0, 1.2, 2 * 101325.)),
    (cons(1.0, 0.3, 0.1, -0.2, 1.0), cons(0.8, 0.1, 0.0, 0.4, 0.6)),
    (cons(1.0, 0.75, -0.1, 0.2, 1.0), cons(0.125, 0.0, 0.0, 0.0, 0.1)),    # Sod-like
    (cons(1.0, -2.0, 0.0, 0.0, 0.4), cons(1.0, 2.0, 0.0, 0.0, 0.4)),       # 123 problem
]
NORMALS = [np.array([-0.4, 0.7, 0.2]), np.array([0.0, 0.0, 0.25]), np.array([1.5, 0.0, 0.0])]


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_numerical_flux_consistency(flux):
    # test/test_physics.cpp:479-520: Fhat(U, U, N) = F(U) . N
    for UL, _ in STATES:
        for N in NORMALS:
            Fh = O.flux_numerical(flux, UL, UL, N)
            Fn = O.flux_interior(UL).T @ N
            assert np.allclose(Fh, Fn, rtol=1e-13, atol=1e-13 * np.max(np.abs(Fn)))


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_numerical_flux_conservation(flux):
    # test/test_physics.cpp:522-576: Fhat(UL, UR, N) = -Fhat(UR, UL, -N)
    for UL, UR in STATES:
        for N in NORMALS:
            a = O.flux_numerical(flux, UL, UR, N)
            b = O.flux_numerical(flux, UR, UL, -N)
            assert np.allclose(a, -b, rtol=1e-12, atol=1e-12 * np.max(np.abs(a)))


@pytest.mark.parametrize("flux", ["HLLC", "Roe"])
def test_stationary_contact_is_exact(flux):
    # An isolated stationary contact (u = 0, equal p, rho jump) is resolved
    # exactly by HLLC and Roe: Fhat = (0, p N, 0).  LLF is not.
    N = np.array([0.0, 0.3, 0.0])
    UL, UR = cons(1.0, 0.0, 0.2, -0.1, 0.9), cons(0.3, 0.0, 0.2, -0.1, 0.9)
    UR[2] = 0.3 * 0.2; UR[3] = 0.3 * -0.1   # same tangential velocity
    UR[4] = 0.9 / (GAM - 1) + 0.5 * 0.3 * (0.2 ** 2 + 0.1 ** 2)
    # rotate so that the normal velocity is zero: normal is y, set v = 0
    UL = cons(1.0, 0.2, 0.0, -0.1, 0.9)
    UR = cons(0.3, 0.2, 0.0, -0.1, 0.9)
    Fh = O.flux_numerical(flux, UL, UR, N)
    assert np.allclose(Fh, [0.0, 0.0, 0.9 * 0.3, 0.0, 0.0], atol=1e-14)
    Fllf = O.flux_numerical("LaxFriedrichs", UL, UR, N)
    assert abs(Fllf[0]) > 1e-3


def test_roe_stationary_shock_is_exact():
    # Roe's linearisation captures a single stationary shock exactly:
    # Rankine-Hugoniot states (Mach 2 normal shock) give Fhat = F(UL) N.
    M1, rho1, p1 = 2.0, 1.0, 1.0
    c1 = np.sqrt(GAM * p1 / rho1)
    u1 = M1 * c1
    rho2 = rho1 * (GAM + 1) * M1 ** 2 / ((GAM - 1) * M1 ** 2 + 2)
    p2 = p1 * (2 * GAM * M1 ** 2 - (GAM - 1)) / (GAM + 1)
    u2 = rho1 * u1 / rho2
    UL, UR = cons(rho1, u1, 0, 0, p1), cons(rho2, u2, 0, 0, p2)
    N = np.array([0.5, 0.0, 0.0])
    FL = O.flux_interior(UL).T @ N
    assert np.allclose(O.flux_interior(UR).T @ N, FL, rtol=1e-13)
    Fh = O.flux_numerical("Roe", UL, UR, N)
    assert np.allclose(Fh, FL, rtol=1e-12, atol=1e-12)


def test_llf_textbook_formula():
    # LLF is fully determined: |N| (1/2 (FL + FR).n - 1/2 a (UR - UL))
    for UL, UR in STATES:
        for N in NORMALS:
            nm = np.linalg.norm(N); n = N / nm
            a = max(O.maxwavespeed(UL), O.maxwavespeed(UR))
            ref = nm * (0.5 * (O.flux_interior(UL).T @ n + O.flux_interior(UR).T @ n) - 0.5 * a * (UR - UL))
            assert np.allclose(O.flux_numerical("LaxFriedrichs", UL, UR, N), ref, rtol=1e-13, atol=1e-13 * np.max(np.abs(ref)))


def test_supersonic_upwinding():
    # all three fluxes reduce to the upwind flux when both states are
    # supersonic in the normal direction
    UL, UR = cons(1.0, 3.0, 0.1, 0.0, 1.0), cons(0.9, 2.8, 0.0, 0.1, 0.8)
    N = np.array([0.2, 0.0, 0.0])
    FL = O.flux_interior(UL).T @ N
    for f in ("HLLC", "Roe"):
        assert np.allclose(O.flux_numerical(f, UL, UR, N), FL, rtol=1e-12, atol=1e-14)


def test_vortex3d_ic_matches_formula():
    # src/physics/euler/functions.cpp:147-191 restated independently
    prm = O.make_params(ic="IsentropicVortex", ic_data=(1, 1, 1, 1, 5))
    g = GAM
    for x in ([0.5, 0.2, 0.55], [0.1, 0.9, 0.3], [0.62, 0.0, 0.41]):
        r2 = ((x[0] - 0.5) ** 2 + (x[2] - 0.5) ** 2) / 0.01
        beta = 5 * np.sqrt(2) / (4 * np.pi) * np.exp(0.5)
        Om = beta * np.exp(-0.5 * r2)
        du, dv = -(x[2] - 0.5) / 0.1 * Om, (x[0] - 0.5) / 0.1 * Om
        dT = -(g - 1) / 2 * Om ** 2
        rho = (1 + dT) ** (1 / (g - 1))
        p = (1 + dT) ** (g / (g - 1)) / g
        ref = cons(rho, 1 + du, 0, dv, p)
        assert np.allclose(O.ic_eval(prm, np.array(x)), ref, rtol=1e-14)


def test_extruded_vortex_advects_and_wraps():
    prm = O.make_params(ic="ExtrudedIsentropicVortex", ic_data=(1, 1, 1, 1, 5),
                        lo=(-5, -5, -5), hi=(5, 5, 5))
    x = np.array([0.3, -0.4, 1.7])
    u0 = O.ic_eval(prm, x, 0.0)
    # translation by (u,v) t, and periodic by 10
    assert np.allclose(O.ic_eval(prm, x + np.array([2.0, 2.0, 0.0]), 2.0), u0, rtol=1e-14)
    assert np.allclose(O.ic_eval(prm, x + np.array([10.0, 0.0, 0.0]), 10.0), u0, rtol=1e-12)
    assert u0[3] == 0.0
    # far field is the free stream rho=1, u=v=1, p=1
    far = O.ic_eval(prm, np.array([4.9, 4.9, 0.0]), 0.0)
    assert np.allclose(far, cons(1, 1, 1, 0, 1), atol=2e-6)
    # entropy p/rho^gamma is uniform (isentropic)
    p = O.pressure(u0)
    assert abs(p / u0[0] ** GAM - 1.0) < 1e-13
