"""The oracle's advection flux operators (reference kernels_common.hpp:13-33, 83-88) against what fixes them
without their formulas: the centred forms (h_adv_order 6, v_adv_order 4: up = 0) reconstruct the face value of a
polynomial from its cell averages exactly up to degree 5 (six cells) and 3 (four cells) and no further; the upwind forms
(up = 1) are the textbook fifth- and third-order upwind-biased interpolations, coefficients (2, -13, 47, 27, -3)/60
and (-1, 5, 2)/6 on the upstream side, mirrored for the other wind sign; the folded-constant positive-definite
variants agree with them to rounding; the first-order flux is the Courant-clipped upwind value."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import olib

X6 = np.array([-2.5, -1.5, -0.5, 0.5, 1.5, 2.5])  # cell centres qm3..qp2 around the face at 0
X4 = X6[1:5]


def flux(which, q, ua, up=1.0):
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = C.c_double()
    rc = olib().orc_flux_eval(which, q.ctypes.data_as(C.POINTER(C.c_double)), ua, up, C.byref(out))
    assert rc == 0
    return out.value


def poly(deg, x, rng):
    """cell averages over [x - 1/2, x + 1/2] of a random polynomial of the degree, and its value at the face 0:
    the operators reconstruct face values from cell averages (finite volumes)"""
    c = rng.standard_normal(deg + 1)
    ci = np.polyint(c)
    return np.polyval(ci, x + 0.5) - np.polyval(ci, x - 0.5), np.polyval(c, 0.0)


@pytest.mark.parametrize("which", [0, 2])
def test_sixth_order_centred_is_exact_to_degree_five(which):
    rng = np.random.default_rng(1)
    for deg in range(6):
        q, exact = poly(deg, X6, rng)
        for ua in (1.0, -1.0):
            assert flux(which, q, ua, up=0.0) == pytest.approx(exact, rel=1e-13, abs=1e-13)
    q, exact = poly(6, X6, rng)
    assert abs(flux(which, q, 1.0, up=0.0) - exact) > 1e-6


@pytest.mark.parametrize("which", [1, 3])
def test_fourth_order_centred_is_exact_to_degree_three(which):
    rng = np.random.default_rng(2)
    for deg in range(4):
        q, exact = poly(deg, X4, rng)
        assert flux(which, q, 1.0, up=0.0) == pytest.approx(exact, rel=1e-13, abs=1e-13)
    q, exact = poly(4, X4, rng)
    assert abs(flux(which, q, 1.0, up=0.0) - exact) > 1e-6


@pytest.mark.parametrize("which", [0, 2])
def test_fifth_order_upwind_coefficients(which):
    rng = np.random.default_rng(3)
    up = np.array([2.0, -13.0, 47.0, 27.0, -3.0]) / 60.0
    for _ in range(50):
        q = rng.standard_normal(6)
        assert flux(which, q, 1.0) == pytest.approx(float(up @ q[0:5]), rel=1e-12, abs=1e-14)
        assert flux(which, q, -1.0) == pytest.approx(float(up[::-1] @ q[1:6]), rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("which", [1, 3])
def test_third_order_upwind_coefficients(which):
    rng = np.random.default_rng(4)
    up = np.array([-1.0, 5.0, 2.0]) / 6.0
    for _ in range(50):
        q = rng.standard_normal(4)
        assert flux(which, q, 1.0) == pytest.approx(float(up @ q[0:3]), rel=1e-12, abs=1e-14)
        assert flux(which, q, -1.0) == pytest.approx(float(up[::-1] @ q[1:4]), rel=1e-12, abs=1e-14)


def test_first_order_upwind_with_courant_clipping():
    qm1, q0 = 3.0, 7.0
    for cr in (0.1, 0.3, 0.49):
        assert flux(4, [qm1, q0], cr) == pytest.approx(cr * qm1, rel=1e-15)
        assert flux(4, [qm1, q0], -cr) == pytest.approx(-cr * q0, rel=1e-15)
    for cr in (0.5, 0.8, 2.0):  # at most half a cell's content leaves through one face
        assert flux(4, [qm1, q0], cr) == pytest.approx(0.5 * qm1, rel=1e-15)
        assert flux(4, [qm1, q0], -cr) == pytest.approx(-0.5 * q0, rel=1e-15)
