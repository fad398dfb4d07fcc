"""Pins the oracle's multipole algebra.

1. Against the reference's own moments.c (compiled from its source into
   oracle/_ref; fixture tests/golden/moments_golden.npz written from it).
   moments.c computes in FP64 but with single-precision literals
   (1.0f/3.0f, 0.2f, 0.1f, 1.0f/7.0f, 2.0f/7.0f, 0.4f at moments.c:800-842,
   1006-1025, and onethird = 1.0f/3.0f at :1472); those bound the agreement
   at ~3e-7 relative, so the bar is 2e-6 of the largest component.
2. Against mathematics: the expansion of a point mass converges to 1/|r-x|;
   shifting then evaluating equals making about the new centre; moments are
   trace-free; rescaling round-trips.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "moments_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _close(a, b, tol=2e-6):
    a, b = np.asarray(a), np.asarray(b)
    assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max()), np.abs(a - b).max()


def test_make_matches_reference(gold):
    for inp, ref in zip(gold["make_in"], gold["make_out"]):
        _close(O.mom_make(*inp), ref)


def test_shift_matches_reference(gold):
    for q, ud, ref in zip(gold["shift_q"], gold["shift_ud"], gold["shift_out"]):
        _close(O.mom_shift(q, *ud), ref)


def test_rescale_matches_reference(gold):
    for q, u, ref in zip(gold["shift_q"], gold["rescale_u"], gold["rescale_out"]):
        _close(O.mom_rescale(q, *u), ref, 1e-14)


def test_scaled_add_matches_reference(gold):
    for q, qa, u, ref in zip(gold["shift_q"], gold["sadd_qa"], gold["sadd_u"], gold["sadd_out"]):
        _close(O.mom_scaled_add(q, u[0], qa, u[1]), ref, 1e-14)


def test_momc_matches_reference(gold):
    for q, ref in zip(gold["shift_q"], gold["momc_out"]):
        _close(O.mom_to_momc(q), ref, 1e-15)


def test_eval_matches_reference(gold):
    for q, u, r, ref in zip(gold["shift_q"], gold["eval_u"], gold["eval_r"], gold["eval_out"]):
        pot, a = O.mom_eval(q, u, *r)
        scale = np.abs(ref).max()
        assert abs(pot - ref[0]) <= 2e-6 * scale
        assert np.abs(a - ref[1:]).max() <= 2e-6 * scale


def test_live_reference_if_built():
    """Same comparison against the compiled reference directly (when present)."""
    import ctypes
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(3)
    for _ in range(50):
        m, u = rng.uniform(0.2, 2), rng.uniform(0.5, 2)
        x = rng.normal(size=3)
        ref = np.zeros(22)
        R.momMakeFmomr(P(ref), m, u, *x)
        _close(O.mom_make(m, u, *x), ref)


def test_point_mass_expansion_converges():
    """A few particles about their centre of mass: the hexadecapole expansion
    error falls like (s/r)^5 against the exact sum of 1/|r - x_i|."""
    rng = np.random.default_rng(5)
    x = rng.normal(size=(6, 3)) * 0.1
    m = rng.uniform(0.5, 1.5, 6)
    cm = (m[:, None] * x).sum(0) / m.sum()
    u = 0.3
    q = np.zeros(22)
    for xi, mi in zip(x, m):
        q += O.mom_make(mi, u, *(xi - cm))
    errs = []
    for dist in [1.0, 2.0, 4.0]:
        r = np.array([0.6, -0.3, 0.74]) * dist / np.linalg.norm([0.6, -0.3, 0.74])
        pot, a = O.mom_eval(q, u, *r)
        d = (r + cm) - x
        dn = np.linalg.norm(d, axis=1)
        pe = -(m / dn).sum()
        ae = -(m[:, None] * d / dn[:, None] ** 3).sum(0)
        errs.append((abs(pot - pe) / abs(pe), np.linalg.norm(a - ae) / np.linalg.norm(ae)))
    errs = np.array(errs)
    # each doubling of the distance cuts the error by ~2^5 (order-5 remainder)
    assert (errs[0] / errs[1] > 16).all() and (errs[1] / errs[2] > 16).all(), errs
    assert errs[-1].max() < 1e-6


def test_shift_equals_make_about_new_centre():
    """Moments about the centre of mass shifted to another centre equal the
    moments made about that centre directly (the dipole about the centre of
    mass vanishes, so the shift has nothing to drop)."""
    rng = np.random.default_rng(6)
    x = rng.normal(size=(5, 3)) * 0.2
    m = rng.uniform(0.5, 1.5, 5)
    c_new = np.array([-0.05, 0.04, 0.0])
    u = 0.5
    cm = (m[:, None] * x).sum(0) / m.sum()
    q_cm = sum(O.mom_make(mi, u, *(xi - cm)) for xi, mi in zip(x, m))
    q_new = sum(O.mom_make(mi, u, *(xi - c_new)) for xi, mi in zip(x, m))
    shifted = O.mom_shift(q_cm, u, *(cm - c_new))
    np.testing.assert_allclose(shifted, q_new, rtol=0, atol=1e-12 * np.abs(q_new).max())


def test_moments_trace_free():
    rng = np.random.default_rng(8)
    q = O.mom_make(1.3, 0.7, *rng.normal(size=3))
    mc = O.mom_to_momc(q)
    # MOMC order: m xx yy xy xz yz xxx xyy xxy yyy xxz yyz xyz ... zz xzz yzz zzz xxzz xyzz xzzz yyzz yzzz zzzz
    xx, yy, zz = mc[1], mc[2], mc[22]
    assert abs(xx + yy + zz) < 1e-15
    assert abs(mc[6] + mc[7] + mc[23]) < 1e-15  # xxx + xyy + xzz
    assert abs(mc[26] + mc[29] + mc[31]) < 1e-14  # xxzz + yyzz + zzzz


def test_rescale_round_trip():
    q = np.random.default_rng(9).normal(size=22)
    back = O.mom_rescale(O.mom_rescale(q, 2.5, 1.0), 1.0, 2.5)
    np.testing.assert_allclose(back, q, rtol=1e-14)
