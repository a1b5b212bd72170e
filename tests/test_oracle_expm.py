"""Pins of the oracle's per-cell exponentials (prep) against known answers
and closed forms that do not depend on the oracle's own formulas.

* Burkardt's known-answer exponentials shipped with the reference
  (tests/golden/burkardt_expm_kat.json), 6 significant digits.
* scipy.linalg.expm of the homogeneous (n+1)x(n+1) matrix, for every branch
  of the 2D closed form (discriminant > 0, < 0, == 0) and the series branch
  (singular linear part), and for the 3D series.
* The 1D closed form against the exact flow e^{a t} x + b (e^{a t} - 1) / a.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def homog(A):
    n = A.shape[0]
    M = np.zeros((n + 1, n + 1))
    M[:n, :] = A
    return M


def test_burkardt_kat(orc):
    kat = json.load(open(os.path.join(HERE, "golden", "burkardt_expm_kat.json")))
    checked = 0
    for case in kat["cases"]:
        A = np.asarray(case["A"], float)
        E = np.asarray(case["exact"], float)
        if case.get("transpose_printed_exact"):
            E = E.T
        if A.shape == (2, 2):
            # linear 2x2 block with zero translation: the 2D affine expm
            T = orc.expm(2, np.array([[A[0, 0], A[0, 1], 0.0, A[1, 0], A[1, 1], 0.0]]))[0]
            got = np.array([[T[0], T[1]], [T[3], T[4]]])
            assert np.allclose(T[[2, 5]], 0.0)
        else:
            # 4x4 with a zero last row: a 3D affine-algebra element, series branch
            T = orc.expm(3, A[:3].reshape(1, 12))[0]
            got = np.vstack([T.reshape(3, 4), [0, 0, 0, 1]])
        np.testing.assert_allclose(got, E, rtol=2e-5, atol=2e-5, err_msg="case %d" % case["id"])
        checked += 1
    assert checked == 6


def _cases_2d(rng):
    out = []
    for _ in range(200):  # generic: discriminant of either sign
        out.append(rng.normal(size=6) * rng.uniform(0.01, 2.0))
    out.append(np.array([0.3, 0.0, 0.2, 0.0, 0.3, -0.1]))     # repeated eigenvalue, (a-d)^2+4bc == 0
    out.append(np.array([0.1, 0.5, 0.7, -0.5, 0.1, 0.2]))     # rotation-like: discriminant < 0
    out.append(np.array([0.2, 0.9, -0.3, 0.4, -0.1, 0.05]))   # discriminant > 0
    return out


def test_expm2d_closed_form_vs_scipy(orc):
    rng = np.random.default_rng(1)
    cases = _cases_2d(rng)
    A = np.array(cases)
    T = orc.expm(2, A)
    signs = set()
    for a, t in zip(A, T):
        M = homog(a.reshape(2, 3))
        ref = sl.expm(M)[:2]
        np.testing.assert_allclose(t.reshape(2, 3), ref, rtol=1e-12, atol=1e-12)
        disc = (a[0] - a[4]) ** 2 + 4 * a[1] * a[3]
        signs.add(np.sign(disc))
    assert signs == {-1.0, 0.0, 1.0}


def test_expm2d_series_branch_vs_scipy(orc):
    rng = np.random.default_rng(2)
    for _ in range(100):
        u = rng.normal(size=2)
        w = rng.normal(size=2)
        lin = np.outer(u, w) * rng.uniform(0.001, 0.1)  # rank one: det == 0 up to rounding
        lin[1] = lin[0] * 0.5  # exactly dependent rows: det is exactly 0
        a = np.array([lin[0, 0], lin[0, 1], rng.normal() * 0.05, lin[1, 0], lin[1, 1], rng.normal() * 0.05])
        assert a[0] * a[4] - a[1] * a[3] == 0.0
        t = orc.expm(2, a[None])[0]
        ref = sl.expm(homog(a.reshape(2, 3)))[:2]
        np.testing.assert_allclose(t.reshape(2, 3), ref, rtol=1e-14, atol=1e-15)


def test_expm2d_zero_is_identity(orc):
    t = orc.expm(2, np.zeros((1, 6)))[0]
    assert np.array_equal(t, [1, 0, 0, 0, 1, 0])


def test_expm3d_series_vs_scipy(orc):
    rng = np.random.default_rng(3)
    A = rng.normal(size=(50, 12)) * 0.02  # dt-scaled tables are small
    T = orc.expm(3, A)
    for a, t in zip(A, T):
        ref = sl.expm(homog(a.reshape(3, 4)))[:3]
        np.testing.assert_allclose(t.reshape(3, 4), ref, rtol=1e-13, atol=1e-15)


def test_prep_sign_and_dt(orc):
    rng = np.random.default_rng(4)
    As = rng.normal(size=(16, 6))
    dt = 0.01
    for sign in (1, -1):
        sA, T = orc.prep(2, As, sign, dt)
        assert np.array_equal(sA, As * sign)
        for a, t in zip(As, T):
            ref = sl.expm(homog((sign * a * dt).reshape(2, 3)))[:2]
            np.testing.assert_allclose(t.reshape(2, 3), ref, rtol=1e-12, atol=1e-14)
    # forward then inverse Trel compose to the identity
    _, Tf = orc.prep(2, As, 1, dt)
    _, Ti = orc.prep(2, As, -1, dt)
    for f, i in zip(Tf, Ti):
        Mf, Mi = homog(f.reshape(2, 3)), homog(i.reshape(2, 3))
        Mf[2, 2] = Mi[2, 2] = 1
        np.testing.assert_allclose(Mf @ Mi, np.eye(3), atol=1e-13)


def test_prep_1d_closed_form(orc):
    a = np.array([0.7, -1.3, 0.0, 2.5])
    b = np.array([0.2, 0.4, -0.3, 0.0])
    dt = 0.05
    sA, T = orc.prep(1, np.stack([a, b], 1), 1, dt)
    x = 0.37
    for k in range(4):
        got = T[k, 0] * x + T[k, 1]
        if a[k] != 0:
            exact = np.exp(a[k] * dt) * x + b[k] * (np.exp(a[k] * dt) - 1) / a[k]
        else:
            exact = x + b[k] * dt
        assert abs(got - exact) < 1e-15


@pytest.mark.parametrize("dim", [2, 3])
def test_prep_sign_exact_negation(orc, dim):
    As = np.array([[0.0, -0.0, 1.5, -2.25] + [0.5] * (dim * (dim + 1) - 4)])
    sA, _ = orc.prep(dim, As, -1, 0.1)
    assert np.array_equal(np.signbit(sA), np.signbit(-As))
