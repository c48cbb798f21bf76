"""Pins the oracle (C restatement of the reference's stages + host LAPACK) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a transposed operand, a wrong sign or index) fails it:
agreement with the host LAPACK's own dsytrd / dormtr / dstedc / dsyevd, invariants that the mathematics fixes
(spectra preserved, exact zeros outside the band, orthogonal transformations), closed forms and brute force on tiny n.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import host_lapack as hl
import sbr
import workloads


def sym(n, seed):
    a = np.random.default_rng(seed).standard_normal((n, n))
    return np.asfortranarray(a + a.T)


@pytest.mark.parametrize("n", [1, 2, 5, 31, 33, 64, 97])
def test_sytrd_matches_host_lapack(n):
    a = sym(n, n)
    oa, od, oe, ot = sbr.sytrd_upper(a)
    ha, hd, he, ht = hl.dsytrd_upper(a)
    s = np.max(np.abs(a))
    assert np.max(np.abs(od - hd)) < 1e-12 * s
    assert np.max(np.abs(oe - he), initial=0) < 1e-12 * s
    assert np.max(np.abs(ot - ht), initial=0) < 1e-12
    assert np.max(np.abs(np.triu(oa, 2) - np.triu(ha, 2)), initial=0) < 1e-11


@pytest.mark.parametrize("n", [2, 33, 70])
def test_ormtr_matches_host_lapack_and_diagonalizes(n):
    a = sym(n, 100 + n)
    oa, d, e, tau = sbr.sytrd_upper(a)
    w, zt = sl.eigh_tridiagonal(d, e)
    z = sbr.ormtr_upper(oa, tau, zt)
    zh = hl.dormtr_upper(oa, tau, zt)
    assert np.max(np.abs(z - zh)) < 1e-12
    assert np.linalg.norm(a @ z - z * w) < 1e-12 * n * np.linalg.norm(a, 2)


def explicit_q1(n, b, v1, t1):
    q = np.eye(n)
    return sbr.back_band(v1, t1, q, b)


@pytest.mark.parametrize("n,b", [(8, 2), (9, 3), (40, 4), (131, 8), (200, 32), (65, 64)])
def test_reduce_to_band(n, b):
    a, lam = workloads.standard_matrix(n, 7)
    band, v1, t1 = sbr.reduce_to_band(a, b)
    assert np.all(np.triu(band, b + 1) == 0) and np.all(np.tril(band, -b - 1) == 0)
    assert np.max(np.abs(np.linalg.eigvalsh(band) - lam)) < 1e-13 * n
    q1 = explicit_q1(n, b, v1, t1)
    assert np.max(np.abs(q1.T @ q1 - np.eye(n))) < 1e-13 * n
    assert np.max(np.abs(q1.T @ a @ q1 - band)) < 1e-13 * n


def random_band(n, b, seed):
    r = np.random.default_rng(seed).standard_normal((n, n))
    r = r + r.T
    return np.triu(np.tril(r, 0), -b) + np.tril(np.triu(r, 1), b)


@pytest.mark.parametrize("n,b", [(3, 2), (7, 2), (8, 3), (60, 5), (130, 16), (200, 32), (150, 64)])
def test_chase(n, b):
    full = random_band(n, b, n + b)
    d, e, v2 = sbr.chase_bulges(full, b)
    w = np.linalg.eigvalsh(full)
    assert np.max(np.abs(sl.eigvalsh_tridiagonal(d, e) - w)) < 1e-13 * n * np.max(np.abs(w))
    q2 = sbr.back_bulges(v2, np.eye(n), b)
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.max(np.abs(q2.T @ q2 - np.eye(n))) < 1e-13 * n
    assert np.max(np.abs(q2.T @ full @ q2 - t)) < 1e-13 * n * np.max(np.abs(full))


def test_chase_brute_force_tiny():
    """n = 5, b = 2 by hand: sweep 0 annihilates A[2, 0] with one reflector on rows 1..2, creating a bulge at
    (4, 1) that reflector (0, 1) on rows 3..4 removes; every reflector is orthogonal and the product diagonalizes."""
    n, b = 5, 2
    full = random_band(n, b, 3)
    d, e, v2 = sbr.chase_bulges(full, b)
    # the stored w of every bulge has |w|^2 in {0, 2} (H = I - w w^T orthogonal)
    for s in range(n - 2):
        k = 0
        while s + 1 + k * b < n:
            rs = s + 1 + k * b
            w = v2[rs: min(rs + b, n), s]
            nn = w @ w
            assert abs(nn) < 1e-15 or abs(nn - 2) < 1e-13
            k += 1
    # explicit product in the forward chase order
    q = np.eye(n)
    for s in range(n - 2):
        k = 0
        while s + 1 + k * b < n:
            rs = s + 1 + k * b
            re = min(rs + b, n)
            wv = np.zeros(n)
            wv[rs:re] = v2[rs:re, s]
            q = q @ (np.eye(n) - np.outer(wv, wv))
            k += 1
    t = q.T @ full @ q
    assert np.max(np.abs(np.tril(t, -2))) < 1e-13
    assert np.allclose(np.diag(t), d, atol=1e-13) and np.allclose(np.diag(t, -1), e, atol=1e-13)


def test_dlaed4_closed_forms():
    # n = 1: lambda = d + rho z^2
    lam, delta, info = hl.dlaed4(np.array([0.3]), np.array([1.0]), 0, 0.5)
    assert info == 0 and abs(lam - 0.8) < 1e-15
    # diag(d) + rho z z^T: roots are the eigenvalues
    rng = np.random.default_rng(2)
    d = np.sort(rng.standard_normal(9))
    z = rng.standard_normal(9)
    z /= np.linalg.norm(z)
    rho = 0.7
    w = np.linalg.eigvalsh(np.diag(d) + rho * np.outer(z, z))
    for i in range(9):
        lam, delta, info = hl.dlaed4(d, z, i, rho)
        assert info == 0 and abs(lam - w[i]) < 1e-14
        assert np.max(np.abs(delta - (d - lam))) < 1e-14


def test_host_dstedc_and_sygvd():
    n = 50
    rng = np.random.default_rng(4)
    d, e = rng.standard_normal(n), rng.standard_normal(n - 1)
    w, z, info = hl.dstedc(d, e)
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert info == 0 and np.max(np.abs(w - np.linalg.eigvalsh(t))) < 1e-13
    h, s = workloads.generalized_pair(n, 5)
    w, c, info = hl.dsygvd(h, s)
    assert info == 0 and workloads.passes(workloads.gates(h, s, w, c, sl.eigh(h, s, eigvals_only=True)))


def test_full_two_stage_chain():
    """dense -> band -> tridiagonal -> host dstedc -> both back-transforms reproduces the eigendecomposition."""
    n, b = 150, 16
    a, lam = workloads.standard_matrix(n, 9)
    band, v1, t1 = sbr.reduce_to_band(a, b)
    d, e, v2 = sbr.chase_bulges(band, b)
    w, zt, info = hl.dstedc(d, e)
    z = sbr.back_band(v1, t1, sbr.back_bulges(v2, zt, b), b)
    g = workloads.gates(a, None, w, z, lam)
    assert workloads.passes(g), g
