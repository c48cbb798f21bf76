"""The synthetic inputs have the structure the configurations describe."""
import numpy as np
import pytest

import workloads


def test_spectrum_mix_and_degeneracies():
    n = 412
    lam = workloads.spectrum(n, np.random.default_rng(1000))
    assert lam.size == n
    assert np.sum((lam >= -0.9) & (lam < -0.25)) >= int(round(0.45 * n)) - 20
    assert np.sum(lam >= 0.0) >= n - int(round(0.45 * n)) - int(round(0.10 * n)) - 20
    s = np.sort(lam)
    gaps = np.diff(s)
    assert np.sum(gaps == 0) >= 6
    assert np.sum(np.abs(gaps - 1e-9) < 1e-12) >= 4


@pytest.mark.parametrize("n", [17, 130])
def test_standard_matrix_has_the_spectrum(n):
    h, lam = workloads.standard_matrix(n, 1000)
    assert np.allclose(h, h.T, atol=0)
    assert np.max(np.abs(np.linalg.eigvalsh(h) - lam)) < 1e-13 * n


def test_householder_product_is_orthogonal_and_matches_unblocked():
    n = 40
    q = workloads.householder_product(n, np.random.default_rng(3), block=7)
    assert np.max(np.abs(q.T @ q - np.eye(n))) < 1e-13
    rng = np.random.default_rng(3)
    ref = np.eye(n)
    i = 0
    while i < n:
        k = min(7, n - i)
        v = rng.standard_normal((k, n))
        for r in range(k):
            vv = v[r]
            ref = ref @ (np.eye(n) - 2 * np.outer(vv, vv) / (vv @ vv))
        i += k
    assert np.max(np.abs(q - ref)) < 1e-13


def test_overlap_is_spd():
    s = workloads.overlap_matrix(300, np.random.default_rng(1))
    w = np.linalg.eigvalsh(s)
    assert w[0] > 0.49 and np.allclose(s, s.T)
