"""The padding of the grouped solve, pinned on the oracle's two-stage chain (CPU).

xtbeig_dev_sygvd_grouped solves a problem of size n inside a group of padded size N as H = [H_p 0; 0 D],
S = [S_p 0; 0 I], and relies on every stage leaving the padded block exactly decoupled, so that the problem's own
eigenvectors are exactly zero in the padded rows and can be picked out by that. This test runs the same chain with
the host LAPACK (Cholesky, congruence, divide and conquer) and the oracle's restatement of the reference's band
reduction, bulge chase and back-transformations on a padded problem, and checks each claim exactly (== 0, not to a
tolerance): the factor and the congruence are block diagonal; the band reduction's reflectors vanish in the padded
rows and on the padded columns, and the band keeps D; the chase's reflectors vanish there and the tridiagonal's
coupling entry is exactly zero; the eigenvectors and their back-transformations keep the blocks apart; and the
problem's own eigenpairs pass the FP64 gates against scipy's solve of the unpadded problem.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import host_lapack as hl
import sbr
import workloads


def pad_values(hp, n, N):
    """The padded diagonal as pad_spectrum sets it: distinct values spread between the smallest and the largest
    diagonal entry of the problem's block (a constant diagonal is not met here)."""
    dg = np.diag(hp)[:n]
    lo, hi = dg.min(), dg.max()
    assert hi > lo
    return np.array([lo + (hi - lo) * (k - n + 1) / (N - n + 1) for k in range(n, N)])


@pytest.mark.parametrize("n,N,b", [(50, 64, 8), (37, 48, 4), (61, 64, 8)])
def test_padded_chain_stays_decoupled(n, N, b):
    h, s = workloads.generalized_pair(n, 4000 + n)
    hpad = np.zeros((N, N))
    hpad[:n, :n] = h
    spad = np.eye(N)
    spad[:n, :n] = s
    # Cholesky: U block diagonal, the identity on the padded block
    u, info = hl.dpotrf_upper(spad)
    assert info == 0
    u = np.triu(u)
    assert np.all(u[:n, n:] == 0) and np.array_equal(u[n:, n:], np.eye(N - n))
    # congruence H' = U^-T H U^-1 (upper triangle), mirrored: zero outside the problem's block
    hp, info = hl.dsygst_upper(hpad, u)
    assert info == 0
    hp = np.triu(hp) + np.triu(hp, 1).T
    assert np.all(hp[:n, n:] == 0) and np.all(hp[n:, :n] == 0) and np.all(hp[n:, n:] == 0)
    dvals = pad_values(hp, n, N)
    hp[np.arange(n, N), np.arange(n, N)] = dvals
    # stage 1: every reflector with tau != 0 is zero in the padded rows; the padded columns' reflectors have tau = 0
    # (a zero column) and a zero column of T, so they act as the identity; the band keeps D. Panel j's T sits at
    # column j of t1, so column c's tau is t1[c % b, c].
    a, v1, t1 = sbr.reduce_to_band(hp, b)
    taus = np.array([t1[c % b, c] for c in range(N)])
    assert np.all(v1[n:, taus != 0] == 0)
    assert np.all(taus[n:] == 0) and np.all(t1[:, n:] == 0)
    band = np.tril(a)
    assert np.all(band[n:, :n] == 0)
    assert np.array_equal(np.diag(band)[n:], dvals) and np.all(np.tril(band[n:, n:], -1) == 0)
    # stage 2: the chase's reflectors vanish on the padded rows; the coupling entry of the tridiagonal is zero
    d, e, v2 = sbr.chase_bulges(band, b)
    assert np.all(v2[n:, :] == 0) and np.all(v2[:, n - 1:] == 0)
    assert e[n - 1] == 0 and np.all(e[n:] == 0) and np.array_equal(d[n:], dvals)
    # divide and conquer and both back-transformations: the blocks stay apart
    w, z, info = hl.dstedc(d, e)
    assert info == 0
    own = np.all(z[n:, :] == 0, axis=0)
    assert own.sum() == n
    assert np.all(z[:n, ~own] == 0)
    zt = sbr.back_band(v1, t1, sbr.back_bulges(v2, z, b), b)
    assert np.all(zt[n:, own] == 0) and np.all(zt[:n, ~own] == 0)
    c = hl.dtrsm_left_upper(u, zt)
    assert np.all(c[n:, own] == 0)
    # the problem's own eigenpairs, in order, against the unpadded problem
    wo, co = w[own], c[:n, own]
    assert np.all(np.diff(wo) >= 0)
    g = workloads.gates(h, s, wo, co, sl.eigh(h, s, eigvals_only=True))
    assert workloads.passes(g), g
    # the padded eigenpairs are (D_k, e_k)
    assert np.allclose(np.sort(w[~own]), dvals, rtol=0, atol=1e-14 * np.abs(dvals).max())
