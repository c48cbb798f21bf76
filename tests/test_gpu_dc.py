"""Divide-and-conquer merge parity: the GPU's deflation bookkeeping equals LAPACK dlaed2's on the same inputs
(K, the column-type counts, the INDXC permutation, the secular poles and weights, the deflated eigenvalues), and the
merged eigensystem is the eigendecomposition of the rank-one-modified matrix."""
import numpy as np
import pytest
from scipy.stats import ortho_group

import host_lapack as hl
import xtbeig

pytestmark = pytest.mark.gpu


def make_merge(n, n1, kind, seed):
    rng = np.random.default_rng(seed)
    n2 = n - n1
    d1 = np.sort(rng.uniform(-1, 1, n1))
    d2 = np.sort(rng.uniform(-1, 1, n2))
    if kind == "repeated":
        # some eigenvalues shared exactly across the halves -> Givens deflation
        m = min(n1, n2) // 3
        d2[:m] = d1[:m]
        d2 = np.sort(d2)
    q1 = ortho_group.rvs(n1, random_state=seed) if n1 > 1 else np.ones((1, 1))
    q2 = ortho_group.rvs(n2, random_state=seed + 1) if n2 > 1 else np.ones((1, 1))
    if kind == "smallz":
        # last row of Q1 / first row of Q2 nearly axis-aligned -> many tiny z components
        q1 = np.eye(n1)[:, ::-1].copy()
        q1[:, :2] = ortho_group.rvs(n1, random_state=seed)[:, :2] if n1 > 2 else q1[:, :2]
        q1, _ = np.linalg.qr(q1)
        q2 = np.eye(n2)
    q = np.zeros((n, n))
    q[:n1, :n1] = q1
    q[n1:, n1:] = q2
    d = np.concatenate([d1, d2])
    rho = rng.uniform(0.2, 0.8) * (1 if seed % 2 else -1)
    return d, np.asfortranarray(q), rho


@pytest.mark.parametrize("kind", ["random", "repeated", "smallz"])
@pytest.mark.parametrize("n,n1", [(2, 1), (3, 1), (9, 4), (64, 32), (301, 150), (1204, 602)])
def test_merge_matches_dlaed2(handle, torch_cuda, kind, n, n1):
    torch = torch_cuda
    d, q, rho = make_merge(n, n1, kind, n + 7)
    z = np.concatenate([q[n1 - 1, :n1], q[n1, n1:]])
    indxq = np.concatenate([np.arange(1, n1 + 1), np.arange(1, n - n1 + 1)]).astype(np.int32)
    ref = hl.dlaed2(n1, d, q, rho, z, indxq)
    D = torch.from_numpy(d.copy()).cuda()
    Q = torch.from_numpy(np.ascontiguousarray(q.T)).cuda()
    out = xtbeig.dc_merge(handle, D, Q, n1, rho)
    assert out["status"] == [0, 0, 0, 0]
    k = ref["k"]
    assert out["k"] == k
    assert list(out["ctot"]) == list(ref["ctot"])
    assert out["rho"] == pytest.approx(ref["rho"], rel=1e-15)
    if k > 0:
        assert np.array_equal(out["indxc"], ref["indxc"])
        # values agree to rounding (Givens-deflated poles are recombined in floating point)
        assert np.max(np.abs(out["dlamda"] - ref["dlamda"][:k])) <= 8e-16 * np.max(np.abs(d))
        assert np.max(np.abs(out["w"] - ref["w"][:k])) <= 8e-16
    assert np.max(np.abs(out["d_deflated"][k:] - ref["d"][k:]), initial=0) <= 8e-16 * np.max(np.abs(d))
    # merged eigensystem of the torn matrix M = Q diag(d) Q^T + |rho| u u^T, u = e_{n1 - 1} + sign(rho) e_{n1}
    # (off-diagonal coupling rho, the |rho| taken off both diagonal entries by the tear)
    v = np.zeros(n)
    v[n1 - 1] = 1
    v[n1] = np.sign(rho)
    m = (q * d) @ q.T + abs(rho) * np.outer(v, v)
    w, qm = out["d"], out["q"]
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - np.linalg.eigvalsh(m))) < 1e-14 * n
    assert np.linalg.norm(m @ qm - qm * w) < 1e-14 * n
    assert np.max(np.abs(qm.T @ qm - np.eye(n))) < 1e-14 * n


@pytest.mark.parametrize("n", [3, 20, 300])
def test_secular_roots_match_dlaed4(handle, torch_cuda, n):
    """With no deflation the merged eigenvalues are exactly the secular roots: compare each with host dlaed4."""
    torch = torch_cuda
    d, q, rho = make_merge(n, n // 2, "random", 3 * n)
    z = np.concatenate([q[n // 2 - 1, : n // 2], q[n // 2, n // 2:]])
    indxq = np.concatenate([np.arange(1, n // 2 + 1), np.arange(1, n - n // 2 + 1)]).astype(np.int32)
    ref = hl.dlaed2(n // 2, d, q, rho, z, indxq)
    k = ref["k"]
    roots = np.array([hl.dlaed4(ref["dlamda"][:k], ref["w"][:k], i, ref["rho"])[0] for i in range(k)])
    out = xtbeig.dc_merge(handle, torch.from_numpy(d.copy()).cuda(),
                          torch.from_numpy(np.ascontiguousarray(q.T)).cuda(), n // 2, rho)
    allw = np.sort(np.concatenate([roots, ref["d"][k:]]))
    assert np.max(np.abs(out["d"] - allw)) <= 2e-15 * max(1.0, np.max(np.abs(allw)))


def degenerate_tridiagonal(n, mult, seed):
    """The tridiagonal of a dense matrix whose eigenvalues come in groups of mult equal values (its reduction leaves
    off-diagonals of all sizes; the D&C merges then meet exactly equal poles, Givens deflations and secular roots
    at the deflated values)."""
    import scipy.linalg as sl
    rng = np.random.default_rng(seed)
    base = np.sort(rng.uniform(-1, 1, n // mult + 1))
    lam = np.sort(np.tile(base, mult)[:n])
    qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (qm * lam) @ qm.T
    t = sl.hessenberg((a + a.T) / 2)
    return np.diag(t).copy(), np.diag(t, -1).copy()


@pytest.mark.parametrize("mult,n", [(3, 600), (4, 600), (3, 1204), (2, 1204)])
def test_stedc_degenerate_spectra(handle, torch_cuda, mult, n):
    """Regression: with degenerate spectra the deflated eigenvalues of a merge can be a few ulps out of order (LAPACK's
    dlaed2 leaves them so too); the merges' orderings must still be permutations (a collision there once returned a
    duplicated eigenpair). Eigenvalues against host LAPACK, orthogonality and residual of the vectors."""
    import scipy.linalg as sl
    torch = torch_cuda
    d, e = degenerate_tridiagonal(n, mult, 11 + n)
    dd, ee = torch.from_numpy(d.copy()).cuda(), torch.from_numpy(e.copy()).cuda()
    z = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    handle.dstedc(dd, ee, z)
    assert handle.synchronize() == [0, 0, 0, 0]
    w, zz = dd.cpu().numpy(), z.cpu().numpy().T
    ref = sl.eigh_tridiagonal(d, e, eigvals_only=True)
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    # ascending, up to the rounding that separates the members of a degenerate group (dlamrg keeps such inversions)
    assert np.all(np.diff(w) >= -4 * np.finfo(float).eps * np.max(np.abs(w)))
    assert np.max(np.abs(w - ref)) < 1e-13
    assert np.max(np.abs(zz.T @ zz - np.eye(n))) < 1e-13
    assert np.max(np.abs(t @ zz - zz * w)) < 1e-13


def test_merge_tree_degenerate(handle, torch_cuda):
    """The same tree merge by merge through the one-merge probe (each merge's inputs the probe's own outputs of its
    halves): every merge is an orthonormal eigendecomposition of its torn matrix."""
    torch = torch_cuda
    n = 600
    d, e = degenerate_tridiagonal(n, 3, 11 + n)
    nrm = max(np.abs(d).max(), np.abs(e).max())

    def solve(d, e):
        m = d.size
        if m == 1:
            return d.copy(), np.ones((1, 1))
        n1 = m // 2
        rho = e[n1 - 1]
        d = d.copy()
        d[n1 - 1] -= abs(rho)
        d[n1] -= abs(rho)
        w1, q1 = solve(d[:n1], e[:n1 - 1])
        w2, q2 = solve(d[n1:], e[n1:])
        dd = np.concatenate([w1, w2])
        q = np.zeros((m, m))
        q[:n1, :n1] = q1
        q[n1:, n1:] = q2
        out = xtbeig.dc_merge(handle, torch.from_numpy(dd.copy()).cuda(),
                              torch.from_numpy(np.ascontiguousarray(q.T)).cuda(), n1, rho)
        assert out["status"] == [0, 0, 0, 0], m
        v = np.zeros(m)
        v[n1 - 1] = 1
        v[n1] = np.sign(rho)
        mm = (q * dd) @ q.T + abs(rho) * np.outer(v, v)
        w, qm = out["d"], out["q"]
        assert np.all(np.diff(w) >= -4 * np.finfo(float).eps * np.max(np.abs(w))), m
        assert np.max(np.abs(qm.T @ qm - np.eye(m))) < 1e-13, m
        assert np.max(np.abs(mm @ qm - qm * w)) < 1e-13, m
        return w, qm

    w, _ = solve(d / nrm, e / nrm)
    import scipy.linalg as sl
    assert np.max(np.abs(w * nrm - sl.eigh_tridiagonal(d, e, eigvals_only=True))) < 1e-13


def classic_tridiagonals():
    """Standard hard cases for tridiagonal eigensolvers (LAPACK's test matrices among them): Wilkinson W+ (pairs of
    eigenvalues that agree to many digits), glued Wilkinson blocks (clusters), Clement (integer spectrum), Legendre,
    a split matrix (30 % of the couplings exactly zero), a diagonal, equal diagonal entries with couplings near
    underflow, scales 1e+-300 and a graded matrix."""
    rng = np.random.default_rng(3)
    n = 300
    out = {}

    def wilk(m):
        return np.abs(m - np.arange(2 * m + 1)).astype(float), np.ones(2 * m)

    out["wilkinson201"] = wilk(100)
    d0, e0 = wilk(10)
    for k, delta in ((14, 1e-14), (57, 1e-8)):
        out[f"glued_w21x{k}_{delta:g}"] = (np.tile(d0, k),
                                          np.concatenate([np.concatenate([e0, [delta]]) for _ in range(k)])[:-1])
    out["clement"] = (np.zeros(n), np.sqrt(np.arange(1, n) * (n - np.arange(1, n))))
    i = np.arange(1, n)
    out["legendre"] = (np.zeros(n), i / np.sqrt(4.0 * i * i - 1))
    e = rng.uniform(-1, 1, n - 1)
    e[rng.uniform(size=n - 1) < 0.3] = 0.0
    out["split30"] = (rng.uniform(-1, 1, n), e)
    out["diagonal"] = (rng.uniform(-1, 1, n), np.zeros(n - 1))
    out["const_d_tiny_e"] = (np.ones(n), np.full(n - 1, 1e-300))
    out["scaled1e300"] = (1e300 * rng.uniform(-1, 1, n), 1e300 * rng.uniform(-1, 1, n - 1))
    out["scaled1e-300"] = (1e-300 * rng.uniform(-1, 1, n), 1e-300 * rng.uniform(-1, 1, n - 1))
    out["graded"] = (10.0 ** (-np.arange(n) / 25.0), 0.5 * 10.0 ** (-np.arange(n - 1) / 25.0))
    return out


CLASSIC = classic_tridiagonals()


@pytest.mark.parametrize("name", list(CLASSIC))
def test_stedc_classic_tridiagonals(handle, torch_cuda, name):
    """Host LAPACK's eigenvalues, orthogonality and residual, each relative to the spectrum's largest magnitude."""
    import scipy.linalg as sl
    torch = torch_cuda
    d, e = CLASSIC[name]
    n = d.size
    dd, ee = torch.from_numpy(d.copy()).cuda(), torch.from_numpy(e.copy()).cuda()
    z = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    handle.dstedc(dd, ee, z)
    assert handle.synchronize() == [0, 0, 0, 0]
    w, zz = dd.cpu().numpy(), z.cpu().numpy().T
    ref = sl.eigh_tridiagonal(d, e, eigvals_only=True)
    scale = max(np.abs(ref).max(), np.finfo(float).tiny)
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.max(np.abs(np.sort(w) - ref)) < 1e-13 * scale
    assert np.max(np.abs(t @ zz - zz * w)) < 1e-13 * scale
    assert np.max(np.abs(zz.T @ zz - np.eye(n))) < 1e-13
