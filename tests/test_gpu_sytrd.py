"""One-stage reduction (xtb_{d,s}sytrd_upper, xtb_dsytrd_full, xtbeig_{d,s}sytrd) and its back-transformation
(xtb_{d,s}ormtr_upper) on the GPU against the oracle's restatement of the reference (oracle/sbr_oracle.c:
orc_sytrd_upper / orc_ormtr_upper, themselves pinned to the host LAPACK's dsytrd / dormtr in test_oracle.py), and
against invariants: Q^T A Q = T, Q^T Q = I, the spectrum, and an untouched lower triangle."""
import numpy as np
import pytest

import sbr
import workloads
import xtbeig

pytestmark = pytest.mark.gpu

SIZES = [2, 5, 17, 31, 32, 33, 64, 65, 131, 412]


def sym(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return (a + a.T) / 2


def tridiag(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


@pytest.mark.parametrize("n", SIZES)
def test_dsytrd_upper_matches_oracle(torch_cuda, n):
    a0 = sym(n, 2000 + n)
    # a lower triangle that differs from the upper: the 'U' entry point must neither read nor write it
    a_in = np.triu(a0) + np.tril(np.full((n, n), 7.0), -1)
    a, d, e, tau = xtbeig.sytrd_upper(a_in)
    ao, do, eo, to = sbr.sytrd_upper(a0.copy())
    nrm = np.linalg.norm(a0, 2)
    tol = 1e-13 * n * nrm
    assert np.max(np.abs(d - do)) < tol
    assert np.max(np.abs(np.abs(e) - np.abs(eo))) < tol
    assert np.max(np.abs(e - eo)) < tol
    assert np.max(np.abs(tau - to)) < 1e-12 * n
    # Householder vectors above the superdiagonal
    iu = np.triu_indices(n, 2)
    assert np.max(np.abs(a[iu] - ao[iu]), initial=0) < 1e-11 * n
    # strictly lower triangle untouched, bit for bit
    il = np.tril_indices(n, -1)
    assert np.array_equal(a[il], a_in[il])
    # spectrum of T
    assert np.max(np.abs(np.linalg.eigvalsh(tridiag(d, e)) - np.linalg.eigvalsh(a0))) < tol


@pytest.mark.parametrize("n", [33, 131, 412])
def test_dsytrd_full_same_reduction(torch_cuda, n):
    a0 = sym(n, 3000 + n)
    a, d, e, tau = xtbeig.sytrd_upper(a0, full=True)
    au, du, eu, tu = xtbeig.sytrd_upper(a0)
    tol = 1e-13 * n * np.linalg.norm(a0, 2)
    assert np.max(np.abs(d - du)) < tol and np.max(np.abs(e - eu)) < tol
    assert np.max(np.abs(tau - tu)) < 1e-12 * n
    iu = np.triu_indices(n, 2)
    assert np.max(np.abs(a[iu] - au[iu])) < 1e-11 * n


@pytest.mark.parametrize("n", [17, 64, 131, 412])
def test_ssytrd_upper_against_oracle(torch_cuda, n):
    a0 = sym(n, 4000 + n)
    a, d, e, tau = xtbeig.sytrd_upper(a0, dtype=np.float32)
    ao, do, eo, to = sbr.sytrd_upper(a0.copy())
    nrm = np.linalg.norm(a0, 2)
    # the reduction is backward stable: eigenvalues of the FP32 tridiagonal within FP32 rounding of ||A||
    wd = np.linalg.eigvalsh(tridiag(d.astype(np.float64), e.astype(np.float64)))
    assert np.max(np.abs(wd - np.linalg.eigvalsh(a0))) < 2e-6 * n * nrm
    assert np.max(np.abs(d - do)) < 1e-4 * n * nrm


@pytest.mark.parametrize("n,m", [(2, 2), (17, 17), (33, 5), (131, 131), (412, 200)])
def test_dormtr_upper_matches_oracle_and_reduces(torch_cuda, n, m):
    a0 = sym(n, 5000 + n)
    a, d, e, tau = xtbeig.sytrd_upper(a0)
    rng = np.random.default_rng(n)
    z = rng.standard_normal((n, m))
    zg = xtbeig.ormtr_upper(a, tau, z)
    zo = sbr.ormtr_upper(a, tau, z.copy())
    assert np.max(np.abs(zg - zo)) < 1e-12 * n
    q = xtbeig.ormtr_upper(a, tau, np.eye(n))
    assert np.max(np.abs(q.T @ q - np.eye(n))) < 1e-13 * n
    assert np.max(np.abs(q.T @ a0 @ q - tridiag(d, e))) < 1e-13 * n * np.linalg.norm(a0, 2)


def test_sormtr_upper(torch_cuda):
    n = 131
    a0 = sym(n, 6131)
    a, d, e, tau = xtbeig.sytrd_upper(a0, dtype=np.float32)
    q = xtbeig.ormtr_upper(a, tau, np.eye(n, dtype=np.float32), dtype=np.float32).astype(np.float64)
    assert np.max(np.abs(q.T @ q - np.eye(n))) < 1e-5 * n
    assert np.max(np.abs(q.T @ a0 @ q - tridiag(d, e))) < 1e-5 * n * np.linalg.norm(a0, 2)


def test_one_stage_eigensolve_chain(torch_cuda):
    """sytrd -> stedc -> ormtr, the chain xTB's fact_solve runs through the sytrd.c symbols, on the C0 matrix."""
    n = 412
    h, _ = workloads.standard_matrix(n, workloads.SEED_BASE + 0)
    a, d, e, tau = xtbeig.sytrd_upper(h)
    w, zt, info = xtbeig.stedc(d, e)
    assert info == 0
    z = xtbeig.ormtr_upper(a, tau, zt)
    wref = np.linalg.eigvalsh(h)
    nrm = np.linalg.norm(h, 2)
    assert np.max(np.abs(w - wref)) < 1e-12 * nrm
    assert np.linalg.norm(h @ z - z * w) / (n * nrm) < 1e-13
    assert np.max(np.abs(z.T @ z - np.eye(n))) < 1e-12 * n


def test_lapack_sytrd_entry_point(torch_cuda):
    n = 65
    a0 = sym(n, 7065)
    a, d, e, tau, info = xtbeig.sytrd(a0)
    assert info == 0
    au, du, eu, tu = xtbeig.sytrd_upper(a0)
    assert np.array_equal(d, du) and np.array_equal(e, eu) and np.array_equal(tau, tu)


@pytest.mark.parametrize("n", [64, 412])
def test_ssygvd(torch_cuda, n):
    import scipy.linalg as sl
    h, s = workloads.generalized_pair(n, workloads.SEED_BASE + 1)
    w, c, u, info = xtbeig.ssygvd(h, s)
    assert info == 0 and w.dtype == np.float32
    wref = sl.eigh(h, s, eigvals_only=True)
    assert np.max(np.abs(w - wref)) < 1e-6 * np.linalg.norm(h, 2) * 4
    c = c.astype(np.float64)
    assert np.max(np.abs(c.T @ s @ c - np.eye(n))) < 1e-5


def test_one_stage_chain_strided_host_arrays(torch_cuda):
    """The sytrd.c symbol set on strided caller arrays large enough for the pinned staging path (n = 600,
    lda = n + 5, ldz = n + 3): xtb_dsytrd_upper, xtb_dstedc, xtb_dormtr_upper give A Z = Z diag(d) with orthonormal
    Z; the padding rows past n and the strictly lower triangle of A come back untouched."""
    n, lda, ldz = 600, 605, 603
    a0 = sym(n, 4242)
    a = np.full((lda, n), 5.0, order="F")
    a[:n] = np.triu(a0) + np.tril(np.full((n, n), 7.0), -1)
    a_in = a.copy()
    d, e, tau, work = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(32 * n)
    L = xtbeig.lib()
    L.xtb_dsytrd_upper(xtbeig._i(n), xtbeig._ptr(a), xtbeig._i(lda), xtbeig._ptr(d), xtbeig._ptr(e),
                       xtbeig._ptr(tau), xtbeig._ptr(work))
    assert np.all(a[n:] == 5.0)
    il = np.tril_indices(n, -1)
    assert np.array_equal(a[:n][il], a_in[:n][il])
    z = np.full((ldz, n), 3.0, order="F")
    lw, liw = 1 + 4 * n + n * n, 3 + 5 * n
    w2, iw = np.zeros(lw), np.zeros(liw, dtype=np.int32)
    info = xtbeig.ctypes.c_int(0)
    L.xtb_dstedc(xtbeig._i(n), xtbeig._ptr(d), xtbeig._ptr(e), xtbeig._ptr(z), xtbeig._i(ldz), xtbeig._ptr(w2),
                 xtbeig._i(lw), xtbeig._ptr(iw), xtbeig._i(liw), xtbeig.ctypes.byref(info))
    assert info.value == 0
    assert np.all(z[n:] == 3.0)
    L.xtb_dormtr_upper(xtbeig._i(n), xtbeig._i(n), xtbeig._ptr(a), xtbeig._i(lda), xtbeig._ptr(tau), xtbeig._ptr(z),
                       xtbeig._i(ldz), xtbeig._ptr(work))
    assert np.all(z[n:] == 3.0)
    zz = z[:n]
    nrm = np.linalg.norm(a0, 2)
    assert np.max(np.abs(a0 @ zz - zz * d)) < 1e-13 * n * nrm
    assert np.max(np.abs(zz.T @ zz - np.eye(n))) < 1e-13 * n
    assert np.max(np.abs(d - np.linalg.eigvalsh(a0))) < 1e-13 * n * nrm
