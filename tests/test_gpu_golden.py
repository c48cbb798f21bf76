"""The GPU paths on the golden fixtures (tests/golden): eigenvalues within 1e-12 ||H|| of the stored host-LAPACK
ones, eigenvectors equal to the stored ones up to sign where the eigenvalue is isolated, the contract's gates."""
import json
import os

import numpy as np
import pytest

import hsbin
import workloads
import xtbeig

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check(h, s, w, c, e, cref):
    nh = np.linalg.norm(h, 2)
    assert np.max(np.abs(w - e)) <= 1e-12 * nh
    gap = np.minimum(np.diff(e, prepend=-np.inf), np.diff(e, append=np.inf))
    sep = gap > 1e-6 * nh
    dots = np.sum(c * (s @ cref), axis=0)
    assert np.all(np.abs(np.abs(dots[sep]) - 1) < 1e-9)
    assert workloads.passes(workloads.gates(h, s, w, c, e))


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("n", [5, 17, 64, 131])
def test_sygvd_on_golden(torch_cuda, n, mixed):
    h, s, e, cref = hsbin.read(os.path.join(GOLDEN, f"hs_{n}.bin"))
    w, c, u, info = xtbeig.dsygvd(h, s, mixed=mixed)
    assert info == 0
    check(h, s, w, c, e, cref)
    # b on exit, as LAPACK's dsygvd leaves it: the Cholesky factor U (S = U^T U) in the upper triangle, the strictly
    # lower triangle of S untouched
    uu = np.triu(u)
    assert np.all(np.diag(uu) > 0)
    assert np.max(np.abs(uu.T @ uu - s)) < 1e-14 * n * np.linalg.norm(s, 2)
    assert np.array_equal(np.tril(u, -1), np.tril(s, -1))


def test_sygvd_lower_factor_on_golden(torch_cuda):
    """uplo 'L': the lower triangles are referenced, b returns L = U^T (S = L L^T) with the upper triangle of S
    untouched."""
    n = 64
    h, s, e, cref = hsbin.read(os.path.join(GOLDEN, f"hs_{n}.bin"))
    w, c, ll, info = xtbeig.dsygvd(h, s, uplo="L")
    assert info == 0
    check(h, s, w, c, e, cref)
    lo = np.tril(ll)
    assert np.max(np.abs(lo @ lo.T - s)) < 1e-14 * n * np.linalg.norm(s, 2)
    assert np.array_equal(np.triu(ll, 1), np.triu(s, 1))


def test_syevd_on_c0_golden(torch_cuda):
    rec = json.load(open(os.path.join(GOLDEN, "c0_412.json")))
    h, _ = workloads.standard_matrix(rec["n"], rec["seed"])
    w, z, info = xtbeig.syevd(h)
    assert info == 0
    assert np.max(np.abs(w - np.array(rec["eigenvalues"]))) <= 1e-12 * np.linalg.norm(h, 2)
