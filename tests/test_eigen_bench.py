"""The dump-driven bench (eigen_bench.py): its "random:N" inputs reproduce the reference's matrices (drand48 checked
against the C library's own drand48), and on the GPU every variant runs on a golden dump and a random matrix with the
reference's accuracy quantities small and each two-stage variant's (d, e) hash the same on every run."""
import ctypes
import os

import numpy as np
import pytest

import eigen_bench

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_drand48_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    for seed in (0, 1, 412, 1204, 2 ** 31 + 5):
        libc.srand48(ctypes.c_long(seed))
        want = np.array([libc.drand48() for _ in range(50)])
        assert np.array_equal(eigen_bench.drand48_stream(seed, 50), want), seed


def test_random_matrix_layout():
    a = eigen_bench.random_matrix(5)
    v = eigen_bench.drand48_stream(5, 15)
    assert np.array_equal(a, a.T)
    assert a[0, 0] == 2 * v[0] - 1 and a[0, 1] == 2 * v[1] - 1 and a[1, 1] == 2 * v[2] - 1 and a[0, 2] == 2 * v[3] - 1


@pytest.mark.gpu
def test_eigen_bench_runs_every_variant(torch_cuda):
    rows = eigen_bench.main([os.path.join(GOLDEN, "hs_131.bin"), "random:97", "--bands", "8,16,32", "--reps", "2"])
    names = {(r["input"], r["variant"]) for r in rows}
    for inp in ("hs_131.bin", "random:97"):
        for v in ("1stage-f64", "2stage-f64-b8", "2stage-f64-b16", "2stage-f64-b32", "syevd-f64", "1stage-f32",
                  "2stage-f32-b8", "syevd-f32"):
            assert (inp, v) in names, (inp, v)
    for r in rows:
        tol = 1e-12 if "f64" in r["variant"] else 2e-5
        assert r["max_dw"] < tol * 10 and r["residual"] < tol and r["orth"] < tol * 10, r
        if "2stage" in r["variant"]:
            assert r["hash_stable"], r
            assert len(r["tridiagonal_fnv1a"]) == 16
