"""Writes the golden fixtures of tests/golden (run in the build container; the files are committed):
  hs_{n}.bin, n in {5, 17, 64, 131}: the generalized problem workloads.generalized_pair(n, 1001 + n) in the
      reference's HS.bin layout, followed by e and C from the host LAPACK's dsygvd(1, 'V', 'U') (scipy's OpenBLAS);
  c0_412.json: the n = 412 standard problem workloads.standard_matrix(412, 1000): SHA-256 of H's bytes (so that a
      test can regenerate H and know it is the same matrix) and its eigenvalues from the host LAPACK's dsyevd.
Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "xtb-diagonalization_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import hsbin  # noqa: E402
import host_lapack as hl  # noqa: E402
import workloads  # noqa: E402

SIZES = (5, 17, 64, 131)


def main():
    for n in SIZES:
        h, s = workloads.generalized_pair(n, 1001 + n)
        e, c, info = hl.dsygvd(h, s)
        assert info == 0
        hsbin.write(os.path.join(HERE, f"hs_{n}.bin"), h, s, e, c)
    h, lam = workloads.standard_matrix(412, 1000)
    w, info = hl.dsyevd_values(h)
    assert info == 0
    rec = {"n": 412, "seed": 1000, "h_sha256": hashlib.sha256(np.asfortranarray(h).tobytes(order="F")).hexdigest(),
           "eigenvalues": [float(x) for x in w], "source": "host LAPACK dsyevd (scipy OpenBLAS 0.3.29)"}
    with open(os.path.join(HERE, "c0_412.json"), "w") as f:
        json.dump(rec, f)


if __name__ == "__main__":
    main()
