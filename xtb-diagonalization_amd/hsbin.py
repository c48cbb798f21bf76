"""HS.bin: the reference's dump format of one generalized eigenproblem (Sources/EigenBench/main.swift:151-175,
scripts/kernels/refine.f90:64-83): int32 n, then H and S (n x n float64, column-major), optionally followed by the
eigenvalues e (n) and eigenvectors C (n x n, column-major). Little-endian, no padding."""
import numpy as np


def write(path, h, s, e=None, c=None):
    h = np.asarray(h, dtype=np.float64)
    n = h.shape[0]
    with open(path, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.asfortranarray(h).tobytes(order="F"))
        f.write(np.asfortranarray(np.asarray(s, dtype=np.float64)).tobytes(order="F"))
        if e is not None:
            f.write(np.asarray(e, dtype=np.float64).tobytes())
            f.write(np.asfortranarray(np.asarray(c, dtype=np.float64)).tobytes(order="F"))


def read(path):
    """Returns (h, s, e, c); e and c are None when the file holds only H and S."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    body = np.frombuffer(raw[4:], dtype=np.float64)
    nn = n * n
    if body.size not in (2 * nn, 3 * nn + n):
        raise ValueError(f"{path}: {body.size} doubles after n = {n}, expected {2 * nn} or {3 * nn + n}")
    h = body[:nn].reshape((n, n), order="F").copy(order="F")
    s = body[nn:2 * nn].reshape((n, n), order="F").copy(order="F")
    if body.size == 2 * nn:
        return h, s, None, None
    e = body[2 * nn:2 * nn + n].copy()
    c = body[2 * nn + n:].reshape((n, n), order="F").copy(order="F")
    return h, s, e, c
