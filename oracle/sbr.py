"""ORACLE — test infrastructure only. ctypes access to liboracle.so (sbr_oracle.c, the C restatement of the
reference's band reduction, bulge chase, back-transformations and 1-stage reduction). Built by oracle/Makefile."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None


def lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
        _L = ctypes.CDLL(path)
    return _L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def reduce_to_band(a, b):
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    v1 = np.zeros((n, n), order="F")
    t1 = np.zeros((b, n), order="F")
    lib().orc_reduce_to_band(n, b, _p(a), n, _p(v1), _p(t1))
    return a, v1, t1


def chase_bulges(band_full, b):
    a = np.array(np.tril(band_full), dtype=np.float64, order="F")
    n = a.shape[0]
    v2 = np.zeros((n, n), order="F")
    d, e = np.zeros(n), np.zeros(max(n - 1, 1))
    lib().orc_chase_bulges(n, b, _p(a), n, _p(v2), _p(d), _p(e))
    return d, e[: n - 1], v2


def back_bulges(v2, z, b):
    z = np.array(z, dtype=np.float64, order="F")
    n, m = z.shape
    lib().orc_back_bulges(n, b, _p(np.asfortranarray(v2)), _p(z), n, m)
    return z


def back_band(v1, t1, z, b):
    z = np.array(z, dtype=np.float64, order="F")
    n, m = z.shape
    lib().orc_back_band(n, b, _p(np.asfortranarray(v1)), _p(np.asfortranarray(t1)), _p(z), n, m)
    return z


def sytrd_upper(a):
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    d, e, tau = np.zeros(n), np.zeros(max(n - 1, 1)), np.zeros(max(n - 1, 1))
    lib().orc_sytrd_upper(n, _p(a), n, _p(d), _p(e), _p(tau))
    return a, d, e[: n - 1], tau[: n - 1]


def ormtr_upper(a, tau, z):
    z = np.array(z, dtype=np.float64, order="F")
    n, m = z.shape
    lib().orc_ormtr_upper(n, m, _p(np.asfortranarray(a)), n, _p(np.ascontiguousarray(tau)), _p(z), n)
    return z
