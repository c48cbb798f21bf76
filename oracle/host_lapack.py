"""ORACLE — test infrastructure only. The host LAPACK (OpenBLAS, found at run time) through ctypes: the FP64 path
the reference's own oracle is built from (LAPACK dsytrd / dstedc / dormtr / dsygvd, and the divide-and-conquer
internals dlaed2 / dlaed4 / dlamrg that the reference calls per merge, Stages/ParallelDivideAndConquer.swift:89-145).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module.
"""
import ctypes
import glob
import os

import numpy as np

_L = None
_prefix = ""
LIBRARY = None


def _find():
    cands = []
    try:
        import scipy
        base = os.path.dirname(os.path.dirname(scipy.__file__))
        cands += glob.glob(os.path.join(base, "scipy.libs", "libscipy_openblas-*.so"))
    except Exception:
        pass
    cands += ["liblapack.so.3", "libopenblas.so.0"]
    for c in cands:
        try:
            L = ctypes.CDLL(c)
        except OSError:
            continue
        for pre in ("scipy_", ""):
            if hasattr(L, pre + "dlaed4_"):
                return L, pre, c
    raise RuntimeError("no host LAPACK found")


def lib():
    global _L, _prefix, LIBRARY
    if _L is None:
        _L, _prefix, LIBRARY = _find()
    return _L


def fn(name):
    return getattr(lib(), _prefix + name + "_")


def set_threads(n):
    lib()
    f = getattr(_L, _prefix + "openblas_set_num_threads", None)
    if f is not None:
        f(int(n))


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dlaed4(d, z, i, rho):
    """Root i (0-based) of the secular equation: (lambda, delta)."""
    n = d.size
    d = np.ascontiguousarray(d, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    delta = np.zeros(n)
    dlam = ctypes.c_double(0)
    info = ctypes.c_int(0)
    fn("dlaed4")(_i(n), _i(i + 1), _p(d), _p(z), _p(delta), ctypes.byref(ctypes.c_double(rho)), ctypes.byref(dlam),
                 ctypes.byref(info))
    return dlam.value, delta, info.value


def dlaed2(n1, d, q, rho, z, indxq):
    """LAPACK dlaed2 on copies of its inputs. Returns a dict of all its outputs (index arrays 1-based)."""
    n = d.size
    d = np.array(d, dtype=np.float64)
    q = np.array(q, dtype=np.float64, order="F")
    z = np.array(z, dtype=np.float64)
    indxq = np.array(indxq, dtype=np.int32)
    dlamda = np.zeros(n)
    w = np.zeros(n)
    q2 = np.zeros(n * n + 1)
    indx = np.zeros(n, dtype=np.int32)
    indxc = np.zeros(n, dtype=np.int32)
    indxp = np.zeros(n, dtype=np.int32)
    coltyp = np.zeros(max(n, 4), dtype=np.int32)  # dlaed2 returns the four counts in COLTYP(1:4)
    k = ctypes.c_int(0)
    r = ctypes.c_double(rho)
    info = ctypes.c_int(0)
    fn("dlaed2")(ctypes.byref(k), _i(n), _i(n1), _p(d), _p(q), _i(n), _p(indxq), ctypes.byref(r), _p(z), _p(dlamda),
                 _p(w), _p(q2), _p(indx), _p(indxc), _p(indxp), _p(coltyp), ctypes.byref(info))
    return dict(k=k.value, d=d, q=q, rho=r.value, z=z, dlamda=dlamda, w=w, q2=q2, indx=indx, indxc=indxc,
                indxp=indxp, ctot=coltyp[:4].copy(), indxq=indxq, info=info.value)


def dlamrg(n1, n2, a, s1, s2):
    a = np.ascontiguousarray(a, dtype=np.float64)
    idx = np.zeros(n1 + n2, dtype=np.int32)
    fn("dlamrg")(_i(n1), _i(n2), _p(a), _i(s1), _i(s2), _p(idx))
    return idx


def dstedc(d, e):
    n = d.size
    d = np.array(d, dtype=np.float64)
    e = np.array(np.concatenate([e, [0.0]]), dtype=np.float64)
    z = np.zeros((n, n), order="F")
    lw = 1 + 4 * n + n * n
    liw = 3 + 5 * n
    work = np.zeros(lw)
    iwork = np.zeros(liw, dtype=np.int32)
    info = ctypes.c_int(0)
    fn("dstedc")(b"I", _i(n), _p(d), _p(e), _p(z), _i(n), _p(work), _i(lw), _p(iwork), _i(liw), ctypes.byref(info))
    return d, z, info.value


def dsytrd_upper(a):
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    d, e, tau = np.zeros(n), np.zeros(max(n - 1, 1)), np.zeros(max(n - 1, 1))
    lw = 64 * n
    work = np.zeros(lw)
    info = ctypes.c_int(0)
    fn("dsytrd")(b"U", _i(n), _p(a), _i(n), _p(d), _p(e), _p(tau), _p(work), _i(lw), ctypes.byref(info))
    return a, d, e[: n - 1], tau[: n - 1]


def dormtr_upper(a, tau, z):
    a = np.array(a, dtype=np.float64, order="F")
    z = np.array(z, dtype=np.float64, order="F")
    n, m = z.shape
    lw = 64 * max(m, 1)
    work = np.zeros(lw)
    info = ctypes.c_int(0)
    fn("dormtr")(b"L", b"U", b"N", _i(n), _i(m), _p(a), _i(n), _p(np.ascontiguousarray(tau)), _p(z), _i(n),
                 _p(work), _i(lw), ctypes.byref(info))
    return z


def dsygvd(h, s):
    a = np.array(h, dtype=np.float64, order="F")
    b = np.array(s, dtype=np.float64, order="F")
    n = a.shape[0]
    w = np.zeros(n)
    lw = 1 + 6 * n + 2 * n * n
    liw = 3 + 5 * n
    work = np.zeros(lw)
    iwork = np.zeros(liw, dtype=np.int32)
    info = ctypes.c_int(0)
    fn("dsygvd")(_i(1), b"V", b"U", _i(n), _p(a), _i(n), _p(b), _i(n), _p(w), _p(work), _i(lw), _p(iwork), _i(liw),
                 ctypes.byref(info))
    return w, a, info.value


def dsyevd_values(h):
    """Eigenvalues of the symmetric h by the host LAPACK's dsyevd('N', 'U')."""
    a = np.array(h, dtype=np.float64, order="F")
    n = a.shape[0]
    w = np.zeros(n)
    lw = 2 * n + 1
    liw = 1
    work = np.zeros(lw)
    iwork = np.zeros(liw, dtype=np.int32)
    info = ctypes.c_int(0)
    fn("dsyevd")(b"N", b"U", _i(n), _p(a), _i(n), _p(w), _p(work), _i(lw), _p(iwork), _i(liw), ctypes.byref(info))
    return w, info.value


def dpotrf_upper(s):
    b = np.array(s, dtype=np.float64, order="F")
    n = b.shape[0]
    info = ctypes.c_int(0)
    fn("dpotrf")(b"U", _i(n), _p(b), _i(n), ctypes.byref(info))
    return b, info.value


def dsygst_upper(h, u):
    a = np.array(h, dtype=np.float64, order="F")
    n = a.shape[0]
    info = ctypes.c_int(0)
    fn("dsygst")(_i(1), b"U", _i(n), _p(a), _i(n), _p(np.asfortranarray(u)), _i(n), ctypes.byref(info))
    return a, info.value


def dtrsm_left_upper(u, c):
    """c <- U^{-1} c"""
    c = np.array(c, dtype=np.float64, order="F")
    n, m = c.shape
    one = ctypes.c_double(1.0)
    fn("dtrsm")(b"L", b"U", b"N", b"N", _i(n), _i(m), ctypes.byref(one), _p(np.asfortranarray(u)), _i(n), _p(c),
                _i(n))
    return c
