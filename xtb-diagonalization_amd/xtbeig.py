"""Host-side interface to libxtbeig.so (the HIP eigensolver), through its C ABI (include/xtbeig.h).

Two layers, mirroring the library:
  * LAPACK-style host calls on numpy arrays: dsygvd / dsygvd_mixed / dsyevd / ssyevd / dstedc / sstedc /
    dsymmetrize_pad, with the same argument meaning and info convention as the Fortran-ABI entry points xTB links.
  * Handle: device-resident calls on torch tensors (torch is used only for device memory and the stream).

The library must be built (xtb-diagonalization_amd/build.py); loading fails loudly otherwise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxtbeig.so")

FP64 = 0
FP32_REFINE = 1

_lib = None

c_int_p = ctypes.POINTER(ctypes.c_int)
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_flt_p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p
ci = ctypes.c_int
ll = ctypes.c_longlong

# symbol -> (restype, argtypes)
SIGNATURES = {
    "xtb_dstedc": (None, [c_int_p, vp, vp, vp, c_int_p, vp, c_int_p, vp, c_int_p, c_int_p]),
    "xtb_sstedc": (None, [c_int_p, vp, vp, vp, c_int_p, vp, c_int_p, vp, c_int_p, c_int_p]),
    "xtb_dsymmetrize_pad": (None, [c_int_p, vp, c_int_p, vp, c_int_p]),
    "xtb_dcongruence_upper": (None, [c_int_p, vp, c_int_p, vp, c_int_p, vp, c_int_p]),
    "xtbeig_hsbin_info": (ci, [ctypes.c_char_p, c_int_p, c_int_p]),
    "xtbeig_hsbin_read": (ci, [ctypes.c_char_p, ci, vp, ci, vp, ci, vp, vp, ci]),
    "xtbeig_hsbin_write": (ci, [ctypes.c_char_p, ci, vp, ci, vp, ci, vp, vp, ci]),
    "xtbeig_hsbin_error": (ctypes.c_char_p, []),
    "xtbeig_fnv1a": (ctypes.c_ulonglong, [vp, ll, ctypes.c_ulonglong]),
    "xtb_dtrmm_back": (None, [c_int_p, c_int_p, vp, c_int_p, vp, c_int_p]),
    "xtbeig_dev_dsyevd_partial": (ci, [vp, ci, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_set_geometry": (ci, [vp, ci, vp, ci]),
    "xtbeig_dev_congruence": (ci, [vp, ci, vp, ci, vp, ci]),
    "xtbeig_dev_back_solve": (ci, [vp, ci, ci, vp, ci]),
    "xtbeig_dev_sygvd_geometry": (ci, [vp, ci, ci, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dsyevd": (None, [ctypes.c_char_p, ctypes.c_char_p, c_int_p, vp, c_int_p, vp, vp, c_int_p, vp, c_int_p,
                             c_int_p]),
    "xtbeig_ssyevd": (None, [ctypes.c_char_p, ctypes.c_char_p, c_int_p, vp, c_int_p, vp, vp, c_int_p, vp, c_int_p,
                             c_int_p]),
    "xtbeig_dsygvd": (None, [c_int_p, ctypes.c_char_p, ctypes.c_char_p, c_int_p, vp, c_int_p, vp, c_int_p, vp, vp,
                             c_int_p, vp, c_int_p, c_int_p]),
    "xtbeig_dsygvd_mixed": (None, [c_int_p, ctypes.c_char_p, ctypes.c_char_p, c_int_p, vp, c_int_p, vp, c_int_p, vp,
                                   vp, c_int_p, vp, c_int_p, c_int_p]),
    "xtbeig_ssygvd": (None, [c_int_p, ctypes.c_char_p, ctypes.c_char_p, c_int_p, vp, c_int_p, vp, c_int_p, vp, vp,
                             c_int_p, vp, c_int_p, c_int_p]),
    "xtb_dsytrd_upper": (None, [c_int_p, vp, c_int_p, vp, vp, vp, vp]),
    "xtb_ssytrd_upper": (None, [c_int_p, vp, c_int_p, vp, vp, vp, vp]),
    "xtb_dsytrd_full": (None, [c_int_p, vp, c_int_p, vp, vp, vp, vp]),
    "xtb_dormtr_upper": (None, [c_int_p, c_int_p, vp, c_int_p, vp, vp, c_int_p, vp]),
    "xtb_sormtr_upper": (None, [c_int_p, c_int_p, vp, c_int_p, vp, vp, c_int_p, vp]),
    "xtbeig_dsytrd": (None, [ctypes.c_char_p, c_int_p, vp, c_int_p, vp, vp, vp, vp, c_int_p, c_int_p]),
    "xtbeig_ssytrd": (None, [ctypes.c_char_p, c_int_p, vp, c_int_p, vp, vp, vp, vp, c_int_p, c_int_p]),
    "xtbeig_dev_dsytrd_upper": (ci, [vp, ci, vp, ci, vp, vp, vp, ci]),
    "xtbeig_dev_ssytrd_upper": (ci, [vp, ci, vp, ci, vp, vp, vp, ci]),
    "xtbeig_dev_dormtr_upper": (ci, [vp, ci, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_sormtr_upper": (ci, [vp, ci, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_density": (ci, [vp, ci, ci, vp, ll, vp, ci, ll, ctypes.c_double, ctypes.c_double, vp, ci, vp, ci, ll,
                                vp]),
    "xtbeig_dev_density_backward": (ci, [vp, ci, ci, vp, ll, vp, ci, ll, ctypes.c_double, vp, ci, vp, vp, ci, ll, vp,
                                         ci, ll, vp, ci, ll]),
    "xtbeig_dev_density_spin": (ci, [vp, ci, ci, vp, ll, vp, ci, ll, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, vp, ci, vp, ci, ll, vp]),
    "xtbeig_dev_density_backward_spin": (ci, [vp, ci, ci, vp, ll, vp, ci, ll, ctypes.c_double, vp, ci, vp, vp, ci, ll,
                                              vp, ci, ll, vp, ci, ll]),
    "xtbeig_dev_density_backward_recompute": (ci, [vp, ci, ci, vp, ci, ll, vp, ci, ll, ci, ctypes.c_double,
                                                   ctypes.c_double, ci, vp, ci, ll, vp, ci, ll, vp, ci, ll]),
    "xtbeig_set_scc_precision": (ci, [vp, ctypes.c_double, ci, ci]),
    "xtbeig_scc_begin": (ci, [vp]),
    "xtbeig_dev_scc_sygvd": (ci, [vp, ci, vp, ci, vp, ci, vp, vp, ci, ctypes.c_double, c_int_p]),
    "xtbeig_get_scc_precision": (ci, [vp, ctypes.POINTER(ctypes.c_double), c_int_p, c_int_p, c_int_p, c_int_p,
                                      c_int_p]),
    "xtbeig_dev_sygvd_graph": (ci, [vp, ci, ci, vp, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_sygvd_batched": (ci, [vp, ci, ci, c_int_p, ctypes.POINTER(vp), c_int_p, ctypes.POINTER(vp), c_int_p,
                                      ctypes.POINTER(vp), ctypes.POINTER(vp), c_int_p, ci, ci,
                                      ctypes.POINTER(ctypes.c_float)]),
    "xtbeig_dev_sygvd_grouped": (ci, [vp, ci, ci, c_int_p, ctypes.POINTER(vp), c_int_p, ctypes.POINTER(vp), c_int_p,
                                      ctypes.POINTER(vp), ctypes.POINTER(vp), c_int_p, ci, ci, c_int_p,
                                      ctypes.POINTER(ctypes.c_float)]),
    "xtbeig_autotune": (ci, [vp, ci, ci, ci, c_int_p, c_int_p, ctypes.POINTER(ctypes.c_float), c_int_p]),
    "xtbeig_get_band_table": (ci, [vp, ci, c_int_p, c_int_p]),
    "xtbeig_set_band_table": (ci, [vp, ci, c_int_p]),
    "xtbeig_band_buckets": (ci, []),
    "xtbeig_create": (ci, [ctypes.POINTER(vp), ci]),
    "xtbeig_destroy": (ci, [vp]),
    "xtbeig_set_stream": (ci, [vp, vp]),
    "xtbeig_get_stream": (vp, [vp]),
    "xtbeig_set_tuning": (ci, [vp, ci, ci]),
    "xtbeig_default_band": (ci, [ci, ci]),
    "xtbeig_set_test_limits": (ci, [vp, ci]),
    "xtbeig_synchronize": (ci, [vp, c_int_p]),
    "xtbeig_refine_check": (ci, [vp, c_dbl_p]),
    "xtbeig_last_error": (ctypes.c_char_p, []),
    "xtbeig_set_gemm_tile": (ci, [ci]),
    "xtbeig_chase_in_lds": (ci, [ci, ci, ci]),
    "xtbeig_gemm_tile_width": (ci, [ci, ci, ci, ci]),
    "xtbeig_gemm_tile_buckets": (ci, []),
    "xtbeig_get_gemm_tile_table": (ci, [ci, c_int_p, c_int_p]),
    "xtbeig_set_gemm_tile_table": (ci, [ci, c_int_p]),
    "xtbeig_autotune_gemm": (ci, [vp, ci, ci, ci, c_int_p, c_int_p, c_int_p]),
    "xtbeig_dev_sygvd": (ci, [vp, ci, ci, vp, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_dsyevd": (ci, [vp, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_ssyevd": (ci, [vp, ci, vp, ci, vp, vp, ci]),
    "xtbeig_dev_dstedc": (ci, [vp, ci, vp, vp, vp, ci]),
    "xtbeig_dev_dgemm": (ci, [vp, ctypes.c_char, ctypes.c_char, ci, ci, ci, ctypes.c_double, vp, ci, vp, ci,
                              ctypes.c_double, vp, ci]),
    "xtbeig_dev_sgemm": (ci, [vp, ctypes.c_char, ctypes.c_char, ci, ci, ci, ctypes.c_float, vp, ci, vp, ci,
                              ctypes.c_float, vp, ci]),
    "xtbeig_dev_potrf_inverse": (ci, [vp, ci, vp, ci, vp, ci, vp, ci]),
    "xtbeig_dev_potrf_inverse_noalloc": (ci, [vp, ci, vp, ci, vp, ci, vp, ci]),
    "xtbeig_band_elems": (ctypes.c_longlong, [ci, ci]),
    "xtbeig_dev_dreduce_to_band": (ci, [vp, ci, ci, vp, ci, vp, vp, vp]),
    "xtbeig_dev_sreduce_to_band": (ci, [vp, ci, ci, vp, ci, vp, vp, vp]),
    "xtbeig_dev_dchase": (ci, [vp, ci, ci, vp, vp, vp, vp]),
    "xtbeig_dev_schase": (ci, [vp, ci, ci, vp, vp, vp, vp]),
    "xtbeig_dev_dback_bulges": (ci, [vp, ci, ci, vp, vp, ci, ci]),
    "xtbeig_dev_sback_bulges": (ci, [vp, ci, ci, vp, vp, ci, ci]),
    "xtbeig_dev_dc_merge": (ci, [vp, ci, ci, ctypes.c_double, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp]),
    "xtbeig_dev_dback_band": (ci, [vp, ci, ci, vp, vp, vp, ci, ci]),
    "xtbeig_dev_sback_band": (ci, [vp, ci, ci, vp, vp, vp, ci, ci]),
}


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run xtb-diagonalization_amd/build.py (no CPU fallback)")
        # torch's wheel carries its own HIP and HSA runtimes; loading it first makes libxtbeig.so bind to those
        # same copies (by soname), so that torch's device pointers and streams are valid in the library's calls.
        # Loaded the other way round, the process would hold two runtimes and the second would see no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _ptr(a):
    return a.ctypes.data_as(vp)


def set_gemm_tile(width):
    """Force one GEMM tile width (64 or 128) for every product of the process; 0 restores the per-product choice."""
    if lib().xtbeig_set_gemm_tile(int(width)) != 0:
        raise ValueError(last_error())


def gemm_tile_width(m, n, batch=1, precision=0):
    return lib().xtbeig_gemm_tile_width(int(m), int(n), int(batch), int(precision))


def gemm_tile_table(precision=0):
    """[(first 64 x 64 tile count of the bucket, width or 0)] for every bucket."""
    nb = lib().xtbeig_gemm_tile_buckets()
    st = (ctypes.c_int * nb)()
    wd = (ctypes.c_int * nb)()
    if lib().xtbeig_get_gemm_tile_table(int(precision), st, wd) != 0:
        raise RuntimeError(last_error())
    return [(st[q], wd[q]) for q in range(nb)]


def set_gemm_tile_table(precision, widths):
    nb = lib().xtbeig_gemm_tile_buckets()
    w = (ctypes.c_int * nb)(*[int(x) for x in widths])
    if lib().xtbeig_set_gemm_tile_table(int(precision), w) != 0:
        raise ValueError(last_error())


def last_error():
    return lib().xtbeig_last_error().decode()


# ---- LAPACK-style host calls -----------------------------------------------------------------------------------

def _f64(a):
    return np.array(a, dtype=np.float64, order="F", copy=True)


def dsygvd(h, s, mixed=False, uplo="U"):
    """Generalized eigenpairs of (h, s) through xtbeig_dsygvd[_mixed]. Returns (w, c, u, info): c has S-orthonormal
    columns, u is s's Cholesky factor as LAPACK leaves it in b."""
    a = _f64(h)
    b = _f64(s)
    n = a.shape[0]
    w = np.zeros(n)
    lw = 1 + 6 * n + 2 * n * n
    liw = 3 + 5 * n
    work = np.zeros(max(lw, 1))
    iwork = np.zeros(max(liw, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    f = lib().xtbeig_dsygvd_mixed if mixed else lib().xtbeig_dsygvd
    f(_i(1), b"V", uplo.encode(), _i(n), _ptr(a), _i(n), _ptr(b), _i(n), _ptr(w), _ptr(work), _i(lw), _ptr(iwork),
      _i(liw), ctypes.byref(info))
    return w, a, b, info.value


def ssygvd(h, s, uplo="U"):
    """FP32 interface of the generalized solve (xtbeig_ssygvd). Returns (w, c, u, info) in float32."""
    a = np.array(h, dtype=np.float32, order="F", copy=True)
    b = np.array(s, dtype=np.float32, order="F", copy=True)
    n = a.shape[0]
    w = np.zeros(n, dtype=np.float32)
    lw = 1 + 6 * n + 2 * n * n
    liw = 3 + 5 * n
    work = np.zeros(max(lw, 1), dtype=np.float32)
    iwork = np.zeros(max(liw, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    lib().xtbeig_ssygvd(_i(1), b"V", uplo.encode(), _i(n), _ptr(a), _i(n), _ptr(b), _i(n), _ptr(w), _ptr(work),
                        _i(lw), _ptr(iwork), _i(liw), ctypes.byref(info))
    return w, a, b, info.value


def sytrd_upper(a_in, dtype=np.float64, full=False):
    """xtb_{d,s}sytrd_upper (full: xtb_dsytrd_full) on a copy of a_in. Returns (a, d, e, tau)."""
    a = np.array(a_in, dtype=dtype, order="F", copy=True)
    n = a.shape[0]
    d = np.zeros(max(n, 1), dtype=dtype)
    e = np.zeros(max(n, 1), dtype=dtype)
    tau = np.zeros(max(n, 1), dtype=dtype)
    work = np.zeros(max(32 * n, 1), dtype=dtype)
    if full:
        if dtype != np.float64:
            raise ValueError("xtb_dsytrd_full is FP64 only")
        f = lib().xtb_dsytrd_full
    else:
        f = lib().xtb_dsytrd_upper if dtype == np.float64 else lib().xtb_ssytrd_upper
    f(_i(n), _ptr(a), _i(n), _ptr(d), _ptr(e), _ptr(tau), _ptr(work))
    return a, d[:n], e[: max(n - 1, 0)], tau[: max(n - 1, 0)]


def ormtr_upper(a, tau, z_in, dtype=np.float64):
    """xtb_{d,s}ormtr_upper: z = Q z, Q from sytrd_upper's (a, tau)."""
    a = np.asfortranarray(a, dtype=dtype)
    n = a.shape[0]
    z = np.array(z_in, dtype=dtype, order="F", copy=True)
    m = z.shape[1]
    t = np.zeros(max(n, 1), dtype=dtype)
    t[: max(n - 1, 0)] = tau
    work = np.zeros(max(32 * (2 * n + 2 * m + 32), 1), dtype=dtype)
    f = lib().xtb_dormtr_upper if dtype == np.float64 else lib().xtb_sormtr_upper
    f(_i(n), _i(m), _ptr(a), _i(n), _ptr(t), _ptr(z), _i(n), _ptr(work))
    return z


def sytrd(a_in, uplo="U", dtype=np.float64, lwork=None):
    """LAPACK-convention xtbeig_{d,s}sytrd. Returns (a, d, e, tau, info); lwork=-1 returns the query in tau[0]."""
    a = np.array(a_in, dtype=dtype, order="F", copy=True)
    n = a.shape[0]
    d = np.zeros(max(n, 1), dtype=dtype)
    e = np.zeros(max(n, 1), dtype=dtype)
    tau = np.zeros(max(n, 1), dtype=dtype)
    lw = 32 * max(n, 1) if lwork is None else lwork
    work = np.zeros(max(lw, 1), dtype=dtype)
    info = ctypes.c_int(0)
    f = lib().xtbeig_dsytrd if dtype == np.float64 else lib().xtbeig_ssytrd
    f(uplo.encode(), _i(n), _ptr(a), _i(n), _ptr(d), _ptr(e), _ptr(tau), _ptr(work), _i(lw), ctypes.byref(info))
    if lw == -1:
        return work[0], info.value
    return a, d[:n], e[: max(n - 1, 0)], tau[: max(n - 1, 0)], info.value


def syevd(a_in, jobz="V", uplo="U", dtype=np.float64):
    a = np.array(a_in, dtype=dtype, order="F", copy=True)
    n = a.shape[0]
    w = np.zeros(n, dtype=dtype)
    lw = 1 + 6 * n + 2 * n * n
    liw = 3 + 5 * n
    work = np.zeros(max(lw, 1), dtype=dtype)
    iwork = np.zeros(max(liw, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    f = lib().xtbeig_dsyevd if dtype == np.float64 else lib().xtbeig_ssyevd
    f(jobz.encode(), uplo.encode(), _i(n), _ptr(a), _i(n), _ptr(w), _ptr(work), _i(lw), _ptr(iwork), _i(liw),
      ctypes.byref(info))
    return w, a, info.value


def stedc(d_in, e_in, dtype=np.float64):
    d = np.array(d_in, dtype=dtype, copy=True)
    n = d.shape[0]
    e = np.zeros(max(n, 1), dtype=dtype)
    e[: max(n - 1, 0)] = e_in
    z = np.zeros((n, n), dtype=dtype, order="F")
    lw = 1 + 4 * n + n * n
    liw = 3 + 5 * n
    work = np.zeros(max(lw, 1), dtype=dtype)
    iwork = np.zeros(max(liw, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    f = lib().xtb_dstedc if dtype == np.float64 else lib().xtb_sstedc
    f(_i(n), _ptr(d), _ptr(e), _ptr(z), _i(max(n, 1)), _ptr(work), _i(lw), _ptr(iwork), _i(liw), ctypes.byref(info))
    return d, z, info.value


def dcongruence_upper(h, x, z0=None):
    """xtb_dcongruence_upper: Z's upper triangle = X^T H X; Z starts as z0 (its lower triangle is left as it is)."""
    h = np.asfortranarray(h, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64)
    n = h.shape[0]
    z = np.array(z0 if z0 is not None else np.zeros((n, n)), dtype=np.float64, order="F", copy=True)
    lib().xtb_dcongruence_upper(_i(n), _ptr(h), _i(h.shape[0]), _ptr(x), _i(x.shape[0]), _ptr(z), _i(z.shape[0]))
    return z


def dtrmm_back(x, c, m):
    """xtb_dtrmm_back: C(:, :m) = X C(:, :m) on a copy of c."""
    x = np.asfortranarray(x, dtype=np.float64)
    c = np.array(c, dtype=np.float64, order="F", copy=True)
    n = x.shape[0]
    lib().xtb_dtrmm_back(_i(n), _i(m), _ptr(x), _i(x.shape[0]), _ptr(c), _i(c.shape[0]))
    return c


FNV_OFFSET = 0xcbf29ce484222325


def fnv1a(*arrays, hash=FNV_OFFSET):
    """xtbeig_fnv1a over the bytes of the arrays in turn (the eigen-bench fingerprint of a tridiagonal's d, e)."""
    for a in arrays:
        a = np.ascontiguousarray(a)
        hash = lib().xtbeig_fnv1a(a.ctypes.data_as(vp), a.nbytes, hash)
    return hash


def hsbin_read(path):
    """(h, s, e, c) through the library's C reader; e and c are None when the file holds only H and S."""
    n, sol = ctypes.c_int(0), ctypes.c_int(0)
    if lib().xtbeig_hsbin_info(str(path).encode(), ctypes.byref(n), ctypes.byref(sol)):
        raise ValueError(lib().xtbeig_hsbin_error().decode())
    n = n.value
    h = np.zeros((n, n), order="F")
    s = np.zeros((n, n), order="F")
    e = np.zeros(n) if sol.value else None
    c = np.zeros((n, n), order="F") if sol.value else None
    rc = lib().xtbeig_hsbin_read(str(path).encode(), n, _ptr(h), n, _ptr(s), n,
                                 _ptr(e) if e is not None else None, _ptr(c) if c is not None else None, n)
    if rc:
        raise ValueError(lib().xtbeig_hsbin_error().decode())
    return h, s, e, c


def hsbin_write(path, h, s, e=None, c=None):
    h = np.asfortranarray(h, dtype=np.float64)
    s = np.asfortranarray(s, dtype=np.float64)
    n = h.shape[0]
    ea = np.ascontiguousarray(e, dtype=np.float64) if e is not None else None
    cf = np.asfortranarray(c, dtype=np.float64) if c is not None else None
    if lib().xtbeig_hsbin_write(str(path).encode(), n, _ptr(h), n, _ptr(s), n, _ptr(ea) if ea is not None else None,
                                _ptr(cf) if cf is not None else None, n):
        raise ValueError(lib().xtbeig_hsbin_error().decode())


def dsymmetrize_pad(src, ldd=None, n=None):
    """xtb_dsymmetrize_pad on the leading n x n block of src (lds = src.shape[0]); dst has ldd rows."""
    src = np.asfortranarray(src, dtype=np.float64)
    n = src.shape[1] if n is None else n
    ldd = ldd or n
    dst = np.zeros((ldd, n), order="F")
    lib().xtb_dsymmetrize_pad(_i(n), _ptr(src), _i(src.shape[0]), _ptr(dst), _i(ldd))
    return dst[:n, :n]


# ---- device-resident calls ---------------------------------------------------------------------------------------

def _dev(t, dtype, numel, what):
    """Refuses a device argument the C entry point would misread: a host tensor, another element type (an FP32
    eigenvalue buffer handed to an entry point that writes FP64 ones would be overrun by half its size) or fewer
    elements than the call writes or reads."""
    import torch
    dt = {"f64": torch.float64, "f32": torch.float32}[dtype]
    if not t.is_cuda:
        raise ValueError(f"{what}: expected a device tensor")
    if t.dtype != dt:
        raise ValueError(f"{what}: expected {dt}, got {t.dtype}")
    if t.numel() < numel:
        raise ValueError(f"{what}: expected at least {numel} elements, got {t.numel()}")


def _prec(t, what):
    """The element type of a stage call, from its first tensor: "f64" or "f32"."""
    import torch
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise ValueError(f"{what}: expected torch.float64 or torch.float32, got {t.dtype}")


def _dev_square(t, dtype, n, what):
    _dev(t, dtype, n * n, what)
    if n > 1 and (t.dim() != 2 or t.stride(1) != 1 or t.stride(0) < n):
        raise ValueError(f"{what}: expected a row-major (n, ld >= n) tensor holding the column-major matrix")


class Handle:
    """A device handle; run on the given torch stream (default: the handle's own stream)."""

    def __init__(self, device=0, stream=None):
        self._h = vp()
        if lib().xtbeig_create(ctypes.byref(self._h), device) != 0:
            raise RuntimeError(last_error())
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if self._h:
            lib().xtbeig_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(last_error())

    @staticmethod
    def _args_sygvd(n, h, s, w, c):
        _dev_square(h, "f64", n, "sygvd h")
        _dev_square(s, "f64", n, "sygvd s")
        _dev(w, "f64", n, "sygvd w")
        _dev_square(c, "f64", n, "sygvd c")

    def set_stream(self, stream):
        raw = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
        self._check(lib().xtbeig_set_stream(self._h, vp(raw)))

    def stream_ptr(self):
        return lib().xtbeig_get_stream(self._h)

    def set_tuning(self, band=0, refine_iterations=2):
        self._check(lib().xtbeig_set_tuning(self._h, band, refine_iterations))

    def set_test_limits(self, secular_maxit=-1):
        self._check(lib().xtbeig_set_test_limits(self._h, secular_maxit))

    def synchronize(self):
        """Wait; the four status words [Cholesky info, chase stall, D&C info, refinement check], cleared."""
        st = (ctypes.c_int * 4)()
        rc = lib().xtbeig_synchronize(self._h, st)
        if rc == 1:
            raise RuntimeError(last_error())
        return list(st)

    def refine_check(self):
        """The last refinement check's (residual, orthogonality) ratios; a solve passes with both <= 1."""
        r = (ctypes.c_double * 2)()
        self._check(lib().xtbeig_refine_check(self._h, r))
        return list(r)

    def sygvd(self, h, s, w, c, precision=FP64):
        n = h.shape[0]
        self._args_sygvd(n, h, s, w, c)
        self._check(lib().xtbeig_dev_sygvd(self._h, precision, n, vp(h.data_ptr()), h.stride(0),
                                           vp(s.data_ptr()), s.stride(0), vp(w.data_ptr()), vp(c.data_ptr()),
                                           c.stride(0)))

    def dsyevd_partial(self, a, w, z, m):
        n = a.shape[0]
        _dev_square(a, "f64", n, "dsyevd_partial a")
        _dev(w, "f64", n, "dsyevd_partial w")
        _dev(z, "f64", n * m, "dsyevd_partial z")
        self._check(lib().xtbeig_dev_dsyevd_partial(self._h, n, m, vp(a.data_ptr()), n, vp(w.data_ptr()),
                                                    vp(z.data_ptr()), z.stride(0) if z.dim() == 2 else n))

    def set_geometry(self, s):
        n = s.shape[0]
        self._check(lib().xtbeig_dev_set_geometry(self._h, n, vp(s.data_ptr()), s.stride(0)))

    def congruence(self, h, z):
        n = h.shape[0]
        self._check(lib().xtbeig_dev_congruence(self._h, n, vp(h.data_ptr()), h.stride(0), vp(z.data_ptr()),
                                                z.stride(0)))

    def back_solve(self, c, m):
        """c: (cols, n) tensor holding the column-major n x cols matrix; the first m columns are transformed."""
        n = c.shape[1]
        self._check(lib().xtbeig_dev_back_solve(self._h, n, m, vp(c.data_ptr()), c.stride(0)))

    def sygvd_geometry(self, h, w, c, m, precision=FP64):
        n = h.shape[0]
        self._check(lib().xtbeig_dev_sygvd_geometry(self._h, precision, n, m, vp(h.data_ptr()), h.stride(0),
                                                    vp(w.data_ptr()), vp(c.data_ptr()), c.stride(0)))

    def dsyevd(self, a, w, z):
        n = a.shape[0]
        _dev_square(a, "f64", n, "dsyevd a")
        _dev(w, "f64", n, "dsyevd w")
        _dev_square(z, "f64", n, "dsyevd z")
        self._check(lib().xtbeig_dev_dsyevd(self._h, n, vp(a.data_ptr()), n, vp(w.data_ptr()), vp(z.data_ptr()), n))

    def ssyevd(self, a, w, z):
        """FP32 a and z; the eigenvalues w are FP64 (the divide and conquer runs in FP64)."""
        n = a.shape[0]
        _dev_square(a, "f32", n, "ssyevd a")
        _dev(w, "f64", n, "ssyevd w")
        _dev_square(z, "f32", n, "ssyevd z")
        self._check(lib().xtbeig_dev_ssyevd(self._h, n, vp(a.data_ptr()), n, vp(w.data_ptr()), vp(z.data_ptr()), n))

    def dstedc(self, d, e, z):
        n = d.shape[0]
        _dev(d, "f64", n, "dstedc d")
        _dev(e, "f64", n - 1, "dstedc e")
        _dev_square(z, "f64", n, "dstedc z")
        self._check(lib().xtbeig_dev_dstedc(self._h, n, vp(d.data_ptr()), vp(e.data_ptr()), vp(z.data_ptr()), n))

    def gemm(self, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        import torch
        dt = _prec(c, "gemm c")
        ra, ca = (m, k) if ta == "N" else (k, m)
        rb, cb = (k, n) if tb == "N" else (n, k)
        _dev(a, dt, lda * (ca - 1) + ra if ca > 0 else 0, "gemm a")
        _dev(b, dt, ldb * (cb - 1) + rb if cb > 0 else 0, "gemm b")
        _dev(c, dt, ldc * (n - 1) + m if n > 0 else 0, "gemm c")
        f = lib().xtbeig_dev_dgemm if c.dtype == torch.float64 else lib().xtbeig_dev_sgemm
        self._check(f(self._h, ta.encode(), tb.encode(), m, n, k, alpha, vp(a.data_ptr()), lda, vp(b.data_ptr()),
                      ldb, beta, vp(c.data_ptr()), ldc))

    def potrf_inverse(self, s, u, x):
        n = s.shape[0]
        for t, nm in ((s, "s"), (u, "u"), (x, "x")):
            _dev_square(t, "f64", n, "potrf_inverse " + nm)
        self._check(lib().xtbeig_dev_potrf_inverse(self._h, n, vp(s.data_ptr()), n, vp(u.data_ptr()), n,
                                                   vp(x.data_ptr()), n))

    def reduce_to_band(self, a, b, v1, t1, ab):
        import torch
        n = a.shape[0]
        dt = _prec(a, "reduce_to_band a")
        _dev_square(a, dt, n, "reduce_to_band a")
        _dev(v1, dt, n * n, "reduce_to_band v1")
        _dev(t1, dt, n * b, "reduce_to_band t1")
        _dev(ab, dt, band_elems(n, b), "reduce_to_band ab")
        f = lib().xtbeig_dev_dreduce_to_band if a.dtype == torch.float64 else lib().xtbeig_dev_sreduce_to_band
        self._check(f(self._h, n, b, vp(a.data_ptr()), n, vp(v1.data_ptr()), vp(t1.data_ptr()), vp(ab.data_ptr())))

    def chase(self, n, b, ab, v2, d, e):
        import torch
        dt = _prec(ab, "chase ab")
        _dev(ab, dt, band_elems(n, b), "chase ab")
        _dev(v2, dt, n * n, "chase v2")
        _dev(d, dt, n, "chase d")
        _dev(e, dt, max(n - 1, 0), "chase e")
        f = lib().xtbeig_dev_dchase if ab.dtype == torch.float64 else lib().xtbeig_dev_schase
        self._check(f(self._h, n, b, vp(ab.data_ptr()), vp(v2.data_ptr()), vp(d.data_ptr()), vp(e.data_ptr())))

    def back_bulges(self, n, b, v2, z, m):
        import torch
        dt = _prec(z, "back_bulges z")
        _dev(v2, dt, n * n, "back_bulges v2")
        _dev(z, dt, n * m, "back_bulges z")
        f = lib().xtbeig_dev_dback_bulges if z.dtype == torch.float64 else lib().xtbeig_dev_sback_bulges
        self._check(f(self._h, n, b, vp(v2.data_ptr()), vp(z.data_ptr()), n, m))

    def back_band(self, n, b, v1, t1, z, m):
        import torch
        dt = _prec(z, "back_band z")
        _dev(v1, dt, n * n, "back_band v1")
        _dev(t1, dt, n * b, "back_band t1")
        _dev(z, dt, n * m, "back_band z")
        f = lib().xtbeig_dev_dback_band if z.dtype == torch.float64 else lib().xtbeig_dev_sback_band
        self._check(f(self._h, n, b, vp(v1.data_ptr()), vp(t1.data_ptr()), vp(z.data_ptr()), n, m))

    # ---- graphs, batches, autotuning ----
    def sygvd_graph(self, h, s, w, c, precision=FP64):
        n = h.shape[0]
        self._args_sygvd(n, h, s, w, c)
        self._check(lib().xtbeig_dev_sygvd_graph(self._h, precision, n, vp(h.data_ptr()), h.stride(0),
                                                 vp(s.data_ptr()), s.stride(0), vp(w.data_ptr()), vp(c.data_ptr()),
                                                 c.stride(0)))

    def sygvd_batched(self, hs, ss, ws, cs, precision=FP64, slots=8, use_graph=False):
        """Lists of device tensors (each h[q] n x n with ld = stride(0)); returns the per-solve latencies (ms)."""
        if not (len(hs) == len(ss) == len(ws) == len(cs)):
            raise ValueError("sygvd_batched: hs, ss, ws, cs must have one entry per problem")
        for h, s_, w, c in zip(hs, ss, ws, cs):
            self._args_sygvd(h.shape[0], h, s_, w, c)
        cnt = len(hs)
        arr_i = ctypes.c_int * max(cnt, 1)
        arr_p = vp * max(cnt, 1)
        ns = arr_i(*[t.shape[0] for t in hs])
        ldh = arr_i(*[t.stride(0) for t in hs])
        lds = arr_i(*[t.stride(0) for t in ss])
        ldc = arr_i(*[t.stride(0) for t in cs])
        ph = arr_p(*[t.data_ptr() for t in hs])
        ps = arr_p(*[t.data_ptr() for t in ss])
        pw = arr_p(*[t.data_ptr() for t in ws])
        pc = arr_p(*[t.data_ptr() for t in cs])
        lat = (ctypes.c_float * max(cnt, 1))()
        self._check(lib().xtbeig_dev_sygvd_batched(self._h, precision, cnt, ns, ph, ldh, ps, lds, pw, pc, ldc,
                                                   slots, 1 if use_graph else 0, lat))
        return [lat[q] for q in range(cnt)]

    def sygvd_grouped(self, hs, ss, ws, cs, precision=FP64, granule=0, slots=0):
        """Lists of device tensors as sygvd_batched; every kernel runs once per group of equal padded size. Returns
        (per-problem status words [count][4], per-problem latencies in ms)."""
        if not (len(hs) == len(ss) == len(ws) == len(cs)):
            raise ValueError("sygvd_grouped: hs, ss, ws, cs must have one entry per problem")
        for h, s_, w, c in zip(hs, ss, ws, cs):
            self._args_sygvd(h.shape[0], h, s_, w, c)
        cnt = len(hs)
        arr_i = ctypes.c_int * max(cnt, 1)
        arr_p = vp * max(cnt, 1)
        ns = arr_i(*[t.shape[0] for t in hs])
        ldh = arr_i(*[t.stride(0) for t in hs])
        lds = arr_i(*[t.stride(0) for t in ss])
        ldc = arr_i(*[t.stride(0) for t in cs])
        ph = arr_p(*[t.data_ptr() for t in hs])
        ps = arr_p(*[t.data_ptr() for t in ss])
        pw = arr_p(*[t.data_ptr() for t in ws])
        pc = arr_p(*[t.data_ptr() for t in cs])
        st = (ctypes.c_int * (4 * max(cnt, 1)))()
        lat = (ctypes.c_float * max(cnt, 1))()
        self._check(lib().xtbeig_dev_sygvd_grouped(self._h, precision, cnt, ns, ph, ldh, ps, lds, pw, pc, ldc,
                                                   granule, slots, st, lat))
        return [[st[4 * q + k] for k in range(4)] for q in range(cnt)], [lat[q] for q in range(cnt)]

    def autotune(self, precision, nmin, nmax):
        """Returns [(bucket_start_n, band, ms)] for the buckets tuned."""
        nb = lib().xtbeig_band_buckets()
        st = (ctypes.c_int * nb)()
        bd = (ctypes.c_int * nb)()
        ms = (ctypes.c_float * nb)()
        cnt = ctypes.c_int(0)
        self._check(lib().xtbeig_autotune(self._h, precision, nmin, nmax, st, bd, ms, ctypes.byref(cnt)))
        return [(st[q], bd[q], ms[q]) for q in range(cnt.value)]

    def autotune_gemm(self, precision, nmin, nmax):
        """Times the GEMM tile widths per bucket of tile counts reached by n in [nmin, nmax] and stores the faster;
        returns [(first 64 x 64 tile count of the bucket, width)]."""
        nb = lib().xtbeig_gemm_tile_buckets()
        st = (ctypes.c_int * nb)()
        wd = (ctypes.c_int * nb)()
        cnt = ctypes.c_int(0)
        self._check(lib().xtbeig_autotune_gemm(self._h, precision, nmin, nmax, st, wd, ctypes.byref(cnt)))
        return [(st[q], wd[q]) for q in range(cnt.value)]

    def band_table(self, precision):
        nb = lib().xtbeig_band_buckets()
        st = (ctypes.c_int * nb)()
        bd = (ctypes.c_int * nb)()
        self._check(lib().xtbeig_get_band_table(self._h, precision, st, bd))
        return list(st), list(bd)

    def set_band_table(self, precision, bands):
        nb = lib().xtbeig_band_buckets()
        bd = (ctypes.c_int * nb)(*bands)
        self._check(lib().xtbeig_set_band_table(self._h, precision, bd))

    # ---- SCC precision schedule ----
    def set_scc_precision(self, fp32_rmsq, fp32_maxiter=160, fp32_refine=0):
        """Returns 0, or raises ValueError with the library's message (settings unchanged)."""
        rc = lib().xtbeig_set_scc_precision(self._h, float(fp32_rmsq), int(fp32_maxiter), int(fp32_refine))
        if rc:
            raise ValueError(last_error())

    def scc_begin(self):
        self._check(lib().xtbeig_scc_begin(self._h))

    def scc_sygvd(self, h, s, w, c, rmsq):
        """One SCC cycle's solve; returns the precision used (FP64 or FP32_REFINE)."""
        used = ctypes.c_int(-1)
        n = h.shape[0]
        self._check(lib().xtbeig_dev_scc_sygvd(self._h, n, vp(h.data_ptr()), h.stride(0), vp(s.data_ptr()),
                                               s.stride(0), vp(w.data_ptr()), vp(c.data_ptr()), c.stride(0),
                                               float(rmsq), ctypes.byref(used)))
        return used.value

    def get_scc_precision(self):
        """(settings, counters, rc): settings (fp32_rmsq, fp32_maxiter, fp32_refine), counters dict, rc != 0 when
        no cycle has run."""
        r = ctypes.c_double()
        i = [ctypes.c_int() for _ in range(5)]
        rc = lib().xtbeig_get_scc_precision(self._h, ctypes.byref(r), *[ctypes.byref(v) for v in i])
        return ((r.value, i[0].value, i[1].value),
                dict(fp32_cycles=i[2].value, fp64_restart=i[3].value, scc_cycles=i[4].value), rc)

    # ---- density matrix and its backward (batches of equal-n problems; tensors (batch, n, n) hold C^T etc.) ----
    # Matrices are passed as torch tensors whose last two dimensions are a column-major n x n matrix stored
    # row-major-transposed (tensor[p].T is the matrix), i.e. tensor[p] is the matrix's transpose in C order, the
    # layout the rest of this module uses; w is (batch, n), mu (batch,), occ (batch, 3n).
    def density(self, w, c, kT, nel, mu, bisect=True, p=None, occ=None):
        import torch
        batch, n = w.shape
        if p is None:
            p = torch.empty((batch, n, n), dtype=torch.float64, device=w.device)
        if occ is None:
            occ = torch.empty((batch, 3 * n), dtype=torch.float64, device=w.device)
        self._check(lib().xtbeig_dev_density(self._h, batch, n, vp(w.data_ptr()), n, vp(c.data_ptr()), n, n * n,
                                             float(kT), float(nel), vp(mu.data_ptr()), 1 if bisect else 0,
                                             vp(p.data_ptr()), n, n * n, vp(occ.data_ptr())))
        return p, occ

    def density_backward(self, w, c, kT, mu, occ, gbar, mu_fixed=False):
        import torch
        batch, n = w.shape
        hbar = torch.empty((batch, n, n), dtype=torch.float64, device=w.device)
        sbar = torch.empty_like(hbar)
        self._check(lib().xtbeig_dev_density_backward(self._h, batch, n, vp(w.data_ptr()), n, vp(c.data_ptr()), n,
                                                      n * n, float(kT), vp(mu.data_ptr()), 1 if mu_fixed else 0,
                                                      vp(occ.data_ptr()), vp(gbar.data_ptr()), n, n * n,
                                                      vp(hbar.data_ptr()), n, n * n, vp(sbar.data_ptr()), n, n * n))
        return hbar, sbar

    def density_spin(self, w, c, kT, nel_alpha, nel_beta, mu, bisect=True):
        """Open-shell density: mu (batch, 2) alpha / beta chemical potentials; returns (P, occ (batch, 6 n))."""
        import torch
        batch, n = w.shape
        p = torch.empty((batch, n, n), dtype=torch.float64, device=w.device)
        occ = torch.empty((batch, 6 * n), dtype=torch.float64, device=w.device)
        self._check(lib().xtbeig_dev_density_spin(self._h, batch, n, vp(w.data_ptr()), n, vp(c.data_ptr()), n, n * n,
                                                  float(kT), float(nel_alpha), float(nel_beta), vp(mu.data_ptr()),
                                                  1 if bisect else 0, vp(p.data_ptr()), n, n * n, vp(occ.data_ptr())))
        return p, occ

    def density_backward_spin(self, w, c, kT, mu, occ, gbar, mu_fixed=False):
        import torch
        batch, n = w.shape
        hbar = torch.empty((batch, n, n), dtype=torch.float64, device=w.device)
        sbar = torch.empty_like(hbar)
        self._check(lib().xtbeig_dev_density_backward_spin(
            self._h, batch, n, vp(w.data_ptr()), n, vp(c.data_ptr()), n, n * n, float(kT), vp(mu.data_ptr()),
            1 if mu_fixed else 0, vp(occ.data_ptr()), vp(gbar.data_ptr()), n, n * n, vp(hbar.data_ptr()), n, n * n,
            vp(sbar.data_ptr()), n, n * n))
        return hbar, sbar

    def density_backward_recompute(self, h, s, kT, nel, gbar, precision=0, mu_fixed=False):
        import torch
        batch, n, _ = h.shape
        hbar = torch.empty((batch, n, n), dtype=torch.float64, device=h.device)
        sbar = torch.empty_like(hbar)
        self._check(lib().xtbeig_dev_density_backward_recompute(
            self._h, batch, n, vp(h.data_ptr()), n, n * n, vp(s.data_ptr()), n, n * n, precision, float(kT),
            float(nel), 1 if mu_fixed else 0, vp(gbar.data_ptr()), n, n * n, vp(hbar.data_ptr()), n, n * n,
            vp(sbar.data_ptr()), n, n * n))
        return hbar, sbar


def dc_merge(handle, d, q, n1, rho):
    """One device merge on torch tensors d (n), q (n x n column-major as (n, n) row-major of q^T). Returns dlaed2's
    outputs and the merged (d, q) as numpy arrays."""
    import torch
    n = d.shape[0]
    dd = torch.empty(n, dtype=torch.float64, device=d.device)
    q2 = torch.zeros(n * n, dtype=torch.float64, device=d.device)
    kc = torch.zeros(5, dtype=torch.int32, device=d.device)
    dl = torch.zeros(n, dtype=torch.float64, device=d.device)
    w = torch.zeros(n, dtype=torch.float64, device=d.device)
    ix = torch.zeros(n, dtype=torch.int32, device=d.device)
    ro = torch.zeros(1, dtype=torch.float64, device=d.device)
    handle._check(lib().xtbeig_dev_dc_merge(handle._h, n, n1, float(rho), vp(d.data_ptr()), vp(q.data_ptr()), n,
                                            vp(dd.data_ptr()), vp(q2.data_ptr()), vp(kc.data_ptr()),
                                            vp(dl.data_ptr()), vp(w.data_ptr()), vp(ix.data_ptr()),
                                            vp(ro.data_ptr())))
    st = handle.synchronize()
    k = int(kc[0])
    return dict(k=k, ctot=kc[1:].cpu().numpy(), d_deflated=dd.cpu().numpy(), q2=q2.cpu().numpy(),
                dlamda=dl.cpu().numpy()[:k], w=w.cpu().numpy()[:k], indxc=ix.cpu().numpy(), rho=float(ro[0]),
                d=d.cpu().numpy(), q=q.cpu().numpy().T, status=st)


def default_band(n, fp32=False):
    """The library's built-in band width for an n x n solve."""
    return int(lib().xtbeig_default_band(int(n), 1 if fp32 else 0))


def band_elems(n, b):
    return int(lib().xtbeig_band_elems(n, b))


def band_to_dense(ab, n, b):
    """Band storage (element (i, j), 0 <= i - j < 2b, at i + (2b - 1) j) -> dense symmetric numpy matrix."""
    ld = 2 * b - 1 if b > 1 else 1
    a = np.zeros((n, n))
    flat = np.asarray(ab).ravel()
    for j in range(n):
        for dd in range(min(ld + 1, n - j)):
            i = j + dd
            v = flat[i + ld * j]
            a[i, j] = v
            a[j, i] = v
    return a
