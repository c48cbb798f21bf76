"""The C-ABI library loads and exports every function include/xtbeig.h declares (no GPU needed)."""
import os
import re

import xtbeig

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "xtbeig.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(xtb[a-z0-9_]*)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_reference_symbol_set():
    names = declared()
    for s in ["xtb_dsytrd_upper", "xtb_ssytrd_upper", "xtb_dsytrd_full", "xtb_dormtr_upper", "xtb_sormtr_upper",
              "xtb_dstedc", "xtb_sstedc", "xtb_dsymmetrize_pad", "xtbeig_dsytrd", "xtbeig_ssytrd", "xtbeig_dsyevd",
              "xtbeig_ssyevd", "xtbeig_dsygvd", "xtbeig_ssygvd"]:
        assert s in names


def test_library_exports_every_declared_symbol():
    L = xtbeig.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert set(declared()) <= set(xtbeig.SIGNATURES)


def test_workspace_queries_without_gpu():
    """lwork = -1 queries are answered on the host (no device touched)."""
    import ctypes
    import numpy as np
    n = 100
    work = np.zeros(1)
    iwork = np.zeros(1, dtype=np.int32)
    info = ctypes.c_int(7)
    xtbeig.lib().xtbeig_dsygvd(xtbeig._i(1), b"V", b"U", xtbeig._i(n), None, xtbeig._i(n), None, xtbeig._i(n), None,
                               xtbeig._ptr(work), xtbeig._i(-1), xtbeig._ptr(iwork), xtbeig._i(-1), ctypes.byref(info))
    assert info.value == 0 and work[0] == 1 + 6 * n + 2 * n * n and iwork[0] == 3 + 5 * n
    xtbeig.lib().xtb_dstedc(xtbeig._i(n), None, None, None, xtbeig._i(n), xtbeig._ptr(work), xtbeig._i(-1),
                            xtbeig._ptr(iwork), xtbeig._i(-1), ctypes.byref(info))
    assert info.value == 0 and work[0] == 1 + 4 * n + n * n and iwork[0] == 3 + 5 * n
    # illegal argument: itype 2 is rejected with info = -1
    xtbeig.lib().xtbeig_dsygvd(xtbeig._i(2), b"V", b"U", xtbeig._i(n), None, xtbeig._i(n), None, xtbeig._i(n), None,
                               xtbeig._ptr(work), xtbeig._i(-1), xtbeig._ptr(iwork), xtbeig._i(-1), ctypes.byref(info))
    assert info.value == -1


def test_sytrd_query_and_argument_checks_without_gpu():
    """xtbeig_dsytrd: lwork = -1 answers n * 32; uplo 'L' and a short lda are rejected (no device touched)."""
    import numpy as np
    n = 50
    q, info = xtbeig.sytrd(np.zeros((n, n)), lwork=-1)
    assert info == 0 and q == 32 * n
    import ctypes
    work = np.zeros(1)
    info = ctypes.c_int(0)
    xtbeig.lib().xtbeig_dsytrd(b"L", xtbeig._i(n), None, xtbeig._i(n), None, None, None, xtbeig._ptr(work),
                               xtbeig._i(-1), ctypes.byref(info))
    assert info.value == -1
    xtbeig.lib().xtbeig_dsytrd(b"U", xtbeig._i(n), None, xtbeig._i(n - 1), None, None, None, xtbeig._ptr(work),
                               xtbeig._i(-1), ctypes.byref(info))
    assert info.value == -4


class _FakeDeviceTensor:
    """Stands in for a device tensor (the checks read only these attributes), so that the host-side argument checks
    of xtbeig.Handle run without a GPU."""

    def __init__(self, dtype, shape):
        self.is_cuda = True
        self.dtype = dtype
        self.shape = shape
        self._numel = 1
        for s in shape:
            self._numel *= s

    def numel(self):
        return self._numel

    def dim(self):
        return len(self.shape)

    def stride(self, k):
        return self.shape[1] if (k == 0 and len(self.shape) == 2) else 1


def test_device_call_argument_checks():
    # the checks xtbeig.Handle runs before any launch: device memory, the entry point's element type, enough elements
    import pytest
    import torch

    n = 8
    with pytest.raises(ValueError, match="device tensor"):
        xtbeig._dev(torch.zeros(n, dtype=torch.float64), "f64", n, "w")
    with pytest.raises(ValueError, match="float64"):
        xtbeig._dev(_FakeDeviceTensor(torch.float32, (n,)), "f64", n, "ssyevd w")
    with pytest.raises(ValueError, match="at least 8"):
        xtbeig._dev(_FakeDeviceTensor(torch.float64, (n - 1,)), "f64", n, "w")
    xtbeig._dev(_FakeDeviceTensor(torch.float64, (n,)), "f64", n, "w")
    with pytest.raises(ValueError, match="row-major"):
        xtbeig._dev_square(_FakeDeviceTensor(torch.float64, (n * n,)), "f64", n, "a")
    xtbeig._dev_square(_FakeDeviceTensor(torch.float64, (n, n)), "f64", n, "a")
    with pytest.raises(ValueError, match="sygvd w"):
        xtbeig.Handle._args_sygvd(n, _FakeDeviceTensor(torch.float64, (n, n)), _FakeDeviceTensor(torch.float64, (n, n)),
                                  _FakeDeviceTensor(torch.float32, (n,)), _FakeDeviceTensor(torch.float64, (n, n)))
