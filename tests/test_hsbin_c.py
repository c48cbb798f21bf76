"""The library's C-side HS.bin reader / writer and FNV-1a hash (host code: runs without a GPU).

Pins: the golden fixtures through both readers (the C one and hsbin.py's numpy one) give the same arrays; the
published FNV-1a 64-bit test vectors; a file of the wrong size is refused with a message."""
import os

import numpy as np
import pytest

import hsbin
import xtbeig

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("n", [5, 17, 64, 131])
def test_c_reader_matches_python_reader(n):
    path = os.path.join(GOLDEN, f"hs_{n}.bin")
    h, s, e, c = xtbeig.hsbin_read(path)
    h2, s2, e2, c2 = hsbin.read(path)
    assert np.array_equal(h, h2) and np.array_equal(s, s2)
    assert np.array_equal(e, e2) and np.array_equal(c, c2)


def test_c_writer_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(1)
    h, s = rng.standard_normal((9, 9)), rng.standard_normal((9, 9))
    p = tmp_path / "a.bin"
    xtbeig.hsbin_write(p, h, s)
    h2, s2, e2, c2 = hsbin.read(p)
    assert np.array_equal(h, h2) and np.array_equal(s, s2) and e2 is None
    e, c = rng.standard_normal(9), rng.standard_normal((9, 9))
    xtbeig.hsbin_write(p, h, s, e, c)
    h3, s3, e3, c3 = xtbeig.hsbin_read(p)
    assert np.array_equal(h3, h) and np.array_equal(e3, e) and np.array_equal(c3, c)
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-1])
    with pytest.raises(ValueError, match="bytes, expected"):
        xtbeig.hsbin_read(p)
    with pytest.raises(ValueError, match="cannot open"):
        xtbeig.hsbin_read(tmp_path / "missing.bin")


def test_fnv1a_published_vectors():
    L = xtbeig.lib()
    for text, want in ((b"", 0xcbf29ce484222325), (b"a", 0xaf63dc4c8601ec8c), (b"foobar", 0x85944171f73967e8)):
        buf = np.frombuffer(text, dtype=np.uint8) if text else np.zeros(0, dtype=np.uint8)
        assert xtbeig.fnv1a(buf) == want, text
    d, e = np.arange(5, dtype=np.float64), np.arange(4, dtype=np.float32)
    assert xtbeig.fnv1a(d, e) == xtbeig.fnv1a(np.frombuffer(d.tobytes() + e.tobytes(), dtype=np.uint8))
