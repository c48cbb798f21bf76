"""GEMM tile width controls: validation and the untuned round model on the host (no GPU); on the GPU, results
bit-identical across tile widths and after the tile autotuner, against an FP64 product to rounding."""
import numpy as np
import pytest

import xtbeig


def test_width_controls_validate_without_gpu():
    with pytest.raises(ValueError):
        xtbeig.set_gemm_tile(96)
    nb = xtbeig.lib().xtbeig_gemm_tile_buckets()
    with pytest.raises(ValueError):
        xtbeig.set_gemm_tile_table(0, [32] * nb)
    assert xtbeig.gemm_tile_width(0, 5) == -1
    starts = [q[0] for q in xtbeig.gemm_tile_table(0)]
    assert starts[0] == 1 and all(b > a for a, b in zip(starts, starts[1:]))
    # half-octave buckets of 64 x 64 tile counts
    assert starts[6] == 8 and starts[8] == 16 and starts[12] == 64


def test_forced_width_and_table_take_precedence():
    nb = xtbeig.lib().xtbeig_gemm_tile_buckets()
    saved = [w for _, w in xtbeig.gemm_tile_table(0)]
    try:
        xtbeig.set_gemm_tile(128)
        assert xtbeig.gemm_tile_width(1204, 1204, 1, 0) == 128
        xtbeig.set_gemm_tile(0)
        xtbeig.set_gemm_tile_table(0, [128] * nb)
        assert xtbeig.gemm_tile_width(300, 300, 1, 0) == 128
        xtbeig.set_gemm_tile_table(0, [64] * nb)
        assert xtbeig.gemm_tile_width(300, 300, 1, 0) == 64
    finally:
        xtbeig.set_gemm_tile(0)
        xtbeig.set_gemm_tile_table(0, saved)


def _gemm(hnd, torch, ta, tb, m, n, k, dt, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    a = torch.randn(m * k, dtype=dt, device="cuda", generator=g)
    b = torch.randn(k * n, dtype=dt, device="cuda", generator=g)
    c = torch.zeros(m * n, dtype=dt, device="cuda")
    hnd.gemm(ta, tb, m, n, k, 1.0, a, k if ta == "T" else m, b, n if tb == "T" else k, 0.0, c, m)
    hnd.synchronize()
    ad = a.double().view(k, m).t() if ta == "N" else a.double().view(m, k)
    bd = b.double().view(n, k).t() if tb == "N" else b.double().view(k, n)
    return c, ad @ bd


@pytest.mark.gpu
@pytest.mark.parametrize("dt_name", ["float64", "float32"])
@pytest.mark.parametrize("ta,tb,m,n,k", [("N", "N", 1204, 1204, 1204), ("T", "N", 600, 1788, 300),
                                         ("N", "T", 37, 1300, 129), ("T", "T", 1000, 70, 1204)])
def test_widths_give_identical_products(handle, torch_cuda, dt_name, ta, tb, m, n, k):
    torch = torch_cuda
    dt = getattr(torch, dt_name)
    out = []
    try:
        for w in (64, 128):
            xtbeig.set_gemm_tile(w)
            c, ref = _gemm(handle, torch, ta, tb, m, n, k, dt, 11)
            out.append(c)
    finally:
        xtbeig.set_gemm_tile(0)
    assert torch.equal(out[0], out[1])
    got = out[0].double().view(n, m).t()
    tol = (1e-13 if dt == torch.float64 else 2e-5) * k ** 0.5
    assert float((got - ref).abs().max() / ref.abs().max()) < tol


@pytest.mark.gpu
def test_autotune_gemm_keeps_solves_bit_identical(torch_cuda):
    import workloads
    torch = torch_cuda
    n = 600
    h, s = workloads.generalized_pair(n, workloads.SEED_BASE + 3)
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    hnd = xtbeig.Handle(0)
    saved = {p: [w for _, w in xtbeig.gemm_tile_table(p)] for p in (xtbeig.FP64, xtbeig.FP32_REFINE)}
    try:
        res = []
        for step in range(2):
            if step == 1:
                for prec in (xtbeig.FP64, xtbeig.FP32_REFINE):
                    tuned = hnd.autotune_gemm(prec, 124, 1204)
                    assert tuned and all(w in (64, 128) for _, w in tuned)
                    table = dict(xtbeig.gemm_tile_table(prec))
                    assert all(table[t] == w for t, w in tuned)
            for prec in (xtbeig.FP64, xtbeig.FP32_REFINE):
                w = torch.zeros(n, dtype=torch.float64, device="cuda")
                c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
                hnd.sygvd(hd, sd, w, c, prec)
                assert hnd.synchronize() == [0, 0, 0, 0]
                res.append((w, c))
        for (w0, c0), (w1, c1) in zip(res[:2], res[2:]):
            assert torch.equal(w0, w1) and torch.equal(c0, c1)
    finally:
        for p, ws in saved.items():
            xtbeig.set_gemm_tile_table(p, ws)
        hnd.close()
