"""ec_filter_coarse_noalloc (the filter entry point that allocates nothing,
for callers under graph capture or that forbid hipMalloc) against
ec_filter_coarse, through the C ABI: two calls in a row on the shared
scratch, the second on a smaller grid, bit for bit where both take the
separable passes and to 1e-13 where ec_filter_coarse takes its dense launch;
sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _quantities(abi, n, nbar, outs, keep, wide):
    c = n // nbar
    g = np.exp(-0.5 * (np.arange(-4, 5) / 1.7) ** 2)
    g /= g.sum()
    box = np.full(c, 1.0 / c)
    wq = np.convolve(box, g) if wide else box
    srcs = [("u1", 0), ("s12", 6), ("r22", 10), ("p", 12), ("k3", 15)]
    qs = (abi.FilterQuantity * len(srcs))()
    for m, (_, code) in enumerate(srcs):
        qs[m].source = code
        for d in range(3):
            w = np.ascontiguousarray(wq if d != 2 or m % 2 == 0 else g)
            keep.append(w)
            qs[m].len[d] = len(w)
            qs[m].off[d] = -(len(w) // 2) - (d + m) % 3
            qs[m].w[d] = w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        qs[m].out = outs[m].data_ptr()
    return qs


def _run(gpu, entry, n, nbar, wide, dtype, seed):
    abi = gpu._abi
    gen = torch.Generator(device="cuda").manual_seed(seed)
    u = [torch.randn((n, n, n), device="cuda", dtype=dtype, generator=gen) for _ in range(3)]
    p = torch.randn((n, n, n), device="cuda", dtype=dtype, generator=gen)
    outs = [torch.full((nbar,) * 3, float("nan"), dtype=torch.float64, device="cuda") for _ in range(5)]
    keep = []
    qs = _quantities(abi, n, nbar, outs, keep, wide)
    code = abi.EC_F64 if dtype == torch.float64 else abi.EC_F32
    abi.context().call(entry, code, n, nbar, 1.0, 0.01, abi.P3(*[x.data_ptr() for x in u]),
                       abi.c_void_p(p.data_ptr()), 5, qs)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_filter_matches_allocating_entry(gpu, dtype):
    for n, nbar, wide, seed in ((96, 16, True, 1), (45, 9, True, 2), (60, 12, False, 3), (24, 8, False, 4)):
        want = _run(gpu, "ec_filter_coarse", n, nbar, wide, dtype, seed)
        got = _run(gpu, "ec_filter_coarse_noalloc", n, nbar, wide, dtype, seed)
        for g, w in zip(got, want):
            assert np.isfinite(g).all()
            if wide:
                assert np.array_equal(g, w), (n, nbar)
            else:
                assert np.abs(g - w).max() <= 1e-13 * np.abs(w).max(), (n, nbar)


def test_noalloc_filter_refuses_sizes_beyond_scratch(gpu):
    abi = gpu._abi
    n, nbar = 256, 64  # 64 * 256^2 + 64^2 * 256 doubles > 4 Mi
    u = [torch.zeros((n, n, n), device="cuda") for _ in range(3)]
    out = torch.zeros((nbar,) * 3, dtype=torch.float64, device="cuda")
    keep = []
    qs = _quantities(abi, n, nbar, [out] * 5, keep, False)
    with pytest.raises(abi.ECError, match="scratch"):
        abi.context().call("ec_filter_coarse_noalloc", abi.EC_F32, n, nbar, 1.0, 0.01,
                           abi.P3(*[x.data_ptr() for x in u]), abi.c_void_p(u[0].data_ptr()), 5, qs)
