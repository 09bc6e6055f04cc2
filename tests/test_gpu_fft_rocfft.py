"""The hand-written r2c/c2r FFT of libexactclosure against rocFFT (torch.fft on
ROCm) on the device, the FFT the reference runs through CUFFT/FFTW
(plan_rfft / ldiv!, src/N.jl:217,267,301): forward and normalised inverse at
radix-2/3/5 sizes up to 1024^3, FP64 and FP32, with both timed."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("n,dtype", [(60, torch.float64), (64, torch.float64), (96, torch.float64),
                                     (250, torch.float64), (256, torch.float64), (512, torch.float64),
                                     (64, torch.float32), (120, torch.float32), (512, torch.float32),
                                     (1024, torch.float32)])
def test_fft_matches_rocfft(gpu, n, dtype):
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, dtype)
    gen = torch.Generator(device="cuda").manual_seed(n)
    x = torch.randn((n, n, n), dtype=dtype, device="cuda", generator=gen)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    X = NS.rfft(x, ps)                       # [kz, ky, kx], kx < n/2 + 1
    R = torch.fft.rfftn(x)                   # rocFFT, same index order and sign
    assert X.shape == R.shape
    layout = (R.is_contiguous(), R.stride())
    err_f = _rel(X, R)
    err_own = _rel(NS.irfft(X, ps), x)
    Rc = R.contiguous()                      # the library takes packed spectra
    if not R.is_contiguous():
        with pytest.raises(ValueError):
            NS.irfft(R, ps)
    back = NS.irfft(Rc, ps)
    err_b = _rel(back, x)
    ref_b = torch.fft.irfftn(R, s=(n, n, n))
    err_rb = _rel(back, ref_b)
    del R, Rc, ref_b
    ms = {}
    if n >= 512:
        for name, fn in (("ec", lambda: NS.rfft(x, ps)), ("rocfft", lambda: torch.fft.rfftn(x))):
            fn()
            torch.cuda.synchronize()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(3):
                fn()
            t1.record()
            t1.synchronize()
            ms[name] = t0.elapsed_time(t1) / 3
    print(f"\n[fft n={n} {dtype}] forward vs rocFFT {err_f:.2e}, own round trip {err_own:.2e}, inverse of "
          f"rocFFT's spectrum {err_b:.2e} (vs rocFFT's inverse {err_rb:.2e}); rocFFT output contiguous, strides "
          f"{layout}; ms {ms}")
    assert err_f < tol and err_own < tol and err_b < tol and err_rb < tol
