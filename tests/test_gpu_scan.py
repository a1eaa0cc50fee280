"""The FFT instrument scan on the GPU (clblm_fft_scan) against the oracle's scan,
itself pinned against the reference's compiled clblm_FFT.f90: identical output grid,
FP64 values within 1e-10 of the largest value, and the FP32 transform's error
against FP64 within 1e-5 of the largest value."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _spectrum(n=20001, v1=600.0, dv=0.01, seed=3):
    rng = np.random.default_rng(seed)
    v = v1 + dv * np.arange(n)
    return 1.0 + 0.5 * np.sin(v / 3.0) + 0.3 * np.exp(-((v - 700.0) / 0.05) ** 2) + 0.1 * rng.normal(size=n)


def _run(ctx, sp, v1, dv, v1o, v2o, dvo, fid, hw, fill, prec):
    import torch
    x = torch.tensor(sp, dtype=torch.float64, device="cuda")
    y, a, b = ctx.fft_scan(x, v1, dv, v1o, v2o, dvo, fid, hw, fill, prec)
    torch.cuda.synchronize()
    return y.cpu().numpy(), a, b


@pytest.mark.parametrize("fid", range(2, 9))
@pytest.mark.parametrize("v1o,v2o,dvo,hw,fill", [(620.0, 780.0, 0.05, 0.6034, 0.0), (650.0, 700.0, 0.0, 0.25, 1.0)])
def test_scan_fp64_matches_oracle(ctx, fid, v1o, v2o, dvo, hw, fill):
    from clblm_amd._lib import FP64
    from oracle.py import fft_scan
    sp = _spectrum()
    ref, va, da = fft_scan(sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, fill)
    y, vb, db = _run(ctx, sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, fill, FP64)
    assert y.size == ref.size and va == vb and da == db
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("fid %d: %d points, max err / max %.2e" % (fid, y.size, err))
    assert err <= 1e-10


def test_scan_regrids_coarse_input(ctx):
    from clblm_amd._lib import FP64
    from oracle.py import fft_scan
    v = 600.0 + 0.2 * np.arange(1001)
    sp = 1.0 + 0.5 * np.sin(v / 3.0)
    ref, va, _ = fft_scan(sp, 600.0, 0.2, 650.0, 750.0, 0.1, 5, 0.6034, 0.0)
    y, vb, _ = _run(ctx, sp, 600.0, 0.2, 650.0, 750.0, 0.1, 5, 0.6034, 0.0, FP64)
    assert y.size == ref.size and va == vb
    assert np.abs(y - ref).max() <= 1e-10 * np.abs(ref).max()


def test_scan_large_fp32_vs_fp64(ctx):
    """A 1.4 M-point spectrum (transform of 8 M points): FP32 transform error."""
    from clblm_amd._lib import FP32C, FP64
    sp = _spectrum(n=1400001, v1=500.0, dv=0.0025, seed=5)
    y64, a, b = _run(ctx, sp, 500.0, 0.0025, 600.0, 3900.0, 0.3, 5, 0.6034, 0.0, FP64)
    y32, a2, b2 = _run(ctx, sp, 500.0, 0.0025, 600.0, 3900.0, 0.3, 5, 0.6034, 0.0, FP32C)
    assert y64.size == y32.size and a == a2 and b == b2
    err = np.abs(y32 - y64).max() / np.abs(y64).max()
    print("FP32 transform: %d points, max err / max %.2e" % (y64.size, err))
    assert err <= 1e-5
