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


# (functID, v1out, v2out, dvout, hwhm, preBoxHW, deconvPreBox), as pinned against the
# reference in tests/test_oracle_scan.py
SCAN_MODES = [
    (1, 620., 780., 0., 0.05, -1., 0),      # boxcar scan
    (1, 620., 780., 0.03, 0.2, -1., 0),
    (9, 620., 780., 0., 0.6, -1., 0),       # Norton-Beer weak, medium, strong
    (10, 620., 780., 0.05, 0.6, -1., 0),
    (11, 620., 780., 0., 0.6, -1., 0),
    (2, 690., 710., 0., 1.0, -1., 0),       # below CRATIO: transform of the scanning function
    (3, 698., 702., 0., 1.0, -1., 0),
    (3, 698., 702., 0.01, 1.0, 0.02, 1),
    (5, 620., 780., 0., 0.6, 0.05, 1),      # boxcar pre-scan, with and without deconvolution
    (5, 620., 780., 0., 0.6, 0.05, 0),
    (3, 620., 780., 0.02, 0.3, 0.03, 1),
    (105, 620., 780., 0., 0.6, 0.02, 0),    # pre-scanned elsewhere: deconvolution
    # deconvolution only: without an apodization the factor Z/sin(Z) is bounded only
    # for a boxcar narrower than DV (here Z <= pi/2), else FP32 noise is amplified
    (-1, 620., 780., 0., 0.6, 0.005, 0),
]


@pytest.mark.parametrize("fid,v1o,v2o,dvo,hw,pb,dec", SCAN_MODES)
def test_scan_modes_match_oracle(ctx, fid, v1o, v2o, dvo, hw, pb, dec):
    """FP64 within 1e-10 of the largest value (the boxcar scan, which has no transform,
    bit for bit); the FP32 transform within 1e-5 of FP64."""
    import torch
    from clblm_amd._lib import FP32C, FP64
    from oracle.py import fft_scan
    sp = _spectrum()
    ref, va, da = fft_scan(sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, 0.0, pb, dec)
    x = torch.tensor(sp, dtype=torch.float64, device="cuda")
    out = {}
    for prec in (FP64, FP32C):
        y, vb, db = ctx.fft_scan(x, 600.0, 0.01, v1o, v2o, dvo, fid, hw, 0.0, prec, pre_box_hw=pb,
                                 deconv_pre_box=bool(dec))
        torch.cuda.synchronize()
        out[prec] = y.cpu().numpy()
        assert out[prec].size == ref.size and vb == va and db == da
    err = np.abs(out[FP64] - ref).max() / np.abs(ref).max()
    err32 = np.abs(out[FP32C] - out[FP64]).max() / np.abs(out[FP64]).max()
    print("fid %d: %d points, FP64 err %.2e, FP32 vs FP64 %.2e" % (fid, ref.size, err, err32))
    assert err <= 1e-10 and err32 <= 1e-5
    if fid == 1 and dvo == 0.:
        assert np.array_equal(out[FP64], ref)


@pytest.mark.parametrize("fid,pb,fill", [(12, -1., 0.), (13, -1., 0.), (14, -1., 0.), (5, -1., 1.5)])
def test_scan_rejects_what_the_reference_cannot_run(ctx, fid, pb, fill):
    import torch
    from clblm_amd._lib import ClblmError, FP64
    x = torch.ones(20001, dtype=torch.float64, device="cuda")
    with pytest.raises(ClblmError):
        ctx.fft_scan(x, 600.0, 0.01, 620.0, 780.0, 0.0, fid, 0.6, fill, FP64, pre_box_hw=pb)


RECT_CASES = [(20001, 0.01, 0.05), (20001, 0.01, 0.253), (9001, 0.002, 0.0155), (7001, 0.01, 0.07),
              (30001, 0.004, 0.61), (5001, 0.01, 0.0151), (3001, 0.01, 0.012), (12345, 0.0031, 0.04)]


@pytest.mark.parametrize("n,dv,hw", RECT_CASES)
def test_rect_scan_matches_oracle(ctx, n, dv, hw):
    """The boxcar pre-average of Jacobian arrays (clblm_SCANFN functID 1) on the GPU:
    the oracle's grid and bit-identical values (the oracle is itself bit-identical to
    the reference's routine), across 2400-point input panels, with the interpolation
    onto HWHM/12 (hw 0.07 at dv 0.01) and without a scan (hw 0.012 at dv 0.01)."""
    import torch
    from oracle.py import prebox
    rng = np.random.default_rng(11)
    v = 600.0 + dv * np.arange(n)
    x = 1.0 + 0.5 * np.sin(v / 3.0) + 0.3 * np.exp(-((v - 620.0) / 0.05) ** 2) + 0.1 * rng.normal(size=n)
    want, w1, wd = prebox(x, 600.0, dv, hw)
    got, g1, gd = ctx.prebox(torch.tensor(x, device="cuda"), 600.0, dv, hw)
    got = got.cpu().numpy()
    assert (got.size, g1, gd) == (want.size, w1, wd), ((got.size, g1, gd), (want.size, w1, wd))
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def _conv_gpu(ctx, x, v1, dv, a, b, dvo, fid, hw):
    import torch
    y, g1, gd = ctx.scanfn(torch.tensor(x, device="cuda"), v1, dv, a, b, dvo, fid, hw)
    torch.cuda.synchronize()
    return y.cpu().numpy(), g1, gd


def _conv_cases():
    from test_oracle_scan import CONV_CASES
    return CONV_CASES


@pytest.mark.parametrize("n,dv,v1o,v2o,dvo,fid,hw,tail", _conv_cases())
def test_conv_scan_matches_oracle_bitwise(ctx, n, dv, v1o, v2o, dvo, fid, hw, tail):
    """Triangle, Gaussian, sinc**2 and sinc scans (clblm_SCANFN functIDs 2..5, -4,
    -5) on the GPU: the oracle's grid and count and bit-identical values, the oracle
    being bit-identical to the reference's routine on these cases."""
    from oracle.py import scanfn
    from test_oracle_scan import _conv_input
    x = _conv_input(n, dv, tail)
    print("case", n, dv, v1o, v2o, dvo, fid, hw, tail, flush=True)
    want, w1, wd = scanfn(x, 600.0, dv, v1o, v2o, dvo, fid, hw)
    got, g1, gd = _conv_gpu(ctx, x, 600.0, dv, v1o, v2o, dvo, fid, hw)
    assert (got.size, g1, gd) == (want.size, w1, wd), ((got.size, g1, gd), (want.size, w1, wd))
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_conv_scan_random_cases_bitwise(ctx):
    """Thirty seeded random scans, each bit-identical to the oracle (count and grid
    included): zero runs, zero tails, all-zero inputs, every function and sampling."""
    from oracle.py import scanfn
    rng = np.random.default_rng(77)
    done = 0
    while done < 30:
        fid = int(rng.choice([2, 3, 4, 5, -4, -5]))
        n = int(rng.integers(50, 60000))
        dv = float(rng.choice([0.001, 0.002, 0.0037, 0.01, 0.05]))
        hw = float(rng.uniform(0.02, 1.5))
        v1 = 600.0 + float(rng.uniform(-1, 1))
        span = dv * (n - 1)
        a = v1 + float(rng.uniform(-0.2, 0.8)) * span
        b = a + float(rng.uniform(0.05, 1.2)) * span
        dvo = float(rng.choice([0.0, hw / rng.uniform(1.6, 5.0)]))
        x = 1.0 + 0.5 * np.sin(np.arange(n) * dv / 0.7) + 0.1 * rng.normal(size=n)
        kind = int(rng.integers(0, 4))
        if kind == 1:
            x[-int(rng.integers(1, n // 2 + 1)):] = 0.0
        elif kind == 2:
            x[rng.uniform(size=n) < 0.3] = 0.0
        elif kind == 3 and rng.uniform() < 0.3:
            x[:] = 0.0
        try:
            want, w1, wd = scanfn(x, v1, dv, a, b, dvo, fid, hw)
        except RuntimeError:
            continue
        print("case", fid, n, dv, hw, a, b, dvo, kind, flush=True)
        got, g1, gd = _conv_gpu(ctx, x, v1, dv, a, b, dvo, fid, hw)
        assert (got.size, g1, gd) == (want.size, w1, wd), (fid, n, dv, hw, a, b, dvo)
        assert np.array_equal(got, want), (fid, n, dv, hw, a, b, dvo, float(np.abs(got - want).max()))
        done += 1


def test_conv_scan_rows(ctx):
    """Three rows scanned in one call equal three single scans; rows whose counts
    differ (the input ends before V2 + bound and one row ends in zeros) are refused."""
    import torch
    from clblm_amd._lib import ClblmError
    from oracle.py import scanfn
    rng = np.random.default_rng(8)
    n, dv = 40001, 0.002
    xs = 1.0 + 0.5 * np.sin(np.arange(n)[None, :] * dv / np.array([[0.7], [0.3], [1.9]])) + \
        0.05 * rng.normal(size=(3, n))
    y, g1, gd = ctx.scanfn(torch.tensor(xs, device="cuda"), 600.0, dv, 610.0, 670.0, 0.02, 5, 0.08)
    y = y.cpu().numpy()
    for r in range(3):
        want, w1, wd = scanfn(xs[r], 600.0, dv, 610.0, 670.0, 0.02, 5, 0.08)
        assert (y.shape[1], g1, gd) == (want.size, w1, wd)
        assert np.array_equal(y[r], want)
    xs[1, -n // 2:] = 0.0
    with pytest.raises(ClblmError):
        ctx.scanfn(torch.tensor(xs, device="cuda"), 600.0, dv, 610.0, 700.0, 0.02, 3, 0.08)


def test_conv_scan_large(ctx):
    """A 2 M-point spectrum at 0.001 cm-1 scanned by the sinc (HWHM 0.5, 72 zero
    crossings each side): bit-identical to the oracle."""
    from oracle.py import scanfn
    rng = np.random.default_rng(12)
    n, dv = 2000001, 0.001
    v = 500.0 + dv * np.arange(n)
    x = 1.0 + 0.5 * np.sin(v / 3.0) + 0.2 * np.exp(-((v - 1500.0) / 0.05) ** 2) + 0.05 * rng.normal(size=n)
    want, w1, wd = scanfn(x, 500.0, dv, 600.0, 2400.0, 0.125, 5, 0.5)
    got, g1, gd = _conv_gpu(ctx, x, 500.0, dv, 600.0, 2400.0, 0.125, 5, 0.5)
    assert (got.size, g1, gd) == (want.size, w1, wd)
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_filter_matches_oracle_bitwise(ctx):
    """The SRF filter (clblm_filter: clblm_FLTRFN per channel) on the GPU, all the
    seeded channels of the oracle's pin in one call per spectrum, bit for bit; the
    channels where the reference stops (CNVFLT's subscript check) are refused."""
    import torch
    from clblm_amd._lib import ClblmError
    from oracle.py import fltrfn
    from test_oracle_scan import filter_cases
    n_ok = 0
    for x, v1, dv, v1f, dvf, xf in filter_cases():
        try:
            want = fltrfn(x, v1, dv, v1f, dvf, xf)
        except RuntimeError:
            with pytest.raises(ClblmError):
                ctx.filter(torch.tensor(x, device="cuda"), v1, dv, [(v1f, dvf, xf)])
            continue
        got = ctx.filter(torch.tensor(x, device="cuda"), v1, dv, [(v1f, dvf, xf)]).cpu().numpy()
        assert got[0] == want, (x.size, dv, v1f, dvf)
        n_ok += 1
    assert n_ok >= 40
    # many channels over one spectrum in one call; every 24th of these ends where the
    # reference stops (CNVFLT's subscript check): those are kept for a call of their own
    rng = np.random.default_rng(5)
    n, dv = 200001, 0.005
    x = 1 + 0.3 * np.sin(np.arange(n) * dv / 0.7) + 0.05 * rng.normal(size=n)
    chans, stop = [], []
    for k in range(1500):
        c = (620.0 + 0.5 * k, -0.01, np.exp(-np.linspace(-2.5, 2.5, 201) ** 2))
        try:
            chans.append((c, fltrfn(x, 600.0, dv, *c)))
        except RuntimeError:
            stop.append(c)
    assert len(chans) > 1400 and stop
    xt = torch.tensor(x, device="cuda")
    got = ctx.filter(xt, 600.0, dv, [c for c, _ in chans]).cpu().numpy()
    assert np.array_equal(got, np.array([w for _, w in chans]))
    with pytest.raises(ClblmError):
        ctx.filter(xt, 600.0, dv, [chans[0][0], stop[0]])
