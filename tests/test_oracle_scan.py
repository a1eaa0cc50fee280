"""Pins the oracle's instrument scan (FFTSCN) against the reference's own compiled
clblm_FFT.f90 / clblm_ScanFilter.f90 (oracle/_ref), and against closed forms.

* Intpdr/clblm_INTERP (4-point interpolation with the sequentially accumulated
  fraction and 2400-point panels) is compared bit for bit.
* clblm_FFTSCN_array is compared for every analytic-apodization scanning function
  (triangle, Gaussian, sinc**2, sinc, Beer, Hamming, Hanning): identical output
  grid (count, V1, DV) and values within 1e-10 of the largest value; the
  remaining difference is the reference's transform (twiddles by recurrence)
  against numpy's.
* Closed forms: a constant spectrum scanned with the same constant as fill is
  multiplied by the apodization's value at zero path difference, and the sinc scan
  of a narrow Gaussian line is the sinc of the stated half width broadened by the
  line (the transform of a rectangle times a Gaussian)."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libclblm_ref.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")

D = ctypes.c_double
I = ctypes.c_int
P = ctypes.c_void_p


@pytest.fixture(scope="module")
def ref():
    lib = ctypes.CDLL(REF, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    lib.ref_intpdr.argtypes = [I, P, D, D, D, D, D, I, P, ctypes.POINTER(I), ctypes.POINTER(D)]
    lib.ref_fftscn.argtypes = [I, P, D, D, D, D, D, D, I, D, D, I, P, ctypes.POINTER(I), ctypes.POINTER(D),
                               ctypes.POINTER(D), ctypes.POINTER(D)]
    return lib


def _ref_intpdr(lib, x, v1, dv, vv1, vv2, dvv):
    out = np.zeros(200000)
    n, a = I(), D()
    lib.ref_intpdr(x.size, x.ctypes.data, v1, dv, vv1, vv2, dvv, out.size, out.ctypes.data, ctypes.byref(n),
                   ctypes.byref(a))
    return out[:n.value], a.value


def _ref_scan(lib, x, v1, dv, v1o, v2o, dvo, fid, hw, fill):
    out = np.zeros(200000)
    n, a, b, c = I(), D(), D(), D()
    lib.ref_fftscn(x.size, x.ctypes.data, v1, v1 + dv * (x.size - 1), dv, v1o, v2o, dvo, fid, hw, fill, out.size,
                   out.ctypes.data, ctypes.byref(n), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return out[:n.value], a.value, c.value


def _spectrum(n=20001, seed=3):
    rng = np.random.default_rng(seed)
    v = 600.0 + 0.01 * np.arange(n)
    return v, 1.0 + 0.5 * np.sin(v / 3.0) + 0.3 * np.exp(-((v - 700.0) / 0.05) ** 2) + 0.1 * rng.normal(size=n)


@needs_ref
@pytest.mark.parametrize("vv1,vv2,dvv", [(620.0, 780.0, 0.05), (590.0, 810.0, 0.037), (600.5, 799.9, 0.0031),
                                         (650.0, 650.5, 0.01), (601.234, 700.111, 0.017)])
def test_scan_interp_bitwise(ref, vv1, vv2, dvv):
    from oracle.py import scan_interp
    x = np.cumsum(np.random.default_rng(3).normal(size=20001)) * 0.01 + 2.0
    a, va = scan_interp(x, 600.0, 0.01, vv1, vv2, dvv)
    b, vb = _ref_intpdr(ref, x, 600.0, 0.01, vv1, vv2, dvv)
    assert va == vb and a.size == b.size
    assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("fid", range(2, 9))
@pytest.mark.parametrize("v1o,v2o,dvo,hw,fill", [(620.0, 780.0, 0.05, 0.6034, 0.0), (650.0, 700.0, 0.0, 0.25, 1.0)])
def test_fft_scan_matches_reference(ref, fid, v1o, v2o, dvo, hw, fill):
    from oracle.py import fft_scan
    _, sp = _spectrum()
    a, va, da = fft_scan(sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, fill)
    b, vb, db = _ref_scan(ref, sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, fill)
    assert a.size == b.size and va == vb and da == db
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


@needs_ref
def test_fft_scan_coarse_input_is_regridded_first(ref):
    """DV > 2 HWHM/12: the reference interpolates onto DVN before the transform."""
    from oracle.py import fft_scan
    v = 600.0 + 0.2 * np.arange(1001)
    sp = 1.0 + 0.5 * np.sin(v / 3.0)
    a, va, _ = fft_scan(sp, 600.0, 0.2, 650.0, 750.0, 0.1, 5, 0.6034, 0.0)
    b, vb, _ = _ref_scan(ref, sp, 600.0, 0.2, 650.0, 750.0, 0.1, 5, 0.6034, 0.0)
    assert a.size == b.size and va == vb
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


@pytest.mark.parametrize("fid", range(2, 9))
def test_constant_spectrum_is_scaled_by_the_zero_path_gain(fid):
    """A constant spectrum (filled with the same constant) is multiplied by the
    apodization at zero path difference: 1, except Hamming's 0.53835685 (1 + 0.857504)
    (clblm_FFT.f90:1690)."""
    from oracle.py import fft_scan
    sp = np.full(20001, 0.8)   # FILL must lie in [0, 1] (clblm_FFT.f90:428-431)
    y, _, _ = fft_scan(sp, 600.0, 0.01, 650.0, 750.0, 0.0, fid, 0.3, 0.8)
    gain = 0.53835685 * (1.0 + 0.857504) if fid == 7 else 1.0
    np.testing.assert_allclose(y, 0.8 * gain, rtol=1e-12)


def test_sinc_scan_of_a_narrow_line_is_the_instrument_line_shape():
    """A Gaussian line of area S and width sigma << HWHM scans to
    S int_{-L}^{L} exp(-2 pi^2 sigma^2 x^2) cos(2 pi (v - v0) x) dx, the sinc of maximum
    path difference L = 1/(C4 HWHM) broadened by the line; its half maximum lies at
    v0 +- HWHM."""
    from oracle.py import fft_scan
    dv, hw, S, v0, sig = 0.01, 0.6034, 2.0, 700.003, 0.02
    v = 600.0 + dv * np.arange(20001)
    sp = S * np.exp(-0.5 * ((v - v0) / sig) ** 2) / (sig * np.sqrt(2 * np.pi))
    y, v1, dvy = fft_scan(sp, 600.0, dv, 650.0, 750.0, 0.0, 5, hw, 0.0)
    L = 1.0 / (3.3148 * hw)
    vy = v1 + dvy * np.arange(y.size)
    near = np.abs(vy - v0) < 8.0
    x = np.linspace(-L, L, 20001)
    wgt = np.exp(-2 * np.pi ** 2 * sig ** 2 * x ** 2)
    ils = np.array([S * np.trapezoid(wgt * np.cos(2 * np.pi * (vv - v0) * x), x) for vv in vy[near]])
    assert np.abs(y[near] - ils).max() <= 2e-3 * ils.max()
    half = vy[near][y[near] >= 0.5 * y.max()]
    assert abs(0.5 * (half.max() - half.min()) - hw) < 0.03 * hw


def _ref_scan2(lib, x, v1, dv, v1o, v2o, dvo, fid, hw, fill, pb, dec):
    lib.ref_fftscn2.argtypes = [I, P, D, D, D, D, D, D, I, D, D, D, I, P, I, P, ctypes.POINTER(I), ctypes.POINTER(D),
                                ctypes.POINTER(D), ctypes.POINTER(D)]
    out = np.zeros(400000)
    n, a, b, c = I(), D(), D(), D()
    parm = np.zeros(3)
    lib.ref_fftscn2(x.size, x.ctypes.data, v1, v1 + dv * (x.size - 1), dv, v1o, v2o, dvo, fid, hw, fill, pb, dec,
                    parm.ctypes.data, out.size, out.ctypes.data, ctypes.byref(n), ctypes.byref(a), ctypes.byref(b),
                    ctypes.byref(c))
    return out[:n.value], a.value, c.value


# (functID, v1out, v2out, dvout, hwhm, preBoxHW, deconvPreBox)
SCAN_MODES = [
    (1, 620., 780., 0., 0.05, -1., 0),      # boxcar scan
    (1, 620., 780., 0.03, 0.2, -1., 0),
    (9, 620., 780., 0., 0.6, -1., 0),       # Norton-Beer weak, medium, strong
    (10, 620., 780., 0.05, 0.6, -1., 0),
    (11, 620., 780., 0., 0.6, -1., 0),
    (2, 690., 710., 0., 1.0, -1., 0),       # triangle, Gaussian below CRATIO: transform of the scanning function
    (3, 698., 702., 0., 1.0, -1., 0),
    (5, 620., 780., 0., 0.6, 0.05, 1),      # boxcar pre-scan, with and without deconvolution
    (5, 620., 780., 0., 0.6, 0.05, 0),
    (3, 620., 780., 0.02, 0.3, 0.03, 1),
    (105, 620., 780., 0., 0.6, 0.02, 0),    # pre-scanned elsewhere: deconvolution
    (-1, 620., 780., 0., 0.6, 0.02, 0),     # deconvolution only
]


@needs_ref
@pytest.mark.parametrize("fid,v1o,v2o,dvo,hw,pb,dec", SCAN_MODES)
def test_fft_scan_modes_match_reference(ref, fid, v1o, v2o, dvo, hw, pb, dec):
    from oracle.py import fft_scan
    _, sp = _spectrum()
    a, va, da = fft_scan(sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, 0.0, pb, dec)
    b, vb, db = _ref_scan2(ref, sp, 600.0, 0.01, v1o, v2o, dvo, fid, hw, 0.0, pb, dec)
    assert a.size == b.size and va == vb and da == db
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    if fid == 1:
        assert np.array_equal(a, b)        # no transform: the boxcar means are bit-identical


def test_boxcar_scan_is_the_mean_of_m_points():
    """functID 1: M = int(2 HWHM / DV) consecutive points averaged, the grid shifted by
    (M-1) DV / 2 (clblm_FFT.f90:689-706)."""
    from oracle.py import fft_scan
    _, sp = _spectrum()
    y, v1, dvy = fft_scan(sp, 600.0, 0.01, 620.0, 780.0, 0.0, 1, 0.05, 0.0)
    assert abs(dvy - 0.1) < 1e-12
    k0 = int(round((v1 - 0.045 - 600.0) / 0.01))
    means = sp[k0:k0 + 10 * y.size].reshape(y.size, 10).mean(axis=1)
    np.testing.assert_allclose(y, means, rtol=1e-13)


@pytest.mark.parametrize("fid,v1o,v2o,hw", [(2, 690., 710., 1.0), (3, 698., 702., 1.0)])
def test_scanning_function_transform_has_unit_gain(fid, v1o, v2o, hw):
    """IVX = -1: the transform of the scanning function is divided by its area, so a
    constant spectrum padded with the same constant is unchanged."""
    from oracle.py import fft_scan
    y, _, _ = fft_scan(np.full(20001, 0.75), 600.0, 0.01, v1o, v2o, 0.0, fid, hw, 0.75)
    np.testing.assert_allclose(y, 0.75, rtol=1e-12)


@pytest.mark.parametrize("fid", [12, 13, 14])
def test_functions_the_reference_cannot_run_are_rejected(fid):
    """Brault and Kaiser-Bessel stop in the reference (they need an optical path
    difference, clblm_FFT.f90:480-489); Kiruna's multi-record scanning function reads a
    variable that is never set (Scntrn :2136)."""
    from oracle.py import fft_scan
    with pytest.raises(RuntimeError):
        fft_scan(np.ones(20001), 600.0, 0.01, 620.0, 780.0, 0.0, fid, 0.6)


def _ref_scanfn(lib, x, v1, dv, v1o, v2o, dvo, fid, hw):
    lib.ref_scanfn.argtypes = [I, P, D, D, D, D, D, D, I, D, I, P, ctypes.POINTER(I), ctypes.POINTER(D),
                               ctypes.POINTER(D), ctypes.POINTER(D)]
    out = np.zeros(400000)
    n, a, b, c = I(), D(), D(), D()
    lib.ref_scanfn(x.size, x.ctypes.data, v1, v1 + dv * (x.size - 1), dv, v1o, v2o, dvo, fid, hw, out.size,
                   out.ctypes.data, ctypes.byref(n), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return out[:n.value], a.value, c.value


RECT_CASES = [  # (points, dv, boxcar half width): boxcar pre-averages of Jacobian arrays
    (20001, 0.01, 0.05), (20001, 0.01, 0.253), (9001, 0.002, 0.0155), (7001, 0.01, 0.07),
    (30001, 0.004, 0.61), (5001, 0.01, 0.0151), (3001, 0.01, 0.012), (12345, 0.0031, 0.04),
]


@needs_ref
@pytest.mark.parametrize("n,dv,hw", RECT_CASES)
def test_rect_scan_matches_reference(ref, n, dv, hw):
    """The rectangle scan of output_JacArray's pre-average (clblm_SCANFN, functID 1,
    CNVRCT) against the reference's own routine: identical grid and bit-identical
    values, across input panels of 2400 points, with the interpolation onto HWHM/12
    when the input is coarser (hw = 0.07 at dv = 0.01) and without a scan when the
    boxes hold fewer than two points (hw = 0.012 at dv = 0.01)."""
    from oracle.py import prebox
    x = _spectrum(n, seed=11)[1]
    v1 = 600.0
    m = int(2.0 * hw / dv)
    v2 = v1 + dv * (n - 1)
    got, g1, gd = prebox(x, v1, dv, hw)
    want, w1, wd = _ref_scanfn(ref, x, v1, dv, v1 + (m - 1) * dv / 2.0, v2 - (m - 1) * dv / 2.0, (m - 1) * dv, 1, hw)
    assert (got.size, g1, gd) == (want.size, w1, wd), ((got.size, g1, gd), (want.size, w1, wd))
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_rect_scan_closed_forms():
    """A constant spectrum stays constant inside the data (the weights sum to the
    count).  On a binary-exact grid (dv = 0.0625, half width 1.25, boxes of 40 points
    whose edges fall on cell boundaries) every box is the plain mean of its 40
    points, so a linear ramp keeps its value at each box centre, across the
    2400-point input panels as well."""
    from oracle.py import prebox
    n, dv, hw = 20001, 0.01, 0.1
    v1 = 600.0
    got, g1, gd = prebox(np.full(n, 3.25), v1, dv, hw)
    inner = got[5:-5]
    assert np.allclose(inner, 3.25, rtol=1e-14, atol=0)
    n, dv, hw = 8001, 0.0625, 1.25
    ramp = 2.0 + 0.5 * (v1 + dv * np.arange(n))
    got, g1, gd = prebox(ramp, v1, dv, hw)
    assert gd == 2.4375 and g1 == 601.21875
    centres = g1 + gd * np.arange(got.size)
    ok = (centres > v1 + 2 * hw) & (centres < v1 + dv * (n - 1) - 2 * hw)
    assert ok.sum() > 150
    assert np.allclose(got[ok], 2.0 + 0.5 * centres[ok], rtol=1e-14, atol=0)


# (points, dv, v1out, v2out, dvout, functID, hwhm, tail): the spectral-domain scans
# with a tabulated scanning function.  tail: "zeros" ends the input in zeros (the
# reference's count then follows the last nonzero box), "short" leaves a last input
# panel shorter than one SHRKSC box.
CONV_CASES = [
    (20001, 0.01, 620., 780., 0.05, 2, 0.2, None),
    (20001, 0.01, 620., 780., 0.05, 3, 0.2, None),
    (20001, 0.01, 620., 780., 0.05, 4, 0.2, None),
    (20001, 0.01, 620., 780., 0.05, 5, 0.2, None),          # input ends before V2 + bound
    (20001, 0.01, 620., 780., 0.0, 5, 0.6, None),           # default sampling
    (20001, 0.01, 620., 780., 0.1, 3, 1.0, None),           # SHRKSC ratio 8
    (20001, 0.01, 650., 750., 0.0, -5, 0.6, None),          # first zero crossing given
    (20001, 0.01, 650., 750., 0.0, -4, 0.6, None),
    (20001, 0.01, 550., 850., 0.05, 3, 0.25, None),         # output range wider than the input
    (5001, 0.2, 650., 1500., 0.3, 5, 1.2, None),            # coarse input interpolated onto HWHM/12
    (30001, 0.01, 620., 900., 0.05, 2, 0.2, "zeros"),
    (72003, 0.001, 600.5, 675.003, 0.025, 2, 0.1, "short"),
    (72008, 0.001, 600.5, 675.008, 0.025, 5, 0.1, None),
    (9001, 0.002, 601., 615., 0.0, 3, 0.02, None),          # no compression (IRATSH 1)
]


def _conv_input(n, dv, tail, seed=5):
    rng = np.random.default_rng(seed)
    x = 1.0 + 0.5 * np.sin(np.arange(n) * dv / 0.7) + 0.1 * rng.normal(size=n)
    if tail == "zeros":
        x[-n // 3:] = 0.0
        x[rng.uniform(size=n) < 0.2] = 0.0
    return x


@needs_ref
@pytest.mark.parametrize("n,dv,v1o,v2o,dvo,fid,hw,tail", CONV_CASES)
def test_conv_scan_matches_reference_bitwise(ref, n, dv, v1o, v2o, dvo, fid, hw, tail):
    """clblm_SCANFN for the triangle, Gaussian, sinc**2 and sinc against the
    reference's own routine: identical grid and count and bit-identical values,
    through SHRKSC's boxes carried across input panels, CONVSC's sequential ZF and
    R1 shifts, and the reference's count rules at the end of the input."""
    from oracle.py import scanfn
    x = _conv_input(n, dv, tail)
    got, g1, gd = scanfn(x, 600.0, dv, v1o, v2o, dvo, fid, hw)
    want, w1, wd = _ref_scanfn(ref, x, 600.0, dv, v1o, v2o, dvo, fid, hw)
    assert (got.size, g1, gd) == (want.size, w1, wd)
    assert np.array_equal(got, want), float(np.abs(got - want).max())


@needs_ref
def test_conv_scan_random_cases_match_reference(ref):
    """Forty seeded random scans (function, half width, grids, output range, zero runs
    and all-zero inputs): every one bit-identical to the reference."""
    from oracle.py import scanfn
    rng = np.random.default_rng(2024)
    done = 0
    while done < 40:
        fid = int(rng.choice([2, 3, 4, 5, -4, -5]))
        n = int(rng.integers(50, 30000))
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
            got, g1, gd = scanfn(x, v1, dv, a, b, dvo, fid, hw)
        except RuntimeError:
            continue    # outside R1 or the table: the reference overruns its arrays there
        want, w1, wd = _ref_scanfn(ref, x, v1, dv, a, b, dvo, fid, hw)
        assert (got.size, g1, gd) == (want.size, w1, wd), (fid, n, dv, hw, a, b, dvo)
        assert np.array_equal(got, want), (fid, n, dv, hw, a, b, dvo)
        done += 1


@pytest.mark.parametrize("jfn", [1, 2, 3, 4])
def test_scan_tables_closed_forms(jfn):
    """The tabulated functions every 0.02 half widths: the triangle 0.5 (1 - x/2) and
    the Gaussian sqrt(ln2/pi) exp(-ln2 x^2) (both zero at their last point), the
    sinc**2 and sinc XSCALE/pi (sin(xXSCALE)/(xXSCALE))**k; each is 1/2 of its peak at
    one half width, has unit area, and is zero past NJ."""
    from oracle.py import scan_shape
    f = scan_shape(jfn)
    nf = {1: 101, 2: 201, 3: 2710, 4: 5968}[jfn]
    x = 0.02 * np.arange(nf)
    xs = {3: 1.391557377, 4: 1.89549425}.get(jfn)
    if jfn == 1:
        want = 0.5 * (1 - 0.5 * x)
    elif jfn == 2:
        want = np.sqrt(np.log(2) / np.pi) * np.exp(-np.log(2) * x * x)
        want[-1] = 0.0
    else:
        t = np.where(x == 0, 1.0, x * xs)
        s = np.where(x == 0, 1.0, np.sin(t) / t)
        want = xs / np.pi * (s * s if jfn == 3 else s)
    np.testing.assert_allclose(f[:nf], want, rtol=1e-14, atol=1e-17)
    assert np.all(f[nf:] == 0.0)
    assert abs(f[50] / f[0] - 0.5) < 1e-6                # x = 1: half maximum
    area = 0.02 * (2.0 * f.sum() - f[0])
    assert abs(area - 1.0) < {1: 1e-12, 2: 1e-5, 3: 5e-3, 4: 5e-3}[jfn]   # the tails past HWJ are cut


@pytest.mark.parametrize("fid,hw", [(2, 0.1), (3, 0.1), (3, 0.3)])
def test_conv_scan_of_a_narrow_line(fid, hw):
    """A Gaussian line of width sigma << HWHM and area A scans (triangle, Gaussian)
    to the scanning function of the same area: the output integrates to A, peaks
    at the line, and its half maximum lies at the line +- HWHM (sigma adds in
    quadrature for the Gaussian)."""
    from oracle.py import scanfn
    dv, A, v0, sig = 0.001, 3.0, 650.0037, 0.002
    n = 100001
    v = 600.0 + dv * np.arange(n)
    x = A * np.exp(-0.5 * ((v - v0) / sig) ** 2) / (sig * np.sqrt(2 * np.pi))
    y, y1, dy = scanfn(x, 600.0, dv, 640.0, 660.0, hw / 20.0, fid, hw)
    vy = y1 + dy * np.arange(y.size)
    assert abs(y.sum() * dy - A) < 1e-3 * A
    assert abs(vy[np.argmax(y)] - v0) <= dy
    half = vy[y >= 0.5 * y.max()]
    width = 0.5 * (half.max() - half.min() + dy)
    expect = hw if fid == 2 else np.sqrt(hw ** 2 + 2 * np.log(2) * sig ** 2)
    assert abs(width - expect) < 0.06 * hw


def _ref_fltrfn(lib, x, v1, dv, v1f, dvf, xf):
    lib.ref_fltrfn.argtypes = [I, P, D, D, D, D, I, P, ctypes.POINTER(D)]
    o = D()
    xf = np.ascontiguousarray(xf, np.float64)
    lib.ref_fltrfn(x.size, x.ctypes.data, v1, dv, v1f, dvf, xf.size, xf.ctypes.data, ctypes.byref(o))
    return o.value


def filter_cases(seed=1, count=60):
    """Seeded SRF filter cases: (spectrum, v1, dv, v1f, dvf, filter values)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        n = int(rng.integers(100, 30000))
        dv = float(rng.choice([0.001, 0.0037, 0.01, 0.02]))
        v1 = 600.0 + float(rng.uniform(-1, 1))
        x = 1 + 0.5 * np.sin(np.arange(n) * dv / 0.3) + 0.1 * rng.normal(size=n)
        nf = int(rng.integers(3, 400)) | 1
        dvf = float(rng.uniform(0.002, 0.05))
        v1f = v1 + float(rng.uniform(-0.1, 0.9)) * dv * (n - 1)
        if rng.uniform() < 0.3:
            dvf = -dvf
        xf = np.exp(-np.linspace(-2, 2, nf) ** 2) + 0.05 * rng.uniform(size=nf)
        out.append((x, v1, dv, v1f, dvf, xf))
    return out


@needs_ref
def test_filter_matches_reference_bitwise(ref):
    """clblm_FLTRFN / CNVFLT against the reference's own routine on sixty seeded
    channels: filters inside, across and past the ends of the spectrum, given by
    their first point or their centre, over 2400-point input panels."""
    from oracle.py import fltrfn
    done = 0
    for x, v1, dv, v1f, dvf, xf in filter_cases():
        try:
            got = fltrfn(x, v1, dv, v1f, dvf, xf)
        except RuntimeError:
            continue       # CNVFLT's subscript check: the reference stops there
        assert got == _ref_fltrfn(ref, x, v1, dv, v1f, dvf, xf), (x.size, dv, v1f, dvf, xf.size)
        done += 1
    assert done >= 40


def test_filter_closed_forms():
    """A constant spectrum filters to itself once the filter lies inside the data and
    its points fall on the spectral grid (the trapezoid weights of both sums agree);
    a linear spectrum through a symmetric filter gives its value at the centre."""
    from oracle.py import fltrfn
    dv = 0.01
    x = np.full(20001, 2.5)
    xf = np.exp(-np.linspace(-3, 3, 301) ** 2)
    assert abs(fltrfn(x, 600.0, dv, 650.0, dv, xf) - 2.5) < 1e-3
    v = 600.0 + dv * np.arange(20001)
    y = fltrfn(1.0 + 0.01 * v, 600.0, dv, 700.0, -dv, xf)   # centred at 700
    assert abs(y - (1.0 + 0.01 * 700.0)) < 1e-3 * 8.0
