"""Line coupling in the oracle (LNCORQ clblm_ODLAY.f90:1821-1940, CNVFNQ :2195-2235,
LINF4Q clblm_ODLAY_LineF4.f90:860-988,1130-1131, SHRINQ :1563-1574, CNVF4Q :388-462).
The reference's ODLAY needs netCDF and is not built here, so these checks pin the
restatement by what the formulas fix:
- zero coefficients give the uncoupled result bit for bit (IFLAG 1 and 3);
- G scales a lone line by 1 + G(T) (P/P0)^2, with G(T) interpolated on the bracket
  each routine picks (nodes chosen non-linear, so a wrong bracket shows);
- IFLAG 3 with GAMMA1 = 1/2 at P = P0 equals the line with both widths halved, bit
  for bit;
- Y adds a term linear in Y, odd about the centre (positive above it for Y > 0) and
  equal to SPPSP (v - v0) / alpha_V times the line's own optical depth: the ratio to
  (v - v0) OD is one constant from the line core (CNVFNQ's CLC ZF factors) out to the
  far wings (CNVF4Q's XM DPTRAT), two independent pieces of code."""
import numpy as np
import pytest

from clblm_amd import workloads as W
from oracle.py import Oracle


def _od(lines, layer, v1=600.0, v2=700.0, dv=None):
    dv = dv or W.calculate_dv(0.5 * (v1 + v2), layer.pave, layer.tave, 0.01)
    o = Oracle(lines)
    o.jrad = -1          # no radiation term: every contribution is linear in the strength
    od, dvu = o.odlay(layer, v1, v2, dv)
    return od, dvu


def _lone(v0=650.0):
    """One strong water line at v0 and a far one (the upload needs two lines)."""
    L = W.make_lines(2, 100.0, 200.0, seed=3)
    L.nu[:] = [v0, 1500.0]
    L.s[:] = np.float32(1e-20)
    L.a_air[:] = np.float32(0.07)
    L.a_self[:] = np.float32(0.35)
    L.e_lo[:] = np.float32(200.0)
    L.tmpalf[:] = np.float32(0.25)
    L.shift[:] = np.float32(0.0)
    return L


def _layer(p=1013.25, t=270.0):
    return W.Layer(pave=p, tave=t, wtot=1e24, w={1: 1e22})


@pytest.mark.parametrize("iflag", [1, 3])
def test_zero_coefficients_are_the_plain_lines(iflag):
    lines = W.make_lines(3000, 560.0, 740.0, seed=5)
    idx = np.arange(0, lines.n, 7)
    c = W.couple_lines(lines, idx, iflag, [0, 0, 0, 0], [0, 0, 0, 0])
    p = W.couple_lines(lines, idx, iflag, [0, 0, 0, 0], [0, 0, 0, 0])
    p.iflg[p.iflg == iflag] = 0
    lay = _layer(506.625, 260.0)
    a, _ = _od(c, lay)
    b, _ = _od(p, lay)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("t", [215.0, 250.0, 270.0, 300.0, 330.0])
def test_g_scales_a_lone_line(t):
    # at 200, 250, 296, 340 K; stored as REAL*4 in the coefficient record
    g_nodes = np.array([0.0, 0.2, 0.1, 0.3], np.float32).astype(np.float64)
    lay = _layer(1013.25, t)                     # P = P0: factor 1 + G(T)
    base, _ = _od(W.couple_lines(_lone(), [0], 1, [0, 0, 0, 0], [0, 0, 0, 0]), lay)
    od, _ = _od(W.couple_lines(_lone(), [0], 1, [0, 0, 0, 0], g_nodes), lay)
    # the bracket containing t (both routines' brackets interpolate the same line)
    T = np.array([200.0, 250.0, 296.0, 340.0])
    k = min(max(int(np.searchsorted(T, t, side="right")) - 1, 0), 2)
    g = g_nodes[k] + (g_nodes[k + 1] - g_nodes[k]) / (T[k + 1] - T[k]) * (t - T[k])
    m = base > 1e-6 * base.max()
    ratio = od[m] / base[m]
    assert np.allclose(ratio, 1.0 + g, rtol=1e-12, atol=0), (t, g, ratio.min(), ratio.max())


def test_reduced_width_halves_the_widths():
    lay = _layer(1013.25, 280.0)
    c = W.couple_lines(_lone(), [0], 3, [0.5] * 4, [0.0] * 4)
    h = _lone()
    h.a_air[0] = np.float32(h.a_air[0] * 0.5)
    h.a_self[0] = np.float32(h.a_self[0] * 0.5)
    h = W.couple_lines(h, [0], 3, [0.0] * 4, [0.0] * 4)
    a, _ = _od(c, lay)
    b, _ = _od(h, lay)
    assert np.array_equal(a, b)


def test_y_term_is_linear_odd_and_proportional_to_the_offset():
    lay = _layer(506.625, 270.0)
    v0 = 650.0
    plain, dv = _od(W.couple_lines(_lone(v0), [0], 1, [0.0] * 4, [0.0] * 4), lay)
    y = 0.02
    d1 = _od(W.couple_lines(_lone(v0), [0], 1, [y] * 4, [0.0] * 4), lay)[0] - plain
    d2 = _od(W.couple_lines(_lone(v0), [0], 1, [2 * y] * 4, [0.0] * 4), lay)[0] - plain
    assert np.abs(d2 - 2 * d1).max() <= 1e-9 * np.abs(d2).max()
    v = 600.0 + dv * np.arange(plain.size)
    x = v - v0
    # odd about the centre, positive above it for Y > 0
    near = np.abs(x) < 2.0
    assert abs(d1[near].sum()) < 1e-2 * np.abs(d1[near]).sum()
    assert d1[(x > 0.05) & (x < 1.0)].min() > 0 and d1[(x < -0.05) & (x > -1.0)].max() < 0
    # d1 = SPPSP (v - v0) / alpha_V * OD_line, from the R1/R2/R3 parts near the centre
    # (CLC ZF) and from the F4 wings (XM DPTRAT) alike: one ratio in every zone
    means = []
    for lo, hi in ((0.02, 0.3), (0.3, 1.0), (1.0, 3.0), (3.0, 10.0), (10.0, 24.0)):
        zone = (np.abs(x) > lo) & (np.abs(x) < hi)
        r = d1[zone] / (x[zone] * plain[zone])
        assert r.std() / abs(r.mean()) < 2e-3, (lo, hi, r.min(), r.max())
        means.append(r.mean())
    assert (max(means) - min(means)) / abs(np.mean(means)) < 1e-3, means
