"""Pins the oracle's line sum by what the mathematics fixes (CPU only).

The reference's ODLAY cannot be built here (it needs netCDF), so the oracle's
nested-grid line sum is checked against properties a dropped term, a wrong sign
or a wrong index would break: the table identities, the Voigt width rule against
the published Olivero-Longbothum formula, the Faddeeva Voigt profile for isolated
lines and for the single-layer workload, area conservation, the Lorentz and
Doppler limits, linearity in the column, and the two radiation-term forms."""
import numpy as np
import pytest
from scipy.special import wofz

from clblm_amd import workloads as W

C2 = 1.4387752


def test_shape_tables_identities():
    from oracle.py import tables
    f1, f2, f3, fg, xv = tables()
    recpi = 1.0 / np.pi
    A = lambda z: (1 + 2 * z * z) / (1 + z * z) ** 2
    B = lambda z: -1 / (1 + z * z) ** 2
    # at common nodes x = 0.032 k (every 16th F1 node, every 4th F2 node, each F3 node)
    for k in range(0, 125):
        x = 0.032 * k
        s = f1[1 + 16 * k] + f2[1 + 4 * k] + f3[1 + k]
        lor = recpi * (1 / (1 + x * x) - (A(64.) + B(64.) * x * x))
        assert abs(s - lor) < 1e-13, (k, s, lor)
    # beyond 4 halfwidths F1 is zero and F2+F3 = Lorentz - Q3
    for k in range(126, 500):
        x = 0.032 * k
        s = f2[1 + 4 * k] + f3[1 + k] if 4 * k < 2000 else f3[1 + k]
        ref = recpi * (1 / (1 + x * x) - (A(64.) + B(64.) * x * x)) if x < 16 else None
        if ref is not None and x < 15.9:
            assert abs(s - ref) < 1e-13
    # Gaussian normalisation: 2 * integral_0^inf FG dz = 1 (trapezoid on 0..4 plus the
    # analytic tail beyond 4 halfwidths)
    from scipy.special import erfc
    z = 0.002 * np.arange(2001)
    g = fg[1:2002].copy()
    g[-1] = np.sqrt(np.log(2) / np.pi) * np.exp(-np.log(2) * 16)
    tail = 0.5 * erfc(4.0 * np.sqrt(np.log(2)))
    assert abs(2 * (np.trapezoid(g, z) + tail) - 1.0) < 1e-8
    # quadratic continuation meets the Lorentzian at 64 halfwidths (F4 and F3 boundaries)
    zb = 64.0
    assert abs((A(zb) + B(zb) * zb * zb) - 1 / (1 + zb * zb)) < 1e-18


def test_avrat_matches_olivero_longbothum():
    """Voigt HWHM from AVRAT(zeta) vs 0.5346 aL + sqrt(0.2166 aL^2 + aD^2) (J.Q.S.R.T. 17, 1977)."""
    from oracle.py import lib
    import ctypes
    lib().orc_avrat.argtypes = [ctypes.c_int]
    for iz in range(1, 101):
        zeta = (iz - 1) / 100.0
        ratio = lib().orc_avrat(iz)
        ol = 0.5346 * zeta + np.sqrt(0.2166 * zeta ** 2 + (1 - zeta) ** 2)
        assert abs(ratio - ol) / ol < 2.5e-3, (iz, ratio, ol)


def _voigt_od(lines, layer, grid, nu_idx=None):
    """S(T) W V(x) with the 25 cm-1 cutoff and the F4 bound subtracted, times the
    radiation term, for the lines of molecule 1 (independent of the oracle)."""
    T = layer.tave
    ln2 = np.log(2)
    fad = ln2 * 2 * 6.02214199e23 * 1.3806503e-16 / (2.99792458e10 ** 2)
    rhorat = (layer.pave / 1013.25) * (296.0 / T)
    rhoslf = rhorat * layer.w[1] / layer.wtot
    n = 1.0 - lines.tmpalf.astype(float)
    scale = (T / 296.0) ** (1.0 - n)
    aL = (lines.a_air.astype(float) * (rhorat - rhoslf) + lines.a_self.astype(float) * rhoslf) * scale
    aD = lines.nu * np.sqrt(fad * T / 18.01)
    S = lines.s.astype(float) * layer.w[1] * (1 + np.exp(-C2 * lines.nu / T))
    out = np.zeros_like(grid)
    for i in range(lines.n) if nu_idx is None else nu_idx:
        lo, hi = np.searchsorted(grid, lines.nu[i] - 25), np.searchsorted(grid, lines.nu[i] + 25)
        x = grid[lo:hi] - lines.nu[i]
        sig = aD[i] / np.sqrt(2 * ln2)
        V = wofz((x + 1j * aL[i]) / (sig * np.sqrt(2))).real / (sig * np.sqrt(2 * np.pi))
        out[lo:hi] += S[i] * (V - aL[i] / np.pi / (aL[i] ** 2 + 625.0))
    xkt = T / C2
    return out * grid * (1 - np.exp(-grid / xkt)) / (1 + np.exp(-grid / xkt))


@pytest.mark.parametrize("p,tol_core", [(1013.25, 2e-2), (100.0, 2e-2), (3.0, 3e-2)])
def test_isolated_line_voigt(p, tol_core):
    """One line at 850.3 cm-1 (neighbours 0 strength): Lorentz, mixed and Doppler regimes."""
    from oracle.py import Oracle
    lines = W.make_lines(3, 800.0, 900.0, seed=11, mols=(1,))
    lines.nu[:] = [820.0, 850.3, 880.0]
    lines.s[:] = [0.0, np.float32(1e-21), 0.0]
    layer = W.Layer(pave=p, tave=296.0, wtot=1e24, w={1: 1e22})
    dv = W.calculate_dv(850.0, p, 296.0, 0.01, alfal0=float(lines.a_air[1]) * p / 1013.25 * 1013.25 / p)
    dv = W.calculate_dv(850.0, p, 296.0, 0.01)
    od, _ = Oracle(lines).odlay(layer, 830.0, 870.0, dv)
    grid = 830.0 + dv * np.arange(od.size)
    ref = _voigt_od(lines, layer, grid, nu_idx=[1])
    core = np.abs(grid - 850.3) < 0.5
    rel = np.abs(od[core] - ref[core]) / ref[core].max()
    assert rel.max() < tol_core, rel.max()
    inside = np.abs(grid - 850.3) < 20
    area = od[inside].sum() / ref[inside].sum()
    assert abs(area - 1) < 2e-3, area


def test_single_layer_workload_vs_faddeeva():
    from oracle.py import Oracle
    lines, layers, cfg = W.single_layer()
    od, _ = Oracle(lines).odlay(layers[0], cfg["v1"], cfg["v2"], cfg["dv"])
    grid = cfg["v1"] + cfg["dv"] * np.arange(od.size)
    ref = _voigt_od(lines, layers[0], grid)
    rel = np.abs(od / ref - 1)
    assert np.median(rel) < 2e-3
    assert abs(od.sum() / ref.sum() - 1) < 1e-4


def test_linear_in_column():
    from oracle.py import Oracle
    lines = W.make_lines(4000, 600.0, 700.0, seed=12, mols=(1,))
    a = W.Layer(pave=300.0, tave=296.0, wtot=1e24, w={1: 1e22})
    b = W.Layer(pave=300.0, tave=296.0, wtot=2e24, w={1: 2e22})
    dv = W.calculate_dv(650.0, 300.0, 296.0, 0.01)
    oa, _ = Oracle(lines).odlay(a, 620.0, 680.0, dv)
    ob, _ = Oracle(lines).odlay(b, 620.0, 680.0, dv)
    np.testing.assert_allclose(ob, 2 * oa, rtol=1e-12)


def test_radiation_term_forms_agree():
    """JRAD=1 (radiation factor per line, set when V1 >= 5 kT) vs JRAD=0 (per panel)."""
    from oracle.py import Oracle
    lines = W.make_lines(3000, 1450.0, 1600.0, seed=13, mols=(1,))
    lay = W.Layer(pave=500.0, tave=280.0, wtot=1e24, w={1: 1e22})
    dv = W.calculate_dv(1525.0, 500.0, 280.0, 0.01)
    o1 = Oracle(lines)
    od1, _ = o1.odlay(lay, 1500.0, 1550.0, dv)
    assert o1.jrad == 1
    o0 = Oracle(lines)
    o0.jrad = 0
    od0, _ = o0.odlay(lay, 1500.0, 1550.0, dv)
    assert np.median(np.abs(od1 / od0 - 1)) < 2e-3


def test_calculate_dv_limits():
    from oracle.py import lib
    L = lib()
    # collision limit: DV -> ALFAL0 (p/p0) sqrt(T0/T) (1+4 x_h2o) / SAMPLE
    v = L.orc_calculate_dv(1000.0, 10 * 1013.25, 296.0, 0.0, 0.04, 4.0)
    assert abs(v / (0.04 * 10 / 4) - 1) < 1e-4
    # Doppler limit: DV -> 3.58115e-7 vbar sqrt(T/36) / SAMPLE
    v = L.orc_calculate_dv(1000.0, 1e-6, 250.0, 0.0, 0.04, 4.0)
    assert abs(v / (3.58115e-7 * 1000 * np.sqrt(250 / 36.0) / 4) - 1) < 1e-6


def test_xint_exact_for_linear():
    from oracle.py import xint
    v1a, dva = 500.0, 1.0
    a = np.concatenate([[0.0], 3.0 + 0.25 * (v1a + dva * np.arange(100))])
    r3 = np.zeros(200)
    xint(v1a, v1a + 99.0, dva, a, 510.3, 0.0123, r3, 1, 155)
    vi = 510.3 + 0.0123 * np.arange(155)
    np.testing.assert_allclose(r3[1:156], 3.0 + 0.25 * vi, rtol=1e-13)
