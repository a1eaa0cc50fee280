"""The oracle's merge-down radiance (mode1_lookDn_mergeDn / layerMerge_mode1 with
addThermalBoundary and addSolarBoundary, clblm_noScattRT.f90:42-842, 1945-2148),
CPU only.  The reference's driver needs ODLAY and netCDF scenes and cannot be built
here; the restatement is pinned by what the recursion fixes:
* an isothermal atmosphere over a surface at the same temperature radiates
  B(T) (1 - TxA (1 - eps) Tdn) for linInTau 1 and 2 (for 0 the up emission is the
  down path's, lyrEmUp = lyrEmDn, and the closed form carries that term);
* on one grid it equals the merge-up recursion (orc_rt_up) to rounding, the down path
  being the view path times the diffusivity factor;
* the solar term is TxA refl Tsun S0 with the path transmittances' closed forms;
* an observer inside the atmosphere sees the layers above it only through the
  down-welling radiance (the same as zero view-path OD above it, linInTau 1 and 2);
* on nested per-layer grids the result lands on the finest grid and agrees with the
  one-grid result of the same smooth ODs to the interpolation's accuracy."""
import numpy as np
import pytest

from oracle import py as O
from tests_helpers_rt import rt_up

C1, C2 = 1.191042722E-12, 1.4387752


def _planck(v, t):
    return C1 * v ** 3 / np.expm1(C2 * v / t)


def _atm(nlay=12, n=3001, v1=600.0, dv=0.05, seed=4, iso=None):
    rng = np.random.default_rng(seed)
    v = v1 + dv * np.arange(n)
    ods = [np.abs(0.3 * rng.uniform(0.2, 1.5) * (1.0 + 0.8 * np.sin(v / (1.0 + l))) + 0.01) for l in range(nlay)]
    t_lev = np.linspace(290.0, 215.0, nlay + 1) if iso is None else np.full(nlay + 1, iso)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1]) if iso is None else np.full(nlay, iso)
    return ods, t_lay, t_lev, v1, v1 + dv * (n - 1), dv


@pytest.mark.parametrize("lin", [0, 1, 2])
@pytest.mark.parametrize("emis", [1.0, 0.9])
def test_isothermal(lin, emis):
    ods, t_lay, t_lev, v1, v2, dv = _atm(iso=260.0)
    nlay = len(ods)
    sec = 1.66
    rad, tx, dvo = O.rt_mode1(ods, [dv] * nlay, v1, v2, t_lay, t_lev, fac_dn=np.full(nlay, sec), lin=lin,
                              emis=emis, tskin=260.0)
    v = v1 + dvo * np.arange(rad.size)
    tup = np.exp(-np.sum(ods, axis=0))
    tdn = np.exp(-np.sum([o * sec for o in ods], axis=0))
    B = _planck(v, 260.0)
    if lin == 0:   # the up emission is the down path's (lyrEmUp = lyrEmDn, :785)
        above = np.exp(-np.cumsum(np.array(ods)[::-1], axis=0))[::-1]     # Tx from the top to the layer's base
        above = np.vstack([above[1:], np.ones((1, v.size))])              # ... to its top
        emit = np.sum(above * (1 - np.exp(-np.array(ods) * sec)), axis=0)
        want = B * (emit + tup * (emis + (1 - emis) * (1 - tdn)))
    else:
        want = B * (1 - tup * (1 - emis) * tdn)
    np.testing.assert_allclose(rad, want, rtol=1e-11)
    np.testing.assert_allclose(tx, tup, rtol=1e-12)


@pytest.mark.parametrize("lin", [0, 1, 2])
def test_one_grid_equals_merge_up(lin):
    ods, t_lay, t_lev, v1, v2, dv = _atm()
    nlay = len(ods)
    rad, _, dvo = O.rt_mode1(ods, [dv] * nlay, v1, v2, t_lay, t_lev, fac_dn=np.full(nlay, 1.5), lin=lin, emis=0.95,
                             tskin=293.0)
    want = rt_up(np.array(ods), t_lay, t_lev, v1, dv, emis=0.95, tskin=293.0, lin=lin, sec_dn=1.5)
    assert dvo == dv and rad.size == want.size
    np.testing.assert_allclose(rad, want, rtol=1e-12)


def test_solar_term():
    ods, t_lay, t_lev, v1, v2, dv = _atm()
    nlay = len(ods)
    n = ods[0].size
    rng = np.random.default_rng(2)
    fu, fs = rng.uniform(1.0, 1.3, nlay), rng.uniform(1.1, 3.0, nlay)
    s0 = 1e-3 * (1.0 + 0.1 * np.cos(np.arange(n) / 50.0))
    kw = dict(fac_up=fu, fac_dn=np.full(nlay, 1.66), lin=1, emis=0.9, tskin=288.0)
    th, _, _ = O.rt_mode1(ods, [dv] * nlay, v1, v2, t_lay, t_lev, **kw)
    both, tx, _ = O.rt_mode1(ods, [dv] * nlay, v1, v2, t_lay, t_lev, fac_sun=fs, solar=True, refl=0.1, sol=s0,
                             sol_dv=dv, **kw)
    tup = np.prod([np.exp(-(o * f)) for o, f in zip(ods, fu)], axis=0)
    tsun = np.prod([np.exp(-(o * f)) for o, f in zip(ods, fs)], axis=0)
    np.testing.assert_allclose(both - th, tup * 0.1 * tsun * s0, rtol=1e-9, atol=1e-30)
    np.testing.assert_allclose(tx, tup, rtol=1e-12)


@pytest.mark.parametrize("lin", [1, 2])
def test_observer_inside(lin):
    ods, t_lay, t_lev, v1, v2, dv = _atm()
    nlay, k = len(ods), 7
    a, _, _ = O.rt_mode1(ods, [dv] * nlay, v1, v2, t_lay, t_lev, obs_level=k, lin=lin, emis=0.9, tskin=290.0)
    fu = np.where(np.arange(nlay) < k, 1.0, 0.0)
    b, _, _ = O.rt_mode1(ods, [dv] * nlay, v1, v2, t_lay, t_lev, fac_up=fu, lin=lin, emis=0.9, tskin=290.0)
    assert np.array_equal(a, b)


def test_nested_grids():
    """Layer DVs growing downward by the ratios the fixed-ratio grids use (the top
    finest): the result is on the top layer's grid and within 2e-6 of the one-grid
    result on that grid (the smooth ODs sampled on each layer's own grid)."""
    nlay, v1, v2 = 10, 600.0, 700.0
    dv_top = 0.004
    ratios = [1, 1, 4 / 3, 3 / 2, 2, 2, 3 / 2, 4 / 3, 2, 3 / 2]
    dvs = []
    d = dv_top
    for r in ratios:
        d = d * r
        dvs.append(d)
    dvs = dvs[::-1]                       # layer 0 (surface) coarsest
    t_lev = np.linspace(288.0, 220.0, nlay + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1])
    f = lambda v, l: 0.2 + 0.15 * np.sin(v / (2.0 + 0.3 * l)) ** 2 * (1 + 0.1 * l)
    ods = []
    for l, d in enumerate(dvs):
        n = int(np.ceil((v2 - v1) / d)) + 1
        ods.append(f(v1 + d * np.arange(n), l))
    rad, _, dvo = O.rt_mode1(ods, dvs, v1, v2, t_lay, t_lev, fac_dn=np.full(nlay, 1.66), lin=1, emis=0.95,
                             tskin=290.0)
    assert dvo == min(dvs)
    n = rad.size
    vf = v1 + dvo * np.arange(n)
    fine = np.array([f(vf, l) for l in range(nlay)])
    want = rt_up(fine, t_lay, t_lev, v1, dvo, emis=0.95, tskin=290.0, lin=1, sec_dn=1.66)
    inner = vf < v2 - 0.1
    np.testing.assert_allclose(rad[inner], want[inner], rtol=2e-6)


def test_reflecting_surface_needs_the_down_path():
    ods, t_lay, t_lev, v1, v2, dv = _atm(nlay=3)
    with pytest.raises(RuntimeError):
        O.rt_mode1(ods, [dv] * 3, v1, v2, t_lay, t_lev, thermal=False, solar=True, emis=0.9, sol=np.ones(3001),
                   sol_dv=dv)
