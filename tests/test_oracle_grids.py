"""Spectra on different grids and the merge-up on the layers' own grids (CPU only).

* orc_interp_to_finest equals the reference's own interpToFinestGrid
  (clblm_Spectrum.f90:799-866, oracle/_ref) bit for bit, for spectra whose DVs
  differ by the aligned ratios n/(n+1) and (n+1)/n, by arbitrary ratios, by nothing
  but their length, or not at all;
* orc_rt_up_grids (layerMerge_mode2 with interpToFinestGrid, clblm_noScattRT.f90:
  1700-1794) equals the single-grid merge orc_rt_up bit for bit when every layer is
  on one grid, gives B(T) for an isothermal atmosphere over a black surface at the
  same temperature on any grids, and lands on the finest grid of the set; its
  pieces, layerEmis and interpSpectrum, are pinned against the reference elsewhere
  (test_rt_down.py, test_oracle_ref.py)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libclblm_ref.so")


def _ref_finest(spectra, dvs, v1):
    ref = ctypes.CDLL(REF, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    n = len(spectra)
    nl = np.array([s.size for s in spectra], np.int64)
    off = np.concatenate([[0], np.cumsum(nl)[:-1]]).astype(np.int64)
    data = np.concatenate(spectra).astype(np.float64)
    cap = int(nl.max()) + 8
    out = np.zeros((n, cap))
    nout, dvout = np.zeros(n, np.int64), np.zeros(n)
    P = ctypes.c_void_p
    ref.ref_interp_to_finest.argtypes = [ctypes.c_int, ctypes.c_double, P, P, P, P, ctypes.c_int, P, P, P]
    dv = np.ascontiguousarray(dvs, np.float64)
    ref.ref_interp_to_finest(n, v1, dv.ctypes.data, nl.ctypes.data, off.ctypes.data, data.ctypes.data, cap,
                             out.ctypes.data, nout.ctypes.data, dvout.ctypes.data)
    return [out[k, :nout[k]] for k in range(n)], dvout


CASES = {
    "halves": ([0.01, 0.005, 0.01], [1000, 1999, 1000]),
    "ratios": ([0.006, 0.004, 0.0048, 0.005, 0.003], [500, 749, 624, 599, 999]),
    "general": ([0.0123, 0.00457, 0.0071], [300, 806, 519]),
    "lengths": ([0.002, 0.002, 0.002], [800, 803, 798]),
    "same": ([0.004, 0.004], [600, 600]),
}


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("case", sorted(CASES))
def test_interp_to_finest_matches_the_reference(case):
    dvs, nls = CASES[case]
    rng = np.random.default_rng(len(case))
    v1 = 612.25
    spectra = [np.cumsum(rng.normal(0.0, 1.0, n)) + 50.0 for n in nls]
    want, dvw = _ref_finest(spectra, dvs, v1)
    got, loc = O.interp_to_finest(spectra, dvs, v1)
    for k in range(len(spectra)):
        assert got[k].size == want[k].size, (k, got[k].size, want[k].size)
        assert np.array_equal(got[k], want[k]), (case, k, np.abs(got[k] - want[k]).max())
        assert dvw[k] == dvs[loc]


def _atm(nlay, rng):
    t_lev = np.linspace(290.0, 215.0, nlay + 1) + rng.normal(0.0, 1.5, nlay + 1)
    t_lay = 0.5 * (t_lev[:-1] + t_lev[1:])
    return t_lay, t_lev


@pytest.mark.parametrize("lin", [0, 1, 2])
def test_one_grid_equals_the_single_grid_merge(lin):
    from tests_helpers_rt import rt_up
    rng = np.random.default_rng(5 + lin)
    nlay, n = 9, 700
    od = 10.0 ** rng.uniform(-3.0, 0.8, (nlay, n))
    t_lay, t_lev = _atm(nlay, rng)
    a, dvo = O.rt_up_grids(list(od), [0.01] * nlay, 700.0, t_lay, t_lev, lin=lin, sec_dn=1.4)
    b = rt_up(od, t_lay, t_lev, 700.0, 0.01, lin=lin, sec_dn=1.4)
    assert dvo == 0.01 and np.array_equal(a, b)


@pytest.mark.parametrize("lin", [0, 1, 2])
@pytest.mark.parametrize("dvs,tol", [([0.012, 0.006, 0.004, 0.004, 0.002], 1e-12),
                                     ([0.012, 0.0071, 0.0049, 0.0037, 0.002], 2e-6),
                                     ([0.006, 0.012, 0.004, 0.0048, 0.002], 1e-4)])
def test_isothermal_black_surface_on_mixed_grids(lin, dvs, tol):
    """Every layer at T over a surface of emissivity 1 at T.  Where the grids only get
    finer upward (the usual atmosphere), only the merged arrays are interpolated,
    R = B(T) stays smooth and comes out as B(T) to the accuracy of interpSpectrum on
    it: exact to rounding for the aligned ratios (4-point Lagrange), about 1e-6 for
    arbitrary ratios (the fraction quantised to 1/200).  A layer coarser than the
    merged grid has its transmittance and emission interpolated separately, which the
    reference does too: (1 - interp tx) B(v) and interp((1 - tx) B) then differ by
    about B'(v) DV, so B(T) holds only to that."""
    rng = np.random.default_rng(11)
    v1, span = 800.0, 6.0
    ods = [10.0 ** rng.uniform(-2.0, 0.5, int(span / d) + 1) for d in dvs]
    T = 250.0
    rad, dvo = O.rt_up_grids(ods, dvs, v1, np.full(len(dvs), T), np.full(len(dvs) + 1, T), emis=1.0, tskin=T,
                             lin=lin)
    v = v1 + dvo * np.arange(rad.size)
    B = np.array([O.lib().orc_planck(float(x), T) for x in v])
    assert dvo == 0.002 and rad.size == ods[-1].size
    err = np.abs(rad / B - 1.0).max()
    assert err < tol, err
    if tol > 1e-5:
        assert err > 1e-7      # the separate interpolation is visible


def test_finest_grid_and_length():
    rng = np.random.default_rng(3)
    dvs = [0.01, 0.005, 0.0025, 0.004]
    ods = [rng.uniform(0.0, 1.0, int(5.0 / d) + k) for k, d in enumerate(dvs)]
    t_lay, t_lev = _atm(len(dvs), rng)
    rad, dvo = O.rt_up_grids(ods, dvs, 900.0, t_lay, t_lev)
    assert dvo == 0.0025 and rad.size == ods[2].size
