"""Layer grids (setDV, src/clblm_src/clblm_setDV.f90:114-382) from the library's host
code (no GPU): calculateDV against the oracle's restatement and the Python one (bit for
bit) and against its limits (pinned in test_oracle_physics.py); fixDVRatio on exact DVs
built to hit each of its branches, with the results worked by hand; the nesting that
is its purpose (every adjacent ratio one of interpSpectrum's aligned ratios); DVSET
and the uniform output DV; the reference's stop on a ratio above 2.5 as an error."""
import ctypes

import numpy as np
import pytest

from clblm_amd import _lib
from clblm_amd import workloads as W

D = ctypes.c_double


def _arr(x):
    return np.ascontiguousarray(x, np.float64)


def _set_dv(pave, tave, wtot, wh2o, v1, v2, grid_type, user_dv=-1.0, alfal0=0.04, alfal0_fine=0.005, sample=4.0):
    lib = _lib.load()
    n = len(pave)
    a = [_arr(x) for x in (pave, tave, wtot, wh2o)]
    dvn, dvf, dvo = np.zeros(n), np.zeros(n), ctypes.c_double()
    rc = lib.clblm_set_dv(n, *[x.ctypes.data for x in a], v1, v2, alfal0, alfal0_fine, sample, grid_type, user_dv,
                          dvn.ctypes.data, dvf.ctypes.data, ctypes.byref(dvo))
    return rc, dvn, dvf, dvo.value


def _profile():
    lay = W.atmosphere(60, 0.1)
    return ([l.pave for l in lay], [l.tave for l in lay], [l.wtot for l in lay], [l.w[1] for l in lay])


def test_calculate_dv_three_ways_bit_for_bit():
    from oracle.py import lib as orc
    p, t, wt, wh = _profile()
    out = np.zeros(len(p))
    a = [_arr(x) for x in (p, t, wt, wh)]
    _lib.check(_lib.load().clblm_calculate_dv(len(p), *[x.ctypes.data for x in a], 1500.0, 0.04, 4.0,
                                              out.ctypes.data))
    for k in range(len(p)):
        f = wh[k] / wt[k]
        assert out[k] == W.calculate_dv(1500.0, p[k], t[k], f)
        assert out[k] == orc().orc_calculate_dv(1500.0, p[k], t[k], f, 0.04, 4.0)


def test_exact_grid():
    p, t, wt, wh = _profile()
    rc, dvn, dvf, dvo = _set_dv(p, t, wt, wh, 500.0, 2500.0, _lib.GRID_EXACT, user_dv=0.001)
    assert rc == 0 and dvo == -999.9
    for k in range(len(p)):
        f = wh[k] / wt[k]
        assert dvn[k] == W.calculate_dv(1500.0, p[k], t[k], f, 0.04)
        assert dvf[k] == W.calculate_dv(1500.0, p[k], t[k], f, 0.005)


def _layers_for(dvc, alfal0=0.04, sample=4.0):
    """Layers whose exact DV is dvc: at 296 K, no water and V1 = V2 = 0 the Doppler term
    vanishes and DV = ALFAL0 (p / 1013.25) / SAMPLE."""
    p = [d * sample * 1013.25 / alfal0 for d in dvc]
    n = len(p)
    return p, [296.0] * n, [1e24] * n, [0.0] * n


def test_fixed_ratio_branches_by_hand():
    """Each next exact DV is the previous fixed DV over a chosen ratio, so each layer
    takes one branch of fixDVRatio."""
    want = [1e-4 * 124.0]   # log10(0.0123456) - 3 = -4.9 -> -4; 123.456 + 0.5 -> 123, odd -> 124
    dvc = [0.0123456]
    steps = [(1.999, 1.0 / 2.0),   # ITYPE int(1/0.999 + .5) = 1: x 1/2
             (1.512, 2.0 / 3.0),   # int(1.953 + .5) = 2: x 2/3
             (1.3, 2.0 / 3.0),     # int(3.33 + .5) = 3, taken as 2
             (1.22, 5.0 / 6.0),    # int(4.55 + .5) = 5: x 5/6
             (1.1, 1.0),           # below 1.2: unchanged
             (0.7, 1.0)]           # the exact DV grows: unchanged
    for ratio, f in steps:
        dvc.append(want[-1] / ratio)
        num = {0.5: 1.0, 2.0 / 3.0: 2.0, 5.0 / 6.0: 5.0}.get(f)
        want.append(want[-1] * num / (num + 1.0) if num else want[-1])
    rc, dvn, dvf, dvo = _set_dv(*_layers_for(dvc), 0.0, 0.0, _lib.GRID_FIXED_RATIO)
    assert rc == 0 and dvo == -999.9
    assert list(dvn) == want, (list(dvn), want)


def test_fixed_ratio_grids_nest():
    """On the 60-layer profile every adjacent ratio of the fixed-ratio DVs is 1 or n/(n+1),
    n in {1, 2, 4, 5}: the ratios interpSpectrum interpolates on its aligned branch."""
    p, t, wt, wh = _profile()
    rc, dvn, dvf, _ = _set_dv(p, t, wt, wh, 500.0, 2500.0, _lib.GRID_FIXED_RATIO)
    assert rc == 0
    ok = [1.0, 1 / 2, 2 / 3, 4 / 5, 5 / 6]
    for d in (dvn, dvf):
        r = d[1:] / d[:-1]
        assert all(min(abs(x - o) for o in ok) < 1e-12 for x in r), r
        s = d[0] / 10.0 ** np.floor(np.log10(d[0]) - 2)     # three significant figures, even
        assert abs(s - round(s)) < 1e-9 and int(round(s)) % 2 == 0 and 100 <= round(s) < 1000
        assert np.all(d <= np.maximum.accumulate(d))


def test_ratio_above_2_5_is_refused():
    rc, *_ = _set_dv(*_layers_for([0.01, 0.0039]), 0.0, 0.0, _lib.GRID_FIXED_RATIO)
    assert rc == _lib.ERR_ARG
    assert b"2.5" in _lib.load().clblm_last_error()


@pytest.mark.parametrize("grid_type", [0, 1])
def test_user_dv_rescales_within_20_percent(grid_type):
    p, t, wt, wh = _profile()
    _, base, basef, _ = _set_dv(p, t, wt, wh, 500.0, 2500.0, grid_type)
    user = base[-1] * 1.15
    rc, dvn, dvf, dvo = _set_dv(p, t, wt, wh, 500.0, 2500.0, grid_type, user_dv=user)
    assert rc == 0
    assert np.array_equal(dvn, base * (user / base[-1]))
    assert dvn[-1] == pytest.approx(user, rel=1e-15)
    # the narrow-line DVs are rescaled by the same rule, on their own last DV
    if abs(1.0 - user / basef[-1]) < 0.2:
        assert np.array_equal(dvf, basef * (user / basef[-1]))
    else:
        assert np.array_equal(dvf, basef)
    assert dvo == (user if grid_type == 1 else -999.9)
    rc, dvn, _, _ = _set_dv(p, t, wt, wh, 500.0, 2500.0, grid_type, user_dv=base[-1] * 1.25)
    assert rc == 0 and np.array_equal(dvn, base)


def test_uniform_output_dv():
    p, t, wt, wh = _profile()
    rc, dvn, dvf, dvo = _set_dv(p, t, wt, wh, 500.0, 2500.0, _lib.GRID_UNIFORM)
    assert rc == 0 and dvo == dvf.min()
