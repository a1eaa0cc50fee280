"""Spectra on different grids on the GPU (FP64, through the C-ABI) against the oracle:
interpSpectrum for every branch (equal, aligned n/(n+1) and (n+1)/n, arbitrary
ratios); interpToFinestGrid over a mixed set; and the merge-up radiance on the layers'
own grids (clblm_rt_up_grids) with layer ODs from clblm_od_batch on the fixed-ratio
and the exact grids of setDV, for every linInTau."""
import ctypes

import numpy as np
import pytest

from clblm_amd import _lib
from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _t(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("dvin,dvout", [(0.004, 0.004), (0.004, 0.002), (0.002, 0.004), (0.006, 0.004),
                                        (0.004, 0.006), (0.005, 0.004), (0.0048, 0.004), (0.0123, 0.0041),
                                        (0.0037, 0.0049), (0.01, 0.00033)])
def test_interp_spectrum_matches_oracle(ctx, dvin, dvout):
    from oracle.py import interp
    rng = np.random.default_rng(int(dvin * 1e5 + dvout * 1e4))
    nin = int(40.0 / dvin) + 1
    x = np.cumsum(rng.normal(0.0, 1.0, nin)) + 10.0
    nout = int(40.0 / dvout) + 3          # two points past the end: the extrapolated ones
    got = ctx.interp_spectrum(_t(x), 731.5, dvin, dvout, nout).cpu().numpy()
    want = interp(x, dvin, 731.5, dvout, 731.5, nout)
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_interp_to_finest_matches_oracle(ctx):
    from oracle.py import interp_to_finest
    rng = np.random.default_rng(9)
    dvs = [0.006, 0.004, 0.0048, 0.005, 0.0037, 0.004]
    nls = [500, 749, 624, 599, 809, 752]
    xs = [np.cumsum(rng.normal(0.0, 1.0, n)) + 30.0 for n in nls]
    got, dvo = ctx.interp_to_finest([_t(x) for x in xs], dvs, 612.0)
    want, loc = interp_to_finest(xs, dvs, 612.0)
    assert dvo == dvs[loc] == 0.0037
    for g, w in zip(got, want):
        assert np.array_equal(g.cpu().numpy(), w)


def _set_dv(layers, v1, v2, grid_type):
    lib = _lib.load()
    n = len(layers)
    a = [np.ascontiguousarray(x, np.float64) for x in ([l.pave for l in layers], [l.tave for l in layers],
                                                        [l.wtot for l in layers], [l.w[1] for l in layers])]
    dvn, dvf, dvo = np.zeros(n), np.zeros(n), ctypes.c_double()
    _lib.check(lib.clblm_set_dv(n, *[x.ctypes.data for x in a], v1, v2, 0.04, 0.005, 4.0, grid_type, -1.0,
                                dvn.ctypes.data, dvf.ctypes.data, ctypes.byref(dvo)))
    return dvn


@pytest.mark.parametrize("grid_type,lin", [(_lib.GRID_FIXED_RATIO, 1), (_lib.GRID_FIXED_RATIO, 0),
                                           (_lib.GRID_EXACT, 1), (_lib.GRID_EXACT, 2)])
def test_radiance_on_layer_grids_matches_oracle(ctx, grid_type, lin):
    """30 layers of the 60-layer profile's lower half and the upper 20 (the DVs shrink
    upward), OD from clblm_od_batch on each layer's own grid over 8 cm-1."""
    from clblm_amd.jacobian import upwelling_radiance_grids
    from clblm_amd.odlay import Context
    from oracle.py import rt_up_grids
    layers = W.atmosphere(60, 0.1)[::2]
    t_lev = W.level_temperatures(60, 0.1)[::2]
    t_lay = [l.tave for l in layers]
    lines = W.profile_lines(6000, 880.0, 960.0, seed=1001)
    v1, v2 = 916.0, 924.0
    dvs = _set_dv(layers, v1, v2, grid_type)
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    od, off = ctx.od_batch(h, layers, [Context.params(v1, v2, d) for d in dvs])
    ctx.sync()
    rad, dvo = upwelling_radiance_grids(ctx, od, off, dvs, v1, t_lay, t_lev, lin_in_tau=lin, sec_dn=1.3)
    odh = od.cpu().numpy()
    want, dvw = rt_up_grids([odh[off[k]:off[k + 1]] for k in range(len(layers))], dvs, v1, t_lay, t_lev, lin=lin,
                            sec_dn=1.3)
    ctx.free_lines(h)
    got = rad.cpu().numpy()
    assert dvo == dvw == dvs.min() and got.shape == want.shape
    err = np.abs(got - want).max() / np.abs(want).max()
    print("grid type %d lin %d: %d points, max err %.2e" % (grid_type, lin, got.size, err))
    assert err < 1e-13
