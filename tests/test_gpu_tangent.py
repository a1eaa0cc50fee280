"""The tangent batch (clblm_od_tangent) against the oracle's derivatives
(orc_odlay_tan, pinned in tests/test_oracle_tangent.py), FP64 within 1e-10 of each
channel's peak: the OD of the smooth model and its derivatives with respect to the
layer temperature, the broadening column and each molecule's column.

Cases: the radiation term in panelOut (JRAD 0) and in the line strengths and on R4
(JRAD 1), the Van Vleck-Weisskopf mirror, coupled lines with pressure shifts, the
continuum with its derivative and the DVout re-grid, and a batch of layers of
different pressures, temperatures and grids in one call.
"""
import dataclasses

import numpy as np
import pytest

from clblm_amd import workloads as W
from clblm_amd.odlay import Context
from oracle.py import Oracle

pytestmark = pytest.mark.gpu

BAR = 1e-10


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _lines(n, v1, v2, seed, mols):
    L = W.make_lines(n, v1, v2, seed=seed, mols=mols)
    rng = np.random.default_rng(seed + 1)
    L.shift = rng.uniform(-0.01, 0.01, L.n).astype(np.float32)
    return L


def _dirs(mols):
    return [(1.0, 0.0, {}), (0.0, 1.0, {})] + [(0.0, 0.0, {m: 1.0}) for m in mols]


def _compare(gpu, ref, what):
    assert gpu.shape == ref.shape, (what, gpu.shape, ref.shape)
    for c in range(ref.shape[0]):
        scale = np.max(np.abs(ref[c]))
        err = np.max(np.abs(gpu[c] - ref[c])) / scale
        print("%s channel %d: %d points, max |gpu - oracle| / peak %.3e" % (what, c, ref.shape[1], err))
        assert err < BAR, (what, c, err)


CASES = {
    "jrad0_panel": ((560.0, 700.0), (1, 2), (500.0, 250.0), (600.0, 650.0)),
    "jrad1_lines_r4": ((1950.0, 2100.0), (1, 2, 3), (300.0, 240.0), (2000.0, 2040.0)),
    "vvw_mirror": ((0.5, 90.0), (1, 3), (800.0, 280.0), (20.0, 45.0)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_tangent_matches_oracle(ctx, name):
    (l1, l2), mols, (p, t), (v1, v2) = CASES[name]
    lines = _lines(2500, l1, l2, 11, mols)
    layer = W.Layer(pave=p, tave=t, wtot=p * W.AIR_COLUMN_PER_HPA * 0.01,
                    w={m: p * W.AIR_COLUMN_PER_HPA * 0.01 * (1e-3 if m == 1 else 3e-4) for m in mols})
    dv = W.calculate_dv(0.5 * (v1 + v2), p, t, 0.01)
    dirs = _dirs(mols)
    o = Oracle(lines)
    ref, _ = o.odlay_tan(layer, v1, v2, dv, dirs)
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    out, off = ctx.od_tangent(h, [layer], [Context.params(v1, v2, dv)], dirs)
    assert ctx.jrad == o.jrad
    _compare(out.cpu().numpy()[:, off[0]:off[1]], ref, name)
    ctx.free_lines(h)


def test_tangent_coupled_lines(ctx):
    base = _lines(1200, 760.0, 840.0, 21, (2,))
    idx = list(range(5, 1200, 37))
    lines = W.couple_lines(base, idx, [1 if k % 2 else 3 for k in range(len(idx))],
                           [0.02, 0.015, 0.011, 0.008], [0.004, 0.003, 0.0025, 0.002])
    layer = W.Layer(pave=700.0, tave=265.0, wtot=2e25, w={2: 8e21})
    dv = W.calculate_dv(800.0, layer.pave, layer.tave, 0.0)
    dirs = _dirs((2,))
    ref, _ = Oracle(lines).odlay_tan(layer, 790.0, 810.0, dv, dirs)
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    out, off = ctx.od_tangent(h, [layer], [Context.params(790.0, 810.0, dv)], dirs)
    _compare(out.cpu().numpy()[:, off[0]:off[1]], ref, "coupled")
    ctx.free_lines(h)


def test_tangent_continuum_and_regrid(ctx):
    lines = _lines(2000, 880.0, 1020.0, 31, (1,))
    layer = W.Layer(pave=900.0, tave=285.0, wtot=2.2e25, w={1: 4e22})
    v1, v2 = 920.0, 960.0
    dv = W.calculate_dv(940.0, layer.pave, layer.tave, 0.01)
    coef, v1abs, npt = W.continuum_absrb(dataclasses.replace(layer, w={1: 1.0}), v1, v2)
    dirs = [(1.0, 0.0, {}), (0.0, 0.0, {1: 1.0})]
    cont = coef * layer.w[1]
    dcont = np.stack([np.zeros(npt), cont])
    ref, _ = Oracle(lines).odlay_tan(layer, v1, v2, dv, dirs, dvout=0.75 * dv, continuum=cont, dcontinuum=dcont)
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    out, off = ctx.od_tangent(h, [layer], [Context.params(v1, v2, dv, dvout=0.75 * dv, continuum=cont)], dirs,
                              dcontinuum=[dcont])
    _compare(out.cpu().numpy()[:, off[0]:off[1]], ref, "continuum+regrid")
    ctx.free_lines(h)


def test_tangent_batch_of_layers(ctx):
    """Eight layers of one profile in one call: each layer's channels against the
    oracle's, with JRAD carried from layer to layer as successive calls carry it."""
    lines = _lines(6000, 560.0, 760.0, 41, (1, 2, 3))
    prof = W.level_pressures(8, 50.0)
    tl = W.level_temperatures(8, 50.0)
    layers = []
    for k in range(8):
        p = 0.5 * (prof[k] + prof[k + 1])
        col = (prof[k] - prof[k + 1]) * W.AIR_COLUMN_PER_HPA
        layers.append(W.Layer(pave=p, tave=0.5 * (tl[k] + tl[k + 1]), wtot=col,
                              w={1: col * 2e-3 * (k + 1) ** -1.5, 2: col * 4e-4, 3: col * 1e-6 * (k + 1)}))
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, 0.01) for l in layers]
    dirs = [(1.0, -1.0 / layers[0].tave, {}), (0.0, 0.0, {1: 1.0}), (0.0, 0.0, {3: 1.0})]
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    out, off = ctx.od_tangent(h, layers, [Context.params(620.0, 680.0, dv) for dv in dvs], dirs)
    o = Oracle(lines)
    got = out.cpu().numpy()
    for k, (lay, dv) in enumerate(zip(layers, dvs)):
        ref, _ = o.odlay_tan(lay, 620.0, 680.0, dv, dirs)
        _compare(got[:, off[k]:off[k + 1]], ref, "layer %d" % k)
    ctx.free_lines(h)
