"""The MT_CKD water-vapour continuum on the GPU (clblm_mtckd_absrb and the
CLBLM_OD_MTCKD path of the OD) against the oracle: the continuum itself bit for bit,
the OD with the continuum within the FP64 tolerance of the other OD tests, and
the OD the same continuum gives when the caller passes it in."""
import os

import numpy as np
import pytest

from clblm_amd import workloads as W
from clblm_amd.continuum import load_mtckd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def table():
    return load_mtckd(os.path.join(HERE, "golden", "mtckd_h2o.npz"))


@pytest.fixture(scope="module")
def ctx(table):
    from clblm_amd.odlay import Context
    c = Context(0)
    c.mtckd_load(table)
    yield c
    c.close()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("jrad", [1, 0, -1])
def test_absrb_bitwise(ctx, table, jrad):
    from oracle.py import mtckd_h2o
    layers = W.atmosphere()[::7]
    bands = [(500.0, 2500.0), (10.3, 350.7), (2499.5, 3100.2)]
    for v1, v2 in bands:
        got = ctx.mtckd_absrb(layers, v1, v2, jrad)
        for lay, g in zip(layers, got):
            want = mtckd_h2o(table, lay.pave, lay.tave, lay.w.get(1, 0.0), lay.wtot, v1, v2, jrad)
            g = g.cpu().numpy()
            assert g.size == want.size
            assert np.array_equal(g, want), (v1, v2, lay.pave, _rel(g, want))


@pytest.mark.parametrize("k", [0, 20, 45])
def test_od_with_mtckd_matches_oracle(ctx, table, k):
    from oracle.py import Oracle, mtckd_h2o
    from test_gpu_od_fp64 import _cmp
    lines = W.profile_lines(20000, 500.0, 800.0, seed=1001)
    lay = W.atmosphere()[k]
    dv = W.calculate_dv(650.0, lay.pave, lay.tave, lay.w[1] / lay.wtot)
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    od, _ = ctx.odlay(h, lay, 600.0, 700.0, dv, mtckd=True)
    orc = Oracle(lines)
    ref, _ = orc.odlay(lay, 600.0, 700.0, dv, mtckd=table)
    print("layer %d: max rel err %.3e" % (k, _cmp(od, ref)))
    # the same continuum passed in by the caller gives the same OD (to the rounding of
    # pass 1's atomic sums, whose order varies from run to run)
    ctx.jrad = 1
    jr = 0 if 600.0 / (lay.tave / 1.4387752) < 5.0 else 1
    cont = mtckd_h2o(table, lay.pave, lay.tave, lay.w[1], lay.wtot, 600.0, 700.0, jr)
    od2, _ = ctx.odlay(h, lay, 600.0, 700.0, dv, continuum=cont)
    assert _rel(od, od2) <= 1e-13
    ctx.free_lines(h)


def test_batch_mixes_caller_continuum_and_mtckd(ctx, table):
    """A batch where layers carry the MT_CKD continuum, a caller continuum, both (the
    caller's first, then MT_CKD added to it) or neither, in FP64 and FP32C."""
    import torch
    from clblm_amd._lib import FP32C, FP64
    from clblm_amd.odlay import Context
    from oracle.py import Oracle
    lines = W.profile_lines(20000, 500.0, 800.0, seed=1001)
    layers = W.atmosphere()[::15]
    h = ctx.upload_lines(lines)
    modes = [(True, False), (False, True), (True, True), (False, False)]
    for prec in (FP64, FP32C):
        pars, refs = [], []
        orc = Oracle(lines)
        ctx.jrad = 1
        for lay, (mt, cc) in zip(layers, modes):
            dv = W.calculate_dv(650.0, lay.pave, lay.tave, lay.w[1] / lay.wtot)
            cont = W.continuum_absrb(lay, 600.0, 700.0)[0] if cc else None
            pars.append(Context.params(600.0, 700.0, dv, precision=prec, continuum=cont, mtckd=mt))
            refs.append(orc.odlay(lay, 600.0, 700.0, dv, continuum=cont, mtckd=table if mt else None)[0])
        out, off = ctx.od_batch(h, layers, pars)
        torch.cuda.synchronize()
        out = out.cpu().numpy()
        for l, ref in enumerate(refs):
            od = out[off[l]:off[l + 1]]
            err = _rel(od, ref)
            print("prec %d layer %d modes %s: max err / max %.3e" % (prec, l, modes[l], err))
            assert err <= (1e-10 if prec == FP64 else 1e-5)
    ctx.free_lines(h)


def test_mtckd_needs_a_table():
    from clblm_amd._lib import ClblmError
    from clblm_amd.odlay import Context
    c = Context(0)
    try:
        lines = W.profile_lines(2000, 500.0, 800.0, seed=1)
        h = c.upload_lines(lines)
        lay = W.atmosphere()[3]
        with pytest.raises(ClblmError):
            c.odlay(h, lay, 600.0, 700.0, 0.002, mtckd=True)
    finally:
        c.close()


# ---- CONTNM's other bands (CO2, O3, O2, N2, Rayleigh) ------------------------------
CN_BANDS = [(500.0, 2500.0), (20.3, 340.7), (1300.0, 1900.0), (1990.0, 2910.0), (2300.0, 2480.0), (830.0, 1000.0),
            (4300.0, 4950.0), (7500.0, 8550.0), (9050.0, 11050.0), (12900.0, 13300.0), (12970.0, 13210.0),
            (13000.0, 18000.0), (27300.0, 30000.0), (39000.0, 42000.0), (39000.0, 40801.0), (35900.0, 38000.0),
            (56700.0, 58000.0), (57013.0, 58000.0)]


@pytest.mark.parametrize("jrad", [1, 0, -1])
def test_contnm_absrb_bitwise(ctx, table, jrad):
    """clblm_contnm_absrb against the oracle: the other bands alone, and after the water
    vapour as CONTNM adds them, bit for bit."""
    from clblm_amd import _lib
    from oracle.py import contnm_ir, mtckd_h2o
    layers = W.atmosphere()[::9]
    for v1, v2 in CN_BANDS:
        alone = ctx.contnm_absrb(layers, v1, v2, jrad, _lib.OD_CONTNM)
        both = ctx.contnm_absrb(layers, v1, v2, jrad, _lib.OD_MTCKD | _lib.OD_CONTNM)
        for lay, g1, g2 in zip(layers, alone, both):
            w1, _ = contnm_ir(lay.pave, lay.tave, lay.wtot, lay.w, v1, v2, jrad)
            h2o = mtckd_h2o(table, lay.pave, lay.tave, lay.w.get(1, 0.0), lay.wtot, v1, v2, jrad)
            w2, _ = contnm_ir(lay.pave, lay.tave, lay.wtot, lay.w, v1, v2, jrad, absrb=h2o)
            assert np.array_equal(g1.cpu().numpy(), w1), (v1, v2, lay.pave)
            assert np.array_equal(g2.cpu().numpy(), w2), (v1, v2, lay.pave)


def test_contnm_scales(ctx):
    from clblm_amd import _lib
    from oracle.py import contnm_ir
    layers = W.atmosphere()[::20]
    sc = (0.5, 1.0, 2.0, 0.0, 3.0)
    ctx.contnm_scales(*sc)
    try:
        got = ctx.contnm_absrb(layers, 100.0, 2600.0, 1, _lib.OD_CONTNM)
    finally:
        ctx.contnm_scales()
    for lay, g in zip(layers, got):
        want, _ = contnm_ir(lay.pave, lay.tave, lay.wtot, lay.w, 100.0, 2600.0, 1, sc)
        assert np.array_equal(g.cpu().numpy(), want)


@pytest.mark.parametrize("k", [0, 30, 59])
@pytest.mark.parametrize("band", [(2000.0, 2600.0, 2250.0, 2450.0), (12800.0, 13400.0, 12950.0, 13150.0),
                                  (9300.0, 9700.0, 9400.0, 9550.0)])
def test_od_with_all_continua_matches_oracle(ctx, table, k, band):
    """The OD with CONTNM's water vapour and other bands against the oracle's OD over the
    N2 fundamental and the CO2 band head, the O2 A-band, and the 1.06 um O2 band with the
    ozone Chappuis band."""
    from oracle.py import Oracle
    from test_gpu_od_fp64 import _cmp
    lo, hi, v1, v2 = band
    lines = W.profile_lines(20000, lo, hi, seed=1001)
    lay = W.atmosphere()[k]
    dv = W.calculate_dv(0.5 * (v1 + v2), lay.pave, lay.tave, lay.w[1] / lay.wtot)
    h = ctx.upload_lines(lines)
    ctx.jrad = 1
    od, _ = ctx.odlay(h, lay, v1, v2, dv, mtckd=True, other_continua=True)
    orc = Oracle(lines)
    ref, _ = orc.odlay(lay, v1, v2, dv, mtckd=table, contnm=(1.0, 1.0, 1.0, 1.0, 1.0))
    err = _cmp(od, ref)
    print("layer %d, %g-%g cm-1: max rel err %.3e" % (k, v1, v2, err))
    ctx.free_lines(h)


def test_contnm_absrb_longer_than_the_reference_is_refused(ctx):
    """CONTNM's arrays hold 5050 points (N_ABSRB): a longer ABSRB is an argument error,
    a range just inside it is not."""
    from clblm_amd import _lib
    from clblm_amd._lib import ClblmError
    lay = W.atmosphere()[0]
    with pytest.raises(ClblmError, match="5050"):
        ctx.contnm_absrb([lay], 1000.0, 6100.0, 1, _lib.OD_CONTNM)
    got = ctx.contnm_absrb([lay], 1000.0, 6000.0, 1, _lib.OD_CONTNM)
    assert got[0].numel() == 5007
