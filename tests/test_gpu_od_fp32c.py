"""FP32-compensated GPU path vs the FP64 GPU path: relative OD error <= 1e-5."""
import numpy as np
import pytest

from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu
RTOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _both(ctx, h, lay, v1, v2, dv, **kw):
    from clblm_amd._lib import FP32C, FP64
    ctx.jrad = 1
    a, _ = ctx.odlay(h, lay, v1, v2, dv, precision=FP64, **kw)
    ctx.jrad = 1
    b, _ = ctx.odlay(h, lay, v1, v2, dv, precision=FP32C, **kw)
    err = np.abs(b - a) / np.abs(a)
    return float(err.max()), a.size


def test_single_layer_case(ctx):
    lines, layers, cfg = W.single_layer()
    h = ctx.upload_lines(lines)
    e, n = _both(ctx, h, layers[0], cfg["v1"], cfg["v2"], cfg["dv"])
    print("single layer fp32c: %d points, max rel err %.3e" % (n, e))
    assert e <= RTOL


@pytest.mark.parametrize("k", [0, 20, 40, 59])
def test_profile_layers(ctx, k):
    lines = W.profile_lines(40000, 500.0, 800.0, seed=1001)
    lay = W.atmosphere()[k]
    dv = W.calculate_dv(650.0, lay.pave, lay.tave, lay.w[1] / lay.wtot)
    cont, _, _ = W.continuum_absrb(lay, 600.0, 700.0)
    h = ctx.upload_lines(lines)
    e, n = _both(ctx, h, lay, 600.0, 700.0, dv, continuum=cont)
    print("layer %d fp32c: %d points, max rel err %.3e" % (k, n, e))
    assert e <= RTOL


def test_scene_batch_fp32c_vs_fp64(ctx):
    """The scene batch (four scenes, 207 layers, one batch each) in FP32C against FP64."""
    import torch
    from clblm_amd._lib import FP32C, FP64
    from clblm_amd.odlay import Context
    scenes = W.scene_batch(4, seed=1003)
    lines = W.profile_lines(40000, 500.0, 800.0, seed=1001)
    layers = [l for s in scenes for l in s.layers]
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    h = ctx.upload_lines(lines)
    res = {}
    for p in (FP64, FP32C):
        ctx.jrad = 1
        out, off = ctx.od_batch(h, layers, [Context.params(600.0, 700.0, dv, precision=p) for dv in dvs])
        torch.cuda.synchronize()
        res[p] = out.cpu().numpy()
    e = float((np.abs(res[FP32C] - res[FP64]) / np.abs(res[FP64])).max())
    print("scene batch fp32c: %d layers, %d points, max rel err %.3e" % (len(layers), off[-1], e))
    assert e <= RTOL


def test_low_wavenumber_fp32c(ctx):
    """Large-scene layers over 10-60 cm-1 (Van Vleck-Weisskopf mirrors, RADFNI ramp) in
    FP32C against FP64; the denominator is floored at 1e-4 of the maximum where the
    spectrum crosses zero."""
    from clblm_amd._lib import FP32C, FP64
    lines = W.make_lines(50000, 10.0, 100.0, seed=1004, mols=(1, 2, 3, 4, 5, 6, 7))
    h = ctx.upload_lines(lines)
    for k in (0, 40, 79):
        lay = W.large_scene_atmosphere()[k]
        dv = W.large_scene_dv(lay)
        ctx.jrad = 1
        a, _ = ctx.odlay(h, lay, 10.0, 60.0, dv, precision=FP64)
        ctx.jrad = 1
        b, _ = ctx.odlay(h, lay, 10.0, 60.0, dv, precision=FP32C)
        e = float((np.abs(b - a) / np.maximum(np.abs(a), 1e-4 * np.abs(a).max())).max())
        print("low-wavenumber layer %d fp32c: %d points, max rel err %.3e" % (k, a.size, e))
        assert e <= RTOL


@pytest.mark.parametrize("k", [30, 59])
def test_far_field_nodes_match_every_point(ctx, k):
    """FP32C evaluates far-wing F4 records at 16 Chebyshev nodes per tile and
    interpolates; the same layer with every record summed at every point agrees to
    1e-6 relative (the interpolation itself is good to ~1e-8 for records half a tile
    away, ~5e-7 of the record's peak for a third of a tile; the rest is FP32)."""
    from clblm_amd._lib import FP32C, TUNE_F4_FAR_FIELD
    lines = W.profile_lines(40000, 500.0, 800.0, seed=1001)
    lay = W.atmosphere()[k]
    dv = W.calculate_dv(650.0, lay.pave, lay.tave, lay.w[1] / lay.wtot)
    h = ctx.upload_lines(lines)
    out = {}
    for mode in (1, 2, 3):
        ctx.tuning(TUNE_F4_FAR_FIELD, mode)
        ctx.jrad = 1
        out[mode], _ = ctx.odlay(h, lay, 600.0, 700.0, dv, precision=FP32C)
    ctx.tuning(TUNE_F4_FAR_FIELD, 1)
    for mode in (1, 3):
        e = float((np.abs(out[mode] - out[2]) / np.abs(out[2])).max())
        print("layer %d far-field nodes (setting %d) vs every point: %d points, max rel diff %.3e" %
              (k, mode, out[1].size, e))
        assert e <= 1e-6
