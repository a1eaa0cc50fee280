"""Launch-shape settings and the autotuner: every setting gives the same optical depth
(FP64 within the LDS-atomic reordering of pass 1; FP32C within its 1e-5 bar of FP64),
and the autotuner leaves the context on candidate values."""
import numpy as np
import pytest

from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _batch(ctx, h, layers, precision):
    import torch
    from clblm_amd.odlay import Context
    pars = [Context.params(600.0, 700.0, W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot), precision=precision)
            for l in layers]
    ctx.jrad = 1
    out, off = ctx.od_batch(h, layers, pars)
    torch.cuda.synchronize()
    return out.cpu().numpy(), pars


def test_every_launch_shape_gives_the_same_od(ctx):
    from clblm_amd import _lib
    lines = W.profile_lines(30000, 500.0, 800.0, seed=1001)
    layers = W.atmosphere()[::10]
    h = ctx.upload_lines(lines)
    ref64, _ = _batch(ctx, h, layers, _lib.FP64)
    for key, vals, prec in ((_lib.TUNE_F4_POINTS_FP64, (2, 3, 4), _lib.FP64),
                            (_lib.TUNE_CONV_UNROLL, (2, 4, 8), _lib.FP64),
                            (_lib.TUNE_F4_XCD_CHUNK, (1, 5, 64), _lib.FP64),
                            (_lib.TUNE_F4_STREAM, (1, 2), _lib.FP64),
                            (_lib.TUNE_F4_POINTS_FP32C, (2, 3, 4), _lib.FP32C),
                            (_lib.TUNE_F4_FAR_FIELD, (1, 2, 3), _lib.FP32C)):
        old = ctx.tuning(key)
        for v in vals:
            ctx.tuning(key, v)
            od, _ = _batch(ctx, h, layers, prec)
            tol = 1e-13 if prec == _lib.FP64 else 1e-5
            err = float((np.abs(od - ref64) / np.abs(ref64)).max())
            print("key %d value %d: max rel diff %.3e" % (key, v, err))
            assert err <= tol, (key, v, err)
        ctx.tuning(key, old)


def test_autotune_picks_candidates(ctx):
    from clblm_amd import _lib
    lines = W.profile_lines(30000, 500.0, 800.0, seed=1001)
    layers = W.atmosphere()[::6]
    h = ctx.upload_lines(lines)
    ref, pars = _batch(ctx, h, layers, _lib.FP32C)
    chosen, ms = ctx.autotune(h, layers, pars, reps=2)
    print("autotune: %r, %.3f ms per batch" % (chosen, ms))
    assert ms > 0.0
    assert chosen[_lib.TUNE_F4_POINTS_FP32C] in (2, 3, 4)
    assert chosen[_lib.TUNE_CONV_UNROLL] in (2, 4, 8)
    assert chosen[_lib.TUNE_F4_XCD_CHUNK] in (4, 16, 64)
    assert chosen[_lib.TUNE_F4_STREAM] in (1, 2)
    od, _ = _batch(ctx, h, layers, _lib.FP32C)
    assert float((np.abs(od - ref) / np.abs(ref)).max()) <= 1e-6
