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
