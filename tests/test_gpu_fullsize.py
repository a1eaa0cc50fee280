"""Full-size cases, against the oracle and through size-independent properties.

* The full-profile workload (60 layers, 200 000 lines of 7 species, 500-2500 cm-1,
  spline continuum): FP64 against the oracle within 1e-10 on every layer at the full
  range, FP32C against FP64 within the 1e-5 bar at every point, and the FP64 optical
  depth linear in the absorber columns (all columns and the broadening column doubled
  give twice the line optical depth).
* The large scene (2 000 000 lines over 10-3500 cm-1, 80 layers to 0.01 hPa): FP32C
  against FP64 for every layer, in batches of 8 layers, and FP64 against the oracle at
  the full range on sampled layers (bottom, middle, top).
"""
import numpy as np
import pytest

from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu
V1, V2 = 500.0, 2500.0


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _run(ctx, h, layers, dvs, precision, v1, v2, conts=None):
    import torch
    from clblm_amd.odlay import Context
    pars = [Context.params(v1, v2, dv, precision=precision, continuum=None if conts is None else conts[k])
            for k, dv in enumerate(dvs)]
    ctx.jrad = 1
    out, off = ctx.od_batch(h, layers, pars)
    torch.cuda.synchronize()
    ctx.sync()
    return out.cpu().numpy(), off


def _rel(a, b, floor):
    """max |a - b| / max(|b|, floor * max|b|) per layer segment"""
    den = np.maximum(np.abs(b), floor * np.abs(b).max())
    return float((np.abs(a - b) / den).max())


def test_full_profile_fp32c_within_1e5_of_fp64(ctx):
    from clblm_amd._lib import FP32C, FP64
    lines = W.profile_lines(200000, V1, V2, seed=1001)
    layers = W.atmosphere(60, 0.1)
    dvs = [W.calculate_dv(0.5 * (V1 + V2), l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    conts = [W.continuum_absrb(l, V1, V2)[0] for l in layers]
    h = ctx.upload_lines(lines)
    a, off = _run(ctx, h, layers, dvs, FP64, V1, V2, conts)
    b, _ = _run(ctx, h, layers, dvs, FP32C, V1, V2, conts)
    worst = max(float((np.abs(b[off[k]:off[k + 1]] - a[off[k]:off[k + 1]]) / np.abs(a[off[k]:off[k + 1]])).max())
                for k in range(len(layers)))
    print("full profile: %d points, FP32C max rel err vs FP64 %.3e" % (off[-1], worst))
    assert worst <= 1e-5


def test_full_profile_fp64_linear_in_columns(ctx):
    from clblm_amd._lib import FP64
    lines = W.profile_lines(200000, V1, V2, seed=1001)
    layers = W.atmosphere(60, 0.1)[::6]
    dvs = [W.calculate_dv(0.5 * (V1 + V2), l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    doubled = [W.Layer(pave=l.pave, tave=l.tave, wtot=2.0 * l.wtot, w={m: 2.0 * v for m, v in l.w.items()})
               for l in layers]
    h = ctx.upload_lines(lines)
    a, off = _run(ctx, h, layers, dvs, FP64, V1, V2)
    b, _ = _run(ctx, h, doubled, dvs, FP64, V1, V2)
    err = _rel(b, 2.0 * a, 1e-12)
    print("full profile, columns doubled: %d points, max rel deviation from 2x %.3e" % (off[-1], err))
    assert err <= 1e-12


def test_large_scene_fp32c_within_1e5_of_fp64(ctx):
    from clblm_amd._lib import FP32C, FP64
    lines = W.large_scene_lines()
    layers = W.large_scene_atmosphere()
    h = ctx.upload_lines(lines)
    worst, npts = 0.0, 0
    for k0 in range(0, len(layers), 8):
        sub = layers[k0:k0 + 8]
        dvs = [W.large_scene_dv(l) for l in sub]
        a, off = _run(ctx, h, sub, dvs, FP64, W.LARGE_V1, W.LARGE_V2)
        b, _ = _run(ctx, h, sub, dvs, FP32C, W.LARGE_V1, W.LARGE_V2)
        for k in range(len(sub)):
            e = _rel(b[off[k]:off[k + 1]], a[off[k]:off[k + 1]], 1e-4)
            worst = max(worst, e)
        npts += int(off[-1])
        print("large scene layers %d-%d: %d points, worst so far %.3e" % (k0, k0 + len(sub) - 1, off[-1], worst),
              flush=True)
    print("large scene: %d lines, %d points, FP32C max rel err vs FP64 %.3e" % (lines.n, npts, worst))
    assert worst <= 1e-5


def test_full_profile_fp64_matches_oracle_every_layer(ctx):
    """All 60 layers of the full-profile workload, 500-2500 cm-1, the oracle's layers in
    worker processes (about a minute of CPU)."""
    from clblm_amd._lib import FP64
    from tests_helpers_pool import oracle_layers
    lines = W.profile_lines(200000, V1, V2, seed=1001)
    layers = W.atmosphere(60, 0.1)
    dvs = [W.calculate_dv(0.5 * (V1 + V2), l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    conts = [W.continuum_absrb(l, V1, V2)[0] for l in layers]
    refs = oracle_layers("profile", [(l, V1, V2, dv, c) for l, dv, c in zip(layers, dvs, conts)])
    h = ctx.upload_lines(lines)
    a, off = _run(ctx, h, layers, dvs, FP64, V1, V2, conts)
    worst = 0.0
    for k, ref in enumerate(refs):
        od = a[off[k]:off[k + 1]]
        assert od.size == ref.size, k
        err = float((np.abs(od - ref) / np.maximum(np.abs(ref), 1e-300)).max())
        worst = max(worst, err)
        assert err <= 1e-10, (k, err)
    print("full profile FP64 vs oracle: %d points, max rel err %.3e" % (off[-1], worst))
    ctx.free_lines(h)


def test_large_scene_fp64_matches_oracle_sampled_layers(ctx):
    from clblm_amd._lib import FP64
    from tests_helpers_pool import oracle_layers
    lines = W.large_scene_lines()
    layers = W.large_scene_atmosphere()
    pick = [layers[k] for k in (0, 40, 79)]
    dvs = [W.large_scene_dv(l) for l in pick]
    refs = oracle_layers("large", [(l, W.LARGE_V1, W.LARGE_V2, dv, None) for l, dv in zip(pick, dvs)], workers=3)
    h = ctx.upload_lines(lines)
    a, off = _run(ctx, h, pick, dvs, FP64, W.LARGE_V1, W.LARGE_V2)
    for k, ref in enumerate(refs):
        od = a[off[k]:off[k + 1]]
        assert od.size == ref.size, k
        rel = np.abs(od - ref) / np.maximum(np.abs(ref), 1e-300)
        i = int(np.argmax(rel))
        print("large scene layer %d: %d points, FP64 max rel err vs oracle %.3e at OD %.3e (max OD %.3e); "
              "max abs err / max OD %.3e; rel err where OD >= 1e-6 max: %.3e"
              % (k, od.size, rel[i], ref[i], np.abs(ref).max(), np.abs(od - ref).max() / np.abs(ref).max(),
                 rel[np.abs(ref) >= 1e-6 * np.abs(ref).max()].max()))
        assert _rel(od, ref, 1e-6) <= 1e-10, k
    ctx.free_lines(h)
