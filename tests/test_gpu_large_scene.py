"""The large scene end to end: 2 000 000 lines over 10-3500 cm-1, 80 layers to 0.01 hPa.

The whole noScattJacob pipeline over the full range in spectral blocks (the layer ODs
recomputed per block instead of stored: each block's (3 + 7) x 80 OD variants, the
radiance, dR/dT_lev and dR/dln q_lev of every level and molecule), then the FFT scan of
the radiance with the sinc-apodized instrument function (0.5 cm maximum path
difference), in FP32C against FP64: radiance and scan within 1e-5 relative, the
Jacobians within 1e-4 of each quantity's maximum over the scene.  The comparisons are
reduced on the GPU block by block.  The scene's 2 000 000 lines leave no window: every
molecule's dR/dln q stays many orders below the radiance, so its bar is taken against
the larger of its own maximum and 1e-6 of the radiance's (the FP32C/FP64 difference of
a quantity at rounding level compares noise).

Each block's layer DVs come from calculateDV at the block's centre (setDV's rule for a
run over that block).  One DV per layer for the whole 10-3500 cm-1 (the mid-range
rule) leaves the lines below ~500 cm-1 up to 16x narrower than DV: CNVFNQ then clamps
ALFV to DV while CNVF4Q keeps the true width, the two halves of a line disagree between
60 and 131 DV by about 6e-5 of the line's peak, negative, and the strongest lines
(peak OD 1e5 at 0.01-1 hPa) make the layer OD -15 there, in the oracle as on the GPU;
the radiance of such a layer has no meaning.  The blocks' radiances are re-gridded
(interpSpectrum) onto one grid for the scan.
"""
import numpy as np
import pytest

from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu

# block edges: narrow at low wavenumber, where the Doppler widths (and DV) are small
EDGES = [10.0, 25.0, 50.0, 100.0, 175.0, 250.0, 350.0, 500.0] + list(np.arange(600.0, 3500.0 + 1.0, 100.0))
DV_SCAN = 0.004   # the scan grid: 1/250 of the instrument function's 1.2 cm-1 width


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def test_large_scene_pipeline_fp32c_vs_fp64(ctx):
    import torch
    from clblm_amd._lib import FP32C, FP64
    from clblm_amd.jacobian import UpwellingJacobian
    lines = W.large_scene_lines()
    layers = W.large_scene_atmosphere()
    nlay = len(layers)
    h = ctx.upload_lines(lines)
    worst = {"rad": 0.0, "jac_t": 0.0, "jac_lnq": np.zeros(7)}
    radmax = 0.0
    peak = {"jac_t": 0.0, "jac_lnq": np.zeros(7)}
    scan_in = {FP64: [], FP32C: []}
    npts = 0
    edges = [W.LARGE_V1 + round((e - W.LARGE_V1) / DV_SCAN) * DV_SCAN for e in EDGES]
    for a, b in zip(edges[:-1], edges[1:]):
        dvs = [W.calculate_dv(0.5 * (a + b), l.pave, l.tave, l.w.get(1, 0.0) / l.wtot) for l in layers]
        res = {}
        for prec in (FP64, FP32C):
            jb = UpwellingJacobian(ctx, h, layers, W.level_temperatures(nlay, 0.01), dvs,
                                   p_lev=W.level_pressures(nlay, 0.01), q_lev=W.level_densities(nlay, 0.01),
                                   precision=prec, continuum=W.continuum_absrb)
            res[prec] = jb.block(a, b)
            # onto the scan grid: the block's points a + k DV_SCAN, k < its share
            ns = int(round((b - a) / DV_SCAN))
            scan_in[prec].append(ctx.interp_spectrum(res[prec]["rad"], a, jb.dvout, DV_SCAN, ns))
        r64, r32 = res[FP64], res[FP32C]
        worst["rad"] = max(worst["rad"], float(((r32["rad"] - r64["rad"]).abs() / r64["rad"].abs()).max()))
        worst["jac_t"] = max(worst["jac_t"], float((r32["jac_t"] - r64["jac_t"]).abs().max()))
        peak["jac_t"] = max(peak["jac_t"], float(r64["jac_t"].abs().max()))
        worst["jac_lnq"] = np.maximum(worst["jac_lnq"],
                                      (r32["jac_lnq"] - r64["jac_lnq"]).abs().amax(dim=(1, 2)).cpu().numpy())
        peak["jac_lnq"] = np.maximum(peak["jac_lnq"], r64["jac_lnq"].abs().amax(dim=(1, 2)).cpu().numpy())
        radmax = max(radmax, float(r64["rad"].abs().max()))
        npts += r64["n"]
        print("block %.3f-%.3f: %d points (DVout %.3g); worst so far: radiance %.2e" %
              (a, b, r64["n"], jb.dvout, worst["rad"]), flush=True)
        del res, r64, r32
    rt = worst["jac_t"] / peak["jac_t"]
    rq = worst["jac_lnq"] / np.maximum(peak["jac_lnq"], 1e-6 * radmax)
    print("large scene: %d points per spectrum, %d levels; FP32C vs FP64: radiance %.2e, dR/dT_lev %.2e of its max, "
          "dR/dln q %s (largest dR/dln q over the largest radiance: %s)" %
          (npts, nlay + 1, worst["rad"], rt, ["%.1e" % x for x in rq], ["%.1e" % x for x in peak["jac_lnq"] / radmax]))
    assert worst["rad"] < 1e-5
    assert rt < 1e-4
    assert np.all(rq < 1e-4)
    x64, x32 = torch.cat(scan_in[FP64]), torch.cat(scan_in[FP32C])
    s64, v1s, dvs_ = ctx.fft_scan(x64, edges[0], DV_SCAN, 20.0, 3490.0, 0.0, funct_id=5, hwhm=0.6034, precision=FP64)
    s32, _, _ = ctx.fft_scan(x32, edges[0], DV_SCAN, 20.0, 3490.0, 0.0, funct_id=5, hwhm=0.6034, precision=FP32C)
    es = float((s32 - s64).abs().max() / s64.abs().max())
    print("sinc scan: %d points in, %d out (v1 %.4f, dv %.4g); FP32C vs FP64 %.2e of the max" %
          (x64.numel(), s64.numel(), v1s, dvs_, es))
    assert es < 1e-5
    ctx.free_lines(h)
