"""A batch captured in a hipGraph (clblm_od_graph_*): every replay gives the optical
depth of the eager batch (FP64 within the LDS-atomic reordering of pass 1, FP32C
within its 1e-5 bar), also with other eager batches, torch work and host
synchronisations between the replays, on another stream, and with the F4 wing chain
on its own stream or on the batch's stream."""
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


def _pars(layers, precision, v1=600.0, v2=700.0):
    from clblm_amd.odlay import Context
    return [Context.params(v1, v2, W.calculate_dv(0.5 * (v1 + v2), l.pave, l.tave, l.w[1] / l.wtot),
                           precision=precision) for l in layers]


def _rel(a, b):
    return float((np.abs(a - b) / np.maximum(np.abs(b), 1e-300)).max())


@pytest.mark.parametrize("precision,stream_mode", [(0, 1), (0, 2), (1, 1)])
def test_graph_replays_match_eager(ctx, precision, stream_mode):
    import torch
    from clblm_amd import _lib
    lines = W.profile_lines(30000, 500.0, 800.0, seed=1001)
    h = ctx.upload_lines(lines)
    layers = W.atmosphere()[::8]
    pars = _pars(layers, precision)
    old = ctx.tuning(_lib.TUNE_F4_STREAM, stream_mode)
    try:
        ctx.jrad = 1
        ref, off = ctx.od_batch(h, layers, pars)
        torch.cuda.synchronize()
        ref = ref.cpu().numpy()
        ctx.jrad = 1
        g = ctx.od_graph(h, layers, pars)
        assert np.array_equal(g.offsets, off)
        # other work of the program between replays: another eager batch, torch ops
        other_layers = W.atmosphere()[3::9]
        other_pars = _pars(other_layers, precision, 640.0, 720.0)
        tol = 1e-13 if precision == _lib.FP64 else 1e-5
        side = torch.cuda.Stream()
        for rep in range(4):
            g.out.fill_(float("nan"))
            ctx.jrad = 1
            junk, _ = ctx.od_batch(h, other_layers, other_pars)
            junk.mul_(2.0)
            torch.cuda.synchronize()
            if rep == 2:
                with torch.cuda.stream(side):
                    g.launch()
                side.synchronize()
            else:
                g.launch()
                torch.cuda.synchronize()
            g.check()
            od = g.out[: int(off[-1])].cpu().numpy()
            assert np.isfinite(od).all(), rep
            err = _rel(od, ref)
            print("precision %d stream mode %d replay %d: max rel diff %.3e" % (precision, stream_mode, rep, err))
            assert err <= tol, (rep, err)
        # back-to-back replays without a host wait end in the same state
        for _ in range(3):
            g.launch()
        g.check()
        assert _rel(g.out[: int(off[-1])].cpu().numpy(), ref) <= tol
        g.close()
        ctx.sync()
    finally:
        ctx.tuning(_lib.TUNE_F4_STREAM, old)
        ctx.free_lines(h)

