"""Line coupling on the GPU (IFLG 1: Y and G; IFLG 3: reduced widths) against the
oracle: a list of 7 species in which every fifth line is coupled and every
thirteenth has reduced widths, each with its own coefficients at 200, 250, 296 and
340 K, over layers whose temperatures fall in every bracket of the interpolation.
FP64 within 1e-10 of the oracle (relative, floor 1e-4 of the layer's peak), FP32C
within 1e-5 of FP64."""
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


def coupled_list(seed=11):
    lines = W.make_lines(20000, 560.0, 740.0, seed=seed, mols=(1, 2, 3, 4, 5, 6, 7))
    rng = np.random.default_rng(seed)
    idx1 = np.arange(2, lines.n, 5)
    idx3 = np.array([i for i in range(3, lines.n, 13) if i % 5 != 2])
    idx = np.concatenate([idx1, idx3])
    order = np.argsort(idx)
    idx = idx[order]
    flags = np.concatenate([np.full(idx1.size, 1), np.full(idx3.size, 3)])[order]
    a = np.where(flags[:, None] == 1, rng.uniform(-0.05, 0.05, (idx.size, 4)), rng.uniform(0.0, 0.1, (idx.size, 4)))
    b = np.where(flags[:, None] == 1, rng.uniform(-0.01, 0.01, (idx.size, 4)), rng.uniform(0.0, 0.05, (idx.size, 4)))
    return W.couple_lines(lines, idx, flags, a, b)


def _layers():
    atm = W.atmosphere()
    lays = [atm[k] for k in (0, 3, 8, 14, 20, 30, 45, 59)]
    # temperatures in each interpolation bracket and at the nodes
    for t in (195.0, 250.0, 296.0, 345.0):
        l = atm[5]
        lays.append(W.Layer(pave=l.pave, tave=t, wtot=l.wtot, w=dict(l.w)))
    return lays


def test_coupled_lines_match_oracle(ctx):
    import torch
    from clblm_amd.odlay import Context
    from oracle.py import Oracle
    lines = coupled_list()
    print("lines %d, IFLG 1: %d, IFLG 3: %d" % (lines.n, int((lines.iflg == 1).sum()), int((lines.iflg == 3).sum())))
    h = ctx.upload_lines(lines)
    layers = _layers()
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    res = {}
    for prec in (0, 1):
        ctx.jrad = 1
        out, off = ctx.od_batch(h, layers, [Context.params(600.0, 700.0, dv, precision=prec) for dv in dvs])
        torch.cuda.synchronize()
        ctx.sync()
        res[prec] = (out.cpu().numpy(), off)
    od64, off = res[0]
    od32, _ = res[1]
    orc = Oracle(lines)
    orc.jrad = 1
    worst, worst32 = 0.0, 0.0
    for k, (lay, dv) in enumerate(zip(layers, dvs)):
        ref, _ = orc.odlay(lay, 600.0, 700.0, dv)
        got = od64[off[k]:off[k + 1]]
        assert ref.size == got.size
        fl = 1e-4 * np.abs(ref).max()
        worst = max(worst, float((np.abs(got - ref) / np.maximum(np.abs(ref), fl)).max()))
        g32 = od32[off[k]:off[k + 1]]
        worst32 = max(worst32, float((np.abs(g32 - got) / np.maximum(np.abs(got), fl)).max()))
    assert ctx.jrad == orc.jrad
    print("coupled lines: FP64 vs oracle %.3e, FP32C vs FP64 %.3e" % (worst, worst32))
    assert worst <= 1e-10
    assert worst32 <= 1e-5
    ctx.free_lines(h)


def test_upload_checks_coefficient_records(ctx):
    from clblm_amd._lib import ClblmError
    lines = W.make_lines(600, 600.0, 700.0, seed=2)
    bad = W.couple_lines(lines, [10], 1, [0.0] * 4, [0.0] * 4)
    bad.iflg[11] = 0        # the coefficient record of line 10 is gone
    with pytest.raises(ClblmError):
        ctx.upload_lines(bad)
    ok = W.couple_lines(lines, [249, 300], 1, [0.0] * 4, [0.0] * 4)   # 249: needs a filler first
    assert ok.iflg[249] == -1 and ok.iflg[250] == 1 and ok.iflg[251] == -1
    h = ctx.upload_lines(ok)
    ctx.free_lines(h)
