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


def broadened_list(seed=12):
    """Every line with species data: each of the 7 broadeners flagged with
    probability 1/2 (the line's own molecule -1 for one line in 17), widths,
    exponents and shifts drawn per broadener; every fifth line also coupled."""
    lines = W.make_lines(20000, 560.0, 740.0, seed=seed, mols=(1, 2, 3, 4, 5, 6, 7))
    rng = np.random.default_rng(seed)
    n = lines.n
    # O2's air width is rescaled to a foreign width on read, (a_air - 0.21 a_self) / 0.79;
    # with self widths 5x air it turns negative, and the species-weighted width of
    # such a line can reach zero, where the line shape is undefined.  Give O2 lines
    # their air width as self width, as in the line lists.
    o2 = lines.mol % 100 == 7
    lines.a_self[o2] = lines.a_air[o2]
    flg = (rng.uniform(size=(n, 7)) < 0.5).astype(np.int32)
    own = (lines.mol % 100) - 1
    minus = np.arange(n) % 17 == 3
    flg[np.arange(n)[minus], own[minus]] = -1
    lines.brd_flg = flg
    lines.brd_hw = rng.uniform(0.02, 0.12, (n, 7)).astype(np.float32)
    lines.brd_tmp = rng.uniform(0.4, 0.9, (n, 7)).astype(np.float32)
    lines.brd_shft = rng.uniform(-0.01, 0.01, (n, 7)).astype(np.float32)
    idx = np.arange(2, n, 5)
    return W.couple_lines(lines, idx, 1, rng.uniform(-0.05, 0.05, (idx.size, 4)),
                          rng.uniform(-0.01, 0.01, (idx.size, 4)))


@pytest.mark.parametrize("species_broad", [True, False])
def test_species_broadening_matches_oracle(ctx, species_broad):
    import torch
    from clblm_amd.odlay import Context
    from oracle.py import Oracle
    lines = broadened_list()
    h = ctx.upload_lines(lines)
    layers = _layers()
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    res = {}
    for prec in (0, 1):
        ctx.jrad = 1
        pars = [Context.params(600.0, 700.0, dv, precision=prec, species_broad=species_broad) for dv in dvs]
        out, off = ctx.od_batch(h, layers, pars)
        torch.cuda.synchronize()
        ctx.sync()
        res[prec] = (out.cpu().numpy(), off)
    od64, off = res[0]
    od32, _ = res[1]
    orc = Oracle(lines)
    orc.jrad = 1
    worst, worst32 = 0.0, 0.0
    for k, (lay, dv) in enumerate(zip(layers, dvs)):
        ref, _ = orc.odlay(lay, 600.0, 700.0, dv, species_broad=species_broad)
        got = od64[off[k]:off[k + 1]]
        assert ref.size == got.size
        fl = 1e-4 * np.abs(ref).max()
        worst = max(worst, float((np.abs(got - ref) / np.maximum(np.abs(ref), fl)).max()))
        g32 = od32[off[k]:off[k + 1]]
        worst32 = max(worst32, float((np.abs(g32 - got) / np.maximum(np.abs(got), fl)).max()))
    print("species broadening %s: FP64 vs oracle %.3e, FP32C vs FP64 %.3e" % (species_broad, worst, worst32))
    assert worst <= 1e-10
    assert worst32 <= 1e-5
    ctx.free_lines(h)


def test_nonpositive_width_is_reported(ctx):
    """An O2 line whose foreign width reads negative ((a_air - 0.21 a_self) / 0.79 with
    a_self = 5 a_air) and whose species data give it no other width has a Lorentz
    width below zero: the batch reports it instead of reading past the tables."""
    import torch
    from clblm_amd._lib import ClblmError
    from clblm_amd.odlay import Context
    lines = W.make_lines(600, 600.0, 700.0, seed=9, mols=(1, 7))
    n = lines.n
    lines.brd_flg = np.zeros((n, 7), np.int32)
    o2 = np.where(lines.mol % 100 == 7)[0]
    lines.brd_flg[o2, 6] = 1                      # broadened by O2 only, with width 0
    lines.brd_hw = np.zeros((n, 7), np.float32)
    lines.brd_tmp = np.full((n, 7), 0.7, np.float32)
    lines.brd_shft = np.zeros((n, 7), np.float32)
    h = ctx.upload_lines(lines)
    lay = W.Layer(pave=500.0, tave=250.0, wtot=1e24, w={1: 1e22, 7: 2.1e23})
    ctx.jrad = 1
    ctx.od_batch(h, [lay], [Context.params(640.0, 660.0, 0.002, species_broad=True)])
    torch.cuda.synchronize()
    with pytest.raises(ClblmError, match="Lorentz width"):
        ctx.sync()
    ctx.jrad = 1   # without species broadening the same lines are valid
    ctx.od_batch(h, [lay], [Context.params(640.0, 660.0, 0.002, species_broad=False)])
    ctx.sync()
    ctx.free_lines(h)


def test_tape3_file_gives_the_same_od(ctx, tmp_path):
    """A TAPE3 file (short panels, coupling, species data) loaded by the library's
    reader gives the optical depth of the same records uploaded from memory (up to
    the order of pass 1's atomics) and of the oracle."""
    import torch
    from clblm_amd.odlay import Context
    from clblm_amd.tape3 import read_tape3, write_tape3
    lines = broadened_list()
    sizes, a = [], 0
    while a < lines.n:
        k = min(200 if (len(sizes) % 3) else 250, lines.n - a)
        if a + k < lines.n and lines.iflg[a + k - 1] in (1, 3):
            k -= 1
        sizes.append(k)
        a += k
    p = tmp_path / "tape3"
    write_tape3(p, lines, panel_sizes=sizes)
    h_file = ctx.load_tape3(p, 560.0, 740.0)
    h_mem = ctx.upload_lines(read_tape3(p, 560.0, 740.0))
    layers = _layers()[:6]
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    pars = [Context.params(600.0, 700.0, dv, species_broad=True) for dv in dvs]
    outs = []
    for h in (h_file, h_mem):
        ctx.jrad = 1
        out, off = ctx.od_batch(h, layers, pars)
        torch.cuda.synchronize()
        ctx.sync()
        outs.append(out.cpu().numpy())
    # the same records and kernels: equal up to the order of pass 1's LDS atomics
    d = np.abs(outs[0] - outs[1]) / np.maximum(np.abs(outs[1]), 1e-4 * np.abs(outs[1]).max())
    assert float(d.max()) <= 1e-13
    from oracle.py import Oracle
    orc = Oracle(read_tape3(p, 560.0, 740.0))
    orc.jrad = 1
    worst = 0.0
    for k, (lay, dv) in enumerate(zip(layers, dvs)):
        ref, _ = orc.odlay(lay, 600.0, 700.0, dv, species_broad=True)
        got = outs[0][off[k]:off[k + 1]]
        fl = 1e-4 * np.abs(ref).max()
        worst = max(worst, float((np.abs(got - ref) / np.maximum(np.abs(ref), fl)).max()))
    print("TAPE3 file with short panels: FP64 vs oracle %.3e" % worst)
    assert worst <= 1e-10
    ctx.free_lines(h_file)
    ctx.free_lines(h_mem)
