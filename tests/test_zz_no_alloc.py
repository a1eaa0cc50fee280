"""clblm_rt_jac_noalloc (no allocation, static device scratch) against clblm_rt_jac
through the C-ABI: the same kernel on the same inputs, so every output must be equal
bit for bit, on two calls in a row (the second smaller, so it runs on the states the
first left in the shared array) and for each linInTau; a call beyond the scratch
size is refused with CLBLM_ERR_CAPACITY before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = ctypes.c_double


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _case(seed, nlay, n):
    rng = np.random.default_rng(seed)
    od = 10.0 ** rng.uniform(-4.0, 0.7, (nlay, n))
    t_lev = np.linspace(290.0, 215.0, nlay + 1) + rng.normal(0.0, 1.0, nlay + 1)
    t_lay = 0.5 * (t_lev[:-1] + t_lev[1:]) + rng.normal(0.0, 0.5, nlay)
    up, lo = rng.uniform(0.3, 0.7, nlay), rng.uniform(0.3, 0.7, nlay)
    return od, t_lay, t_lev, up, lo


def _call(ctx, fn, od, t_lay, t_lev, up, lo, lin, sec):
    import torch
    L, n = od.shape
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    # outputs start as NaN so that a point left unwritten cannot match
    e = lambda *sh: torch.full(sh, float("nan"), dtype=torch.float64, device="cuda")
    outs = [e(n), e(L, n), e(L + 1, n), e(n), e(n)]
    ins = [dev(od), dev(t_lay), dev(t_lev), dev(up), dev(lo)]
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = fn(ctx.ptr, L, n, D(640.0), D(0.011), *[p(t) for t in ins], D(0.97), D(293.0), lin, D(sec),
            *[p(t) for t in outs], ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, [t.cpu().numpy() for t in outs]


@pytest.mark.parametrize("lin,sec", [(1, 1.66), (2, 1.0), (0, 1.3)])
def test_noalloc_equals_rt_jac(ctx, lin, sec):
    from clblm_amd._lib import check
    for k, (nlay, n) in enumerate(((30, 40000), (12, 5000))):
        case = _case(100 * lin + k, nlay, n)
        print("case: lin %d, %d layers x %d points" % (lin, nlay, n), flush=True)
        rc, want = _call(ctx, ctx.lib.clblm_rt_jac, *case, lin, sec)
        check(rc)
        rc, got = _call(ctx, ctx.lib.clblm_rt_jac_noalloc, *case, lin, sec)
        check(rc)
        for name, a, b in zip(("rad", "dR/dtau", "dR/dT_lev", "dR/dTskin", "dR/demis"), got, want):
            assert not np.isnan(b).any(), name
            assert np.array_equal(a, b), (name, np.nanmax(np.abs(a - b)))


def test_noalloc_refuses_beyond_its_scratch(ctx):
    from clblm_amd import _lib
    cap = 64 << 20
    nlay = 60
    n = cap // (8 * (3 * nlay + 2)) + 1          # one point more than fits
    null = ctypes.c_void_p(0)
    dummy = ctypes.c_void_p(256)                  # never dereferenced: the size check comes first
    rc = ctx.lib.clblm_rt_jac_noalloc(ctx.ptr, nlay, n, D(640.0), D(0.011), dummy, dummy, dummy, dummy, dummy,
                                      D(0.97), D(293.0), 1, D(1.0), dummy, dummy, dummy, null, null, null)
    assert rc == _lib.ERR_CAPACITY, rc
    assert b"scratch" in ctx.lib.clblm_last_error()
