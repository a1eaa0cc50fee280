"""Radiance recursion and its reverse sweep on the GPU: forward vs the oracle, derivatives vs
central differences, Jacobian blocks FP32C vs FP64."""
import ctypes

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


def _case(seed=21, nlay=12, n=2000):
    rng = np.random.default_rng(seed)
    od = 10 ** rng.uniform(-4, 1.5, size=(nlay, n))
    t_lev = np.linspace(290.0, 215.0, nlay + 1) + rng.normal(0, 2, nlay + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1]) + rng.normal(0, 1, nlay)
    return od, t_lay, t_lev


def _run(ctx, od, t_lay, t_lev, v1, dv, lin, sec, grads=True, emis=0.98, tskin=295.0):
    import torch
    L, n = od.shape
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    tod, tl, tv = d(od), d(t_lay), d(t_lev)
    rad = torch.empty(n, dtype=torch.float64, device="cuda")
    outs = [torch.empty(L, n, dtype=torch.float64, device="cuda") for _ in range(2)] if grads else [None, None]
    sfc = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(2)] if grads else [None, None]
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    from clblm_amd._lib import check
    check(ctx.lib.clblm_rt_up(ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(dv), p(tod), p(tl), p(tv),
                              ctypes.c_double(emis), ctypes.c_double(tskin), lin, ctypes.c_double(sec), p(rad),
                              p(outs[0]), p(outs[1]), p(sfc[0]), p(sfc[1]), None))
    res = [rad.cpu().numpy()]
    if grads:
        res += [o.cpu().numpy() for o in outs] + [s.cpu().numpy() for s in sfc]
    return res


@pytest.mark.parametrize("lin,sec", [(0, 1.0), (1, 1.0), (1, 1.67), (2, 1.0)])
def test_radiance_matches_oracle(ctx, lin, sec):
    from oracle.py import lib
    od, t_lay, t_lev = _case()
    v1, dv = 600.0, 0.37
    rad = _run(ctx, od, t_lay, t_lev, v1, dv, lin, sec, grads=False)[0]
    ref = np.zeros(od.shape[1])
    D = ctypes.c_double
    lib().orc_rt_up.argtypes = [ctypes.c_int, ctypes.c_int64, D, D, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, D, D, ctypes.c_int, D, ctypes.c_void_p]
    odc = np.ascontiguousarray(od)
    lib().orc_rt_up(od.shape[0], od.shape[1], v1, dv, odc.ctypes.data, t_lay.ctypes.data, t_lev.ctypes.data,
                    0.98, 295.0, lin, sec, ref.ctypes.data)
    np.testing.assert_allclose(rad, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("lin,sec", [(0, 1.0), (1, 1.0), (1, 1.67), (2, 1.0)])
def test_reverse_sweep_vs_central_differences(ctx, lin, sec):
    od, t_lay, t_lev = _case(n=500)
    v1, dv = 700.0, 1.1
    rad, g_od, g_t, g_ts, g_e = _run(ctx, od, t_lay, t_lev, v1, dv, lin, sec)
    scale = np.abs(rad).max()
    for l in (0, 3, 7, 11):
        h = np.maximum(1e-5 * od[l], 1e-4)
        a, b = od.copy(), od.copy()
        a[l] += h
        b[l] -= h
        fd = (_run(ctx, a, t_lay, t_lev, v1, dv, lin, sec, False)[0] -
              _run(ctx, b, t_lay, t_lev, v1, dv, lin, sec, False)[0]) / (2 * h)
        np.testing.assert_allclose(g_od[l], fd, rtol=1e-5, atol=1e-9 * scale)
        ta, tb = t_lay.copy(), t_lay.copy()
        ta[l] += 1e-3
        tb[l] -= 1e-3
        fd = (_run(ctx, od, ta, t_lev, v1, dv, lin, sec, False)[0] -
              _run(ctx, od, tb, t_lev, v1, dv, lin, sec, False)[0]) / 2e-3
        np.testing.assert_allclose(g_t[l], fd, rtol=1e-5, atol=1e-9 * scale)
    fd = (_run(ctx, od, t_lay, t_lev, v1, dv, lin, sec, False, tskin=295.001)[0] -
          _run(ctx, od, t_lay, t_lev, v1, dv, lin, sec, False, tskin=294.999)[0]) / 2e-3
    np.testing.assert_allclose(g_ts, fd, rtol=1e-5, atol=1e-9 * scale)
    fd = (_run(ctx, od, t_lay, t_lev, v1, dv, lin, sec, False, emis=0.981)[0] -
          _run(ctx, od, t_lay, t_lev, v1, dv, lin, sec, False, emis=0.979)[0]) / 2e-3
    np.testing.assert_allclose(g_e, fd, rtol=1e-5, atol=1e-9 * scale)


def test_autograd_radiance(ctx):
    import torch
    from clblm_amd.jacobian import upwelling_radiance
    od, t_lay, t_lev = _case(n=300)
    tod = torch.tensor(od, dtype=torch.float64, device="cuda", requires_grad=True)
    tl = torch.tensor(t_lay, dtype=torch.float64, device="cuda", requires_grad=True)
    tv = torch.tensor(t_lev, dtype=torch.float64, device="cuda")
    r = upwelling_radiance(ctx, tod, tl, tv, 650.0, 0.5)
    w = torch.linspace(0.5, 1.5, r.numel(), device="cuda", dtype=torch.float64)
    (r * w).sum().backward()
    _, g_od, g_t, _, _ = _run(ctx, od, t_lay, t_lev, 650.0, 0.5, 1, 1.0)
    wn = w.cpu().numpy()
    np.testing.assert_allclose(tod.grad.cpu().numpy(), g_od * wn[None, :], rtol=1e-12)
    np.testing.assert_allclose(tl.grad.cpu().numpy(), (g_t * wn[None, :]).sum(1), rtol=1e-12)


def _jac_setup(ctx, prec, n_lines=3000, nlay=12):
    # a thin line list and a non-isothermal profile, so that every layer shows through
    from clblm_amd.jacobian import UpwellingJacobian
    lines = W.profile_lines(n_lines, 500.0, 800.0, seed=1001)
    lines.s[:] = lines.s * np.float32(1e-3)
    noise = np.random.default_rng(5).normal(0.0, 3.0, nlay)
    layers = W.atmosphere(nlay, 1.0, t_noise=noise)
    t_lev = np.array([layers[0].tave + 2.0] + [0.5 * (a.tave + b.tave) for a, b in zip(layers[:-1], layers[1:])] +
                     [layers[-1].tave - 1.0])
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    h = ctx.upload_lines(lines)
    return UpwellingJacobian(ctx, h, layers, t_lev, dvs, precision=prec, dt=1.0,
                             continuum=lambda lay, a, b: W.continuum_absrb(lay, a, b))


def test_jacobian_fp32c_vs_fp64(ctx):
    from clblm_amd._lib import FP32C, FP64
    a = _jac_setup(ctx, FP64).block(640.0, 660.0)
    b = _jac_setup(ctx, FP32C).block(640.0, 660.0)
    r64, r32 = a["rad"].cpu().numpy(), b["rad"].cpu().numpy()
    assert np.max(np.abs(r32 - r64) / np.abs(r64)) < 1e-5
    for k in ("jac_t", "jac_lnq"):
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        err = np.abs(y - x).max() / np.abs(x).max()
        print(k, "max |FP32C - FP64| / max |FP64| = %.2e" % err)
        assert err < 1e-4, (k, err)


def test_jacobian_t_vs_pipeline_difference(ctx):
    """dR/dT_l from the block equals the central difference of the whole OD+radiance pipeline."""
    import copy
    from clblm_amd._lib import FP64
    jb = _jac_setup(ctx, FP64)
    jb.dt = 0.1
    base = jb.block(640.0, 650.0)
    for l in (0, 5, 11):
        plus, minus = copy.copy(jb), copy.copy(jb)
        plus.layers = [copy.copy(x) for x in jb.layers]
        minus.layers = [copy.copy(x) for x in jb.layers]
        plus.layers[l].tave += 0.1
        minus.layers[l].tave -= 0.1
        rp, rm = plus.block(640.0, 650.0)["rad"], minus.block(640.0, 650.0)["rad"]
        fd = ((rp - rm) / 0.2).cpu().numpy()
        j = base["jac_t"][l].cpu().numpy()
        assert np.abs(j - fd).max() <= 1e-3 * np.abs(fd).max() + 1e-15, l
