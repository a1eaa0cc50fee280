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


def _d(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _rt_jac(ctx, od, t_lay, t_lev, up, lo, v1, dv, lin, sec, emis=0.98, tskin=295.0):
    import torch
    from clblm_amd._lib import check
    L, n = od.shape
    e = lambda *sh: torch.empty(*sh, dtype=torch.float64, device="cuda")
    outs = [e(n), e(L, n), e(L + 1, n), e(n), e(n)]
    ins = [_d(od), _d(t_lay), _d(t_lev), _d(up), _d(lo)]
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    check(ctx.lib.clblm_rt_jac(ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(dv), *[p(t) for t in ins],
                               ctypes.c_double(emis), ctypes.c_double(tskin), lin, ctypes.c_double(sec),
                               *[p(t) for t in outs], None))
    return [t.cpu().numpy() for t in outs]


@pytest.mark.parametrize("lin,sec,opaque", [(0, 1.0, False), (0, 1.66, False), (1, 1.0, False), (1, 1.66, False),
                                            (2, 1.5, False), (1, 1.66, True), (2, 1.0, True)])
def test_rt_jac_matches_oracle(ctx, lin, sec, opaque):
    """clblm_rt_jac against the oracle's Jacobians restatement, the reference's own
    formulation (radiance, dR/dtau, dR/dT_lev, surface terms); the opaque case takes the
    branch without the reflected downwelling."""
    from oracle.py import jacobians
    od, t_lay, t_lev = _case(seed=31, nlay=15, n=3000)
    if opaque:
        od = od + 2.0
    rng = np.random.default_rng(32)
    up, lo = rng.uniform(0.3, 0.7, 15), rng.uniform(0.3, 0.7, 15)
    got = _rt_jac(ctx, od, t_lay, t_lev, up, lo, 620.0, 0.013, lin, sec)
    ref = jacobians(od, t_lay, t_lev, 620.0, 0.013, emis=0.98, tskin=295.0, lin=lin, sec_dn=sec, dt_dtup=up,
                    dt_dtlo=lo)
    for name, a, b in zip(("rad", "dtau", "dtlev", "dtskin", "demis"), got, ref):
        err = np.abs(a - b).max() / np.abs(b).max()
        print("%s max err / max = %.2e" % (name, err))
        assert err < 1e-12, name


def test_level_jac_matches_oracle(ctx):
    import torch
    from clblm_amd._lib import check
    from oracle.py import level_jac
    rng = np.random.default_rng(41)
    L, M, n = 11, 3, 2500
    od_var = 10 ** rng.uniform(-4, 0.5, (3 + M, L, n))
    dtau, dtlev = rng.normal(0, 1e-6, (L, n)), rng.normal(0, 1e-8, (L + 1, n))
    wtot, tave = rng.uniform(1e21, 1e24, L), rng.uniform(200, 300, L)
    up, lo = rng.uniform(0.3, 0.7, L), rng.uniform(0.3, 0.7, L)
    wu, wl = rng.uniform(0.2, 0.8, (M, L)), rng.uniform(0.2, 0.8, (M, L))
    jt, jq = (torch.empty(L + 1, n, dtype=torch.float64, device="cuda"),
              torch.empty(M, L + 1, n, dtype=torch.float64, device="cuda"))
    ins = [_d(a) for a in (od_var, dtau, dtlev, wtot, tave)]
    fac = [_d(a) for a in (up, lo, wu.reshape(-1), wl.reshape(-1))]
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    check(ctx.lib.clblm_level_jac(ctx.ptr, L, M, n, *[p(t) for t in ins], ctypes.c_double(1.0),
                                  ctypes.c_double(0.01), *[p(t) for t in fac], p(jt), p(jq), None))
    torch.cuda.synchronize()
    rt, rq = level_jac(od_var, dtau, dtlev, wtot, tave, 1.0, 0.01, up, lo, wu, wl)
    np.testing.assert_allclose(jt.cpu().numpy(), rt, rtol=1e-13, atol=1e-13 * np.abs(rt).max())
    np.testing.assert_allclose(jq.cpu().numpy(), rq, rtol=1e-13, atol=1e-13 * np.abs(rq).max())


def _scene(nlay=12, n_lines=3000):
    # a thin line list and a non-isothermal profile, so that every layer shows through
    lines = W.profile_lines(n_lines, 500.0, 800.0, seed=1001)
    lines.s[:] = lines.s * np.float32(1e-3)
    noise = np.random.default_rng(5).normal(0.0, 3.0, nlay)
    layers = W.atmosphere(nlay, 1.0, t_noise=noise)
    t_lev = np.array([layers[0].tave + 2.0] + [0.5 * (a.tave + b.tave) for a, b in zip(layers[:-1], layers[1:])] +
                     [layers[-1].tave - 1.0])
    p_lev = W.level_pressures(nlay, 1.0)
    q_lev = W.level_densities(nlay, 1.0)
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    return lines, layers, t_lev, p_lev, q_lev, dvs


def _jac_setup(ctx, prec, nlay=12, mols=None, boxcar_hw=-1.0):
    from clblm_amd.jacobian import UpwellingJacobian
    lines, layers, t_lev, p_lev, q_lev, dvs = _scene(nlay)
    h = ctx.upload_lines(lines)
    return UpwellingJacobian(ctx, h, layers, t_lev, dvs, p_lev=p_lev, q_lev=q_lev, mols=mols, precision=prec,
                             continuum=lambda lay, a, b: W.continuum_absrb(lay, a, b), boxcar_hw=boxcar_hw)


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


def test_level_jacobians_match_the_oracle_pipeline(ctx):
    """The whole noScattJacob block in FP64 against the oracle: every layer-OD variant
    from the oracle's ODLAY, then its Jacobians and drv_LevelTemp/drv_LevelConc.  The
    layer ODs agree to 1e-10, so differences of ODs (T + 1 K, Wtot + 1 %) agree to
    about 1e-10 / (relative size of the difference)."""
    from clblm_amd._lib import FP64
    from oracle.py import Oracle, jacobians, level_jac
    jb = _jac_setup(ctx, FP64, nlay=8, mols=[1, 2, 3])
    v1, v2 = 640.0, 646.0
    got = jb.block(v1, v2)
    lines, layers, t_lev, p_lev, q_lev, dvs = _scene(8)
    orc = Oracle(lines)
    orc.jrad = 1
    jobs, pars = jb._jobs(v1, v2)
    L, n = len(layers), got["n"]
    od_var = []
    for k, (lay, par) in enumerate(zip(jobs, pars)):
        c = W.continuum_absrb(lay, v1, v2)[0] if (k < 3 * L or jb.mols[(k - 3 * L) // L] == 1) else None
        od, _ = orc.odlay(lay, v1, v2, dvs[k % L], dvout=jb.dvout, continuum=c)
        assert od.size == n
        od_var.append(od)
    od_var = np.array(od_var).reshape(3 + len(jb.mols), L, n)
    np.testing.assert_allclose(got["od_var"].cpu().numpy(), od_var, rtol=1e-10, atol=1e-10 * np.abs(od_var).max())
    t_lay = np.array([l.tave for l in layers])
    rad, dtau, dtlev, dts, de = jacobians(od_var[0], t_lay, t_lev, v1, jb.dvout, emis=jb.emis, tskin=jb.tskin,
                                          lin=jb.lin, sec_dn=jb.sec_dn, dt_dtup=jb.dt_dtup, dt_dtlo=jb.dt_dtlo)
    jt, jq = level_jac(od_var, dtau, dtlev, [l.wtot for l in layers], t_lay, jb.dt, jb.dw_frac, jb.dt_dtup,
                       jb.dt_dtlo, jb.dw_dqup, jb.dw_dqlo)
    for name, a, b, tol in (("rad", got["rad"], rad, 1e-11), ("dr_dtau", got["dr_dtau"], dtau, 1e-9),
                            ("jac_t", got["jac_t"], jt, 1e-7), ("jac_lnq", got["jac_lnq"], jq, 1e-9),
                            ("dr_dtskin", got["dr_dtskin"], dts, 1e-11), ("dr_demis", got["dr_demis"], de, 1e-11)):
        a = a.cpu().numpy()
        err = np.abs(a - b).max() / np.abs(b).max()
        print("%s max err / max = %.2e" % (name, err))
        assert err < tol, (name, err)


@pytest.mark.parametrize("hw_in_dv", [6.3, 40.0])
def test_jacobian_boxcar_pre_average(ctx, hw_in_dv):
    """doPreBox (clblm_noScattJac.f90:187-189, output_JacArray :1341-1409): with a boxcar
    wider than DVout every output array of a block is pre-averaged by the rectangle
    scan.  The pre-averaged block equals the oracle's rectangle scan (itself
    bit-identical to the reference's clblm_SCANFN) applied to the same block's plain
    outputs, bit for bit (the plain block is computed again, FP64)."""
    from clblm_amd._lib import FP64
    from oracle.py import prebox
    plain = _jac_setup(ctx, FP64, nlay=6, mols=[1, 2])
    hw = hw_in_dv * plain.dvout
    boxed = _jac_setup(ctx, FP64, nlay=6, mols=[1, 2], boxcar_hw=hw)
    assert boxed.prebox
    v1, v2 = 640.0, 652.0
    a = plain.block(v1, v2)
    b = boxed.block(v1, v2)
    for name in ("rad", "jac_t", "jac_lnq", "dr_dtskin", "dr_demis"):
        x = a[name].cpu().numpy()
        rows = x.reshape(-1, x.shape[-1])
        want = [prebox(r, v1, plain.dvout, hw) for r in rows]
        g1, gd = b["grid"]
        assert (g1, gd) == (want[0][1], want[0][2]), name
        got = b[name].cpu().numpy().reshape(rows.shape[0], -1)
        # the plain outputs of the two runs agree to the pass-1 atomic reordering
        np.testing.assert_allclose(got, np.array([w[0] for w in want]), rtol=1e-12,
                                   atol=1e-12 * np.abs(rows).max(), err_msg=name)
    print("boxcar %.4g cm-1: %d -> %d points per block" % (hw, a["rad"].numel(), b["rad"].numel()))


def _perturbed_block(ctx, k, dt_lev=0.0, q_fac=None, mol=None, v1=640.0, v2=646.0, nlay=8):
    """The FP64 radiance of _scene with level k's temperature moved by dt_lev (its
    layers' mean temperatures by dT/dT_lev and their broadening columns by
    -Wtot dT/T, the perturbation drv_LevelTemp linearises, clblm_noScattJac.f90:
    1005-1034) or molecule mol's level density scaled by q_fac (its layers' columns by
    q_fac ** dlnW/dlnq, drv_LevelConc :1159-1171)."""
    import copy
    from clblm_amd._lib import FP64
    from clblm_amd.jacobian import UpwellingJacobian
    lines, layers, t_lev, p_lev, q_lev, dvs = _scene(nlay)
    base = _jac_setup(ctx, FP64, nlay=nlay, mols=[mol] if mol else [1])
    layers = [copy.deepcopy(l) for l in layers]
    t_lev = t_lev.copy()
    q_lev = {m: v.copy() for m, v in q_lev.items()}
    mi = base.mols.index(mol) if mol else 0
    for l in (k - 1, k):
        if l < 0 or l >= nlay:
            continue
        upper = l == k - 1          # level k is the upper boundary of layer k - 1
        if dt_lev:
            d = (base.dt_dtup[l] if upper else base.dt_dtlo[l]) * dt_lev
            layers[l].wtot = layers[l].wtot * (1.0 - d / layers[l].tave)
            layers[l].tave = layers[l].tave + d
        if q_fac is not None:
            e = base.dw_dqup[mi, l] if upper else base.dw_dqlo[mi, l]
            layers[l].w[mol] = layers[l].w[mol] * q_fac ** e
    t_lev[k] += dt_lev
    if q_fac is not None:
        q_lev[mol][k] *= q_fac
    h = ctx.upload_lines(lines)
    jb = UpwellingJacobian(ctx, h, layers, t_lev, dvs, p_lev=p_lev, q_lev=q_lev, mols=[mol] if mol else [1],
                           precision=FP64, continuum=lambda lay, a, b: W.continuum_absrb(lay, a, b))
    r = jb.block(v1, v2)["rad"].cpu().numpy()
    ctx.free_lines(h)
    return r


@pytest.mark.parametrize("k", [0, 3, 8])
def test_jacobian_t_vs_pipeline_central_difference(ctx, k):
    """dR/dT_lev (the reference's dtau/dT: a 1 K one-sided difference of the layer OD,
    drv_LevelTemp) against the central difference (h = 0.5 K) of the whole OD + radiance
    pipeline: within 0.15 of the largest value (measured: 0.09 at the surface level,
    where the signal is smallest, 0.02 in the middle, 0.002 at the top).  The one-sided
    difference's O(dT) curvature term bounds the agreement; a wrong level factor, sign
    or level assignment gives errors of order 1.  Steps much below 1 K are dominated by
    the OD's table indices stepping (each step a jump of a table interval) in the line
    wings, in both differences."""
    from clblm_amd._lib import FP64
    jb = _jac_setup(ctx, FP64, nlay=8, mols=[1])
    jac = jb.block(640.0, 646.0)["jac_t"][k].cpu().numpy()
    h = 0.5
    fd = (_perturbed_block(ctx, k, dt_lev=h) - _perturbed_block(ctx, k, dt_lev=-h)) / (2 * h)
    err = np.abs(jac - fd).max() / np.abs(fd).max()
    print("level %d: max |jac - fd| / max |fd| = %.2e" % (k, err))
    assert err < 0.15


@pytest.mark.parametrize("mol,k", [(2, 2), (3, 6), (1, 4)])
def test_jacobian_lnq_vs_pipeline_central_difference(ctx, mol, k):
    """dR/dln q_lev against the central difference (dln q = 1e-4) of the whole pipeline,
    within 1e-2 of the largest value (measured: 4e-3 for CO2, 2e-5 for O3 and H2O).  The
    Jacobian takes the OD of the molecule alone (drv_LevelConc) as its derivative; the
    true OD is not exactly linear in the column (self-broadening, and the F4 wings'
    1000-line buffers, which change when the other molecules' lines drop out)."""
    from clblm_amd._lib import FP64
    jb = _jac_setup(ctx, FP64, nlay=8, mols=[mol])
    jac = jb.block(640.0, 646.0)["jac_lnq"][0][k].cpu().numpy()
    e = 1e-4
    fd = (_perturbed_block(ctx, k, q_fac=np.exp(e), mol=mol) -
          _perturbed_block(ctx, k, q_fac=np.exp(-e), mol=mol)) / (2 * e)
    err = np.abs(jac - fd).max() / np.abs(fd).max()
    print("molecule %d level %d: max |jac - fd| / max |fd| = %.2e" % (mol, k, err))
    assert err < 1e-2
