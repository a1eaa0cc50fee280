"""Down-welling radiance of a look-up path (mode3_lookUp_mergeUp,
src/clblm_src/clblm_noScattRT.f90:843-1108).  The oracle's restatement (orc_rt_down)
is pinned by closed forms: an isothermal path gives B(T)(1 - prod tx) for every
linInTau (the linear-in-tau source is B itself there); the solar term arrives through
the whole path's transmittance; an opaque layer next to the observer hides every
layer beyond it and emits from its lower (observer-side) level; a thin layer with
linInTau 2 gives the explicit (1 - tx)(B(Tbot) + (B(Ttop) - B(Tbot)) h(tau)) with
Planck and h written out here; and the oracle's layer emission equals the
reference's own layerEmis (oracle/_ref, when built).  The GPU kernel (clblm_rt_down)
matches the oracle."""
import ctypes

import numpy as np
import pytest

from oracle.py import lib

D = ctypes.c_double


def orc_down(od, t_lay, t_lev, v1, dv, lin, sol=None):
    L = lib()
    L.orc_rt_down.argtypes = [ctypes.c_int, ctypes.c_int64, D, D] + [ctypes.c_void_p] * 3 + \
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    od = np.ascontiguousarray(od, np.float64)
    n = od.shape[1]
    rad, tx = np.zeros(n), np.zeros(n)
    s = np.ascontiguousarray(sol, np.float64) if sol is not None else None
    L.orc_rt_down(od.shape[0], n, v1, dv, od.ctypes.data, np.ascontiguousarray(t_lay, np.float64).ctypes.data,
                  np.ascontiguousarray(t_lev, np.float64).ctypes.data, lin, s.ctypes.data if s is not None else None,
                  rad.ctypes.data, tx.ctypes.data)
    return rad, tx


def planck(v, t):
    L = lib()
    L.orc_planck.restype = D
    L.orc_planck.argtypes = [D, D]
    return np.array([L.orc_planck(float(x), float(t)) for x in v])


@pytest.mark.parametrize("lin", [0, 1, 2])
def test_isothermal_path(lin):
    rng = np.random.default_rng(1)
    od = rng.uniform(0.0, 2.0, (12, 500))
    v1, dv = 600.0, 0.5
    t = 250.0
    rad, tx = orc_down(od, np.full(12, t), np.full(13, t), v1, dv, lin)
    v = v1 + dv * np.arange(500)
    assert np.allclose(tx, np.exp(-od.sum(axis=0)), rtol=1e-13)
    assert np.allclose(rad, planck(v, t) * (1 - tx), rtol=1e-12)


def test_solar_through_the_path_and_order():
    rng = np.random.default_rng(2)
    od = rng.uniform(0.0, 1.0, (5, 200))
    t_lev = np.linspace(290.0, 220.0, 6)
    t_lay = 0.5 * (t_lev[:-1] + t_lev[1:])
    sol = rng.uniform(0.0, 1e-5, 200)
    r0, tx = orc_down(od, t_lay, t_lev, 800.0, 0.1, 1)
    r1, _ = orc_down(od, t_lay, t_lev, 800.0, 0.1, 1, sol)
    assert np.allclose(r1 - r0, tx * sol, rtol=1e-10, atol=1e-22)
    # a layer's emission is attenuated only by the layers below it
    od2 = od.copy()
    od2[0] = 0.0
    a, _ = orc_down(od2, t_lay, t_lev, 800.0, 0.1, 1)
    b, _ = orc_down(od2[1:], t_lay[1:], t_lev[1:], 800.0, 0.1, 1)
    assert np.allclose(a, b, rtol=1e-13)


RADCN1, RADCN2 = 1.191042722e-12, 1.4387752     # clblm_ConstParam.f90:70-71


def _planck(v, t):
    return RADCN1 * v ** 3 / np.expm1(RADCN2 * v / t)


def _h(tau):
    return 1.0 / tau - 1.0 / np.expm1(tau)


@pytest.mark.parametrize("lin", [0, 1, 2])
def test_opaque_first_layer_hides_the_rest(lin):
    """Layer 0 (next to the observer) with OD 50: the radiance is its own emission,
    B_eff from its lower level, whatever lies beyond it.  Merging in the other order
    or emitting from the upper level would show layer 1 or B(Ttop)."""
    n = 300
    v1, dv = 700.0, 1.0
    v = v1 + dv * np.arange(n)
    od = np.vstack([np.full(n, 50.0), np.full(n, 0.7)])
    t_lev = np.array([290.0, 240.0, 200.0])
    t_lay = np.array([262.0, 218.0])
    a, tx = orc_down(od, t_lay, t_lev, v1, dv, lin)
    od2 = od.copy()
    od2[1] = 3.0
    b, _ = orc_down(od2, t_lay, np.array([290.0, 240.0, 150.0]), v1, dv, lin)
    assert np.array_equal(a, b)
    h = _h(50.0)
    ba, bb, bt = _planck(v, 290.0), _planck(v, 240.0), _planck(v, 262.0)
    want = {0: bt, 1: ba + 2 * (bt - ba) * h, 2: ba + (bb - ba) * h}[lin] * (1 - np.exp(-50.0))
    assert np.allclose(a, want, rtol=1e-12)
    assert np.all(tx < 1e-21)


def test_thin_layer_linear_in_tau_2():
    n = 200
    v1, dv = 500.0, 2.5
    v = v1 + dv * np.arange(n)
    tau = np.linspace(0.002, 0.3, n)
    r, tx = orc_down(tau[None, :], np.array([250.0]), np.array([280.0, 230.0]), v1, dv, 2)
    ba, bb = _planck(v, 280.0), _planck(v, 230.0)
    want = -np.expm1(-tau) * (ba + (bb - ba) * _h(tau))
    assert np.allclose(r, want, rtol=1e-11)
    assert np.allclose(tx, np.exp(-tau), rtol=1e-15)


@pytest.mark.parametrize("lin", [0, 1, 2])
def test_layer_emission_matches_the_reference(lin):
    import os
    ref_so = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                          "libclblm_ref.so")
    if not os.path.exists(ref_so):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref = ctypes.CDLL(ref_so, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    ref.ref_layer_emis.argtypes = [ctypes.c_int, ctypes.c_void_p, D, D, D, D, D, ctypes.c_int, ctypes.c_void_p]
    L = lib()
    L.orc_layer_emis.argtypes = [ctypes.c_int64, ctypes.c_void_p, D, D, D, D, D, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(7 + lin)
    n = 4000
    od = 10.0 ** rng.uniform(-6.0, 1.7, n)       # both rlin limits and the exact form
    for ta, tb, tave in ((285.0, 262.0, 273.0), (215.0, 240.0, 226.5)):
        a, b = np.zeros(n), np.zeros(n)
        ref.ref_layer_emis(n, od.ctypes.data, 600.0, 0.37, ta, tb, tave, lin, a.ctypes.data)
        L.orc_layer_emis(n, od.ctypes.data, 600.0, 0.37, ta, tb, tave, lin, b.ctypes.data)
        assert np.array_equal(a, b), np.abs(a - b).max() / np.abs(a).max()


@pytest.mark.gpu
@pytest.mark.parametrize("lin", [0, 1, 2])
def test_gpu_matches_oracle(lin):
    import torch
    from clblm_amd.jacobian import downwelling_radiance
    from clblm_amd.odlay import Context
    ctx = Context(0)
    rng = np.random.default_rng(3 + lin)
    L, n = 40, 20000
    od = 10.0 ** rng.uniform(-4.0, 1.0, (L, n))
    t_lev = np.linspace(295.0, 210.0, L + 1)
    t_lay = 0.5 * (t_lev[:-1] + t_lev[1:]) + rng.normal(0.0, 1.0, L)
    sol = rng.uniform(0.0, 1e-5, n)
    d = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    rad, tx = downwelling_radiance(ctx, d(od), d(t_lay), d(t_lev), 500.0, 0.01, lin, d(sol))
    torch.cuda.synchronize()
    ref, rtx = orc_down(od, t_lay, t_lev, 500.0, 0.01, lin, sol)
    err = float((np.abs(rad.cpu().numpy() - ref) / np.abs(ref)).max())
    print("down-welling radiance, linInTau %d: max rel diff %.3e" % (lin, err))
    assert err <= 1e-12
    assert np.allclose(tx.cpu().numpy(), rtx, rtol=1e-13, atol=0)
    ctx.close()
