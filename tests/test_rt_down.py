"""Down-welling radiance of a look-up path (mode3_lookUp_mergeUp,
src/clblm_src/clblm_noScattRT.f90:843-1108).  The oracle's restatement (orc_rt_down)
is pinned by closed forms: an isothermal path gives B(T)(1 - prod tx) for every
linInTau (the linear-in-tau source is B itself there), the solar term arrives through
the whole path's transmittance, and an optically thin layer adds its own emission
unattenuated only when it is the lowest.  The layer emission itself (layerEmis) is
pinned against the reference in tests/test_oracle_ref.py.  The GPU kernel
(clblm_rt_down) matches the oracle."""
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
