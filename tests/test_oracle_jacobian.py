"""Pins the oracle's level-Jacobian chain (clblm_noScattJac.f90) by mathematics.

* level2layer (:1208-1279) is the derivative of two closed forms: the layer-mean
  temperature of a profile with exponential density and linear temperature,
  T = alphaT + (Tup*a - Tlo)/(a - 1) with a = rhoU/rhoL and alphaT = -(Tup-Tlo)/ln a,
  and the layer amount of an exponential density profile, W ~ qlo (b - 1)/ln b with
  b = qup/qlo.  Central differences of both match it.
* Jacobians (:346-716) is a second formulation of the radiance: its `rad` equals the
  merge-up recursion (orc_rt_up) whenever the total view transmittance exceeds 1e-6;
  below that it drops the reflected downwelling term, which is checked exactly.  Its
  derivatives match central differences of the merge-up radiance with respect to the
  layer ODs, level temperatures (through the level Planck terms and, for
  linear-in-tau 0 and 1, through the layer-mean temperatures and level2layer), skin
  temperature and emissivity.
* drv_LevelTemp / drv_LevelConc (:899-1203) match the chain rule through a synthetic
  model in which each layer's OD depends on its mean temperature and its molecular
  amounts: level Jacobians equal central differences of the whole radiance with
  respect to each level temperature and ln level density.
"""
import math

import numpy as np
import pytest

K_HPA = 1.3806503e-19


def _tbar(plo, pup, tlo, tup):
    rl, ru = plo / (tlo * K_HPA), pup / (tup * K_HPA)
    a = ru / rl
    return -(tup - tlo) / math.log(a) + (tup * a - tlo) / (a - 1.0)


def _lnw(qlo, qup):
    b = qup / qlo
    return math.log(qlo * (b - 1.0) / math.log(b))


def _profile(nlay, seed):
    rng = np.random.default_rng(seed)
    p = np.exp(np.linspace(math.log(1013.0), math.log(5.0), nlay + 1))
    t = np.linspace(290.0, 215.0, nlay + 1) + rng.normal(0.0, 3.0, nlay + 1)
    q = np.stack([(p / (K_HPA * t)) * vm for vm in (1e-2 * (p / 1013.0) ** 3, 4e-4 * np.ones_like(p))])
    tl = np.array([_tbar(p[l], p[l + 1], t[l], t[l + 1]) for l in range(nlay)])
    return p, t, q, tl


def test_level2layer_is_the_derivative_of_the_layer_means():
    from oracle.py import level2layer
    p, t, q, tl = _profile(8, 3)
    up, lo, wu, wl = level2layer(p, t, tl, q)
    h, e = 1e-4, 1e-6
    for l in range(8):
        f = lambda a, b: _tbar(p[l], p[l + 1], a, b)
        assert abs(lo[l] - (f(t[l] + h, t[l + 1]) - f(t[l] - h, t[l + 1])) / (2 * h)) < 1e-8
        assert abs(up[l] - (f(t[l], t[l + 1] + h) - f(t[l], t[l + 1] - h)) / (2 * h)) < 1e-8
        for m in range(2):
            g = lambda a, b: _lnw(a, b)
            ql, qu = q[m, l], q[m, l + 1]
            assert abs(wl[m, l] - (g(ql * math.exp(e), qu) - g(ql * math.exp(-e), qu)) / (2 * e)) < 1e-7
            assert abs(wu[m, l] - (g(ql, qu * math.exp(e)) - g(ql, qu * math.exp(-e))) / (2 * e)) < 1e-7
    # a zero lower-level density gives zero derivatives (the reference's guard)
    q0 = q.copy()
    q0[1, 3] = 0.0
    _, _, wu0, wl0 = level2layer(p, t, tl, q0)
    assert wu0[1, 3] == 0.0 and wl0[1, 3] == 0.0


def _rt(od, tl, tv, v1, dv, lin, sec, emis=0.98, tskin=295.0):
    import ctypes
    from oracle.py import lib
    D = ctypes.c_double
    L = lib()
    L.orc_rt_up.argtypes = [ctypes.c_int, ctypes.c_int64, D, D, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            D, D, ctypes.c_int, D, ctypes.c_void_p]
    od = np.ascontiguousarray(od, np.float64)
    tl = np.ascontiguousarray(tl, np.float64)
    tv = np.ascontiguousarray(tv, np.float64)
    out = np.zeros(od.shape[1])
    L.orc_rt_up(od.shape[0], od.shape[1], v1, dv, od.ctypes.data, tl.ctypes.data, tv.ctypes.data, emis, tskin, lin,
                sec, out.ctypes.data)
    return out


def _od(nlay, n, seed, hi=0.3):
    return 10 ** np.random.default_rng(seed).uniform(-4, math.log10(hi), size=(nlay, n))


@pytest.mark.parametrize("lin,sec", [(0, 1.0), (1, 1.0), (1, 1.66), (2, 1.0), (2, 1.5)])
def test_radiance_and_od_derivative_match_the_merge_up_form(lin, sec):
    """(linear-in-tau 0 only at sec_dn = 1: the merge-up form's upward emission is its
    downward one there, clblm_noScattRT.f90:1766)"""
    from oracle.py import jacobians
    nlay, n = 7, 40
    p, t, q, tl = _profile(nlay, 11)
    od = _od(nlay, n, 12)
    v1, dv = 700.0, 3.1
    rad, dtau, _, _, _ = jacobians(od, tl, t, v1, dv, lin=lin, sec_dn=sec)
    ref = _rt(od, tl, t, v1, dv, lin, sec)
    assert np.exp(-od.sum(0)).min() > 1e-6
    np.testing.assert_allclose(rad, ref, rtol=1e-13)
    for l in (0, 3, 6):
        h = 1e-5
        a, b = od.copy(), od.copy()
        a[l] += h
        b[l] -= h
        fd = (_rt(a, tl, t, v1, dv, lin, sec) - _rt(b, tl, t, v1, dv, lin, sec)) / (2 * h)
        np.testing.assert_allclose(dtau[l], fd, rtol=2e-6, atol=1e-9 * np.abs(fd).max())


@pytest.mark.parametrize("lin", [0, 1, 2])
def test_level_temperature_derivative(lin):
    """dR/dT_lev through the level Planck terms and the layer means T(T_lev).  For
    linear-in-tau 0 the merge-up form takes the upward emission equal to the downward
    one (clblm_noScattRT.f90:1766), which matches this formulation only at sec_dn = 1."""
    from oracle.py import jacobians, level2layer
    nlay, n = 6, 30
    sec = 1.0 if lin == 0 else 1.3
    p, t, q, tl = _profile(nlay, 21)
    od = _od(nlay, n, 22)
    up, lo, _, _ = level2layer(p, t, tl, q)
    _, _, dtlev, _, _ = jacobians(od, tl, t, 900.0, 1.7, lin=lin, sec_dn=sec, dt_dtup=up, dt_dtlo=lo)

    def rad_of(tv):
        tlay = np.array([_tbar(p[l], p[l + 1], tv[l], tv[l + 1]) for l in range(nlay)])
        return _rt(od, tlay, tv, 900.0, 1.7, lin, sec)

    for k in range(nlay + 1):
        h = 1e-3
        a, b = t.copy(), t.copy()
        a[k] += h
        b[k] -= h
        fd = (rad_of(a) - rad_of(b)) / (2 * h)
        np.testing.assert_allclose(dtlev[k], fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max(), err_msg=str(k))


def test_surface_derivatives():
    from oracle.py import jacobians
    nlay, n = 5, 25
    p, t, q, tl = _profile(nlay, 31)
    od = _od(nlay, n, 32)
    _, _, _, dts, de = jacobians(od, tl, t, 1200.0, 2.3, lin=1, sec_dn=1.66, emis=0.9, tskin=300.0)
    fd_t = (_rt(od, tl, t, 1200.0, 2.3, 1, 1.66, 0.9, 300.01) - _rt(od, tl, t, 1200.0, 2.3, 1, 1.66, 0.9, 299.99)) / 0.02
    fd_e = (_rt(od, tl, t, 1200.0, 2.3, 1, 1.66, 0.9001, 300.0) - _rt(od, tl, t, 1200.0, 2.3, 1, 1.66, 0.8999, 300.0)) \
        / 0.0002
    np.testing.assert_allclose(dts, fd_t, rtol=1e-6)
    np.testing.assert_allclose(de, fd_e, rtol=1e-6)


def test_opaque_path_drops_the_reflected_term():
    """Total view transmittance <= 1e-6: the radiance is the upwelling emission plus
    emis * T * B(Tskin), without the reflected downwelling (:480)."""
    from oracle.py import jacobians, lib
    nlay, n = 5, 20
    p, t, q, tl = _profile(nlay, 41)
    od = _od(nlay, n, 42) + 4.0
    assert np.exp(-od.sum(0)).max() <= 1e-6
    rad = jacobians(od, tl, t, 800.0, 1.0, lin=1, sec_dn=1.66, emis=0.9, tskin=300.0)[0]
    no_refl = _rt(od, tl, t, 800.0, 1.0, 1, 1.66, 1.0, 300.0)
    bsfc = np.array([lib().orc_planck(800.0 + i * 1.0, 300.0) for i in range(n)])
    np.testing.assert_allclose(rad, no_refl - 0.1 * np.exp(-od[::-1].cumsum(0)[-1]) * bsfc, rtol=1e-12)


def test_level_jacobians_follow_the_chain_rule():
    """Synthetic layer ODs tau_l(T_l, W_l,m) = tau_l + dtau/dT (T_l - T_l0) +
    sum_m tau_l,m (W_l,m / W_l,m0 - 1); the variants the reference computes (T + 1 K,
    Wtot + 1 %, each molecule alone) are taken from that model, and the level Jacobians
    must equal central differences of the radiance with respect to each level
    temperature and ln level density."""
    from oracle.py import jacobians, level2layer, level_jac
    nlay, n, lin, sec = 6, 24, 1, 1.66
    rng = np.random.default_rng(51)
    p, t, q, tl0 = _profile(nlay, 52)
    tau = _od(nlay, n, 53, hi=0.2)
    tau_m = np.stack([tau * f for f in (0.6, 0.3)])
    dtdt = rng.uniform(-2e-3, 2e-3, (nlay, n)) * tau
    wtot = rng.uniform(1e22, 1e24, nlay)
    dtdw = rng.uniform(0.0, 1e-3, (nlay, n)) * tau / wtot[:, None]    # dtau/dWtot
    dt, dwf = 1.0, 0.01
    # the variants: T + dt and Wtot + dwf*Wtot, built so that the reference's combination
    # (od_T - od)/dt - (od_W - od)/(dwf Wtot) * Wtot/T recovers dtdt
    od_w = tau + dtdw * (dwf * wtot)[:, None]
    od_t = tau + dt * (dtdt + dtdw * (wtot / tl0)[:, None])
    od_var = np.concatenate([tau[None], od_t[None], od_w[None], tau_m])
    up, lo, wu, wl = level2layer(p, t, tl0, q)
    _, dtau, dtlev, _, _ = jacobians(tau, tl0, t, 650.0, 0.9, lin=lin, sec_dn=sec, dt_dtup=up, dt_dtlo=lo)
    jt, jq = level_jac(od_var, dtau, dtlev, wtot, tl0, dt, dwf, up, lo, wu, wl)
    lnw0 = np.array([[_lnw(q[m, l], q[m, l + 1]) for l in range(nlay)] for m in range(2)])

    def rad_of(tv, qv):
        tlay = np.array([_tbar(p[l], p[l + 1], tv[l], tv[l + 1]) for l in range(nlay)])
        od = tau + dtdt * (tlay - tl0)[:, None]
        for m in range(2):
            r = np.exp(np.array([_lnw(qv[m, l], qv[m, l + 1]) for l in range(nlay)]) - lnw0[m])
            od = od + tau_m[m] * (r - 1.0)[:, None]
        return _rt(od, tlay, tv, 650.0, 0.9, lin, sec)

    for k in range(nlay + 1):
        h = 1e-3
        a, b = t.copy(), t.copy()
        a[k] += h
        b[k] -= h
        fd = (rad_of(a, q) - rad_of(b, q)) / (2 * h)
        np.testing.assert_allclose(jt[k], fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max(), err_msg="T level %d" % k)
        for m in range(2):
            e = 1e-5
            qa, qb = q.copy(), q.copy()
            qa[m, k] *= math.exp(e)
            qb[m, k] *= math.exp(-e)
            fd = (rad_of(t, qa) - rad_of(t, qb)) / (2 * e)
            np.testing.assert_allclose(jq[m, k], fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max() + 1e-30,
                                       err_msg="q level %d mol %d" % (k, m))
