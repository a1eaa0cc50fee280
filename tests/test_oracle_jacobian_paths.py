"""The oracle's Jacobians with every branch (orc_jacobians_paths: air-mass factors
per layer, an observer inside the atmosphere, the up-looking sensor and the solar term
of clblm_noScattJac.f90:346-716), CPU only.  Pinned by:
* the default case (observer at the top, view air mass 1, down air mass sec, no sun)
  being orc_jacobians bit for bit (pinned in test_oracle_jacobian.py);
* its radiance: up-looking, the look-up path radiance orc_rt_down of the layers above
  the observer (pinned in test_rt_down.py); down-looking with an observer inside and the
  sun, the merge-down radiance orc_rt_mode1 (pinned in test_oracle_mode1.py);
* its derivatives: central differences of its own radiance with respect to each layer's
  vertical OD (every path scaled by its air mass), each level temperature (linInTau 2
  through the level Planck terms, 0 and 1 through layer means that are the level
  averages), the skin temperature, the emissivity and the solar reflectance.
"""
import numpy as np
import pytest

from oracle import py as O
from test_rt_down import orc_down

V1, DV, N = 700.0, 0.5, 41


def scene(L=6, seed=3):
    rng = np.random.default_rng(seed)
    od = rng.uniform(0.02, 0.6, (L, N))
    t_lev = np.linspace(290.0, 220.0, L + 1) + rng.normal(0, 3, L + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1])
    return od, t_lay, t_lev


def test_default_case_is_the_down_looking_jacobian():
    od, t_lay, t_lev = scene()
    L = od.shape[0]
    up, lo = np.full(L, 0.5), np.full(L, 0.5)
    for lin in (0, 1, 2):
        a = O.jacobians(od, t_lay, t_lev, V1, DV, emis=0.9, tskin=292.0, lin=lin, sec_dn=1.67, dt_dtup=up,
                        dt_dtlo=lo)
        b = O.jacobians_paths(od, t_lay, t_lev, V1, DV, amf_down=np.full(L, 1.67), emis=0.9, tskin=292.0, lin=lin,
                              dt_dtup=up, dt_dtlo=lo)
        for x, k in zip(a, ("rad", "dtau", "dtlev", "dtskin", "demis")):
            assert np.array_equal(x, b[k]), (lin, k)


@pytest.mark.parametrize("obs", [0, 2, 6])
@pytest.mark.parametrize("lin", [0, 1, 2])
def test_uplooking_radiance_is_the_look_up_path(obs, lin):
    od, t_lay, t_lev = scene()
    r = O.jacobians_paths(od, t_lay, t_lev, V1, DV, lin=lin, obs=obs, uplook=True)
    want, _ = orc_down(od[obs:], t_lay[obs:], t_lev[obs:], V1, DV, lin) if obs < 6 else (np.zeros(N), None)
    np.testing.assert_allclose(r["rad"], want, rtol=1e-13, atol=0)
    assert np.all(r["dtskin"] == 0) and np.all(r["demis"] == 0)
    assert np.all(r["dtau"][:obs] == 0)


@pytest.mark.parametrize("obs", [3, 6])
@pytest.mark.parametrize("lin", [0, 1, 2])
def test_down_looking_radiance_is_the_merge_down(obs, lin):
    od, t_lay, t_lev = scene()
    L = od.shape[0]
    rng = np.random.default_rng(1)
    fu, fd, fs = rng.uniform(1.0, 1.3, L), rng.uniform(1.5, 1.8, L), rng.uniform(1.1, 2.5, L)
    if lin == 0:
        # the merge-down takes a layer's upward emission from its down path for
        # linInTau 0 (lyrEmUp = lyrEmDn); the Jacobians use the view path: the same
        # only where the two paths are
        fd = fu
    sol = 1e-4 * (1.0 + 0.1 * np.sin(np.arange(N)))
    r = O.jacobians_paths(od, t_lay, t_lev, V1, DV, amf_view=fu, amf_down=fd, amf_sun=fs, emis=0.92, tskin=291.0,
                          lin=lin, solar=True, sol_toa=sol, rsfc_sun=0.3, obs=obs)
    v2 = V1 + DV * (N - 1)
    m, _, _ = O.rt_mode1([od[l] for l in range(L)], [DV] * L, V1, v2, t_lay, t_lev, fac_up=fu, fac_dn=fd, fac_sun=fs,
                         obs_level=obs, lin=lin, solar=True, emis=0.92, tskin=291.0, refl=0.3, sol=sol, sol_dv=DV)
    np.testing.assert_allclose(r["rad"], m, rtol=1e-12)


def _fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


@pytest.mark.parametrize("case", ["uplook", "inside_solar", "top_solar"])
@pytest.mark.parametrize("lin", [0, 1, 2])
def test_derivatives_against_central_differences(case, lin):
    od, t_lay, t_lev = scene(L=5, seed=8)
    L = od.shape[0]
    rng = np.random.default_rng(2)
    amf = dict(amf_view=rng.uniform(1.0, 1.4, L), amf_down=rng.uniform(1.5, 1.8, L), amf_sun=rng.uniform(1.1, 2.0, L))
    sol = np.full(N, 2e-4)
    kw = dict(emis=0.9, tskin=289.0, lin=lin, rsfc_sun=0.25, sol_toa=sol, dt_dtup=np.full(L, 0.5),
              dt_dtlo=np.full(L, 0.5), **amf)
    kw.update({"uplook": dict(uplook=True, obs=1), "inside_solar": dict(obs=3, solar=True),
               "top_solar": dict(obs=L, solar=True)}[case])

    def rad(od_=od, t_lev_=t_lev, **over):
        k = dict(kw, **over)
        tl = 0.5 * (t_lev_[1:] + t_lev_[:-1])
        return O.jacobians_paths(od_, tl, t_lev_, V1, DV, **k)["rad"]

    r = O.jacobians_paths(od, t_lay, t_lev, V1, DV, **kw)
    scale = np.abs(r["rad"]).max()
    for l in range(L):
        def f(x, l=l):
            o = od.copy()
            o[l] = x
            return rad(od_=o)
        fd = _fd(f, od[l], 1e-5)
        assert np.abs(r["dtau"][l] - fd).max() <= 1e-6 * scale + 1e-7 * np.abs(fd).max(), (case, lin, l)
    for k in range(L + 1):
        def g(x, k=k):
            tv = t_lev.copy()
            tv[k] = x
            return rad(t_lev_=tv)
        fd = _fd(g, t_lev[k], 1e-3)
        assert np.abs(r["dtlev"][k] - fd).max() <= 1e-6 * scale / 1.0, (case, lin, k)
    if case != "uplook":
        fd = _fd(lambda x: rad(tskin=x), 289.0, 1e-3)
        np.testing.assert_allclose(r["dtskin"], fd, rtol=1e-6, atol=1e-12 * scale)
        fd = _fd(lambda x: rad(emis=x), 0.9, 1e-5)
        np.testing.assert_allclose(r["demis"], fd, rtol=1e-6, atol=1e-9 * scale)
        fd = _fd(lambda x: rad(rsfc_sun=x), 0.25, 1e-3)
        np.testing.assert_allclose(r["drsfc"], fd, rtol=1e-9, atol=1e-14 * scale)
