"""The oracle's fluxes (orc_rt_flux: the flux branches of mode1_lookDn_mergeDn and
mode2_lookDn_mergeUp, src/clblm_src/clblm_noScattRT.f90:133-507, 1214-1450, with the
driver's net flux and heating rate, clblm_Drivers.f90:418-466), CPU only.  The
reference's RT driver needs ODLAY and netCDF scenes, so the restatement is pinned by
what the physics and the quadrature fix:
* an isothermal atmosphere over a black surface at its temperature sends up pi B at
  every level, and down B (1 - exp(-sec tau_above)) summed with the angle weights;
* with no absorption nothing comes down and the surface emission goes up unchanged;
* the 1-3 angle quadrature approximates the diffuse flux of a slab, pi B (1 - 2 E3(tau)),
  to the accuracy of its order (a wrong secant or weight fails);
* with one angle the lambertian and the specular surface are the same;
* the heating rates telescope: sum_j htr_j (p_j - p_j+1) / 8.4391 = net_0 - net_top;
* the output groups hold int(dv_flux / dv) points each from v1.
"""
import numpy as np
import pytest
from scipy.special import expn

from oracle import py as O

C2 = 1.4387752
C1 = 1.191042722e-12


def planck(v, t):
    return C1 * v ** 3 / np.expm1(C2 * v / t)


def grid(v1, v2, dv):
    n = int(round((v2 - v1) / dv)) + 1
    return v1 + dv * np.arange(n)


def groups(x, r, nout):
    return np.array([x[g * r:(g + 1) * r].sum() for g in range(nout)])


V1, V2, DV, DVF = 600.0, 700.0, 2.0 ** -6, 10.0


@pytest.mark.parametrize("nang", [1, 2, 3])
@pytest.mark.parametrize("lin", [0, 1, 2])
def test_isothermal_black_surface(nang, lin):
    v = grid(V1, V2, DV)
    L, T = 5, 260.0
    rng = np.random.default_rng(4)
    od = rng.uniform(0.01, 2.0, (L, v.size))
    p = np.linspace(1000.0, 100.0, L + 1)
    f = O.rt_flux(od, [T] * L, [T] * (L + 1), p, V1, V2, DV, DVF, nang=nang, lin=lin, emis=1.0, tskin=T)
    r = int(DVF / DV)
    nout = int((V2 - V1) / DVF)
    b = planck(v, T)
    piB = np.pi * groups(b, r, nout) * DV * 1e4
    for j in range(L + 1):
        np.testing.assert_allclose(f["up"][j], piB, rtol=1e-6)
    sec = {1: [1 / 0.66666666667], 2: [1 / 0.84494897, 1 / 0.35505103],
           3: [1 / 0.91141204, 1 / 0.59053314, 1 / 0.21234054]}[nang]
    w = {1: [0.5], 2: [0.31804138, 0.18195862], 3: [0.20093191, 0.22924111, 0.06982698]}[nang]
    for j in range(L):
        tau = od[j:].sum(axis=0)
        d = sum(wa * groups(b * (1 - np.exp(-s * tau)), r, nout) for s, wa in zip(sec, w)) * DV * 1e4 * 2 * np.pi
        np.testing.assert_allclose(f["dn"][j], d, rtol=1e-6)
    assert np.all(f["dn"][L] == 0)


def test_transparent_atmosphere():
    v = grid(V1, V2, DV)
    L = 4
    od = np.zeros((L, v.size))
    f = O.rt_flux(od, [250.0] * L, [240.0 + 5 * k for k in range(L + 1)], np.linspace(1000, 200, L + 1),
                  V1, V2, DV, DVF, nang=3, emis=0.9, tskin=300.0)
    r, nout = int(DVF / DV), int((V2 - V1) / DVF)
    up = np.pi * groups(0.9 * planck(v, 300.0), r, nout) * DV * 1e4
    assert np.all(f["dn"] == 0)
    for j in range(L + 1):
        np.testing.assert_allclose(f["up"][j], up, rtol=1e-6)


# the relative error of the 1-3 angle quadrature for a slab (Radsum's angles and
# weights), with a margin: it shrinks with the number of angles and with tau
QUAD_TOL = {(1, 0.05): 0.21, (1, 0.5): 0.055, (1, 2.0): 0.012, (2, 0.05): 0.07, (2, 0.5): 0.005,
            (2, 2.0): 0.001, (3, 0.05): 0.026, (3, 0.5): 0.0025, (3, 2.0): 5e-5}


@pytest.mark.parametrize("nang", [1, 2, 3])
@pytest.mark.parametrize("tau", [0.05, 0.5, 2.0])
def test_quadrature_against_the_diffuse_slab(nang, tau):
    """One isothermal layer of optical depth tau over a cold black surface: the
    quadrature flux at the top against pi B (1 - 2 E3(tau))."""
    tol = QUAD_TOL[(nang, tau)]
    v = grid(V1, V2, DV)
    od = np.full((1, v.size), tau)
    f = O.rt_flux(od, [280.0], [280.0, 280.0], [1000.0, 500.0], V1, V2, DV, DVF, nang=nang, lin=0, emis=1.0,
                  tskin=1.0)
    r, nout = int(DVF / DV), int((V2 - V1) / DVF)
    exact = np.pi * groups(planck(v, 280.0), r, nout) * DV * 1e4 * (1 - 2 * expn(3, tau))
    err = np.abs(f["up"][1] / exact - 1).max()
    assert err < tol, err
    # the downward flux below the layer is the same slab seen from below
    np.testing.assert_allclose(f["dn"][0], f["up"][1], rtol=1e-12)


def test_one_angle_lambertian_is_specular():
    v = grid(V1, V2, DV)
    rng = np.random.default_rng(7)
    L = 3
    od = rng.uniform(0.0, 1.0, (L, v.size))
    args = (od, [250, 240, 230], [255, 245, 235, 225], [1000, 700, 400, 100], V1, V2, DV, DVF)
    a = O.rt_flux(*args, nang=1, emis=0.8, tskin=290.0, lambertian=True)
    b = O.rt_flux(*args, nang=1, emis=0.8, tskin=290.0, lambertian=False)
    for k in ("up", "dn"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-13)
    c = O.rt_flux(*args, nang=3, emis=0.8, tskin=290.0, lambertian=False)
    d = O.rt_flux(*args, nang=3, emis=0.8, tskin=290.0, lambertian=True)
    assert np.abs(c["up"][1:] - d["up"][1:]).max() > 0     # with three angles they differ


def test_heating_rates_telescope():
    v = grid(V1, V2, DV)
    rng = np.random.default_rng(9)
    L = 6
    od = rng.uniform(0.0, 3.0, (L, v.size))
    t_lev = np.linspace(290.0, 220.0, L + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1])
    p = np.geomspace(1000.0, 50.0, L + 1)
    f = O.rt_flux(od, t_lay, t_lev, p, V1, V2, DV, DVF, nang=2, lin=1, emis=0.95, tskin=292.0)
    np.testing.assert_allclose(f["net"], f["up"] - f["dn"], rtol=0, atol=0)
    lhs = (f["heating"][:-1] * (p[:-1] - p[1:])[:, None] / 8.4391).sum(axis=0)
    np.testing.assert_allclose(lhs, f["net"][0] - f["net"][-1], rtol=1e-9)
    assert np.all(f["heating"][-1] == 0)
    # a cooling troposphere: the net upward flux grows with height over this band
    assert np.all(f["net"][-1] > f["net"][0])


def test_output_groups():
    """int(dv_flux / dv) points per group from v1: with no absorption the surface
    emission's group sums pin the grouping, including a ratio that truncates."""
    for dv, dvf in [(2.0 ** -6, 10.0), (0.01, 5.0)]:
        v = grid(V1, V2, dv)
        f = O.rt_flux(np.zeros((1, v.size)), [250.0], [250.0, 250.0], [1000.0, 900.0], V1, V2, dv, dvf, nang=1,
                      emis=1.0, tskin=280.0)
        r = int(dvf / dv)
        nout = int((V2 - V1) / dvf + 1e-4)
        np.testing.assert_allclose(f["up"][1], np.pi * groups(planck(v, 280.0), r, nout) * dv * 1e4, rtol=1e-6)
    with pytest.raises(RuntimeError):
        O.rt_flux(np.zeros((1, 101)), [250.0], [250.0, 250.0], [1000.0, 900.0], V1, V2, 1.0, 30.0)
