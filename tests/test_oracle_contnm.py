"""The oracle's CONTNM bands other than water vapour (orc_contnm_ir: CO2 with FRNCO2
and the XFACCO2 scaling; ozone's Chappuis, Hartley-Huggins and UV bands; the O2
fundamental, 1.27 um, 1.06 um, A-band, visible, Herzberg and far-UV continua; the N2
pure rotation, fundamental and first overtone; Rayleigh extinction;
clblm_ODLAY_Continuum.f90:408-1060), CPU only.  The reference's Module_Continuum uses
mt_ckd_h2o, which reads its table through netCDF-Fortran, absent here, so the
restatement is pinned by what the arithmetic fixes:
* the generated tables (clblm_amd/csrc/clblm_contnm.inc) hold the values of the
  reference's DATA statements (repeat counts and implied-do pieces included), read back
  here by a parser of their own;
* at ABSRB points on a band's table nodes XINT returns the node, so the band there is
  the closed form of the table value at that wavenumber times CONTNM's factors (CO2
  every 2 cm-1 from -4 with XFACCO2 and the band-head exponents; the O2 fundamental,
  N2 rotation, ozone and O2 visible and far-UV tables on their grids); ozone's
  temperature polynomials reduce to the first table at 273.15 K;
* the A-band table sits on half-integer nodes, where XINT is the midpoint rule
  (-a0 + 9 a1 + 9 a2 - a3) / 16, and pre_xint keeps the points outside its table zero;
* around 40800 cm-1 CONTNM's save and restore leave each ozone band on its own side
  when the Hartley-Huggins window reaches past 40815 cm-1, and both bands at 40800 cm-1
  when it does not;
* the 1.06 um band (two damped Lorentzians) integrates, above its upper centre where
  neither is damped, to the Lorentzian areas' closed form; the Herzberg continuum is
  zero up to 36000 cm-1, peaks where its log-normal form does and scales with pressure
  and temperature by HERPRS's factor;
* the N2 fundamental (nodes off the integer grid) scales exactly with the O2
  broadening efficiency 1.294 - 0.4545 T/296, and equals its 272 K table at 272 K;
* Rayleigh extinction is its closed form at every point, for every JRAD;
* JRAD 1 multiplies each node by RADFN, JRAD 0 leaves it out;
* the bands are additive (each factor alone sums to all together), a zero factor or a
  range outside a band adds nothing, and an ABSRB longer than 5050 points is refused.
"""
import os
import re

import numpy as np
import pytest

from oracle import py as O

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "clblm_amd", "csrc", "clblm_contnm.inc")
REF = "/root/reference/src/clblm_src/clblm_ODLAY_Continuum.f90"
C2 = 1.4387752
LOSCHMIDT = 2.68675e19
ALL_OFF = (0.0, 0.0, 0.0, 0.0, 0.0)


def inc_tables():
    txt = open(INC).read()
    out = {}
    for m in re.finditer(r"static const double (\w+)\[(\d+)\] = \{(.*?)\};", txt, re.S):
        vals = [float(x) for x in m.group(3).replace("\n", " ").split(",") if x.strip()]
        assert len(vals) == int(m.group(2))
        out[m.group(1)] = np.array(vals)
    return out


T = inc_tables()


class Layer:
    def __init__(self, pave, tave, wtot, w):
        self.pave, self.tave, self.wtot, self.w = pave, tave, wtot, dict(w)


def layer(pave=1013.0, tave=288.0, h2o=0.01, co2=4e-4, o2=0.209):
    wtot = 2.15e25 * pave / 1013.0
    return Layer(pave, tave, wtot, {1: h2o * wtot, 2: co2 * wtot, 7: o2 * wtot})


def grid(v1, v2):
    a, n = O.absrb_grid(v1, v2)
    return a + np.arange(n)


def radfn(v, t):
    x = v / (t / C2)
    return np.where(x <= 0.01, 0.5 * x * v, np.where(x <= 10.0, v * (1 - np.exp(-x)) / (1 + np.exp(-x)), v))


def amagat(l):
    return (l.pave / 1013.0) * (273.0 / l.tave)


def _ref_subroutines():
    src = open(REF).read()
    text = "\n".join(ln.split("!")[0] for ln in src.splitlines())
    subs = {}
    for m in re.finditer(r"(?im)^\s*(subroutine)\s+(\w+)", text):
        end = text.lower().find("end subroutine", m.end())
        subs[m.group(2).lower()] = text[m.start():end]
    return subs


def _ref_data(body, name, lo, n):
    """The array name(lo:lo+n-1) from the DATA statements of one routine: each statement
    names the array whole, a section name(a:b), or an implied do (name(i),i=a,b); its
    values are separated by commas, `r*x` standing for r copies of x."""
    out = np.full(n, np.nan)
    heads = list(re.finditer(r"(?i)\bdata\s*(\(\s*%s\s*\(\s*\w+\s*\)\s*,\s*\w+\s*=\s*(-?\d+)\s*,\s*(-?\d+)\s*\)"
                             r"|%s\s*\(\s*(\d+)\s*:\s*(\d+)\s*\)|%s)\s*/" % (name, name, name), body))
    assert heads, name
    for h in heads:
        if h.group(2):
            a, b = int(h.group(2)), int(h.group(3))
        elif h.group(4):
            a, b = int(h.group(4)), int(h.group(5))
        else:
            a, b = lo, lo + n - 1
        seg = body[h.end():]
        seg = seg[:seg.index("/")]
        vals = []
        for item in seg.replace("&", " ").split(","):
            item = item.strip()
            if "*" in item:
                r, x = item.split("*")
                vals += [float(x)] * int(r)
            elif item:
                vals.append(float(item))
        assert len(vals) == b - a + 1, (name, a, b, len(vals))
        assert np.all(np.isnan(out[a - lo:b - lo + 1])), (name, a)
        out[a - lo:b - lo + 1] = vals
    assert not np.any(np.isnan(out)), name
    return out


def _ref_grid(body):
    """V1S, V2S, DVS, NPTS of a table routine, from its DATA statement."""
    m = re.search(r"(?i)\bdata\s+v1s\s*,\s*v2s\s*,\s*dvs\s*,\s*npts[\s&]*/([^/]*)/", body)
    return [float(x) for x in m.group(1).replace("&", " ").replace(",", " ").split()]


@pytest.mark.skipif(not os.path.exists(REF), reason="reference sources absent")
def test_tables_hold_the_reference_data():
    """Every value of the generated tables against the reference's DATA text, read by a
    scan of its own."""
    subs = _ref_subroutines()
    want = {"CN_XFACCO2": ("contnm", "XFACCO2", 1), "CN_CO2_S": ("frnco2", "S", 1),
            "CN_CO2_TDEP": ("frnco2", "tdep_bandhead", 1196), "CN_N2R_C296": ("xn2_r", "C_296", 1),
            "CN_N2R_SF296": ("xn2_r", "sf_296", 1), "CN_N2R_C220": ("xn2_r", "C_220", 1),
            "CN_N2R_SF220": ("xn2_r", "sf_220", 1), "CN_N2F_X272": ("n2_ver_1", "xn2_272", 1),
            "CN_N2F_X228": ("n2_ver_1", "xn2_228", 1), "CN_N2F_AH2O": ("n2_ver_1", "a_h2o", 1),
            "CN_O2F_X": ("o2_ver_1", "xo2", 1), "CN_O2F_XT": ("o2_ver_1", "xo2t", 1),
            "CN_N2O_X": ("n2_overtone1", "xn2", 1), "CN_O2I1_X": ("o2inf1", "xo2inf1", 1),
            "CN_O2I3_X": ("o2inf3", "xo2inf3", 1), "CN_O2V_S": ("o2_vis", "S", 1), "CN_O2U_S": ("o2fuv", "S", 1),
            "CN_O3C_X": ("xo3chp", "X", 1), "CN_O3C_Y": ("xo3chp", "Y", 1), "CN_O3C_Z": ("xo3chp", "Z", 1),
            "CN_O3T0_S": ("o3hht0", "S", 1), "CN_O3T1_S": ("o3hht1", "S", 1), "CN_O3T2_S": ("o3hht2", "S", 1),
            "CN_O3UV_S": ("o3hhuv", "S", 1)}
    for cname, (sub, name, lo) in want.items():
        vals = _ref_data(subs[sub], name, lo, T[cname].size)
        assert np.array_equal(vals, T[cname]), cname
    grids = {"CN_CO2_GRID": "frnco2", "CN_O2F_GRID": "o2_ver_1", "CN_N2R_GRID": "xn2_r", "CN_N2F_GRID": "n2_ver_1",
             "CN_N2O_GRID": "n2_overtone1",
             "CN_O2I1_GRID": "o2inf1", "CN_O2I3_GRID": "o2inf3", "CN_O2V_GRID": "o2_vis", "CN_O2U_GRID": "o2fuv",
             "CN_O3C_GRID": "xo3chp", "CN_O3T_GRID": "o3hht0", "CN_O3UV_GRID": "o3hhuv"}
    for cname, sub in grids.items():
        assert _ref_grid(subs[sub]) == list(T[cname]), cname
    # CONTNM indexes O3HHT1's and O3HHT2's windows with O3HHT0's J: one grid for the three
    assert _ref_grid(subs["o3hht1"]) == _ref_grid(subs["o3hht2"]) == list(T["CN_O3T_GRID"])
    assert list(T["CN_CO2_GRID"]) == [-4.0, 10000.0, 2.0, 5003.0]
    assert list(T["CN_O2F_GRID"]) == [1340.0, 1850.0, 5.0, 103.0]
    assert list(T["CN_N2R_GRID"]) == [-10.0, 350.0, 5.0, 73.0]
    assert T["CN_CO2_TEFF"][0] == 246.0
    # every table holds NPTS values
    for tab, grid in [("CN_N2O_X", "CN_N2O_GRID"), ("CN_O2I1_X", "CN_O2I1_GRID"), ("CN_O2I3_X", "CN_O2I3_GRID"),
                      ("CN_O2V_S", "CN_O2V_GRID"), ("CN_O2U_S", "CN_O2U_GRID"), ("CN_O3C_Z", "CN_O3C_GRID"),
                      ("CN_O3T2_S", "CN_O3T_GRID"), ("CN_O3UV_S", "CN_O3UV_GRID")]:
        assert T[tab].size == int(T[grid][3]), tab


@pytest.mark.parametrize("v1,v2,tave", [(600.0, 1100.0, 288.0), (1900.0, 2500.0, 250.0), (2300.0, 2480.0, 246.0),
                                        (2300.0, 2480.0, 296.0)])
def test_co2_at_the_nodes(v1, v2, tave):
    l = layer(tave=tave)
    a, beyond = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 0, (1.0, 0.0, 0.0, 0.0, 0.0))
    assert beyond == 0
    v = grid(v1, v2)
    node = np.mod(v, 2.0) == 0
    k = ((v[node] + 4.0) / 2.0).astype(int)          # 0-based table index of the node
    s = T["CN_CO2_S"][k]
    tcor = np.ones_like(s)
    hb = (k + 1 >= 1196) & (k + 1 <= 1220)
    tcor[hb] = (tave / 246.0) ** T["CN_CO2_TDEP"][k[hb] + 1 - 1196]
    cfac = np.ones_like(s)
    sc = (v[node] >= 2000.0) & (v[node] <= 2998.0)
    cfac[sc] = T["CN_XFACCO2"][((v[node][sc] - 1998.0) / 2.0).astype(int) - 1]
    wco2 = l.w[2] * (l.pave / 1013.0) * (296.0 / l.tave) * 1e-20
    np.testing.assert_allclose(a[node], cfac * tcor * s * wco2, rtol=1e-14, atol=0)
    if v2 > 2434:
        assert np.any(hb) and (tave == 246.0 or np.any(tcor != 1.0))
    # JRAD 1 multiplies each node by RADFN
    b, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, (1.0, 0.0, 0.0, 0.0, 0.0))
    np.testing.assert_allclose(b[node], a[node] * radfn(v[node], tave), rtol=1e-13)


@pytest.mark.parametrize("tave", [296.0, 250.0])
def test_o2_fundamental_at_the_nodes(tave):
    l = layer(tave=tave, co2=0.0)
    v1, v2 = 1330.0, 1870.0
    a, beyond = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 0, (0.0, 0.0, 1.0, 0.0, 0.0))
    assert beyond == 0
    v = grid(v1, v2)
    node = (np.mod(v, 5.0) == 0) & (v >= 1340.0) & (v <= 1850.0)
    k = ((v[node] - 1340.0) / 5.0).astype(int)
    c0 = (1e20 / LOSCHMIDT) * T["CN_O2F_X"][k] * np.exp(T["CN_O2F_XT"][k] * (1 / 296.0 - 1 / tave)) / v[node]
    want = l.w[7] * 1e-20 * amagat(l) * c0
    np.testing.assert_allclose(a[node], want, rtol=1e-13)
    assert a.max() > 0


@pytest.mark.parametrize("tave,which", [(296.0, "296"), (220.0, "220")])
def test_n2_rotation_at_the_reference_temperatures(tave, which):
    l = layer(tave=tave)
    v1, v2 = 20.0, 300.0
    a, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 0, (0.0, 0.0, 0.0, 1.0, 0.0))
    v = grid(v1, v2)
    node = np.mod(v, 5.0) == 0
    k = ((v[node] + 10.0) / 5.0).astype(int)
    c = T["CN_N2R_C" + which][k]
    sf = T["CN_N2R_SF" + which][k]
    x_h2o, x_o2 = l.w[1] / l.wtot, l.w[7] / l.wtot
    x_n2 = 1.0 - x_h2o - x_o2
    fo2 = (sf - 1.0) * 0.79 / 0.21
    want = (x_n2 * l.wtot / LOSCHMIDT) * amagat(l) * c * (x_n2 + fo2 * x_o2 + x_h2o)
    np.testing.assert_allclose(a[node], want, rtol=1e-13)


def test_n2_fundamental_broadening_efficiency():
    """With no water vapour, a layer of 80 % N2 and 20 % O2 against one of pure N2: the
    band scales by the N2 amount (0.8) times the broadening (0.8 + a_O2 0.2) at every
    point, a_O2 = 1.294 - 0.4545 T/296 (the table shape cancels)."""
    v1, v2 = 1995.0, 2900.0
    for tave in (272.0, 240.0):
        base = layer(tave=tave, h2o=0.0, co2=0.0, o2=0.0)
        a0, _ = O.contnm_ir(base.pave, base.tave, base.wtot, base.w, v1, v2, 0, (0.0, 0.0, 0.0, 1.0, 0.0))
        wl = layer(tave=tave, h2o=0.0, co2=0.0, o2=0.2)
        a1, _ = O.contnm_ir(wl.pave, wl.tave, wl.wtot, wl.w, v1, v2, 0, (0.0, 0.0, 0.0, 1.0, 0.0))
        a_o2 = 1.294 - 0.4545 * tave / 296.0
        x_o2 = 0.2
        # a1 = tau(x_n2 = 0.8) (0.8 + a_o2 0.2) c, a0 = tau(x_n2 = 1) (1) c, tau linear in x_n2
        ratio = (0.8 * (0.8 + a_o2 * x_o2)) / 1.0
        live = a0 > 1e-6 * a0.max()
        assert live.sum() > 100
        np.testing.assert_allclose(a1[live], a0[live] * ratio, rtol=1e-12)


def test_n2_fundamental_at_272k_is_its_table():
    """At 272 K the band is x272 / v on its 3.98 cm-1 grid, so the nodes interpolated by
    XINT onto the 1 cm-1 grid of an N2-only layer bracket the table values."""
    l = layer(tave=272.0, h2o=0.0, co2=0.0, o2=0.0)
    a, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, 1995.0, 2900.0, 0, (0.0, 0.0, 0.0, 1.0, 0.0))
    v = grid(1995.0, 2900.0)
    g = T["CN_N2F_GRID"]
    vn = g[0] + g[2] * np.arange(int(g[3]))
    tab = (l.wtot / LOSCHMIDT) * amagat(l) * T["CN_N2F_X272"] / vn
    inner = (v > vn[1]) & (v < vn[-2])
    interp = np.interp(v[inner], vn, tab)
    # the cubic interpolant stays within the local spread of the tabulated values
    assert np.max(np.abs(a[inner] - interp)) < 0.05 * tab.max()
    assert abs(a.max() - tab.max()) < 0.02 * tab.max()


def _layer_o3(tave=288.0, pave=1013.0):
    l = layer(pave=pave, tave=tave)
    l.w[3] = 3e-8 * l.wtot
    return l


def _nodes(grid_name, v):
    """The ABSRB points on the table's nodes, and their 0-based table index."""
    g = T[grid_name]
    k = (v - g[0]) / g[2]
    on = (np.abs(k - np.round(k)) < 1e-9) & (k >= -0.5) & (k <= g[3] - 0.5)
    return on, np.round(k[on]).astype(int)


def _band(l, v1, v2, which, jrad=-1):
    sc = [0.0] * 5
    sc[which] = 1.0
    a, rc = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, jrad, tuple(sc))
    assert rc == 0
    return a


@pytest.mark.parametrize("tave", [273.15, 230.0, 300.0])
def test_ozone_chappuis_at_the_nodes(tave):
    """(X + Y dT + Z dT^2) / v times the O3 column 1e-20 on the 5 cm-1 nodes from 8920
    (the temperature tables are nonzero from 13020 cm-1); at 273.15 K the first table
    alone."""
    l = _layer_o3(tave)
    v1, v2 = 13000.0, 18000.0
    v = grid(v1, v2)
    a = _band(l, v1, v2, 1)
    on, k = _nodes("CN_O3C_GRID", v)
    dt = tave - 273.15
    want = (T["CN_O3C_X"][k] + T["CN_O3C_Y"][k] * dt + T["CN_O3C_Z"][k] * dt * dt) / v[on] * l.w[3] * 1e-20
    if tave == 273.15:
        np.testing.assert_allclose(a[on], T["CN_O3C_X"][k] / v[on] * l.w[3] * 1e-20, rtol=1e-15, atol=0)
    np.testing.assert_allclose(a[on], want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
    assert on.sum() > 900 and np.abs(a).max() > 0
    if tave != 273.15:
        assert np.abs(a[on] / (T["CN_O3C_X"][k] / v[on] * l.w[3] * 1e-20) - 1).max() > 1e-3


@pytest.mark.parametrize("tave", [273.15, 220.0])
def test_ozone_hartley_huggins_at_the_nodes(tave):
    """S0 / v (1 + S1 dT + S2 dT^2) times the O3 column 1e-20 on the 5 cm-1 nodes from 27370."""
    l = _layer_o3(tave)
    v1, v2 = 27300.0, 30000.0
    v = grid(v1, v2)
    a = _band(l, v1, v2, 1)
    on, k = _nodes("CN_O3T_GRID", v)
    tc = tave - 273.15
    want = T["CN_O3T0_S"][k] / v[on] * l.w[3] * 1e-20 * (1 + T["CN_O3T1_S"][k] * tc + T["CN_O3T2_S"][k] * tc * tc)
    np.testing.assert_allclose(a[on], want, rtol=1e-13)
    assert on.sum() > 500


def test_ozone_uv_at_the_nodes():
    """S / v times the O3 column (the table holds cross sections) every 100 cm-1 from 40800."""
    l = _layer_o3()
    v1, v2 = 41000.0, 44000.0
    v = grid(v1, v2)
    a = _band(l, v1, v2, 1)
    on, k = _nodes("CN_O3UV_GRID", v)
    np.testing.assert_allclose(a[on], T["CN_O3UV_S"][k] / v[on] * l.w[3], rtol=1e-14)
    assert on.sum() == 31


def test_ozone_bands_either_side_of_40800():
    """Over 39000-42000 the Hartley-Huggins window reaches past 40815 cm-1: CONTNM keeps the
    points from 40800 cm-1 out of it and the points below out of the UV band, so each
    side holds one band's nodes.  Over 39000-40801 the window ends at 40810 cm-1: both
    bands reach 40800 cm-1."""
    l = _layer_o3()
    v = grid(39000.0, 42000.0)
    a = _band(l, 39000.0, 42000.0, 1)
    w3 = l.w[3]
    lo, klo = _nodes("CN_O3T_GRID", v)
    hh = T["CN_O3T0_S"][klo] / v[lo] * w3 * 1e-20 * (1 + T["CN_O3T1_S"][klo] * 14.85 + T["CN_O3T2_S"][klo] * 14.85 ** 2)
    below = v[lo] < 40800.0
    np.testing.assert_allclose(a[lo][below], hh[below], rtol=1e-13)
    hi, khi = _nodes("CN_O3UV_GRID", v)
    np.testing.assert_allclose(a[hi], T["CN_O3UV_S"][khi] / v[hi] * w3, rtol=1e-14)
    assert v[hi][0] == 40800.0
    b = _band(l, 39000.0, 40801.0, 1)
    vb = grid(39000.0, 40801.0)
    i = int(np.flatnonzero(vb == 40800.0)[0])
    s_hh = T["CN_O3T0_S"][-1] / 40800.0 * w3 * 1e-20 * (1 + T["CN_O3T1_S"][-1] * 14.85 + T["CN_O3T2_S"][-1] * 14.85 ** 2)
    np.testing.assert_allclose(b[i], s_hh + T["CN_O3UV_S"][0] / 40800.0 * w3, rtol=1e-13)


def test_o2_near_infrared_tables_at_the_nodes():
    """The 1.27 um band every 2 cm-1 from 7536 (O2, N2 and H2O broadening 1/0.446, 0.3/0.446,
    1) and the N2 overtone every 3 cm-1 from 4325 (all broadeners 1)."""
    l = layer(tave=260.0)
    x_h2o, x_o2 = l.w[1] / l.wtot, l.w[7] / l.wtot
    x_n2 = 1.0 - x_h2o - x_o2
    v = grid(7500.0, 8550.0)
    a = _band(l, 7500.0, 8550.0, 2)
    on, k = _nodes("CN_O2I1_GRID", v)
    fac = (l.w[7] / LOSCHMIDT) * amagat(l) * (x_o2 / 0.446 + 0.3 / 0.446 * x_n2 + x_h2o)
    np.testing.assert_allclose(a[on], fac * T["CN_O2I1_X"][k] / v[on], rtol=1e-13)
    assert on.sum() > 400
    v = grid(4300.0, 4950.0)
    a = _band(l, 4300.0, 4950.0, 3)
    on, k = _nodes("CN_N2O_GRID", v)
    fac = (x_n2 * l.wtot / LOSCHMIDT) * amagat(l) * (x_n2 + x_o2 + x_h2o)
    np.testing.assert_allclose(a[on], fac * T["CN_N2O_X"][k] / v[on], rtol=1e-13)
    assert on.sum() == 201


def test_o2_a_band_midpoints_and_bounds():
    """The A-band table every 1 cm-1 from 12961.5: at the integer ABSRB points XINT is the
    midpoint rule of the four nodes around; pre_xint keeps the points outside the table
    zero (XINT alone would reach 12961 cm-1 from the padded window)."""
    l = layer()
    v1, v2 = 12900.0, 13300.0
    v = grid(v1, v2)
    a = _band(l, v1, v2, 2)
    g = T["CN_O2I3_GRID"]
    vn = g[0] + np.arange(int(g[3]))
    tau = (l.w[7] / LOSCHMIDT) * amagat(l)
    node = np.concatenate([[0.0, 0.0], tau * T["CN_O2I3_X"] / vn, [0.0, 0.0]])   # two zero pads each side
    inner = (v > g[0] + 1) & (v < g[1] - 1)
    j = (v[inner] - g[0] - 0.5).astype(int) + 2         # node index (padded) just below v
    mid = (-node[j - 1] + 9 * node[j] + 9 * node[j + 1] - node[j + 2]) / 16
    np.testing.assert_allclose(a[inner], mid, rtol=1e-12, atol=1e-15 * node.max())
    assert np.all(a[v <= 12961.0] == 0) and np.all(a[v >= 13222.0] == 0)
    assert a[v == 12962.0][0] != 0 and a[v == 13221.0][0] != 0


def test_o2_visible_and_far_uv_at_the_nodes():
    """The visible table every 10 cm-1 from 15000, in (cm2/molec)/atm(O2) units (factor
    1/(L0 1e-20 (55 273/296)^2 89.5)), times the O2 fraction and the O2 column at the
    layer's density; the far UV every 20 cm-1 from 56740 with the Herzberg continuum,
    which CONTNM adds there too, at the same nodes."""
    l = layer(tave=250.0, pave=500.0)
    v = grid(14900.0, 17000.0)
    a = _band(l, 14900.0, 17000.0, 2)
    on, k = _nodes("CN_O2V_GRID", v)
    factor = 1.0 / ((LOSCHMIDT * 1e-20 * (55.0 * 273.0 / 296.0) ** 2) * 89.5)
    adj = (l.w[7] / l.wtot) * l.w[7] * 1e-20 * (l.pave / 1013.0) * (273.0 / l.tave)
    np.testing.assert_allclose(a[on], factor * T["CN_O2V_S"][k] / v[on] * adj, rtol=1e-13)
    assert on.sum() > 200
    v = grid(56700.0, 58000.0)
    a = _band(l, 56700.0, 58000.0, 2)
    # up to 56740 cm-1 the Herzberg continuum alone (v2 <= 56740: no far UV): pre_xint keeps
    # the far-UV window's padding, which XINT alone would carry down to 56721 cm-1, out
    her = _band(l, 56700.0, 56739.0, 2)
    vh = grid(56700.0, 56739.0)
    sel = (vh >= 56721.0) & (vh <= 56735.0)
    assert np.all(her[sel] > 0)
    assert np.array_equal(a[np.isin(v, vh[sel])], her[sel])
    on, k = _nodes("CN_O2U_GRID", v)
    # the Herzberg term at the nodes, from the same formula evaluated independently below
    y = v[on] / 48811.0
    herz = 6.884e-4 * y * np.exp(-69.738 * np.log(y) ** 2) * (1 + 0.83 * (l.pave / 1013.0) * (273.16 / l.tave))
    want = (T["CN_O2U_S"][k] + herz) / v[on] * l.w[7] * 1e-20
    np.testing.assert_allclose(a[on], want, rtol=1e-12)
    assert on.sum() > 50


def test_o2_herzberg_properties():
    """Zero up to 36000 cm-1; the log-normal peak y exp(-a ln^2 y), y = v/48811, lies at
    ln y = 1/(2a) with height 6.884e-4 y exp(-1/(4a)); HERPRS scales it by
    1 + 0.83 (P/1013)(273.16/T) (the ratio of two layers with the same O2 column)."""
    l = layer(tave=280.0, pave=900.0)
    v = grid(35900.0, 38000.0)
    a = _band(l, 35900.0, 38000.0, 2)
    assert np.all(a[v <= 35990.0] == 0) and np.all(a[v >= 36020.0] > 0)
    v = grid(47000.0, 51000.0)
    a = _band(l, 47000.0, 51000.0, 2)
    prs = 1 + 0.83 * (l.pave / 1013.0) * (273.16 / l.tave)
    h = a * v / (l.w[7] * 1e-20) / prs
    on = np.mod(v, 10.0) == 0
    aa = 69.738
    ypk = np.exp(1.0 / (2 * aa))
    vpk = 48811.0 * ypk
    i = np.argmax(np.where(on, h, -1))
    assert abs(v[i] - vpk) <= 5.0
    np.testing.assert_allclose(h[i], 6.884e-4 * ypk * np.exp(-1.0 / (4 * aa)), rtol=1e-6)
    l2 = layer(tave=220.0, pave=200.0)
    l2.w[7] = l.w[7]
    l2.w[1] = l.w[1]
    b = _band(l2, 47000.0, 51000.0, 2)
    prs2 = 1 + 0.83 * (200.0 / 1013.0) * (273.16 / 220.0)
    np.testing.assert_allclose(b[on], a[on] * prs2 / prs, rtol=1e-13)


def test_o2_one_micron_band_area():
    """Above 9439 cm-1 neither Lorentzian of the 1.06 um band is damped: the band times v
    over the O2 amount integrates (trapezoid on its 2 cm-1 nodes) to
    1.054 (0.31831 pi) sum_i S_i (atan((b - v_i)/HW_i) - atan((a - v_i)/HW_i)) / pi."""
    l = layer()
    v = grid(9050.0, 11050.0)
    a = _band(l, 9050.0, 11050.0, 2)
    rho = (l.pave / 1013.0) * (296.0 / l.tave)
    adj = (l.w[7] / l.wtot) / 0.209 * l.w[7] * 1e-20 * rho
    f = a * v / adj
    lo_, hi_ = 9440.0, 10998.0
    sel = (v >= lo_) & (v <= hi_) & (np.mod(v, 2.0) == 0)
    x = v[sel]
    area = np.sum((f[sel][1:] + f[sel][:-1]) * np.diff(x)) / 2
    want = 0.0
    for s_, v0, hw in [(1.166e-4, 9375.0, 58.96), (3.086e-5, 9439.0, 45.04)]:
        want += s_ * (np.arctan((hi_ - v0) / hw) - np.arctan((lo_ - v0) / hw)) / np.pi
    want *= 1.054 * 0.31831 * np.pi
    np.testing.assert_allclose(area, want, rtol=2e-4)
    # the band is zero at its nodes from 9100 cm-1 down and from 11000 cm-1 up (strict
    # bounds); XINT's cubic reaches one point past them
    assert np.all(a[(v <= 9098.0) | (v >= 11002.0)] == 0) and a[v == 9100.0][0] == 0 and a[v == 11000.0][0] == 0


@pytest.mark.parametrize("which,v1,v2", [(2, 57013.0, 58000.0), (2, 36207.0, 37000.0), (1, 9013.0, 10000.0),
                                          (1, 27391.0, 29000.0), (1, 41234.0, 43000.0), (2, 15017.0, 16000.0),
                                          (2, 7541.0, 8000.0), (2, 9205.0, 10000.0), (2, 12970.0, 13210.0),
                                          (3, 4337.0, 4800.0)])
def test_values_do_not_depend_on_where_the_range_starts(which, v1, v2):
    """A band's value at an ABSRB point comes from the four table nodes around it, so a
    range that starts inside the table (the window cut at the node below V1ABS - DV, by
    the routine's truncation) gives the points it shares with a range starting well
    before the band the same values, bit for bit."""
    l = _layer_o3(tave=250.0)
    a = _band(l, v1, v2, which)
    v = grid(v1, v2)
    b = _band(l, v1 - 180.0, v2, which)
    vb = grid(v1 - 180.0, v2)
    inner = v < v2 - 120.0
    assert np.array_equal(a[inner], b[np.isin(vb, v[inner])])
    assert np.abs(a[inner]).max() > 0


@pytest.mark.parametrize("band", [(1, 13000.0, 13600.0, "CN_O3C_GRID"), (2, 7500.0, 8000.0, "CN_O2I1_GRID"),
                                  (3, 4300.0, 4950.0, "CN_N2O_GRID"), (2, 15000.0, 15600.0, "CN_O2V_GRID")])
def test_jrad_one_multiplies_the_nodes_by_radfn(band):
    which, v1, v2, g = band
    l = _layer_o3(tave=250.0)
    v = grid(v1, v2)
    a = _band(l, v1, v2, which, jrad=-1)
    b = _band(l, v1, v2, which, jrad=1)
    c = _band(l, v1, v2, which, jrad=0)
    on, _ = _nodes(g, v)
    np.testing.assert_allclose(b[on], a[on] * radfn(v[on], l.tave), rtol=1e-13)
    assert np.array_equal(a, c)


@pytest.mark.parametrize("jrad", [-1, 0, 1])
def test_rayleigh_closed_form(jrad):
    l = layer()
    v1, v2 = 830.0, 2500.0
    a, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, jrad, (0.0, 0.0, 0.0, 0.0, 1.0))
    x = grid(v1, v2) / 1e4
    r = x ** 3 / (938.076 - 10.8426 * x ** 2) * (l.wtot * 1e-20 / (2.68675e-1 * 1e5))
    if jrad == 1:
        r = r * x
    elif jrad == 0:
        r = r / 1e4
    np.testing.assert_allclose(a, r, rtol=1e-14)
    b, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, 500.0, 819.0, jrad, (0.0, 0.0, 0.0, 0.0, 1.0))
    assert np.all(b == 0)       # below 820 cm-1 CONTNM leaves Rayleigh out


def test_bands_add_and_gates():
    l = layer(tave=260.0)
    v1, v2 = 100.0, 3000.0
    allb, beyond = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, (1.0, 1.0, 1.0, 1.0, 1.0))
    assert beyond == 0
    parts = np.zeros_like(allb)
    for k in (0, 2, 3, 4):
        s = [0.0] * 5
        s[k] = 1.0
        p, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, tuple(s))
        assert p.max() > 0, k
        parts += p
    np.testing.assert_allclose(allb, parts, rtol=1e-13, atol=1e-300)
    none, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, ALL_OFF)
    assert np.all(none == 0)
    # the same over the near infrared, the visible and the ultraviolet, ozone included
    lo3 = _layer_o3(tave=260.0)
    for v1b, v2b in [(4000.0, 9000.0), (9000.0, 14000.0), (25000.0, 30000.0), (38000.0, 43000.0),
                     (53000.0, 58000.0)]:
        allb, rc = O.contnm_ir(lo3.pave, lo3.tave, lo3.wtot, lo3.w, v1b, v2b, 1, (1.0, 1.0, 1.0, 1.0, 1.0))
        assert rc == 0
        parts = np.zeros_like(allb)
        for k in range(5):
            s = [0.0] * 5
            s[k] = 1.0
            parts += O.contnm_ir(lo3.pave, lo3.tave, lo3.wtot, lo3.w, v1b, v2b, 1, tuple(s))[0]
        np.testing.assert_allclose(allb, parts, rtol=1e-12, atol=1e-300)
    # an ABSRB longer than the reference's 5050 points is refused
    assert O.contnm_ir(l.pave, l.tave, l.wtot, l.w, 1000.0, 6100.0, 1, (1.0, 1.0, 1.0, 1.0, 1.0))[1] == -1
    assert O.contnm_ir(l.pave, l.tave, l.wtot, l.w, 1000.0, 6000.0, 1, (1.0, 1.0, 1.0, 1.0, 1.0))[1] == 0
    # the scale factors scale their band
    one, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, (1.0, 0, 0, 0, 0))
    two, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, (2.0, 0, 0, 0, 0))
    np.testing.assert_allclose(two, 2 * one, rtol=1e-15)
