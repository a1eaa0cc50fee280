"""The oracle's CONTNM bands other than water vapour over the thermal infrared
(orc_contnm_ir: CO2 with FRNCO2 and the XFACCO2 scaling, the O2 collision-induced
fundamental, the N2 collision-induced pure rotation and fundamental, Rayleigh
extinction; clblm_ODLAY_Continuum.f90:408-1060), CPU only.  The reference's
Module_Continuum uses mt_ckd_h2o, which reads its table through netCDF-Fortran, absent
here, so the restatement is pinned by what the arithmetic fixes:
* the generated tables (clblm_amd/csrc/clblm_contnm.inc) hold the values of the
  reference's DATA statements, read back here by a parser of their own;
* at ABSRB points on a band's table nodes XINT returns the node, so the band there is
  the closed form of the table value at that wavenumber times CONTNM's factors (the CO2
  table every 2 cm-1 from -4, XFACCO2 from 2000 cm-1, the band-head temperature
  exponents at 2386-2434 cm-1; the O2 fundamental every 5 cm-1 from 1340; the N2
  rotation every 5 cm-1 from -10, at the two reference temperatures);
* the N2 fundamental (nodes off the integer grid) scales exactly with the O2
  broadening efficiency 1.294 - 0.4545 T/296, and equals its 272 K table at 272 K;
* Rayleigh extinction is its closed form at every point, for every JRAD;
* JRAD 1 multiplies each node by RADFN, JRAD 0 leaves it out;
* the bands are additive (each alone sums to all together) and a zero factor or a
  range outside a band adds nothing; the bands the restatement leaves out are flagged.
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


@pytest.mark.skipif(not os.path.exists(REF), reason="reference sources absent")
def test_tables_hold_the_reference_data():
    """Every value of the generated tables against the reference's DATA text, read by a
    scan of its own: the numbers between `DATA <name>...` and the closing slash."""
    src = open(REF).read()
    lines = [ln.split("!")[0] for ln in src.splitlines()]
    text = "\n".join(lines)
    subs = {}
    for m in re.finditer(r"(?im)^\s*(subroutine)\s+(\w+)", text):
        end = text.lower().find("end subroutine", m.end())
        subs[m.group(2).lower()] = text[m.start():end]
    want = {"CN_XFACCO2": ("contnm", "XFACCO2"), "CN_CO2_S": ("frnco2", "S"),
            "CN_CO2_TDEP": ("frnco2", "tdep_bandhead"), "CN_N2R_C296": ("xn2_r", "C_296"),
            "CN_N2R_SF296": ("xn2_r", "sf_296"), "CN_N2R_C220": ("xn2_r", "C_220"),
            "CN_N2R_SF220": ("xn2_r", "sf_220"), "CN_N2F_X272": ("n2_ver_1", "xn2_272"),
            "CN_N2F_X228": ("n2_ver_1", "xn2_228"), "CN_N2F_AH2O": ("n2_ver_1", "a_h2o"),
            "CN_O2F_X": ("o2_ver_1", "xo2"), "CN_O2F_XT": ("o2_ver_1", "xo2t")}
    for cname, (sub, name) in want.items():
        body = subs[sub]
        k = re.search(r"(?i)data\s*\(?\s*%s(\s*\(|\s*/)" % name, body).start()
        seg = body[body.index("/", k) + 1:]
        seg = seg[:seg.index("/")]
        vals = np.array([float(x) for x in seg.replace("&", " ").replace(",", " ").split()])
        assert vals.size == T[cname].size, cname
        assert np.array_equal(vals, T[cname]), cname
    assert list(T["CN_CO2_GRID"]) == [-4.0, 10000.0, 2.0, 5003.0]
    assert list(T["CN_O2F_GRID"]) == [1340.0, 1850.0, 5.0, 103.0]
    assert list(T["CN_N2R_GRID"]) == [-10.0, 350.0, 5.0, 73.0]
    assert T["CN_CO2_TEFF"][0] == 246.0


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
    # the bands left out are flagged where they would add, and only then
    for v1b, v2b, sc in [(4000.0, 4500.0, (0, 0, 0, 1.0, 0)), (7000.0, 7600.0, (0, 0, 1.0, 0, 0)),
                         (8800.0, 9000.0, (0, 1.0, 0, 0, 0))]:
        assert O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1b, v2b, 1, sc)[1] == 1
        assert O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1b, v2b, 1, ALL_OFF)[1] == 0
    # the scale factors scale their band
    one, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, (1.0, 0, 0, 0, 0))
    two, _ = O.contnm_ir(l.pave, l.tave, l.wtot, l.w, v1, v2, 1, (2.0, 0, 0, 0, 0))
    np.testing.assert_allclose(two, 2 * one, rtol=1e-15)
