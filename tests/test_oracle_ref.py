"""Pins the oracle against the reference's own compiled Fortran (oracle/_ref).

The reference's ODLAY needs netCDF and cannot be built here; the modules that do
build from its sources alone are compiled by `make -C oracle ref` and compared
with the oracle bit for bit: zeta tables, TIPS/MOLEC (partition-function ratio,
Doppler factor, self-broadening fraction), RADFN, RADFNI, XINT, interpSpectrum,
planck, planckFnDrv, rlin.  Skipped where the reference sources are absent."""
import ctypes
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libclblm_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")

D = ctypes.c_double
PD = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def ref():
    # the reference's timing helper uses ITIME, which flang's runtime lacks; it is never called
    lib = ctypes.CDLL(REF, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    lib.ref_radfn.restype = D
    lib.ref_radfn.argtypes = [D, D]
    lib.ref_radfni.restype = D
    lib.ref_radfni.argtypes = [D, D, D, PD, PD, PD]
    lib.ref_planck.restype = D
    lib.ref_planck.argtypes = [D, D]
    lib.ref_planck_drv.argtypes = [D, D, PD, PD]
    lib.ref_rlin.argtypes = [D, PD, PD]
    lib.ref_molec.argtypes = [ctypes.c_int, D, D, D, D, PD, PD, PD]
    lib.ref_xint.argtypes = [D, D, D, ctypes.c_int, PD, D, D, ctypes.c_int, PD, ctypes.c_int, ctypes.c_int]
    lib.ref_interp.argtypes = [ctypes.c_int, PD, D, D, ctypes.c_int, PD, D, D]
    return lib


@pytest.fixture(scope="module")
def orc():
    from oracle import py
    lib = py.lib()
    lib.orc_radfni.restype = D
    lib.orc_radfni.argtypes = [D, D, D, PD, PD, PD]
    lib.orc_planck_drv.argtypes = [D, D, PD, PD]
    lib.orc_rlin.argtypes = [D, PD, PD]
    return lib


def _p(a):
    return a.ctypes.data_as(PD)


def test_zeta_tables_bitwise(ref):
    out = np.zeros((6, 102))
    ref.ref_zeta_tables(_p(out))
    tab = json.load(open(os.path.join(ROOT, "tests", "golden", "tables.json")))["zeta"]
    for k, name in enumerate(["AVRAT", "CGAUSS", "CF1", "CF2", "CF3", "CER"]):
        np.testing.assert_array_equal(np.array(tab[name]), out[k], err_msg=name)


@pytest.mark.parametrize("mol", range(1, 8))
def test_molec_tips_bitwise(ref, mol):
    """SCOR = Q(296)/Q(T), ALFD1 and RHOSLF for every isotopologue of molecules 1-7."""
    from oracle.py import _Layer, lib
    for t in (180.0, 217.0, 250.3, 296.0, 312.7):
        sc, ad, rh = np.zeros(10), np.zeros(10), D()
        ref.ref_molec(mol, 500.0, t, 1e21, 1e24, _p(sc), _p(ad), ctypes.byref(rh))
        lay = _Layer()
        lay.pave, lay.tave, lay.wtot = 500.0, t, 1e24
        lay.w[mol], lay.sel[mol] = 1e21, 1
        osc, orh, oad = np.zeros(48 * 11), np.zeros(48), np.zeros(48 * 11)
        assert lib().orc_molec(ctypes.byref(lay), _p(osc), _p(orh), _p(oad)) == 0
        niso = json.load(open(os.path.join(ROOT, "tests", "golden", "tables.json")))["maxiso"][mol - 1]
        np.testing.assert_array_equal(osc[mol * 11 + 1: mol * 11 + 1 + niso], sc[:niso])
        np.testing.assert_array_equal(oad[mol * 11 + 1: mol * 11 + 1 + niso], ad[:niso])
        assert orh[mol] == rh.value


def test_radfn_radfni_bitwise(ref, orc):
    rng = np.random.default_rng(3)
    for _ in range(300):
        v = float(rng.uniform(1.0, 3500.0))
        t = float(rng.uniform(150.0, 320.0))
        xkt = t / 1.4387752
        assert ref.ref_radfn(v, xkt) == orc.orc_radfn(v, xkt)
        dv = float(10 ** rng.uniform(-4, -1))
        a = [D(v - dv), D(), D(-1.0)]
        b = [D(v - dv), D(), D(-1.0)]
        for _step in range(5):
            ra = ref.ref_radfni(v, dv, xkt, ctypes.byref(a[0]), ctypes.byref(a[1]), ctypes.byref(a[2]))
            rb = orc.orc_radfni(v, dv, xkt, ctypes.byref(b[0]), ctypes.byref(b[1]), ctypes.byref(b[2]))
            assert (ra, a[0].value, a[1].value, a[2].value) == (rb, b[0].value, b[1].value, b[2].value)
            v = a[0].value


def test_xint_bitwise(ref):
    from oracle.py import xint
    rng = np.random.default_rng(4)
    for _ in range(20):
        dva = float(rng.uniform(0.5, 1.5))
        na = 200
        v1a = float(rng.uniform(400, 600))
        v2a = v1a + dva * (na - 1)
        a = rng.normal(size=na)
        dvr3 = dva / float(rng.uniform(3, 40))
        vft = v1a + float(rng.uniform(0, 20))
        r3 = rng.normal(size=200)
        r3_ref = r3.copy()
        ref.ref_xint(v1a, v2a, dva, na, _p(a), vft, dvr3, r3_ref.size, _p(r3_ref), 1, 155)
        a1 = np.concatenate([[0.0], a])
        r3_orc = np.concatenate([[0.0], r3])
        xint(v1a, v2a, dva, a1, vft, dvr3, r3_orc, 1, 155)
        np.testing.assert_array_equal(r3_orc[1:], r3_ref)


@pytest.mark.parametrize("ratio", [1.0, 0.5, 2.0, 2.0 / 3.0, 1.25, 0.37, 1.7315])
def test_interp_bitwise(ref, ratio):
    from oracle.py import interp
    rng = np.random.default_rng(5)
    nin = 6000
    dvin = 0.0123
    x = np.cumsum(rng.normal(size=nin)) + 100.0
    dvout = dvin * ratio
    nout = int((nin - 3) * dvin / dvout)
    out_ref = np.zeros(nout)
    ref.ref_interp(nin, _p(x), dvin, 600.0, nout, _p(out_ref), dvout, 600.0)
    out_orc = interp(x, dvin, 600.0, dvout, 600.0, nout)
    np.testing.assert_array_equal(out_orc, out_ref)


def test_planck_rlin_bitwise(ref, orc):
    rng = np.random.default_rng(6)
    for _ in range(200):
        v = float(rng.uniform(10, 3500))
        t = float(rng.uniform(150, 320))
        assert ref.ref_planck(v, t) == orc.orc_planck(v, t)
        b1, d1, b2, d2 = D(), D(), D(), D()
        ref.ref_planck_drv(v, t, ctypes.byref(b1), ctypes.byref(d1))
        orc.orc_planck_drv(v, t, ctypes.byref(b2), ctypes.byref(d2))
        assert (b1.value, d1.value) == (b2.value, d2.value)
    for od in [0.0, 1e-5, 9e-4, 1e-3, 0.5, 3.0, 19.9, 20.0, 35.0]:
        h1, g1, h2, g2 = D(), D(), D(), D()
        ref.ref_rlin(od, ctypes.byref(h1), ctypes.byref(g1))
        orc.orc_rlin(od, ctypes.byref(h2), ctypes.byref(g2))
        assert (h1.value, g1.value) == (h2.value, g2.value)
