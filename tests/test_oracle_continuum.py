"""The oracle's MT_CKD water-vapour continuum (CONTNM's H2O part with
mt_ckd_h2o_absco, clblm_ODLAY_Continuum.f90:356-404, mt_ckd_h2o_module.f90:27-294),
CPU only.  The reference's module reads its table through netCDF-Fortran, which is
absent here, so it is pinned by what the arithmetic fixes:
* at ABSRB points on the coefficient nodes the 4-point interpolation returns the node,
  so the continuum there is the closed form W_H2O (C_s (T0/T)^n x rho + C_f (1-x) rho)
  times v tanh(c2 v / 2T) (the radiation term, to its 1e-4 branch approximation);
* between nodes the interpolation reproduces linear coefficients exactly (a wrong
  weight or sign fails);
* JRAD 0 drops the radiation term, and JRAD -1 keeps it (CONTNM :365-369);
* CONTNM adds nothing at V1 >= 20000 cm-1 or with both scale factors zero.
The table is the reference's absco-ref_wv-mt-ckd.nc (tests/golden/mtckd_h2o.npz)."""
import os

import numpy as np
import pytest

from clblm_amd.continuum import load_mtckd

HERE = os.path.dirname(os.path.abspath(__file__))
C2 = 1.4387752


@pytest.fixture(scope="module")
def table():
    return load_mtckd(os.path.join(HERE, "golden", "mtckd_h2o.npz"))


def _closed(t, v, pave, tave, w1, wtot, rad=True):
    i = np.searchsorted(t["wavenumbers"], v)
    assert np.all(t["wavenumbers"][i] == v)
    x = w1 / wtot
    rho = (pave / t["ref_press"]) * (t["ref_temp"] / tave)
    s = t["self_absco_ref"][i] * (t["ref_temp"] / tave) ** t["self_texp"][i] * x * rho
    f = t["for_absco_ref"][i] * (1 - x) * rho
    r = v * np.tanh(C2 * v / (2 * tave)) if rad else 1.0
    return w1 * (s + f) * r


@pytest.mark.parametrize("pave,tave,w1,v1,v2", [(1013.0, 288.0, 3e22, 600.0, 1100.0),
                                                (300.0, 230.0, 1e21, 1200.0, 2600.0),
                                                (5.0, 217.0, 1e18, 10.5, 400.0)])
def test_mtckd_at_the_nodes(table, pave, tave, w1, v1, v2):
    from oracle.py import mtckd_h2o
    wtot = 2.1e25 * pave / 1013.0
    a = mtckd_h2o(table, pave, tave, w1, wtot, v1, v2, jrad=1)
    v1abs = float(int(v1)) - 3.0
    v = v1abs + np.arange(a.size)
    assert a.size == int(int(v2 + 3.5) - v1abs + 1.5)
    node = np.mod(v, 10.0) == 0
    want = _closed(table, v[node], pave, tave, w1, wtot)
    np.testing.assert_allclose(a[node], want, rtol=1e-4)
    assert np.all(a > 0)
    # without the radiation term (JRAD 0) the nodes are the closed form to rounding
    b = mtckd_h2o(table, pave, tave, w1, wtot, v1, v2, jrad=0)
    np.testing.assert_allclose(b[node], _closed(table, v[node], pave, tave, w1, wtot, rad=False), rtol=1e-13)


def test_mtckd_interpolation_reproduces_linear_coefficients():
    from oracle.py import mtckd_h2o
    wn = -20.0 + 10.0 * np.arange(2003)
    t = dict(wavenumbers=wn, self_absco_ref=1e-22 * (3.0 + 0.002 * wn), for_absco_ref=1e-23 * (5.0 - 0.0001 * wn),
             self_texp=np.zeros_like(wn), ref_temp=296.0, ref_press=1013.0)
    a = mtckd_h2o(t, 1013.0, 296.0, 1e22, 1e25, 611.3, 1389.9, jrad=0)
    v = float(int(611.3)) - 3.0 + np.arange(a.size)
    x = 1e22 / 1e25
    want = 1e22 * (1e-22 * (3.0 + 0.002 * v) * x + 1e-23 * (5.0 - 0.0001 * v) * (1 - x))
    np.testing.assert_allclose(a, want, rtol=1e-13)


def test_mtckd_jrad_and_gates(table):
    from oracle.py import mtckd_h2o
    args = (table, 800.0, 270.0, 1e22, 1.7e25, 700.0, 900.0)
    one, minus, zero = (mtckd_h2o(*args, jrad=j) for j in (1, -1, 0))
    assert np.array_equal(one, minus)                  # JRAD -1 keeps the radiation term
    assert not np.allclose(one, zero)
    assert np.all(mtckd_h2o(table, 800.0, 270.0, 1e22, 1.7e25, 20000.0, 20010.0, 1) == 0)
    assert np.all(mtckd_h2o(*args, jrad=1, xself=0.0, xfrgn=0.0) == 0)
    s = mtckd_h2o(*args, jrad=1, xself=1.0, xfrgn=0.0)
    f = mtckd_h2o(*args, jrad=1, xself=0.0, xfrgn=1.0)
    np.testing.assert_allclose(s + f, one, rtol=1e-15)
    assert np.all(f > s)           # at a volume mixing ratio of 6e-4 the foreign part dominates


def test_mtckd_magnitude():
    """The self-continuum absorption coefficient of water vapour in the 10 um window at
    296 K is of order 1e-22 cm2 molec-1 at 1 atm of pure vapour (the table's
    coefficients times the radiation term, as its notes say): a pure layer of 1e22 cm-2
    at 1013 hPa gives an OD of order 1 at 1000 cm-1."""
    from oracle.py import mtckd_h2o
    t = load_mtckd(os.path.join(HERE, "golden", "mtckd_h2o.npz"))
    a = mtckd_h2o(t, 1013.0, 296.0, 1e22, 1e22, 999.0, 1001.0, jrad=1)
    c = a[4] / 1e22                                     # v = 1000, self only (x = 1)
    assert 0.5e-22 < c < 5e-22


REF_NC = "/root/reference/clblm_data/spectroscopy/absco-ref_wv-mt-ckd.nc"


@pytest.mark.skipif(not os.path.exists(REF_NC), reason="the reference's data directory is absent")
def test_library_reads_the_reference_netcdf_file():
    """clblm_mtckd_read_nc (the library's own netCDF classic reader) returns the arrays
    and scalars scipy reads from the reference's coefficient file, bit for bit."""
    import ctypes
    from clblm_amd import _lib
    lib = _lib.load()
    t = load_mtckd(REF_NC)
    n = ctypes.c_int()
    assert lib.clblm_mtckd_read_nc(REF_NC.encode(), ctypes.byref(n), None, None, None, None, None, None) == 0
    arr = [np.zeros(n.value) for _ in range(4)]
    rt, rp = ctypes.c_double(), ctypes.c_double()
    assert lib.clblm_mtckd_read_nc(REF_NC.encode(), ctypes.byref(n), *[a.ctypes.data for a in arr],
                                   ctypes.byref(rt), ctypes.byref(rp)) == 0
    for a, k in zip(arr, ("wavenumbers", "self_absco_ref", "for_absco_ref", "self_texp")):
        assert np.array_equal(a, t[k]), k
    assert rt.value == t["ref_temp"] and rp.value == t["ref_press"]
    assert lib.clblm_mtckd_read_nc(os.path.join(HERE, "golden", "mtckd_h2o.npz").encode(), ctypes.byref(n),
                                   None, None, None, None, None, None) != 0
