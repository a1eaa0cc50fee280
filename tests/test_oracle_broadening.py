"""Species-by-species broadening in the oracle (RDLIN clblm_ODLAY.f90:2697-2747, LNCORQ
:1849-1883, RDLNFL / LINF4Q clblm_ODLAY_LineF4.f90:1446-1468,994-1054).  The
reference's ODLAY needs netCDF and is not built here; the checks below pin the
restatement by equivalent plain lines.  Widths and centres that are equal in exact
arithmetic still round differently, and the table indices truncate, so these compare
to 1e-3 of the peak: far below what any dropped or misplaced term changes (the
broadener below carries 30 % of the column)."""
import numpy as np
import pytest

from clblm_amd import workloads as W
from oracle.py import Oracle

P0, T0 = 1013.25, 296.0


def _lone(mol=1, v0=650.0):
    L = W.make_lines(2, 100.0, 200.0, seed=3)
    L.nu[:] = [v0, 1500.0]
    L.s[:] = np.float32(1e-20)
    L.a_air[:] = np.float32(0.07)
    L.a_self[:] = np.float32(0.35)
    L.e_lo[:] = np.float32(200.0)
    L.tmpalf[:] = np.float32(0.25)
    L.shift[:] = np.float32(-0.004)
    L.mol[:] = np.int32(mol + 100)
    return L


def _brd(L, flg=None, hw=None, tmp=None, shft=None):
    n = L.n
    L.brd_flg = np.zeros((n, 7), np.int32) if flg is None else np.asarray(flg, np.int32).reshape(n, 7)
    L.brd_hw = np.zeros((n, 7), np.float32) if hw is None else np.asarray(hw, np.float32).reshape(n, 7)
    L.brd_tmp = np.zeros((n, 7), np.float32) if tmp is None else np.asarray(tmp, np.float32).reshape(n, 7)
    L.brd_shft = np.zeros((n, 7), np.float32) if shft is None else np.asarray(shft, np.float32).reshape(n, 7)
    return L


def _layer(mol=1, p=P0, t=T0):
    # the line molecule with 1 % of the column, CO2 (the broadener) with 30 %
    return W.Layer(pave=p, tave=t, wtot=1e24, w={mol: 1e22, 2: 3e23} if mol != 2 else {2: 1e22})


def _od(L, lay, sb):
    o = Oracle(L)
    o.jrad = -1
    # DV 0.01: ALFMAX = 4 SAMPLE DV = 0.16 stays above every width here
    return o.odlay(lay, 600.0, 700.0, 0.01, species_broad=sb)[0]


def _close(a, b, tol=1e-3):
    return float(np.abs(a - b).max() / np.abs(b).max()) <= tol


def test_zero_flags_change_nothing():
    rng = np.random.default_rng(4)
    lines = W.make_lines(3000, 560.0, 740.0, seed=5, mols=(1, 2, 3, 7))
    n = lines.n
    _brd(lines, None, rng.uniform(0.01, 0.1, (n, 7)), rng.uniform(0.5, 0.8, (n, 7)), rng.uniform(-0.01, 0.01, (n, 7)))
    lay = W.Layer(pave=506.6, tave=260.0, wtot=1e24, w={1: 1e22, 2: 3e23, 3: 1e19, 7: 2e23})
    a = _od(lines, lay, True)
    b = _od(lines, lay, False)
    assert np.array_equal(a, b)


def test_broadener_with_the_air_width_is_the_plain_line():
    L = _lone()
    f = np.zeros((2, 7)); f[0, 1] = 1                  # CO2 broadens line 0
    hw = np.zeros((2, 7)); hw[0, 1] = L.a_air[0]
    tm = np.zeros((2, 7)); tm[0, 1] = L.tmpalf[0]
    sh = np.zeros((2, 7)); sh[0, 1] = L.shift[0]
    lay = _layer(1, 700.0, 250.0)
    plain = _od(_lone(), lay, True)
    got = _od(_brd(L, f, hw, tm, sh), lay, True)
    assert _close(got, plain)
    # a broadener twice as wide widens the line (not the plain result)
    hw2 = hw.copy(); hw2[0, 1] *= 2
    wider = _od(_brd(_lone(), f, hw2, tm, sh), lay, True)
    assert not _close(wider, plain, 1e-2)


def test_species_shift_moves_the_centre():
    lay = _layer(1, P0, T0)                            # RHORAT = 1
    delta = 0.02
    L = _lone()
    f = np.zeros((2, 7)); f[0, 1] = 1
    hw = np.zeros((2, 7)); hw[0, 1] = L.a_air[0]
    tm = np.zeros((2, 7)); tm[0, 1] = L.tmpalf[0]
    sh = np.zeros((2, 7)); sh[0, 1] = np.float32(L.shift[0] + delta)
    got = _od(_brd(L, f, hw, tm, sh), lay, True)
    rho = 0.3                                          # CO2's share of the column
    moved = _lone()
    moved.shift[0] = np.float32(float(moved.shift[0]) + rho * (float(sh[0, 1]) - float(moved.shift[0])))
    assert _close(got, _od(moved, lay, False))
    assert not _close(got, _od(_lone(), lay, False), 1e-2)


def test_flag_minus_one_uses_the_self_width():
    lay = _layer(1, P0, T0)                            # TRATIO = RHORAT = 1: HWHMSI = a_self
    f = np.zeros((2, 7)); f[0, 0] = -1
    got = _od(_brd(_lone(), f), lay, True)
    self_only = _lone()
    self_only.a_air[0] = self_only.a_self[0]           # ALFL = a_self (1 - S) + a_self S
    assert _close(got, _od(self_only, lay, False))
    assert not _close(got, _od(_lone(), lay, False), 1e-2)


def test_o2_shift_is_read_as_the_foreign_shift():
    lay = W.Layer(pave=P0, tave=T0, wtot=1e24, w={7: 2.1e23})
    L = _lone(mol=7)
    f = np.zeros((2, 7)); f[0, 6] = 1
    sh = np.zeros((2, 7)); sh[0, 6] = 0.01
    got = _od(_brd(L, f, None, None, sh), lay, False)  # read-time correction, no broadening
    ref = _lone(mol=7)
    ref.shift[0] = np.float32((float(ref.shift[0]) - 0.21 * float(np.float32(0.01))) / 0.79)
    assert _close(got, _od(ref, lay, False))
    assert not _close(got, _od(_lone(mol=7), lay, False), 1e-2)
