"""Pins of the oracle's layer-OD derivatives (orc_odlay_tan), CPU only.

The oracle differentiates the reference's ODLAY with CNVFNQ's shape tables replaced
by the closed forms they sample (clblm_ODLAY.f90:2804-3016); every other step is the
reference's.  What fixes it:

* central differences of its own OD (channel 0) along each direction -- temperature,
  the broadening column Wtot and each molecule's column -- with small steps, on
  windows that exercise both radiation-term placements (JRAD 0 in panelOut, JRAD 1 in
  the line strengths and on R4), the Van Vleck-Weisskopf mirror of CNVF4Q, line
  coupling (IFLG 1 and 3 with temperature-dependent coefficients), pressure shifts,
  the continuum with its derivative and the DVout re-grid.  The temperature step is
  1e-3 K and the bar 1e-5 of the derivative's peak: a line centre VNU + RHORAT PSHIFT
  near 2000 cm-1 is held to one FP64 ulp (4.5e-13 cm-1, 1e-10 of a grid step), so
  the OD carries a rounding ripple of about 1e-11 that a smaller step would divide
  into the difference; the column directions move no centre (step 1e-6, bar 1e-6);
* closed forms: with self widths equal to air widths a molecule's column does not
  enter the widths, so d tau / d ln W(m) of a one-molecule layer is tau itself and
  d tau / d ln Wtot is zero, exactly; derivatives are linear in the direction;
* the smooth model is the reference's: channel 0 lies within the tables' spacing of
  orc_odlay (the reference's table OD).
"""
import dataclasses

import numpy as np
import pytest

from clblm_amd import workloads as W
from oracle.py import Oracle


def _lines(n, v1, v2, seed, mols, shift=True):
    L = W.make_lines(n, v1, v2, seed=seed, mols=mols)
    if shift:
        rng = np.random.default_rng(seed + 1)
        L.shift = rng.uniform(-0.01, 0.01, L.n).astype(np.float32)
    return L


def _fd(lines, layer, v1, v2, dv, d, h, jrad, **kw):
    """central difference of the smooth OD along d = (dt, dlnwtot, {m: dlnw})"""
    dt, dlw, dq = d

    def moved(s):
        w = {m: v * (1.0 + s * h * dq.get(m, 0.0)) for m, v in layer.w.items()}
        return dataclasses.replace(layer, tave=layer.tave + s * h * dt, wtot=layer.wtot * (1.0 + s * h * dlw), w=w)

    vals = []
    for s in (1.0, -1.0):
        o = Oracle(lines)
        o.jrad = jrad
        vals.append(o.odlay_tan(moved(s), v1, v2, dv, [], **kw)[0][0])
    return (vals[0] - vals[1]) / (2.0 * h)


def _dirs(mols):
    return [(1.0, 0.0, {}), (0.0, 1.0, {})] + [(0.0, 0.0, {m: 1.0}) for m in mols]


CASES = {
    # name: (line range, molecules, layer, window, jrad before the call)
    "jrad0_panel": ((560.0, 700.0), (1, 2), (500.0, 250.0), (600.0, 650.0), 1),
    "jrad1_lines_r4": ((1950.0, 2100.0), (1, 2, 3), (300.0, 240.0), (2000.0, 2040.0), 1),
    "vvw_mirror": ((0.5, 90.0), (1, 3), (800.0, 280.0), (20.0, 45.0), 1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_derivatives_match_central_differences(name):
    (l1, l2), mols, (p, t), (v1, v2), jr = CASES[name]
    lines = _lines(2500, l1, l2, 11, mols)
    layer = W.Layer(pave=p, tave=t, wtot=p * W.AIR_COLUMN_PER_HPA * 0.01,
                    w={m: p * W.AIR_COLUMN_PER_HPA * 0.01 * (1e-3 if m == 1 else 3e-4) for m in mols})
    dv = W.calculate_dv(0.5 * (v1 + v2), p, t, 0.01)
    o = Oracle(lines)
    o.jrad = jr
    dirs = _dirs(mols)
    T, _ = o.odlay_tan(layer, v1, v2, dv, dirs)
    assert (o.jrad == 1) == (name == "jrad1_lines_r4")
    for k, d in enumerate(dirs):
        h, bar = (1e-3, 1e-5) if d[0] else (1e-6, 1e-6)
        fd = _fd(lines, layer, v1, v2, dv, d, h, jr)
        scale = np.max(np.abs(T[1 + k]))
        assert scale > 0.0
        err = np.max(np.abs(fd - T[1 + k])) / scale
        assert err < bar, (name, d, err)


def test_coupled_lines_and_shifts():
    """IFLG 1 (Y, G) and IFLG 3 (GAMMA1, GAMMA2) lines whose coefficients change with
    temperature: the coupling walk, SP/SPP and the reduced widths are differentiated."""
    base = _lines(1200, 760.0, 840.0, 21, (2,))
    idx = list(range(5, 1200, 37))
    lines = W.couple_lines(base, idx, [1 if k % 2 else 3 for k in range(len(idx))],
                           [0.02, 0.015, 0.011, 0.008], [0.004, 0.003, 0.0025, 0.002])
    layer = W.Layer(pave=700.0, tave=265.0, wtot=2e25, w={2: 8e21})
    v1, v2 = 790.0, 810.0
    dv = W.calculate_dv(800.0, layer.pave, layer.tave, 0.0)
    dirs = _dirs((2,))
    o = Oracle(lines)
    T, _ = o.odlay_tan(layer, v1, v2, dv, dirs)
    for k, d in enumerate(dirs):
        h, bar = (1e-3, 1e-5) if d[0] else (1e-6, 1e-6)
        fd = _fd(lines, layer, v1, v2, dv, d, h, 1)
        err = np.max(np.abs(fd - T[1 + k])) / np.max(np.abs(T[1 + k]))
        assert err < bar, (d, err)


def test_continuum_and_regrid():
    """The caller's continuum derivative enters like ABSRB; DVout re-grids every channel."""
    lines = _lines(2000, 880.0, 1020.0, 31, (1,))
    layer = W.Layer(pave=900.0, tave=285.0, wtot=2.2e25, w={1: 4e22})
    v1, v2 = 920.0, 960.0
    dv = W.calculate_dv(940.0, layer.pave, layer.tave, 0.01)
    coef, v1abs, npt = W.continuum_absrb(dataclasses.replace(layer, w={1: 1.0}), v1, v2)
    dirs = [(1.0, 0.0, {}), (0.0, 0.0, {1: 1.0})]
    # the synthetic continuum is linear in the water column and independent of T
    dcont = np.stack([np.zeros(npt), coef * layer.w[1]])
    o = Oracle(lines)
    T, dvu = o.odlay_tan(layer, v1, v2, dv, dirs, dvout=0.75 * dv, continuum=coef * layer.w[1], dcontinuum=dcont)
    assert dvu == 0.75 * dv
    for k, d in enumerate(dirs):
        h, bar = (1e-3, 1e-5) if d[0] else (1e-6, 1e-6)
        dt, _, dq = d
        vals = []
        for s in (1.0, -1.0):
            lay = dataclasses.replace(layer, tave=layer.tave + s * h * dt,
                                      w={1: layer.w[1] * (1.0 + s * h * dq.get(1, 0.0))})
            oo = Oracle(lines)
            vals.append(oo.odlay_tan(lay, v1, v2, dv, [], dvout=0.75 * dv, continuum=coef * lay.w[1])[0][0])
        fd = (vals[0] - vals[1]) / (2.0 * h)
        err = np.max(np.abs(fd - T[1 + k])) / np.max(np.abs(T[1 + k]))
        assert err < bar, (d, err)


def test_column_derivative_closed_form():
    """Self widths equal to air widths: the widths do not depend on the columns, so a
    one-molecule layer has d tau / d ln W = tau and d tau / d ln Wtot = 0 exactly."""
    lines = _lines(3000, 560.0, 700.0, 41, (1,))
    lines.a_self = lines.a_air.copy()
    layer = W.Layer(pave=300.0, tave=230.0, wtot=1e25, w={1: 5e21})
    dv = W.calculate_dv(625.0, layer.pave, layer.tave, 0.0)
    o = Oracle(lines)
    T, _ = o.odlay_tan(layer, 600.0, 650.0, dv, [(0.0, 0.0, {1: 1.0}), (0.0, 1.0, {})])
    scale = np.max(np.abs(T[0]))
    assert np.max(np.abs(T[1] - T[0])) <= 1e-13 * scale
    assert np.max(np.abs(T[2])) <= 1e-13 * scale


def test_linear_in_the_direction():
    lines = _lines(2000, 560.0, 700.0, 51, (1, 2))
    layer = W.Layer(pave=400.0, tave=255.0, wtot=8e24, w={1: 8e21, 2: 3e21})
    dv = W.calculate_dv(625.0, layer.pave, layer.tave, 0.01)
    a, b = 0.7, -2.5
    d1, d2 = (1.0, 0.3, {1: 0.5}), (-0.2, 1.0, {2: 2.0})
    comb = (a * d1[0] + b * d2[0], a * d1[1] + b * d2[1], {1: a * 0.5, 2: b * 2.0})
    T, _ = Oracle(lines).odlay_tan(layer, 600.0, 640.0, dv, [d1, d2, comb])
    scale = np.max(np.abs(T[3]))
    assert np.max(np.abs(a * T[1] + b * T[2] - T[3])) <= 1e-12 * scale


@pytest.mark.parametrize("v1,v2,t", [(600.0, 650.0, 250.0), (2000.0, 2040.0, 240.0)])
def test_smooth_model_within_the_table_spacing_of_the_reference(v1, v2, t):
    """Channel 0 (closed-form shapes at the exact argument) against orc_odlay (the
    reference's nearest-node table lookups): they differ only by the table spacing
    (DXF = 0.002 half-widths in F1), well under 1e-3 of the peak."""
    lines = _lines(2500, v1 - 40.0, v2 + 50.0, 61, (1, 2))
    layer = W.Layer(pave=500.0, tave=t, wtot=1e25, w={1: 1e22, 2: 3e21})
    dv = W.calculate_dv(0.5 * (v1 + v2), layer.pave, layer.tave, 0.01)
    ref, _ = Oracle(lines).odlay(layer, v1, v2, dv)
    T, _ = Oracle(lines).odlay_tan(layer, v1, v2, dv, [])
    assert T.shape[1] == ref.size
    d = np.max(np.abs(T[0] - ref)) / np.max(ref)
    assert 1e-8 < d < 1e-3, d


def test_species_broadening_is_refused():
    lines = _lines(300, 600.0, 700.0, 71, (1,))
    lines.brd_flg = np.ones((lines.n, 7), np.int32)
    lines.brd_hw = np.full((lines.n, 7), 0.05, np.float32)
    lines.brd_tmp = np.full((lines.n, 7), 0.7, np.float32)
    lines.brd_shft = np.zeros((lines.n, 7), np.float32)
    o = Oracle(lines)
    layer = W.Layer(pave=500.0, tave=250.0, wtot=1e25, w={1: 1e22})
    with pytest.raises(RuntimeError, match="species broadening"):
        import ctypes
        from oracle import py as opy
        lay = opy._Layer()
        lay.pave, lay.tave, lay.wtot = 500.0, 250.0, 1e25
        lay.w[1], lay.sel[1] = 1e22, 1
        par = opy._Params(620.0, 640.0, 0.002, 0.002, 4.0, 0.0, 1, 0, None, 0, 1)
        ns, dvu = ctypes.c_int64(), ctypes.c_double()
        if opy.lib().orc_odlay_tan(ctypes.byref(o._L), ctypes.byref(lay), ctypes.byref(par), 0, opy.make_dirs([]),
                                   None, None, ctypes.c_int64(0), ctypes.byref(ns), ctypes.byref(dvu)):
            raise RuntimeError(opy.lib().orc_error().decode())


def test_synthetic_continuum_derivatives():
    """The workloads' continuum states its own derivatives (linear in the H2O column,
    independent of T and Wtot); central differences of the model agree."""
    from clblm_amd.jacobian import continuum_tangents
    lay = W.Layer(pave=600.0, tave=260.0, wtot=1.5e25, w={1: 3e22, 2: 5e21})
    dirs = [(1.0, 0.0, {}), (0.0, 1.0, {}), (0.0, 0.0, {1: 1.0}), (0.0, 0.0, {2: 1.0}), (0.5, -0.2, {1: 2.0})]
    c0, d = W.continuum_absrb_tangents(lay, 700.0, 760.0, dirs)
    c1, dfd = continuum_tangents(lambda l, a, b: W.continuum_absrb(l, a, b), lay, 700.0, 760.0, dirs)
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_allclose(d, dfd, rtol=1e-9, atol=1e-12 * np.abs(c0).max())
