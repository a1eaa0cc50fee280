"""Exact properties of the nested-grid line sum, shared by the oracle pins
(test_oracle_pins.py, CPU) and the same checks on the HIP path (test_gpu_pins.py).

The reference's ODLAY cannot be built here (netCDF), so the parts of the oracle
that only a reference run could compare bit for bit are pinned by what the
algorithm itself fixes exactly:

* translation invariance: on a binary-exact grid, moving V1 by a multiple of 64
  points (one R4 step, DVR4 = 64 DV) keeps every grid point, the R4 grid and
  each line's offsets, but moves every 2400-point panel seam relative to the
  lines.  The OD must come out the same away from the new V1 (bit for bit in the
  sequential oracle), so a wrong home panel, a lost or doubled early-R3/R2 carry
  or a wrong seam interpolation fails (clblm_ODLAY.f90:2067-2159,2289-2545);
* the CO2 far-wing bound (clblm_ODLAY_LineF4.f90:388-427): in the Lorentz zone a
  lone CO2 line minus an identical N2O line (same mass class, so no other
  difference survives beyond the line core) equals -(1 - x^2/625) F4BND, a
  quadratic that XINT and the panel interpolation carry exactly;
* SHRINQ (clblm_ODLAY_LineF4.f90:1516-1647): lines farther apart than DVR4/2, and
  a CO2 line beside a non-CO2 line at the same centre, are never merged, so the
  OD is additive to rounding; two lines of one species closer than DVR4/2 are
  merged at their strength-weighted centre, which leaves an error of second
  order in their separation;
* the O2 air-width rescale on read (RDLIN clblm_ODLAY.f90:2728-2738, RDLNFL
  clblm_ODLAY_LineF4.f90:1463-1468): the Lorentz width read back from the far
  wing of a lone O2 line is (a_air - 0.21 a_self)/0.79 weighted as usual, and
  that of a CO line is not rescaled;
* LINF4Q's buffer (clblm_ODLAY_LineF4.f90:1172-1243): SHRINQ runs on chunks of
  whole TAPE3 records, and a close pair split by a chunk end stays unmerged.

Every case takes ``od(lines, layer, v1, v2, dv)``, a function returning the layer
OD with no radiation term (JRAD = -1), and ``exact``: whether equal sums must be
equal bit for bit (the oracle) or to ``EXACT_GPU`` of the spectrum's maximum (the
HIP pass 1 sums with atomics, so its order of additions varies).
"""
import numpy as np
from scipy.optimize import brentq

from clblm_amd import workloads as W

C2 = 1.4387752
DV = 2.0 ** -8          # binary exact: every V1 + k DV below is exact
EXACT_GPU = 1e-13
FIELDS = ("nu", "s", "a_air", "a_self", "e_lo", "tmpalf", "shift", "mol", "iflg")


def _same(a, b, scale, exact):
    if exact:
        bad = np.nonzero(a != b)[0]
        return bad.size == 0, bad[:5], (np.abs(a - b).max() / scale if bad.size else 0.0)
    d = np.abs(a - b).max() / scale
    return d <= EXACT_GPU, None, d


def spaced(v1, v2, spacing, seed, mols):
    n = int((v2 - v1) / spacing)
    L = W.make_lines(n, v1, v2, seed=seed, mols=mols)
    L.nu = v1 + spacing * np.arange(n) + 0.013
    return L


def translation_invariance(od, p, exact):
    """Lines 0.28 cm-1 apart (> DVR4/2 = 0.125, so SHRINQ merges none), three species
    (CO2 among them), V1 moved by 1216, 6272 and 9664 points: the seams move by
    1216, 1472 and 64 points of a 2400-point panel."""
    lines = spaced(650.0, 810.0, 0.28, 5, (1, 2, 3))
    lay = W.Layer(pave=p, tave=250.0, wtot=1e24, w={1: 1e21, 2: 4e20, 3: 1e18})
    base = od(lines, lay, 700.0, 760.0, DV)
    for k in (1216, 6272, 9664):
        o = od(lines, lay, 700.0 + k * DV, 760.0, DV)
        assert o.size == base.size - k
        # the first R4 step after the new V1 sees the truncated line cores of the
        # first panel (the reference's own start-of-spectrum edge); beyond it the
        # spectra are the same numbers
        ok, where, d = _same(base[k + 64:], o[64:], np.abs(base).max(), exact)
        assert ok, (p, k, where, d)
        edge = np.abs(base[k:k + 64] - o[:64]).max() / np.abs(base).max()
        assert edge < 1e-11, (p, k, edge)


def translation_off_r4_grid(od):
    """A shift that is not a multiple of 64 points moves the R4 grid against the
    lines: the result changes (so the invariance above is not vacuous), but only
    by the far-wing interpolation error."""
    lines = spaced(650.0, 810.0, 0.28, 5, (1, 2, 3))
    lay = W.Layer(pave=100.0, tave=250.0, wtot=1e24, w={1: 1e21, 2: 4e20, 3: 1e18})
    base = od(lines, lay, 700.0, 760.0, DV)
    o = od(lines, lay, 700.0 + 1000 * DV, 760.0, DV)
    d = np.abs(base[1000 + 64:] - o[64:]).max() / np.abs(base).max()
    assert 1e-9 < d < 1e-3, d


def lone(mol, nus, ss, a_air=0.07, a_self=0.3):
    n = len(nus)
    L = W.make_lines(n, 700.0, 750.0, seed=3, mols=(1,))
    L.nu[:] = nus
    L.s[:] = np.asarray(ss, np.float32)
    L.mol[:] = np.asarray(mol, np.int32) + 100
    L.a_air[:] = np.float32(a_air)
    L.a_self[:] = np.float32(a_self)
    L.tmpalf[:] = np.float32(0.25)
    L.e_lo[:] = np.float32(500.0)
    return L


def co2_far_wing_bound(od):
    """CNVF4Q subtracts (2 - x^2/25^2) F4BND from a CO2 line and F4BND from any other,
    F4BND = S a_L / pi / (a_L^2 + 25^2).  At 296 K the line strength and width do not
    depend on the molecule, and beyond the line core only F4 contributes, so
    OD(CO2) - OD(N2O) = -(1 - x^2/625) F4BND exactly, for 3 < |x| < 24 cm-1."""
    nu0, p, T, w = 725.3, 100.0, 296.0, 1e21
    ods = {}
    for mol in (2, 4):
        lay = W.Layer(pave=p, tave=T, wtot=1e24, w={mol: w})
        ods[mol] = od(lone([mol] * 3, [701.0, nu0, 749.0], [0, 1e-21, 0]), lay, 700.5, 750.0, DV)
    x = 700.5 + DV * np.arange(ods[2].size) - nu0
    rho = p / 1013.25
    rs = rho * w / 1e24
    a_l = float(np.float32(0.07)) * (rho - rs) + float(np.float32(0.3)) * rs
    s = float(np.float32(1e-21)) * w * (1.0 + np.exp(-C2 * nu0 / T))
    f4bnd = s * a_l / np.pi / (a_l * a_l + 625.0)
    want = -(1.0 - x * x / 625.0) * f4bnd
    zone = (np.abs(x) > 3.0) & (np.abs(x) < 24.0)
    err = np.abs((ods[2] - ods[4])[zone] - want[zone]).max() / f4bnd
    assert err < 1e-12, err
    # the core differs only through the Doppler widths (masses 43.99 vs 44.00)
    core = np.abs(x) < 1.0
    assert np.abs((ods[2] - ods[4])[core]).max() < 1e-5 * np.abs(ods[2]).max()


def _lay():
    return W.Layer(pave=100.0, tave=250.0, wtot=1e24, w={1: 1e21, 2: 1e21, 4: 1e21})


def _padded(mols, nus, ss):
    # lines of zero strength at 701 and 749 cm-1 are rejected by LNCORQ and LINF4Q
    # alike; they keep every list at two lines or more (the library's minimum)
    return lone([1] + mols + [1], [701.0] + nus + [749.0], [0.0] + ss + [0.0])


def _pair(od, mols, nus, ss):
    lay = _lay()
    ab = od(_padded(mols, nus, ss), lay, 710.0, 740.0, DV)
    a = od(_padded(mols[:1], nus[:1], ss[:1]), lay, 710.0, 740.0, DV)
    b = od(_padded(mols[1:], nus[1:], ss[1:]), lay, 710.0, 740.0, DV)
    return ab, a, b


def shrinq_far_pairs_apart(od, d):
    ab, a, b = _pair(od, [1, 1], [725.3, 725.3 + d], [1e-21, 3e-21])
    assert np.abs(ab - a - b).max() < 1e-14 * np.abs(ab).max()


def shrinq_co2_apart(od):
    ab, a, b = _pair(od, [2, 1], [725.3, 725.3], [1e-21, 3e-21])
    assert np.abs(ab - a - b).max() < 1e-14 * np.abs(ab).max()
    # two non-CO2 species at one centre are merged (their Doppler widths differ)
    ab, a, b = _pair(od, [1, 4], [725.3, 725.3], [1e-21, 3e-21])
    assert np.abs(ab - a - b).max() > 1e-9 * np.abs(ab).max()


def shrinq_merge_second_order(od):
    """Strengths 1:3, equal widths, separations 0.08 .. 0.01 cm-1 (< DVR4/2): the
    merged line sits at the strength-weighted centre, so the far-wing error falls
    by 4 per halving (an unweighted centre would leave a first-order term)."""
    errs = []
    for d in (0.08, 0.04, 0.02, 0.01):
        ab, a, b = _pair(od, [1, 1], [725.3, 725.3 + d], [1e-21, 3e-21])
        x = 710.0 + DV * np.arange(ab.size) - 725.3
        far = np.abs(x) > 2.0
        errs.append(np.abs(ab - a - b)[far].max() / np.abs(ab).max())
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert all(3.6 < r < 4.8 for r in ratios), (errs, ratios)
    assert errs[0] < 1e-7


def o2_air_width_rescale(od, mol, rescaled):
    """The far wing of a lone line at 296 K is S a_L/pi (1/(a_L^2+x^2) - 1/(a_L^2+625)):
    solving it for a_L at two offsets gives the width that was used."""
    nu0, p, T, w = 725.3, 100.0, 296.0, 2e23
    lines = lone([mol] * 3, [701.0, nu0, 749.0], [0, 1e-21, 0], a_air=0.06, a_self=0.05)
    lay = W.Layer(pave=p, tave=T, wtot=1e24, w={mol: w})
    o = od(lines, lay, 700.5, 750.0, DV)
    x = 700.5 + DV * np.arange(o.size) - nu0
    s = float(np.float32(1e-21)) * w * (1.0 + np.exp(-C2 * nu0 / T))
    rho = p / 1013.25
    rs = rho * w / 1e24
    aa, asf = float(np.float32(0.06)), float(np.float32(0.05))
    a_air = (aa - 0.21 * asf) / 0.79 if rescaled else aa
    want = a_air * (rho - rs) + asf * rs
    for xx in (10.0, 18.0):
        i = int(np.argmin(np.abs(x - xx)))
        xv = x[i]
        got = brentq(lambda a: s * a / np.pi * (1 / (a * a + xv * xv) - 1 / (a * a + 625.0)) - o[i], 1e-5, 1.0,
                     xtol=1e-16)
        assert abs(got / want - 1.0) < 2e-5, (mol, xv, got, want)


def linf4q_buffer_chunks(od, rejected, apart):
    """LINF4Q (clblm_ODLAY_LineF4.f90:1172-1243) shrinks its 1000-line buffer in
    chunks of whole TAPE3 records: the first chunk ends at the first record end with
    1000 accepted lines; after SHRINQ the buffer is written once it holds 1000 shrunk
    lines, else the next chunk is appended record by record.  Lines of different
    chunks are never merged.  1400 lines 0.2 cm-1 apart (none merge) carry six pairs
    0.02 cm-1 apart, each opened by a line 0.1 cm-1 below it, so SHRINQ merges a pair
    unless a chunk boundary separates it.
    rejected = 0: chunk 1 is records 1-4 (entries 0-999), two pairs merge, so the
    998 shrunk lines take record 5 as chunk 2 and chunk 3 starts at entry 1250.
    rejected = 3: the first three lines lie below V1L4, records 1-4 hold 997
    accepted lines and chunk 1 runs to entry 1249."""
    n, sp, v0 = 1400, 0.2, 600.0
    L = W.make_lines(n, v0, v0 + n * sp, seed=9, mols=(1,))
    L.nu = v0 + sp * np.arange(n)
    L.s[:] = np.float32(1e-21)
    L.a_air[:] = np.float32(0.05)
    L.a_self[:] = np.float32(0.2)
    L.e_lo[:] = np.float32(300.0)
    pairs = [(249, 250), (600, 601), (999, 1000), (1120, 1121), (1249, 1250), (1300, 1301)]
    for a, b in pairs:
        L.nu[b] = L.nu[a] + 0.02
        L.nu[a - 1] = L.nu[a] - 0.1
    lay = W.Layer(pave=100.0, tave=250.0, wtot=1e24, w={1: 1e21})
    dvr4 = 64 * DV
    # V1L4 = V1 - 3 DVR4 - 25 sits just below entry `rejected`
    v1 = v0 + sp * rejected - 0.1 + 3 * dvr4 + 25.0
    v2 = 875.0
    full = od(L, lay, v1, v2, DV)
    g = v1 + DV * np.arange(full.size)
    # A line of zero strength is not accepted into the buffer (SUI = 0, :955), which
    # would move the chunk ends; one of strength 1e-44 (S W = 1e-23 here) is accepted
    # and grouped as usual but weighs nothing in the OD or in a merged centre.
    tiny = np.float32(1e-44)
    for a, b in pairs:
        without = W.LineList(**{k: getattr(L, k).copy() for k in FIELDS})
        without.s[b] = tiny
        alone = W.LineList(**{k: getattr(L, k).copy() for k in FIELDS})
        alone.s[:] = tiny
        alone.s[b] = L.s[b]
        d = full - od(without, lay, v1, v2, DV) - od(alone, lay, v1, v2, DV)
        near = np.abs(g - L.nu[a]) < 20
        rel = np.abs(d[near]).max() / np.abs(full).max()
        if a in apart:
            assert rel < 1e-14, (a, rel)
        else:
            assert rel > 1e-8, (a, rel)
