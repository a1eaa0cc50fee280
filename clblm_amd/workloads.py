"""Synthetic line lists and atmospheres for the benchmark workloads.

The reference ships no line file (LNFL is external) and no scene, so every
workload is generated from a fixed seed (``numpy.random.default_rng(1000 + k)``).
Line fields are rounded to TAPE3 precision: centre FP64; intensity, widths,
lower-state energy, temperature exponent and shift FP32
(record layout: src/clblm_src/clblm_LineData.f90:60-70).

TAPE3 conventions (LNFL is not in the reference; these follow the reference's
algebra, clblm_ODLAY.f90:1805-1864,1928-1935):
  * stored intensity SP = S_HITRAN(296) / (nu * (1 - exp(-c2 nu / 296)))  so that
    SP * RADFN * (1 + exp(-c2 nu/T)) reproduces S(T) at 296 K;
  * stored TMPALF = 1 - n for a HITRAN-style width exponent n, because the
    reference scales widths by (p/p0)(296/T)(T/296)**TMPALF;
  * molecule code MOL = m + 100*iso.
"""
from __future__ import annotations

import dataclasses
import functools
import math
from typing import List, Optional

import numpy as np

RADCN2 = 1.4387752          # clblm_ConstParam.f90:71
P0 = 1013.25                # clblm_ConstParam.f90:110
T0 = 296.0                  # clblm_ConstParam.f90:112
AIR_COLUMN_PER_HPA = 100.0 / (9.80665 * 0.0289644) * 6.02214199e23 * 1e-4  # molec cm-2 hPa-1

SPECIES = [  # (HITRAN molecule number, name, mass for the synthetic list)
    (1, "H2O", 18), (2, "CO2", 44), (3, "O3", 48), (4, "N2O", 44),
    (5, "CO", 28), (6, "CH4", 16), (7, "O2", 32),
]


@dataclasses.dataclass
class LineList:
    """A TAPE3-equivalent line list, sorted by centre.  Records hold 250 lines."""
    nu: np.ndarray          # float64
    s: np.ndarray           # float32, stored intensity SP
    a_air: np.ndarray       # float32
    a_self: np.ndarray      # float32
    e_lo: np.ndarray        # float32
    tmpalf: np.ndarray      # float32
    shift: np.ndarray       # float32
    mol: np.ndarray         # int32, m + 100*iso
    iflg: np.ndarray        # int32
    # species broadening (BRD_MOL_FLG / BRD_MOL_DAT), (n, 7) each, or None
    brd_flg: Optional[np.ndarray] = None    # int32
    brd_hw: Optional[np.ndarray] = None     # float32
    brd_tmp: Optional[np.ndarray] = None    # float32
    brd_shft: Optional[np.ndarray] = None   # float32

    @property
    def n(self) -> int:
        return int(self.nu.shape[0])


def couple_lines(lines: LineList, coupled, iflag, a_nodes, b_nodes) -> LineList:
    """The line list with the lines at indices ``coupled`` flagged IFLG = ``iflag`` (1:
    Y and G, 3: reduced widths GAMMA1 and GAMMA2; one value or one per coupled line)
    and followed by their coefficient
    records, laid out as LNFL writes them: the record after a coupled line has IFLG -1
    and holds the coefficients at 200, 250, 296 and 340 K, ``a_nodes`` (Y / GAMMA1) in
    its VNU, A_AIR, MOL (as REAL*4 bits) and TMPALF fields and ``b_nodes`` (G / GAMMA2)
    in its S, E_LO, A_SELF and SHIFT fields.  A line and its coefficient record stay in
    one 250-line record: where only one slot is left, a filler record (IFLG -1, not a
    line) is put first.  a_nodes / b_nodes: shape (4,) or (len(coupled), 4)."""
    coupled = list(coupled)
    cset = {int(i): k for k, i in enumerate(coupled)}
    flags = np.broadcast_to(np.asarray(iflag, np.int32), (len(coupled),))
    A = np.broadcast_to(np.asarray(a_nodes, np.float64), (len(coupled), 4))
    Bn = np.broadcast_to(np.asarray(b_nodes, np.float64), (len(coupled), 4))
    fields = ("nu", "s", "a_air", "a_self", "e_lo", "tmpalf", "shift", "mol", "iflg")
    brd = ("brd_flg", "brd_hw", "brd_tmp", "brd_shft")
    if lines.brd_flg is not None:
        fields = fields + brd
    out = {f: [] for f in fields}
    zero7 = {"brd_flg": np.zeros(7, np.int32), "brd_hw": np.zeros(7, np.float32),
             "brd_tmp": np.zeros(7, np.float32), "brd_shft": np.zeros(7, np.float32)}

    def put(vals):
        for f in fields:
            out[f].append(vals[f])

    for i in range(lines.n):
        row = {f: getattr(lines, f)[i] for f in fields}
        if i in cset:
            if len(out["nu"]) % 250 == 249:   # the pair would straddle two records
                filler = dict(row)
                filler["iflg"] = np.int32(-1)
                put(filler)
            k = cset[i]
            row = dict(row)
            row["iflg"] = np.int32(flags[k])
            put(row)
            a, b = A[k], Bn[k]
            rec = {"nu": np.float64(a[0]), "a_air": np.float32(a[1]),
                   "mol": np.float32(a[2]).view(np.int32), "tmpalf": np.float32(a[3]),
                   "s": np.float32(b[0]), "e_lo": np.float32(b[1]), "a_self": np.float32(b[2]),
                   "shift": np.float32(b[3]), "iflg": np.int32(-1)}
            rec.update({f: zero7[f] for f in fields if f in zero7})
            put(rec)
        else:
            put(row)
    dt = dict(nu=np.float64, s=np.float32, a_air=np.float32, a_self=np.float32, e_lo=np.float32,
              tmpalf=np.float32, shift=np.float32, mol=np.int32, iflg=np.int32, brd_flg=np.int32,
              brd_hw=np.float32, brd_tmp=np.float32, brd_shft=np.float32)
    return LineList(**{f: np.asarray(out[f], dtype=dt[f]) for f in fields})


@dataclasses.dataclass
class Layer:
    pave: float             # hPa
    tave: float             # K
    wtot: float             # molec cm-2 (broadening column)
    w: dict                 # molecule number -> column, molec cm-2
    zbot: float = 0.0
    ztop: float = 0.0


def make_lines(n: int, v1: float, v2: float, seed: int, mols=(1,)) -> LineList:
    """Lines from the distributions the workloads name: centres uniform in [v1, v2],
    S_HITRAN log-uniform in [1e-26, 1e-19], air widths uniform in [0.02, 0.10],
    self widths 5x air, width exponent n uniform in [0.5, 0.8], E'' uniform in
    [0, 3000].  Molecules are assigned round-robin over ``mols`` after sorting,
    each line's isotopologue is 1."""
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(v1, v2, n))
    s296 = 10.0 ** rng.uniform(-26.0, -19.0, n)
    a_air = rng.uniform(0.02, 0.10, n)
    a_self = 5.0 * a_air
    nexp = rng.uniform(0.5, 0.8, n)
    e_lo = rng.uniform(0.0, 3000.0, n)
    mol_pick = rng.integers(0, len(mols), n)
    mol = np.asarray(mols, dtype=np.int32)[mol_pick] + 100
    sp = s296 / (nu * (1.0 - np.exp(-RADCN2 * nu / T0)))
    return LineList(
        nu=nu.astype(np.float64),
        s=sp.astype(np.float32),
        a_air=a_air.astype(np.float32),
        a_self=a_self.astype(np.float32),
        e_lo=e_lo.astype(np.float32),
        tmpalf=(1.0 - nexp).astype(np.float32),
        shift=np.zeros(n, np.float32),
        mol=mol.astype(np.int32),
        iflg=np.zeros(n, np.int32),
    )


def calculate_dv(vbar, pave, tave, frh2o, alfal0=0.04, sample=4.0):
    """calculateDV, clblm_setDV.f90:221-271."""
    alfcor = (pave / P0) * math.sqrt(T0 / tave)
    h2oslf = (1.0 - frh2o + 5.0 * frh2o)
    albar = alfal0 * alfcor * h2oslf
    adbar = 3.58115e-07 * vbar * math.sqrt(tave / 36.0)
    avbar = 0.5 * (albar + math.sqrt(albar * albar + 4.0 * adbar * adbar))
    return avbar / sample


# --------------------------------------------------------------------------
# Single-layer case: one layer at 1013.25 hPa, 296 K, 20 000 lines of one species.
# --------------------------------------------------------------------------
def single_layer(n_lines=20000, v1=600.0, v2=1100.0, seed=1000):
    lines = make_lines(n_lines, v1, v2, seed, mols=(1,))
    layer = Layer(pave=1013.25, tave=296.0, wtot=1e24, w={1: 1e22})
    alfal0 = float(np.mean(lines.a_air.astype(np.float64)))
    dv = calculate_dv(0.5 * (v1 + v2), layer.pave, layer.tave, 1e22 / 1e24, alfal0, 4.0)
    return lines, [layer], dict(v1=v1, v2=v2, dv=dv, alfal0=alfal0)


# --------------------------------------------------------------------------
# Full-profile atmosphere.
# --------------------------------------------------------------------------
VMR_SURFACE = {1: 1.0e-2, 2: 400e-6, 3: 3e-8, 4: 320e-9, 5: 100e-9, 6: 1.8e-6, 7: 0.209}


def vmr_profile(m: int, p: float) -> float:
    """Recorded volume-mixing-ratio profiles (p in hPa)."""
    if m == 1:   # H2O: 2 km scale height in pressure terms, floor 4 ppmv
        return max(1.0e-2 * (p / 1013.0) ** 3.5, 4e-6)
    if m == 3:   # O3: peak of 8 ppmv near 10 hPa
        x = math.log(p / 10.0)
        return 3e-8 + 8e-6 * math.exp(-0.5 * (x / 1.2) ** 2)
    return VMR_SURFACE[m]


def temperature_profile(z_km: float) -> float:
    return max(288.0 - 6.5 * z_km, 217.0)


def _levels(n_layers, p_top):
    """Log-spaced pressure levels, their heights (hypsometric) and temperatures."""
    levels = np.exp(np.linspace(math.log(1013.0), math.log(p_top), n_layers + 1))
    z = [0.0]
    t_lev = [temperature_profile(0.0)]
    for k in range(1, n_layers + 1):
        # iterate the hypsometric step with the layer-mean temperature
        t_guess = t_lev[-1]
        for _ in range(3):
            dz = 287.05 * 0.5 * (t_lev[-1] + t_guess) / 9.80665 * math.log(levels[k - 1] / levels[k]) / 1000.0
            t_guess = temperature_profile(z[-1] + dz)
        z.append(z[-1] + dz)
        t_lev.append(t_guess)
    return levels, z, t_lev


def level_temperatures(n_layers=60, p_top=0.1) -> np.ndarray:
    """Level temperatures (n_layers + 1, surface first) of ``atmosphere(n_layers, p_top)``."""
    return np.asarray(_levels(n_layers, p_top)[2], dtype=np.float64)


def level_pressures(n_layers=60, p_top=0.1) -> np.ndarray:
    """Level pressures (hPa, n_layers + 1, surface first) of ``atmosphere(n_layers, p_top)``."""
    return np.asarray(_levels(n_layers, p_top)[0], dtype=np.float64)


BOLTZ_HPA = 1.3806503e-19   # hPa cm3 / K, the constant of clblm_noScattJac.f90:1240


def level_densities(n_layers=60, p_top=0.1, mols=(1, 2, 3, 4, 5, 6, 7), h2o_scale=1.0) -> dict:
    """Number densities (molec cm-3) of each molecule at the levels of
    ``atmosphere(n_layers, p_top)``: the recorded mixing ratio at the level pressure
    times p / (k T)."""
    levels, _, t_lev = _levels(n_layers, p_top)
    out = {}
    for m in mols:
        q = np.array([vmr_profile(m, p) * p / (BOLTZ_HPA * t) for p, t in zip(levels, t_lev)])
        out[m] = q * (h2o_scale if m == 1 else 1.0)
    return out


def atmosphere(n_layers=60, p_top=0.1, mols=(1, 2, 3, 4, 5, 6, 7), t_noise=None,
               h2o_scale=1.0) -> List[Layer]:
    """Layers between log-spaced pressure levels from 1013 hPa to ``p_top``.
    Temperature: 288 K at the surface, -6.5 K/km to 217 K, isothermal above;
    heights from the hypsometric relation.  Layer pressure is the mass-weighted
    mean (dp / dln p), temperature the level mean, columns from dp."""
    levels, z, t_lev = _levels(n_layers, p_top)
    layers = []
    for k in range(n_layers):
        p1, p2 = levels[k], levels[k + 1]
        pave = (p1 - p2) / math.log(p1 / p2)
        tave = 0.5 * (t_lev[k] + t_lev[k + 1])
        if t_noise is not None:
            tave += float(t_noise[k])
        wtot = (p1 - p2) * AIR_COLUMN_PER_HPA
        w = {m: vmr_profile(m, pave) * wtot for m in mols}
        if 1 in w:
            w[1] *= h2o_scale
        layers.append(Layer(pave=float(pave), tave=float(tave), wtot=float(wtot), w=w,
                            zbot=float(z[k]), ztop=float(z[k + 1])))
    return layers


def profile_lines(n_lines=200000, v1=500.0, v2=2500.0, seed=1001):
    """200 000 lines of molecules 1-7, split round-robin by a seeded draw."""
    return make_lines(n_lines, v1, v2, seed, mols=tuple(m for m, _, _ in SPECIES))


def continuum_absrb(layer: Layer, v1: float, v2: float, seed=1001, nknots=50):
    """Smooth continuum on the 1 cm-1 grid OPDPTH uses for ABSRB
    (clblm_ODLAY.f90:881-896): V1ABS = INT(V1)-3, V2ABS = INT(V2+3.5).
    log10 of the coefficient is a natural cubic spline through ``nknots`` seeded
    knots in [-27, -25] cm2/molec; the absorber column is the layer's H2O."""
    v1abs = float(int(v1)) - (1.0 if v1 < 0 else 0.0) - 3.0
    v2abs = float(int(v2 + 3.0 + 0.5))
    npt = int((v2abs - v1abs) / 1.0 + 1.5)
    return _continuum_coef(v1abs, v2abs, npt, seed, nknots) * layer.w.get(1, 0.0), v1abs, npt


def continuum_absrb_tangents(layer: Layer, v1: float, v2: float, dirs, seed=1001, nknots=50):
    """continuum_absrb and its derivatives along dirs ((dt, dlnwtot, {m: dlnw}) each):
    the coefficient does not depend on temperature and the absorber is the H2O
    column, so the derivative along a direction is dlnw[1] times the continuum."""
    c, v1abs, npt = continuum_absrb(layer, v1, v2, seed, nknots)
    d = np.array([c * float((dq or {}).get(1, 0.0)) for _, _, dq in dirs]).reshape(len(dirs), npt)
    return np.asarray(c, np.float64), d


continuum_absrb.tangents = continuum_absrb_tangents


@functools.lru_cache(maxsize=256)
def _continuum_coef(v1abs, v2abs, npt, seed, nknots):
    """The continuum coefficient (cm2/molec) on v1abs + i, i < npt: the seeded spline
    depends on the range and the seed only, so it is evaluated once per block."""
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(seed + 7919)
    knots = np.linspace(v1abs, v2abs, nknots)
    vals = rng.uniform(-27.0, -25.0, nknots)
    spl = CubicSpline(knots, vals, bc_type="natural")
    grid = v1abs + np.arange(npt, dtype=np.float64)
    out = 10.0 ** spl(grid)
    out.setflags(write=False)
    return out


# --------------------------------------------------------------------------
# Batches of scenes: uneven layer counts, perturbed temperatures, scaled water vapour.
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Scene:
    layers: List[Layer]
    t_noise: np.ndarray     # K, added to each layer's mean temperature
    h2o_scale: float        # factor on the water-vapour column


def scene_batch(n_scenes=256, seed=1003, n_min=40, n_max=80, t_sigma=3.0, h2o_lo=0.1, h2o_hi=10.0,
                p_top=0.1) -> List[Scene]:
    """Scenes with N_lay uniform in {n_min..n_max}, layer temperatures perturbed by
    seeded Gaussian noise of ``t_sigma`` K and water-vapour columns scaled
    log-uniformly in [h2o_lo, h2o_hi]; each scene's layers span 1013 hPa to p_top."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_scenes):
        nl = int(rng.integers(n_min, n_max + 1))
        noise = rng.normal(0.0, t_sigma, nl)
        scale = float(10.0 ** rng.uniform(math.log10(h2o_lo), math.log10(h2o_hi)))
        out.append(Scene(layers=atmosphere(nl, p_top, t_noise=noise, h2o_scale=scale), t_noise=noise,
                         h2o_scale=scale))
    return out


# --------------------------------------------------------------------------
# Large single scene: lines over 10-3500 cm-1, 80 layers to 0.01 hPa.
# --------------------------------------------------------------------------
LARGE_V1, LARGE_V2 = 10.0, 3500.0


def large_scene_lines(n_lines=2000000, v1=LARGE_V1, v2=LARGE_V2, seed=1004):
    """Lines of molecules 1-7 over the large scene's range (2 000 000 by default)."""
    return make_lines(n_lines, v1, v2, seed, mols=tuple(m for m, _, _ in SPECIES))


def large_scene_atmosphere():
    """The large scene's 80 layers from 1013 hPa to 0.01 hPa."""
    return atmosphere(80, 0.01)


def large_scene_dv(layer: Layer) -> float:
    """Per-layer DV of the large scene: calculateDV at the range's mid point."""
    return calculate_dv(0.5 * (LARGE_V1 + LARGE_V2), layer.pave, layer.tave, layer.w.get(1, 0.0) / layer.wtot)
