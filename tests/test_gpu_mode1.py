"""The merge-down radiance on the GPU (clblm_rt_mode1) against the oracle's restatement
(orc_rt_mode1, pinned by the properties in test_oracle_mode1.py): per-layer grids,
thermal and solar sources, a reflecting and a black surface, an observer inside the
atmosphere, airmass-scaled paths and every linInTau form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _case(seed, nlay=14, v1=600.0, v2=680.0):
    rng = np.random.default_rng(seed)
    dvs, d = [], 0.003
    for l in range(nlay):                 # top finest, DV growing downward by grid ratios
        dvs.append(d)
        d = d * float(rng.choice([1.0, 4 / 3, 3 / 2, 2.0, 1.37]))
    dvs = dvs[::-1]
    ods = []
    for l, dv in enumerate(dvs):
        n = int(np.ceil((v2 - v1) / dv)) + 1 + int(rng.integers(0, 3))
        v = v1 + dv * np.arange(n)
        ods.append(np.abs(0.2 * (1 + np.sin(v / (1.5 + 0.2 * l))) * rng.uniform(0.3, 2.0) + 1e-3))
    t_lev = np.linspace(290.0, 212.0, nlay + 1) + rng.normal(0, 2, nlay + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1])
    return ods, dvs, t_lay, t_lev, v1, v2


CASES = [  # (seed, lin, thermal, solar, emis, obs_level offset from the top)
    (1, 1, True, False, 0.98, 0), (2, 0, True, False, 0.9, 0), (3, 2, True, True, 0.95, 0),
    (4, 1, True, True, 1.0, 0), (5, 1, True, False, 0.9, 5), (6, 0, False, True, 1.0, 0), (7, 2, True, True, 0.8, 3),
]


@pytest.mark.parametrize("seed,lin,thermal,solar,emis,obs", CASES)
def test_mode1_matches_oracle(ctx, seed, lin, thermal, solar, emis, obs):
    import torch
    from clblm_amd.jacobian import radiance_mode1
    from oracle.py import rt_mode1
    ods, dvs, t_lay, t_lev, v1, v2 = _case(seed)
    nlay = len(ods)
    rng = np.random.default_rng(seed + 100)
    fu, fd, fs = rng.uniform(1.0, 1.2, nlay), rng.uniform(1.5, 1.8, nlay), rng.uniform(1.1, 2.5, nlay)
    nsol = 30001            # a finer grid than the layers', within their range
    sol_dv = (v2 - v1) / (nsol - 1)
    s0 = 1e-4 * (1 + 0.2 * np.cos(np.arange(nsol) / 300.0))
    kw = dict(fac_up=fu, fac_dn=fd, fac_sun=fs, obs_level=nlay - obs, lin=lin, thermal=thermal, solar=solar,
              emis=emis, tskin=291.0, refl=0.2, sol=s0 if solar else None, sol_dv=sol_dv)
    want, wtx, wdv = rt_mode1(ods, dvs, v1, v2, t_lay, t_lev, **kw)
    off = np.concatenate([[0], np.cumsum([o.size for o in ods])])
    od = torch.tensor(np.concatenate(ods), device="cuda")
    print("case", seed, lin, thermal, solar, emis, obs, flush=True)
    rad, tx, dv = radiance_mode1(ctx, od, off, dvs, v1, v2, t_lay, t_lev, fac_up=fu, fac_dn=fd, fac_sun=fs,
                                 obs_level=nlay - obs, lin_in_tau=lin, thermal=thermal, solar=solar, emis=emis,
                                 tskin=291.0, refl=0.2, sol=torch.tensor(s0, device="cuda") if solar else None,
                                 sol_dv=sol_dv)
    rad, tx = rad.cpu().numpy(), tx.cpu().numpy()
    assert dv == wdv and rad.size == want.size
    err = float(np.abs(rad - want).max() / np.abs(want).max())
    print("points %d dv %.4g: max err / max %.2e" % (rad.size, dv, err))
    assert err <= 1e-12
    assert float(np.abs(tx - wtx).max()) <= 1e-12


def test_mode1_rejects_a_reflecting_surface_without_the_down_path(ctx):
    import torch
    from clblm_amd._lib import ClblmError
    from clblm_amd.jacobian import radiance_mode1
    ods, dvs, t_lay, t_lev, v1, v2 = _case(1, nlay=3)
    off = np.concatenate([[0], np.cumsum([o.size for o in ods])])
    od = torch.tensor(np.concatenate(ods), device="cuda")
    with pytest.raises(ClblmError):
        radiance_mode1(ctx, od, off, dvs, v1, v2, t_lay, t_lev, thermal=False, solar=True, emis=0.9,
                       sol=torch.ones(100, dtype=torch.float64, device="cuda"), sol_dv=1.0)
