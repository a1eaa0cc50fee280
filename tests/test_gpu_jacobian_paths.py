"""The Jacobians with every branch on the GPU (clblm_rt_jac_paths through the C-ABI)
against the oracle's orc_jacobians_paths (pinned in test_oracle_jacobian_paths.py):
up-looking observers, down-looking observers inside the atmosphere with the sun, per-layer
air-mass factors, every linInTau, and the default case against clblm_rt_jac."""
import numpy as np
import pytest
import torch

from oracle import py as O

pytestmark = pytest.mark.gpu
V1, DV, N = 700.0, 0.01, 20001


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def scene(L, seed):
    rng = np.random.default_rng(seed)
    od = rng.uniform(0.0, 1.0, (L, N)) * np.geomspace(2.0, 0.01, L)[:, None]
    t_lev = np.linspace(290.0, 215.0, L + 1) + rng.normal(0, 2, L + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1]) + rng.normal(0, 0.5, L)
    return od, t_lay, t_lev


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("lin", [0, 1, 2])
@pytest.mark.parametrize("case", [dict(uplook=True, obs=0), dict(uplook=True, obs=7), dict(obs=5, solar=True),
                                  dict(obs=12, solar=True), dict(obs=12)])
def test_jacobians_paths_match_oracle(ctx, lin, case):
    from clblm_amd.jacobian import jacobians_paths
    L = 12
    od, t_lay, t_lev = scene(L, 5 + lin)
    rng = np.random.default_rng(9)
    av, ad, asun = rng.uniform(1.0, 1.3, L), rng.uniform(1.5, 1.8, L), rng.uniform(1.1, 3.0, L)
    sol = 1e-4 * (1.0 + 0.2 * np.cos(np.arange(N) * 0.01))
    up, lo = rng.uniform(0.3, 0.7, L), rng.uniform(0.3, 0.7, L)
    print("case %s lin %d" % (case, lin), flush=True)
    g = jacobians_paths(ctx, torch.as_tensor(od, device="cuda"), t_lay, t_lev, V1, DV, amf_view=av, amf_down=ad,
                        amf_sun=asun, obs_level=case["obs"], uplook=case.get("uplook", False), lin_in_tau=lin,
                        emis=0.93, tskin=291.0, solar=case.get("solar", False),
                        sol=torch.as_tensor(sol, device="cuda"), rsfc_sun=0.3, dt_dtup=up, dt_dtlo=lo)
    w = O.jacobians_paths(od, t_lay, t_lev, V1, DV, amf_view=av, amf_down=ad, amf_sun=asun, emis=0.93, tskin=291.0,
                          lin=lin, solar=case.get("solar", False), sol_toa=sol, rsfc_sun=0.3, obs=case["obs"],
                          uplook=case.get("uplook", False), dt_dtup=up, dt_dtlo=lo)
    # dR/dtau and dR/dT_lev are differences of terms of the radiance's size; the
    # device's exp and Planck function round differently from the host's in the last
    # place, which those differences magnify (an up-looking layer whose source nearly
    # equals the radiance coming from above it): 1e-10 of each output's maximum for them
    for k in ("rad", "dtau", "dtlev", "dtskin", "demis", "drsfc"):
        got = g[k].cpu().numpy()
        if not np.any(w[k]):
            assert not np.any(got), k
            continue
        bar = 1e-10 if k in ("dtau", "dtlev") else 1e-12
        assert _rel(got, w[k]) <= bar, (k, _rel(got, w[k]))


def test_default_case_is_rt_jac(ctx):
    from clblm_amd import _lib
    from clblm_amd.jacobian import jacobians_paths
    import ctypes
    L = 10
    od, t_lay, t_lev = scene(L, 3)
    up, lo = np.full(L, 0.5), np.full(L, 0.5)
    g = jacobians_paths(ctx, torch.as_tensor(od, device="cuda"), t_lay, t_lev, V1, DV, amf_down=np.full(L, 1.66),
                        lin_in_tau=1, emis=0.9, tskin=292.0, dt_dtup=up, dt_dtlo=lo)
    r = O.jacobians(od, t_lay, t_lev, V1, DV, emis=0.9, tskin=292.0, lin=1, sec_dn=1.66, dt_dtup=up, dt_dtlo=lo)
    for x, k in zip(r, ("rad", "dtau", "dtlev", "dtskin", "demis")):
        assert _rel(g[k].cpu().numpy(), x) <= (1e-10 if k in ("dtau", "dtlev") else 1e-12), k
