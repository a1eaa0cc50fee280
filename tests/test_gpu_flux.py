"""Fluxes on the GPU (clblm_rt_flux through the C-ABI) against the oracle's
orc_rt_flux (pinned by the closed forms and the quadrature in test_oracle_flux.py):
1-3 angles, every linInTau, lambertian and specular surfaces, a group width that
truncates, and a spectrum long enough to take several blocks of groups."""
import numpy as np
import pytest
import torch

from oracle import py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def _case(v1, v2, dv, L, seed):
    rng = np.random.default_rng(seed)
    n = int(round((v2 - v1) / dv)) + 1
    od = rng.uniform(0.0, 2.0, (L, n)) * np.geomspace(3.0, 0.05, L)[:, None]
    t_lev = np.linspace(288.0, 215.0, L + 1) + rng.normal(0, 2, L + 1)
    t_lay = 0.5 * (t_lev[1:] + t_lev[:-1])
    p = np.geomspace(1013.0, 5.0, L + 1)
    return od, t_lay, t_lev, p


@pytest.mark.parametrize("nang,lin,lamb", [(1, 0, True), (2, 1, True), (3, 2, False), (3, 1, True)])
def test_flux_matches_oracle(ctx, nang, lin, lamb):
    from clblm_amd.jacobian import fluxes
    for v1, v2, dv, dvf, L in [(600.0, 700.0, 2.0 ** -6, 10.0, 8), (500.0, 550.0, 0.01, 5.0, 5),
                               (600.0, 700.0, 2.0 ** -12, 10.0, 3)]:
        od, t_lay, t_lev, p = _case(v1, v2, dv, L, 11 + nang)
        print("case %s-%s dv %g nang %d lin %d lambertian %d" % (v1, v2, dv, nang, lin, lamb), flush=True)
        got = fluxes(ctx, torch.as_tensor(od, device="cuda"), t_lay, t_lev, p, v1, v2, dv, dvf, nang=nang,
                     lin_in_tau=lin, emis=0.93, tskin=291.0, lambertian=lamb)
        want = O.rt_flux(od, t_lay, t_lev, p, v1, v2, dv, dvf, nang=nang, lin=lin, emis=0.93, tskin=291.0,
                         lambertian=lamb)
        for k in ("up", "dn", "net"):
            err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
            assert err <= 1e-12, (k, err)
        scale = np.abs(want["heating"]).max()
        assert np.abs(got["heating"] - want["heating"]).max() <= 1e-9 * scale


def test_flux_refuses_inconsistent_groups(ctx):
    from clblm_amd._lib import ClblmError
    from clblm_amd.jacobian import fluxes
    od, t_lay, t_lev, p = _case(600.0, 700.0, 0.5, 2, 3)
    with pytest.raises(ClblmError, match="inconsistent"):
        fluxes(ctx, torch.as_tensor(od, device="cuda"), t_lay, t_lev, p, 600.0, 700.0, 0.5, 30.0)
