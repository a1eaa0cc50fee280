"""The scene batch over its full range with Jacobians, streamed in spectral blocks.

Scenes of the scene-batch workload (40-80 layers each, sharing the 200 000-line list
over 500-2500 cm-1) with radiance, dR/dT_lev and dR/dln q_lev for H2O, CO2 and O3 over
the whole 500-2500 cm-1: every scene's (3 + 3) x L OD variants are computed block by
block (20 blocks of 100 cm-1, each layer's DV from calculateDV at the block's centre,
one DVout per block) and never held for the whole range; 8 scenes at a time (about
2 900 layer-OD jobs, up to ~25 GB of OD at the finest block's DVout) form one batch
(clblm_od_batch into one reused output buffer), then each scene's RT and level assembly
runs on its slice.  Lines only: the host evaluation of the synthetic continuum for
every job of every block would dominate the test.

By default every 8th of the 256 scenes (32 scenes, ~11 600 jobs per block);
CLBLM_ALL_SCENES=1 streams all 256 (~92 500 jobs per block, several minutes).

Every batch runs in FP32C; each group of 8 scenes also runs in FP64 at some blocks
(rotating, so that every block and every scene is compared): radiance within 1e-5
relative, Jacobians within 1e-4 of each scene's maximum of the quantity over its
compared blocks (dR/dln q against the larger of its maximum and 1e-6 of the radiance's,
as in the large scene: a molecule with no band in a block leaves its Jacobian at
rounding level there).  Every FP32C output must be finite.
"""
import os
import time

import numpy as np
import pytest

from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu

V1, V2, BLOCK = 500.0, 2500.0, 100.0
GROUP = 8
MOLS = [1, 2, 3]


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def test_scene_batch_full_range_streamed(ctx):
    import torch
    from clblm_amd import _lib
    from clblm_amd.jacobian import UpwellingJacobian
    scenes = W.scene_batch(256, seed=1003)
    if os.environ.get("CLBLM_ALL_SCENES") != "1":
        scenes = scenes[::8]
    lines = W.profile_lines(200000, V1, V2, seed=1001)
    h = ctx.upload_lines(lines)
    edges = np.arange(V1, V2 + 0.5 * BLOCK, BLOCK)
    nblk, ngrp = len(edges) - 1, (len(scenes) + GROUP - 1) // GROUP
    # FP64 comparisons: group g at blocks b = g mod nblk and (g + nblk/2) mod nblk, or at
    # every block b = g mod ngrp when there are fewer groups than blocks
    if ngrp >= nblk:
        checked = {b: [g for g in range(ngrp) if b in (g % nblk, (g + nblk // 2) % nblk)] for b in range(nblk)}
    else:
        checked = {b: [b % ngrp] for b in range(nblk)}
    assert all(checked[b] for b in range(nblk))
    S = len(scenes)
    err = {"rad": np.zeros(S), "jac_t": np.zeros(S), "jac_lnq": np.zeros((S, len(MOLS)))}
    peak = {"rad": np.zeros(S), "jac_t": np.zeros(S), "jac_lnq": np.zeros((S, len(MOLS)))}
    seen = np.zeros(S, bool)
    npts, njobs = 0, 0
    buf = torch.empty(0, dtype=torch.float64, device="cuda")
    bad = torch.zeros((), dtype=torch.int64, device="cuda")
    t0 = time.time()
    t_lev = [W.level_temperatures(len(s.layers), 0.1) for s in scenes]
    p_lev = [W.level_pressures(len(s.layers), 0.1) for s in scenes]
    q_lev = [W.level_densities(len(s.layers), 0.1, h2o_scale=s.h2o_scale) for s in scenes]
    for b in range(nblk):
        a, z = float(edges[b]), float(edges[b + 1])
        dvs = [[W.calculate_dv(0.5 * (a + z), l.pave, l.tave, l.w[1] / l.wtot) for l in s.layers] for s in scenes]
        dvout = min(min(d) for d in dvs)
        n_blk = 0
        for g in range(ngrp):
            sel = list(range(g * GROUP, min((g + 1) * GROUP, S)))
            precs = [_lib.FP32C] + ([_lib.FP64] if g in checked[b] else [])
            res = {}
            for prec in precs:
                jbs, jobs, pars, first = [], [], [], [0]
                for si in sel:
                    jb = UpwellingJacobian(ctx, h, scenes[si].layers, t_lev[si], dvs[si], p_lev=p_lev[si],
                                           q_lev=q_lev[si], mols=MOLS, precision=prec, dvout=dvout)
                    j, p = jb._jobs(a, z)
                    jbs.append(jb)
                    jobs += j
                    pars += p
                    first.append(first[-1] + len(j))
                lay, par, off = ctx.od_batch_plan(h, jobs, pars)
                n = int(off[1] - off[0])
                assert np.all(np.diff(off) == n)
                if buf.numel() < off[-1]:
                    del buf
                    buf = torch.empty(int(off[-1]), dtype=torch.float64, device="cuda")
                ctx.jrad = 1
                od, _ = ctx.od_batch(h, jobs, pars, out=buf)
                out = []
                for k, si in enumerate(sel):
                    L = len(scenes[si].layers)
                    odv = od[first[k] * n:first[k + 1] * n].view(3 + len(MOLS), L, n)
                    r = jbs[k].assemble(a, odv[0], odv)
                    out.append({q: r[q] for q in ("rad", "jac_t", "jac_lnq")})
                    if prec == _lib.FP32C:
                        for q in out[-1].values():
                            bad += (~torch.isfinite(q)).sum()
                del od, odv, r
                res[prec] = out
                if prec == _lib.FP32C:
                    njobs += len(jobs)
                    n_blk = n
            if _lib.FP64 in res:
                for k, si in enumerate(sel):
                    r32, r64 = res[_lib.FP32C][k], res[_lib.FP64][k]
                    seen[si] = True
                    err["rad"][si] = max(err["rad"][si], float(((r32["rad"] - r64["rad"]).abs() / r64["rad"].abs()).max()))
                    peak["rad"][si] = max(peak["rad"][si], float(r64["rad"].abs().max()))
                    err["jac_t"][si] = max(err["jac_t"][si], float((r32["jac_t"] - r64["jac_t"]).abs().max()))
                    peak["jac_t"][si] = max(peak["jac_t"][si], float(r64["jac_t"].abs().max()))
                    err["jac_lnq"][si] = np.maximum(err["jac_lnq"][si], (r32["jac_lnq"] - r64["jac_lnq"]).abs()
                                                    .amax(dim=(1, 2)).cpu().numpy())
                    peak["jac_lnq"][si] = np.maximum(peak["jac_lnq"][si],
                                                     r64["jac_lnq"].abs().amax(dim=(1, 2)).cpu().numpy())
            del res
        npts += n_blk
        ctx.sync()
        assert int(bad) == 0, "non-finite FP32C output up to block %d" % b
        print("block %.0f-%.0f cm-1 (%.0f s): DVout %.3g, %d points per spectrum; FP64 check of scene groups %s; "
              "worst so far: radiance %.2e" % (a, z, time.time() - t0, dvout, n_blk, checked[b], err["rad"].max()),
              flush=True)
    assert seen.all()
    rt = err["jac_t"] / peak["jac_t"]
    rq = err["jac_lnq"] / np.maximum(peak["jac_lnq"], 1e-6 * peak["rad"][:, None])
    print("%d scenes over %.0f-%.0f cm-1: %d layer-OD jobs, %d points per spectrum; FP32C vs FP64: radiance %.2e, "
          "dR/dT_lev %.2e, dR/dln q %s of each scene's maximum" %
          (S, V1, V2, njobs, npts, err["rad"].max(), rt.max(), ["%.1e" % x for x in rq.max(axis=0)]))
    assert err["rad"].max() < 1e-5
    assert rt.max() < 1e-4
    assert rq.max() < 1e-4
    ctx.free_lines(h)
