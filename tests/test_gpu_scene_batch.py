"""The 256-scene batch: scenes of 40-80 layers (temperatures perturbed by 3 K noise,
water vapour scaled by 0.1-10) sharing the 200 000-line list over 500-2500 cm-1, every
layer of every scene in one batch captured in one hipGraph, launch shapes picked by the
autotuner.  Over a 2 cm-1 window re-gridded to one DVout so that each scene's radiance
follows from its layer ODs:
- FP64 replays against the oracle, layer by layer, on three scenes (1e-10);
- FP32C replays against FP64 (OD within 1e-5 of each layer's peak, radiance 1e-5)."""
import numpy as np
import pytest

from clblm_amd import workloads as W

pytestmark = pytest.mark.gpu

V1, V2 = 1000.0, 1002.0


@pytest.fixture(scope="module")
def ctx():
    from clblm_amd.odlay import Context
    c = Context(0)
    yield c
    c.close()


def test_256_scenes_in_one_graph(ctx):
    import torch
    from clblm_amd import _lib
    from clblm_amd.jacobian import upwelling_radiance
    from clblm_amd.odlay import Context
    from oracle.py import Oracle
    scenes = W.scene_batch(256, seed=1003)
    lines = W.profile_lines(200000, 500.0, 2500.0, seed=1001)
    layers = [l for s in scenes for l in s.layers]
    dvs = [W.calculate_dv(0.5 * (V1 + V2), l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    dvout = min(dvs)
    conts = [W.continuum_absrb(l, V1, V2)[0] for l in layers]
    mk = lambda p: [Context.params(V1, V2, dv, dvout=dvout, precision=p, continuum=c) for dv, c in zip(dvs, conts)]
    h = ctx.upload_lines(lines)
    print("256 scenes: %d layers" % len(layers), flush=True)
    chosen, ms = ctx.autotune(h, layers, mk(_lib.FP32C), reps=1)
    print("autotune: %r, %.2f ms per batch" % (chosen, ms), flush=True)
    res = {}
    for prec in (_lib.FP64, _lib.FP32C):
        ctx.jrad = 1
        g = ctx.od_graph(h, layers, mk(prec))
        off = g.offsets
        n = int(off[1] - off[0])
        assert np.all(np.diff(off) == n)
        for _ in range(2):
            g.out.fill_(float("nan"))
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            g.check()
        od = g.out[: int(off[-1])].clone()
        g.close()
        assert bool(torch.isfinite(od).all())
        rads, k = [], 0
        for s in scenes:
            L = len(s.layers)
            t_lay = torch.tensor([l.tave for l in s.layers], dtype=torch.float64, device="cuda")
            t_lev = torch.tensor(W.level_temperatures(L, 0.1), dtype=torch.float64, device="cuda")
            rads.append(upwelling_radiance(ctx, od[k * n:(k + L) * n].view(L, n), t_lay, t_lev, V1, dvout))
            k += L
        torch.cuda.synchronize()
        res[prec] = (od.cpu().numpy(), torch.stack(rads).cpu().numpy(), off)
        print("precision %d: %d points per layer, %d in all" % (prec, n, off[-1]), flush=True)
    od64, rad64, off = res[_lib.FP64]
    od32, rad32, _ = res[_lib.FP32C]
    worst = 0.0
    for k in range(len(layers)):
        a, b = od32[off[k]:off[k + 1]], od64[off[k]:off[k + 1]]
        worst = max(worst, float(np.abs(a - b).max() / np.abs(b).max()))
    rerr = float((np.abs(rad32 - rad64) / np.abs(rad64)).max())
    print("FP32C vs FP64: OD %.3e of the layer peak, radiance %.3e" % (worst, rerr))
    assert worst <= 1e-5 and rerr <= 1e-5
    # FP64 against the oracle, every layer of three scenes.  The batch is a sequence
    # of ODLAY calls, so a scene sees the SAVEd JRAD its predecessors left: it drops
    # from 1 to 0 at the first layer with V1/XKT < 5, XKT = TAVE/RADCN2
    # (clblm_ODLAY.f90:349-352), here the first layer warmer than 287.8 K
    orc = Oracle(lines)
    starts = np.cumsum([0] + [len(s.layers) for s in scenes])
    ows = 0.0
    for si in (0, 131, 255):
        hot = any(V1 / (l.tave / 1.4387752) < 5.0 for l in layers[:starts[si]])
        orc.jrad = 0 if hot else 1
        for k in range(starts[si], starts[si + 1]):
            ref, _ = orc.odlay(layers[k], V1, V2, dvs[k], dvout=dvout, continuum=conts[k])
            got = od64[off[k]:off[k + 1]]
            assert ref.size == got.size, (k, ref.size, got.size)
            ows = max(ows, float((np.abs(got - ref) / np.maximum(np.abs(ref), 1e-4 * np.abs(ref).max())).max()))
    print("FP64 vs oracle (scenes 0, 131, 255): max rel err %.3e" % ows)
    assert ows <= 1e-10
    ctx.free_lines(h)




def test_256_scenes_jacobians_in_graphs(ctx):
    """Radiance and level Jacobians of all 256 scenes: every scene's (3 + M) x L OD
    variants (M = 3: H2O, CO2, O3; the reference's T + 1 K, Wtot + 1 % and
    molecule-alone re-runs, about 92 000 layer-OD jobs) over a 2 cm-1 window, batched
    32 scenes (about 11 500 jobs) to a hipGraph -- one graph of all 92 000 jobs asks
    for more scratch than the GPU holds: each job's line and panel tables live in the
    graph's arena -- then each scene's RT and level assembly (clblm_rt_jac,
    clblm_level_jac) from its slice of the batch.  Lines only (the synthetic
    continuum's host evaluation for 46 000 jobs would dominate the test).  FP32C
    against FP64 (radiance 1e-5 relative, Jacobians 1e-4 of each scene's maximum of
    the quantity), and the FP64 batch against three scenes computed on their own
    (UpwellingJacobian.block: the same kernels, one scene per batch)."""
    import torch
    from clblm_amd import _lib
    from clblm_amd.jacobian import UpwellingJacobian
    # V1 / XKT < 5 for every surface layer (T > 201 K): JRAD drops to 0 at each scene's
    # first job in the batch as in a scene run on its own
    v1, v2 = 700.0, 702.0
    scenes = W.scene_batch(256, seed=1003)
    lines = W.profile_lines(200000, 500.0, 2500.0, seed=1001)
    h = ctx.upload_lines(lines)
    mols = [1, 2, 3]
    alld = [[W.calculate_dv(0.5 * (v1 + v2), l.pave, l.tave, l.w[1] / l.wtot) for l in s.layers] for s in scenes]
    dvout = min(min(d) for d in alld)

    def jac(s, dvs, prec):
        L = len(s.layers)
        return UpwellingJacobian(ctx, h, s.layers, W.level_temperatures(L, 0.1), dvs, p_lev=W.level_pressures(L, 0.1),
                                 q_lev=W.level_densities(L, 0.1, h2o_scale=s.h2o_scale), mols=mols, precision=prec,
                                 dvout=dvout)

    worst = {"rad": 0.0, "jac_t": 0.0, "jac_lnq": 0.0}
    keep, ms, njobs = {}, {_lib.FP64: 0.0, _lib.FP32C: 0.0}, 0
    for c0 in range(0, len(scenes), 32):
        sel = range(c0, min(c0 + 32, len(scenes)))
        graphs, jbs = {}, {}
        for prec in (_lib.FP64, _lib.FP32C):
            jbs[prec], jobs, pars, first = [], [], [], [0]
            for si in sel:
                jb = jac(scenes[si], alld[si], prec)
                j, p = jb._jobs(v1, v2)
                jbs[prec].append(jb)
                jobs += j
                pars += p
                first.append(first[-1] + len(j))
            ctx.jrad = 1
            g = ctx.od_graph(h, jobs, pars)
            off = g.offsets
            n = int(off[1] - off[0])
            assert np.all(np.diff(off) == n)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            for rep in range(2):
                ev[0].record()
                g.launch(stream=torch.cuda.current_stream().cuda_stream)
                ev[1].record()
                torch.cuda.synchronize()
                g.check()
            ms[prec] += ev[0].elapsed_time(ev[1])
            graphs[prec] = g
        njobs += len(jobs)
        for k, si in enumerate(sel):
            r = {}
            for prec, g in graphs.items():
                L = len(scenes[si].layers)
                odv = g.out[first[k] * n:first[k + 1] * n].view(3 + len(mols), L, n)
                r[prec] = jbs[prec][k].assemble(v1, odv[0], odv)
            a, b = r[_lib.FP32C], r[_lib.FP64]
            worst["rad"] = max(worst["rad"], float(((a["rad"] - b["rad"]).abs() / b["rad"].abs()).max()))
            for q in ("jac_t", "jac_lnq"):
                worst[q] = max(worst[q], float((a[q] - b[q]).abs().max() / b[q].abs().max()))
            if si in (0, 131, 255):
                keep[si] = {q: b[q].cpu().numpy() for q in ("rad", "jac_t", "jac_lnq")}
        for g in graphs.values():
            g.close()
        del graphs, jbs
        print("scenes %d-%d: %d layer-OD jobs of %d points; worst so far radiance %.2e, dR/dT %.2e, dR/dln q %.2e" %
              (sel[0], sel[-1], len(jobs), n, worst["rad"], worst["jac_t"], worst["jac_lnq"]), flush=True)
    print("256 scenes, %d layer-OD jobs: graph replays %.1f ms FP64, %.1f ms FP32C in all; FP32C vs FP64: radiance "
          "%.2e, dR/dT_lev %.2e, dR/dln q %.2e of each scene's maximum" %
          (njobs, ms[_lib.FP64], ms[_lib.FP32C], worst["rad"], worst["jac_t"], worst["jac_lnq"]))
    assert worst["rad"] < 1e-5 and worst["jac_t"] < 1e-4 and worst["jac_lnq"] < 1e-4
    for si, ref in keep.items():
        r = jac(scenes[si], alld[si], _lib.FP64).block(v1, v2)
        for q in ("rad", "jac_t", "jac_lnq"):
            err = float(np.abs(ref[q] - r[q].cpu().numpy()).max() / np.abs(ref[q]).max())
            print("scene %d %s: batch vs alone %.2e" % (si, q, err))
            assert err < 1e-12, (si, q, err)
    ctx.free_lines(h)
