"""noScattJacob (src/clblm_src/clblm_noScattJac.f90:20-336) through the HIP library,
for a nadir, down-looking, thermal sensor above the top layer.

Per spectral block [v1, v2] on the uniform grid DVout (the RT_Jac grid rule,
clblm_setDV.f90:203-208):

1. one clblm_od_batch of (3 + M) x L layer-OD jobs, in this order: the layer ODs
   (pathOD_up_dn_sun, :796-838), the ODs at Tave + dt (drv_LevelTemp, :1005-1009,
   dt = 1 K), the ODs at Wtot + 0.01 Wtot (:1018-1024), and the OD of each Jacobian
   molecule alone (drv_LevelConc, :1156-1166);
2. clblm_rt_jac: the radiance and its per-point derivatives (Jacobians, :346-716);
3. clblm_level_jac: the level Jacobians dR/dT_lev and dR/dln q_lev (:899-1203) with
   level2layer's factors (:1208-1279, clblm_level2layer, computed once);
4. with a boxcar wider than DVout, output_JacArray's pre-average (:1341-1409,
   clblm_scanfn) of the radiance and every Jacobian array of the block.

The down path is the view path scaled by sec_dn (the air-mass-scaled form of
pathOD_up_dn_sun, :826-838).  The OD variants are recomputed per block rather than
stored for the whole range, as the reference does.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib
from ._lib import FP32C, check
from .odlay import Context


def od_variants(layers, mols, dt=1.0, dw_frac=0.01):
    """The (3 + M) x L layer-OD jobs of one Jacobian block, variant-major, as
    (layer, takes the continuum): the layers (pathOD_up_dn_sun :796-838), the layers at
    Tave + dt (drv_LevelTemp :1005-1009), at Wtot + dw_frac Wtot (:1018-1024), and for
    each Jacobian molecule the layers with every other amount zero (drv_LevelConc
    :1156-1166; only water keeps the continuum, whose absorber is water)."""
    out = []
    for lay in layers:
        out.append((lay, True))
    for lay in layers:
        out.append((dataclasses.replace(lay, tave=lay.tave + dt), True))
    for lay in layers:
        out.append((dataclasses.replace(lay, wtot=lay.wtot + dw_frac * lay.wtot), True))
    for m in mols:
        for lay in layers:
            w = {mm: (v if mm == m else 0.0) for mm, v in lay.w.items()}
            out.append((dataclasses.replace(lay, w=w), m == 1))
    return out


def continuum_tangents(continuum, lay, v1, v2, dirs, steps=(0.01, 1e-4)):
    """A caller's continuum model (callable(layer, v1, v2) -> (absrb, v1abs, npt)) on
    the 1 cm-1 ABSRB grid and its derivatives along dirs ((dt, dlnwtot, {m: dlnw})
    each): the model's own `tangents(layer, v1, v2, dirs)` when it has one, else
    central differences of the model with steps (K, ln W).  The line sum's derivatives
    are analytic; the continuum is a smooth function of a few hundred points per layer
    that the caller owns."""
    if hasattr(continuum, "tangents"):   # the model's own derivatives
        return continuum.tangents(lay, v1, v2, dirs)
    c0 = np.asarray(continuum(lay, v1, v2)[0], np.float64)
    ht, hw = steps
    out = []
    for dt_, dlw, dq in dirs:
        h = ht if dt_ else hw
        moved = []
        for sgn in (1.0, -1.0):
            w = {m: v * (1.0 + sgn * h * dq.get(m, 0.0)) for m, v in lay.w.items()}
            lv = dataclasses.replace(lay, tave=lay.tave + sgn * h * dt_, wtot=lay.wtot * (1.0 + sgn * h * dlw), w=w)
            moved.append(np.asarray(continuum(lv, v1, v2)[0], np.float64))
        out.append((moved[0] - moved[1]) / (2.0 * h))
    return c0, np.array(out)


def layer_od(ctx, handle, layers, v1, v2, dvs, t_lay, lnw_tot, lnw, mols, dvout, continuum=None):
    """Differentiable layer optical depths (torch.autograd.Function on ROCm tensors).

    t_lay [L] (K), lnw_tot [L] (ln of the broadening column), lnw [L, M] (ln of the
    columns of ``mols``) are float64 tensors; the other layer fields come from
    ``layers``.  Returns OD [L, N] on v1 + i dvout (every layer re-gridded to dvout):
    the OD of the line sum with closed-form line shapes (clblm_od_tangent channel 0).
    The backward pass recomputes the layers' derivatives with one tangent batch
    instead of storing them, and contracts them with the incoming gradient:
    dL/dT_l = sum_v g[l, v] dOD_l(v)/dT_l and likewise for ln Wtot and ln W(m)."""
    return _LayerOD.apply(t_lay, lnw_tot, lnw, (ctx, handle, list(layers), v1, v2, list(dvs), tuple(mols), dvout,
                                                 continuum))


def _moved_layers(cfg, t_lay, lnw_tot, lnw):
    _, _, layers, _, _, _, mols, _, _ = cfg
    t = t_lay.detach().cpu().numpy()
    lwt = lnw_tot.detach().cpu().numpy()
    lw = lnw.detach().cpu().numpy()
    out = []
    for l, lay in enumerate(layers):
        w = dict(lay.w)
        for k, m in enumerate(mols):
            w[m] = float(np.exp(lw[l, k]))
        out.append(dataclasses.replace(lay, tave=float(t[l]), wtot=float(np.exp(lwt[l])), w=w))
    return out


def _layer_od_run(cfg, lays, dirs, jrad):
    from ._lib import FP64
    ctx, handle, _, v1, v2, dvs, _, dvout, continuum = cfg
    pars, dcont = [], []
    for l, lay in enumerate(lays):
        c = None
        if continuum is not None:
            c, dc = continuum_tangents(continuum, lay, v1, v2, dirs)
            dcont.append(dc)
        pars.append(Context.params(v1, v2, dvs[l], dvout=dvout, precision=FP64, continuum=c))
    ctx.jrad = jrad
    out, off = ctx.od_tangent(handle, lays, pars, dirs, dcontinuum=dcont if (dcont and dirs) else None)
    n = int(off[1] - off[0])
    if not np.all(np.diff(off) == n):
        raise ValueError("layer_od needs every layer on the dvout grid")
    return out.view(len(dirs) + 1, len(lays), n)


class UpwellingJacobian:
    """Radiance and level Jacobians of one scene.

    layers: L layers from the surface up; t_lev, p_lev: the L + 1 level temperatures
    and pressures; dvs: each layer's DV (the OD grids before the re-grid to DVout);
    q_lev: {molecule: L + 1 level number densities} of the Jacobian molecules (the
    reference's path%Q); mols: the Jacobian molecules (default: those of q_lev)."""

    def __init__(self, ctx: Context, handle: int, layers, t_lev, dvs, *, p_lev, q_lev=None, mols=None, emis=0.98,
                 tskin=295.0, lin_in_tau=1, sec_dn=1.0, precision=FP32C, dt=1.0, dw_frac=0.01, continuum=None,
                 dvout=None, boxcar_hw=-1.0, method="variants", continuum_step=(0.01, 1e-4)):
        import torch
        if method not in ("variants", "adjoint"):
            raise ValueError("method must be 'variants' or 'adjoint'")
        self.method = method
        # central-difference steps of the caller's continuum along T (K) and ln W
        self.cont_step = continuum_step
        self.ctx, self.h, self.layers = ctx, handle, list(layers)
        L = len(self.layers)
        self.t_lev = np.ascontiguousarray(t_lev, np.float64)
        self.p_lev = np.ascontiguousarray(p_lev, np.float64)
        if self.t_lev.size != L + 1 or self.p_lev.size != L + 1:
            raise ValueError("t_lev and p_lev need nlay + 1 levels")
        self.dvs = list(dvs)
        self.emis, self.tskin, self.lin, self.sec_dn = emis, tskin, int(lin_in_tau), sec_dn
        self.prec, self.dt, self.dw_frac = precision, float(dt), float(dw_frac)
        self.continuum = continuum   # callable(layer, v1, v2) -> (absrb, v1abs, npt) or None
        self.dvout = min(self.dvs) if dvout is None else dvout
        # doPreBox (clblm_noScattJac.f90:187-189): a boxcar wider than DVout spanning
        # more than one point pre-averages the outputs
        self.boxcar_hw = float(boxcar_hw)
        self.prebox = self.boxcar_hw > 0.0 and self.boxcar_hw > self.dvout and \
            int(2.0 * self.boxcar_hw / self.dvout) > 1
        q_lev = q_lev or {}
        self.mols = sorted(q_lev) if mols is None else list(mols)
        missing = [m for m in self.mols if m not in q_lev]
        if missing:
            raise ValueError("no level densities for molecules %s" % missing)
        M = len(self.mols)
        t_lay = np.array([l.tave for l in self.layers], np.float64)
        q = np.ascontiguousarray(np.array([q_lev[m] for m in self.mols], np.float64).reshape(M, L + 1))
        up, lo = np.zeros(L), np.zeros(L)
        wu, wl = np.zeros((M, L)), np.zeros((M, L))
        check(ctx.lib.clblm_level2layer(L, M, self.p_lev.ctypes.data, self.t_lev.ctypes.data, t_lay.ctypes.data,
                                        q.ctypes.data if M else None, up.ctypes.data, lo.ctypes.data,
                                        wu.ctypes.data if M else None, wl.ctypes.data if M else None))
        self.dt_dtup, self.dt_dtlo, self.dw_dqup, self.dw_dqlo = up, lo, wu, wl
        d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
        self._d = dict(tlay=d(t_lay), tlev=d(self.t_lev), wtot=d([l.wtot for l in self.layers]), up=d(up), lo=d(lo),
                       wu=d(wu.reshape(-1)) if M else None, wl=d(wl.reshape(-1)) if M else None)

    def _variants(self):
        return od_variants(self.layers, self.mols, self.dt, self.dw_frac)

    def _jobs(self, v1, v2):
        jobs, pars = [], []
        L = len(self.layers)
        for k, (lv, cont) in enumerate(self._variants()):
            c = None
            if cont and self.continuum is not None:
                c = self.continuum(lv, v1, v2)[0]
            jobs.append(lv)
            pars.append(Context.params(v1, v2, self.dvs[k % L], dvout=self.dvout, precision=self.prec, continuum=c))
        return jobs, pars

    def _continuum_tangents(self, lay, v1, v2, dirs):
        return continuum_tangents(self.continuum, lay, v1, v2, dirs, self.cont_step)

    def adjoint_dirs(self):
        """The tangent batch's directions: temperature, the broadening column, then each
        Jacobian molecule's column (as ln W)."""
        return [(1.0, 0.0, {}), (0.0, 1.0, {})] + [(0.0, 0.0, {m: 1.0}) for m in self.mols]

    def _block_adjoint(self, v1, v2):
        """One tangent batch gives the layer ODs and dtau/dT, dtau/dln Wtot and
        dtau/dln W(m) exactly (clblm_od_tangent); drv_LevelTemp's combination
        dtau/dT - dtau/dWtot Wtot/Tave (clblm_noScattJac.f90:1027-1031) and drv_LevelConc's
        use of dtau/dln W (:1159-1171) are then assembled by clblm_level_jac with unit
        steps and a zero base block, so that no difference is formed."""
        import torch
        from ._lib import FP64
        L, M = len(self.layers), len(self.mols)
        dirs = self.adjoint_dirs()
        pars, dcont = [], []
        for l, lay in enumerate(self.layers):
            c = None
            if self.continuum is not None:
                c, dc = self._continuum_tangents(lay, v1, v2, dirs)
                dcont.append(dc)
            pars.append(Context.params(v1, v2, self.dvs[l], dvout=self.dvout, precision=FP64, continuum=c))
        self.ctx.jrad = 1
        out, off = self.ctx.od_tangent(self.h, self.layers, pars, dirs, dcontinuum=dcont if dcont else None)
        n = int(off[1] - off[0])
        assert np.all(np.diff(off) == n), "all layers of a block share the DVout grid"
        od = out[0].view(L, n)
        od_var = torch.cat([torch.zeros((1, L, n), dtype=torch.float64, device=out.device),
                            out[1:].view(2 + M, L, n)]).contiguous()
        return od, od_var, n, 1.0, 1.0

    def block(self, v1, v2):
        """Radiance and Jacobians on [v1, v2] (uniform grid DVout).  Returns torch tensors:
        rad [N]; jac_t [L+1, N] = dR/dT_lev; jac_lnq [M, L+1, N] = dR/dln q_lev;
        dr_dtskin, dr_demis [N]; and the per-point derivatives dr_dtau [L, N],
        dr_dtlev [L+1, N] and the layer ODs od [L, N].  grid = (v1, dv) of rad, jac_t,
        jac_lnq, dr_dtskin and dr_demis: the block's grid, or the boxcar's output grid
        when the outputs are pre-averaged."""
        L, M = len(self.layers), len(self.mols)
        if self.method == "adjoint":
            od0, od, n, dt, dwf = self._block_adjoint(v1, v2)
        else:
            jobs, pars = self._jobs(v1, v2)
            self.ctx.jrad = 1
            od, off = self.ctx.od_batch(self.h, jobs, pars)
            n = int(off[1] - off[0])
            assert np.all(np.diff(off) == n), "all jobs of a block share the DVout grid"
            od = od.view(3 + M, L, n)
            od0, dt, dwf = od[0], self.dt, self.dw_frac
        return self.assemble(v1, od0, od, dt, dwf)

    def assemble(self, v1, od0, od, dt=None, dwf=None):
        """The RT and level assembly of a block from its OD variants (as block computes
        them, or from a batch that holds many scenes' variants): od0 [L, N] the layer ODs,
        od [(3 + M), L, N] the variants (or the derivative blocks with unit steps)."""
        import torch
        L, M = len(self.layers), len(self.mols)
        dt = self.dt if dt is None else dt
        dwf = self.dw_frac if dwf is None else dwf
        n = od0.shape[1]
        dev = od.device
        e = lambda *shape: torch.empty(*shape, dtype=torch.float64, device=dev)
        rad, dr_dtau, dr_dtlev, dts, de = e(n), e(L, n), e(L + 1, n), e(n), e(n)
        jac_t, jac_q = e(L + 1, n), (e(M, L + 1, n) if M else None)
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        D = self._d
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        lib = self.ctx.lib
        check(lib.clblm_rt_jac(self.ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(self.dvout), p(od0),
                               p(D["tlay"]), p(D["tlev"]), p(D["up"]), p(D["lo"]), ctypes.c_double(self.emis),
                               ctypes.c_double(self.tskin), self.lin, ctypes.c_double(self.sec_dn), p(rad),
                               p(dr_dtau), p(dr_dtlev), p(dts), p(de), st))
        check(lib.clblm_level_jac(self.ctx.ptr, L, M, n, p(od), p(dr_dtau), p(dr_dtlev), p(D["wtot"]), p(D["tlay"]),
                                  ctypes.c_double(dt), ctypes.c_double(dwf), p(D["up"]), p(D["lo"]),
                                  p(D["wu"]), p(D["wl"]), p(jac_t), p(jac_q), st))
        grid = (float(v1), float(self.dvout))
        if self.prebox:   # output_JacArray (:1341-1409), per block as the reference does
            box = lambda t: self.ctx.prebox(t, v1, self.dvout, self.boxcar_hw)
            rad, g1, gd = box(rad)
            jac_t = box(jac_t)[0]
            jac_q = box(jac_q)[0] if M else None
            dts, de = box(dts)[0], box(de)[0]
            grid = (g1, gd)
        return dict(rad=rad, jac_t=jac_t, jac_lnq=jac_q, dr_dtskin=dts, dr_demis=de, dr_dtau=dr_dtau,
                    dr_dtlev=dr_dtlev, od=od0, od_var=od, n=n, grid=grid)


def upwelling_radiance_grids(ctx, od, offsets, dvs, v1, t_lay, t_lev, emis=0.98, tskin=295.0, lin_in_tau=1,
                             sec_dn=1.0, surf_dv=None, surf_nlim=None):
    """Upwelling radiance of layers each on its own grid (clblm_rt_up_grids: the merge
    of mode2_lookDn_mergeUp with interpToFinestGrid, src/clblm_src/clblm_noScattRT.f90:
    1700-1794).  od: float64 cuda tensor of the layer ODs back to back, layer l at
    offsets[l]..offsets[l+1] on v1 + i*dvs[l] (clblm_od_batch's output without DVout);
    t_lay [L], t_lev [L+1] host sequences (K).  The surface grid defaults to layer 0's.
    Returns (radiance tensor, its DV)."""
    import torch
    off = np.ascontiguousarray(offsets, np.int64)
    L = off.size - 1
    nl = np.ascontiguousarray(np.diff(off), np.int64)
    dv = np.ascontiguousarray(dvs, np.float64)
    tl, tv = np.ascontiguousarray(t_lay, np.float64), np.ascontiguousarray(t_lev, np.float64)
    if dv.size != L or tl.size != L or tv.size != L + 1:
        raise ValueError("dvs and t_lay need one value per layer, t_lev one per level")
    if not od.is_cuda or od.dtype != torch.float64 or od.numel() < off[-1]:
        raise ValueError("od must be a float64 cuda tensor holding every layer")
    sdv = float(dv[0] if surf_dv is None else surf_dv)
    sn = int(nl[0] if surf_nlim is None else surf_nlim)
    dvo, no = ctypes.c_double(), ctypes.c_int64()
    args = [ctx.ptr, L, ctypes.c_double(v1), dv.ctypes.data, nl.ctypes.data, off.ctypes.data]
    check(ctx.lib.clblm_rt_up_grids(*args, None, tl.ctypes.data, tv.ctypes.data, ctypes.c_double(emis),
                                    ctypes.c_double(tskin), ctypes.c_double(sdv), sn, int(lin_in_tau),
                                    ctypes.c_double(sec_dn), None, ctypes.byref(dvo), ctypes.byref(no), None))
    rad = torch.empty(no.value, dtype=torch.float64, device=od.device)
    od = od.contiguous()
    check(ctx.lib.clblm_rt_up_grids(*args, ctypes.c_void_p(od.data_ptr()), tl.ctypes.data, tv.ctypes.data,
                                    ctypes.c_double(emis), ctypes.c_double(tskin), ctypes.c_double(sdv), sn,
                                    int(lin_in_tau), ctypes.c_double(sec_dn), ctypes.c_void_p(rad.data_ptr()),
                                    ctypes.byref(dvo), ctypes.byref(no),
                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return rad, dvo.value


def radiance_mode1(ctx, od, offsets, dvs, v1, v2, t_lay, t_lev, od_dn=None, od_sun=None, fac_up=None, fac_dn=None,
                   fac_sun=None, obs_level=None, lin_in_tau=1, thermal=True, solar=False, emis=0.98, tskin=295.0,
                   refl=0.0, sol=None, sol_dv=None):
    """Radiance at an observer above the layers, merged from the top down
    (clblm_rt_mode1: mode1_lookDn_mergeDn / layerMerge_mode1 with the thermal and solar
    boundary terms, src/clblm_src/clblm_noScattRT.f90:42-842, 1945-2148).  od: float64
    cuda tensor of the view-path layer ODs back to back (layer l at offsets[l], on
    v1 + i*dvs[l], layer 0 at the surface); od_dn / od_sun the down and solar path ODs
    laid out alike (default: od), each scaled by its per-layer factors (default 1).
    sol: cuda tensor of the solar radiance at the top on v1 + i*sol_dv.  Returns
    (radiance, total view transmittance, DV)."""
    import torch
    off = np.ascontiguousarray(offsets, np.int64)
    L = off.size - 1
    nl = np.ascontiguousarray(np.diff(off), np.int64)
    dv = np.ascontiguousarray(dvs, np.float64)
    tl, tv = np.ascontiguousarray(t_lay, np.float64), np.ascontiguousarray(t_lev, np.float64)
    if dv.size != L or tl.size != L or tv.size != L + 1:
        raise ValueError("dvs and t_lay need one value per layer, t_lev one per level")
    for t in (od, od_dn, od_sun):
        if t is not None and (not t.is_cuda or t.dtype != torch.float64 or t.numel() < off[-1]):
            raise ValueError("the ODs must be float64 cuda tensors holding every layer")
    p = _lib.Mode1Params()
    p.nlay, p.obs_level, p.lin_in_tau = L, L if obs_level is None else int(obs_level), int(lin_in_tau)
    p.thermal, p.solar = int(bool(thermal)), int(bool(solar))
    p.v1, p.v2, p.surf_emis, p.t_skin, p.surf_refl = float(v1), float(v2), float(emis), float(tskin), float(refl)
    fac = [None if f is None else np.ascontiguousarray(f, np.float64) for f in (fac_up, fac_dn, fac_sun)]
    fp = [None if f is None else f.ctypes.data for f in fac]
    od = od.contiguous()
    dn = od if od_dn is None else od_dn.contiguous()
    sn = od if od_sun is None else od_sun.contiguous()
    s0 = sol.contiguous() if sol is not None else None
    sdv = float(dv[-1] if sol_dv is None else sol_dv)
    sol_n = 0 if s0 is None else s0.numel()
    dvo, no = ctypes.c_double(), ctypes.c_int64()
    ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    head = [ctx.ptr, ctypes.byref(p), dv.ctypes.data, nl.ctypes.data, off.ctypes.data, ptr(od), ptr(dn), ptr(sn),
            *fp, tl.ctypes.data, tv.ctypes.data, ptr(s0), sol_n, ctypes.c_double(sdv)]
    check(ctx.lib.clblm_rt_mode1(*head, None, None, ctypes.byref(dvo), ctypes.byref(no), None))
    rad = torch.empty(no.value, dtype=torch.float64, device=od.device)
    tx = torch.empty(no.value, dtype=torch.float64, device=od.device)
    check(ctx.lib.clblm_rt_mode1(*head, ptr(rad), ptr(tx), ctypes.byref(dvo), ctypes.byref(no),
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return rad, tx, dvo.value


def jacobians_paths(ctx, od, t_lay, t_lev, v1, dv, amf_view=None, amf_down=None, amf_sun=None, obs_level=None,
                    uplook=False, lin_in_tau=1, emis=0.98, tskin=295.0, solar=False, sol=None, rsfc_sun=0.0,
                    dt_dtup=None, dt_dtlo=None):
    """Per-point Jacobians with every branch of the reference (clblm_rt_jac_paths:
    src/clblm_src/clblm_noScattJac.f90:346-716): per-layer air-mass factors of the view,
    down and solar paths, an observer at any level, up- or down-looking, the solar term.
    od: float64 cuda tensor [L, N] of vertical layer ODs on v1 + i*dv (layer 0 at the
    surface); t_lay [L], t_lev [L+1]; sol: cuda tensor [N] of the solar radiance at the
    top.  Returns a dict of cuda tensors: rad, dtau [L, N], dtlev [L+1, N], dtskin, demis,
    drsfc."""
    import torch
    L, n = od.shape
    if not od.is_cuda or od.dtype != torch.float64:
        raise ValueError("od must be a float64 cuda tensor")
    dev = od.device
    tl = torch.as_tensor(np.asarray(t_lay, np.float64), device=dev)
    tv = torch.as_tensor(np.asarray(t_lev, np.float64), device=dev)
    if tl.numel() != L or tv.numel() != L + 1:
        raise ValueError("t_lay needs one value per layer, t_lev one per level")
    lin = int(lin_in_tau)
    dup = dlo = None
    if lin != 2:
        if dt_dtup is None or dt_dtlo is None:
            raise ValueError("lin_in_tau 0 and 1 need dt_dtup and dt_dtlo")
        dup = torch.as_tensor(np.asarray(dt_dtup, np.float64), device=dev)
        dlo = torch.as_tensor(np.asarray(dt_dtlo, np.float64), device=dev)
        if dup.numel() != L or dlo.numel() != L:
            raise ValueError("dt_dtup and dt_dtlo need one value per layer")
    if solar and not uplook:
        if sol is None or not sol.is_cuda or sol.numel() != n or sol.dtype != torch.float64:
            raise ValueError("the solar term needs sol, a float64 cuda tensor of one value per point")
    keep = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (amf_view, amf_down, amf_sun)]
    for a in keep:
        if a is not None and a.size != L:
            raise ValueError("air-mass factors need one value per layer")
    p = _lib.JacPathsParams()
    p.nlay, p.obs_level, p.uplook, p.lin_in_tau = L, L if obs_level is None else int(obs_level), int(bool(uplook)), lin
    p.solar, p.surf_emis, p.t_skin, p.rsfc_sun = int(bool(solar)), float(emis), float(tskin), float(rsfc_sun)
    p.amf_view, p.amf_down, p.amf_sun = [None if a is None else a.ctypes.data for a in keep]
    od = od.contiguous()
    out = dict(rad=torch.empty(n, dtype=torch.float64, device=dev), dtau=torch.empty_like(od),
               dtlev=torch.empty((L + 1, n), dtype=torch.float64, device=dev),
               dtskin=torch.empty(n, dtype=torch.float64, device=dev), demis=torch.empty(n, dtype=torch.float64, device=dev),
               drsfc=torch.empty(n, dtype=torch.float64, device=dev))
    ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    check(ctx.lib.clblm_rt_jac_paths(ctx.ptr, ctypes.byref(p), n, ctypes.c_double(v1), ctypes.c_double(dv), ptr(od),
                                     ptr(tl), ptr(tv), ptr(dup), ptr(dlo), ptr(sol if (solar and not uplook) else None),
                                     *[ptr(out[k]) for k in ("rad", "dtau", "dtlev", "dtskin", "demis", "drsfc")],
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out


def fluxes(ctx, od, t_lay, t_lev, p_lev, v1, v2, dv, dv_flux, nang=3, lin_in_tau=1, emis=0.98, tskin=295.0,
           lambertian=True):
    """Upwelling and downwelling fluxes, net flux and heating rate at every level
    (clblm_rt_flux: the flux branches of mode1_lookDn_mergeDn and mode2_lookDn_mergeUp,
    src/clblm_src/clblm_noScattRT.f90:133-507, 1214-1450, and the driver's net flux and
    heating rate, clblm_Drivers.f90:418-466).  od: float64 cuda tensor [L, N] of the layer
    ODs on v1 + i*dv (layer 0 at the surface); t_lay [L], t_lev [L+1] (K), p_lev [L+1]
    (hPa).  Returns a dict of numpy arrays [L+1, nbin] (level 0 at the surface, groups of
    dv_flux from v1): up, dn, net (W/m2), heating (K/day)."""
    import torch
    L, n = od.shape
    if not od.is_cuda or od.dtype != torch.float64:
        raise ValueError("od must be a float64 cuda tensor")
    tl = torch.as_tensor(np.asarray(t_lay, np.float64), device=od.device)
    tv = torch.as_tensor(np.asarray(t_lev, np.float64), device=od.device)
    pl = np.ascontiguousarray(p_lev, np.float64)
    if tl.numel() != L or tv.numel() != L + 1 or pl.size != L + 1:
        raise ValueError("t_lay needs one value per layer, t_lev and p_lev one per level")
    p = _lib.FluxParams()
    p.nlay, p.nang, p.lin_in_tau, p.lambertian = L, int(nang), int(lin_in_tau), int(bool(lambertian))
    p.v1, p.v2, p.dv, p.dv_flux, p.surf_emis, p.t_skin = float(v1), float(v2), float(dv), float(dv_flux), \
        float(emis), float(tskin)
    nb = ctypes.c_int32()
    od = od.contiguous()
    args = [ctx.ptr, ctypes.byref(p), ctypes.c_void_p(od.data_ptr()), n, ctypes.c_void_p(tl.data_ptr()),
            ctypes.c_void_p(tv.data_ptr()), pl.ctypes.data]
    check(ctx.lib.clblm_rt_flux(*args, None, None, None, None, ctypes.byref(nb), None))
    out = {k: np.zeros((L + 1, nb.value)) for k in ("up", "dn", "net", "heating")}
    check(ctx.lib.clblm_rt_flux(*args, *[out[k].ctypes.data for k in ("up", "dn", "net", "heating")],
                                ctypes.byref(nb), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out


def upwelling_radiance(ctx, od, t_lay, t_lev, v1, dv, emis=0.98, tskin=295.0, lin_in_tau=1, sec_dn=1.0):
    """Differentiable radiance R(OD, T_lay) (torch.autograd.Function on ROCm tensors)."""
    return _Radiance.apply(od, t_lay, ctx, t_lev, v1, dv, emis, tskin, lin_in_tau, sec_dn)


def _rt(ctx, od, t_lay, t_lev, v1, dv, emis, tskin, lin, sec_dn, grads):
    import torch
    L, n = od.shape
    rad = torch.empty(n, dtype=torch.float64, device=od.device)
    g1 = torch.empty_like(od) if grads else None
    g2 = torch.empty_like(od) if grads else None
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    check(ctx.lib.clblm_rt_up(ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(dv), p(od), p(t_lay), p(t_lev),
                              ctypes.c_double(emis), ctypes.c_double(tskin), lin_in_tau_int(lin),
                              ctypes.c_double(sec_dn), p(rad), p(g1), p(g2), None, None,
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return rad, g1, g2


def lin_in_tau_int(x):
    return int(x)


try:
    import torch

    class _Radiance(torch.autograd.Function):
        @staticmethod
        def forward(fctx, od, t_lay, ctx, t_lev, v1, dv, emis, tskin, lin, sec_dn):
            od = od.contiguous().double()
            t_lay = t_lay.contiguous().double()
            t_lev = t_lev.contiguous().double()
            rad, g_od, g_t = _rt(ctx, od, t_lay, t_lev, v1, dv, emis, tskin, lin, sec_dn, True)
            fctx.save_for_backward(g_od, g_t)
            return rad

        @staticmethod
        def backward(fctx, grad):
            g_od, g_t = fctx.saved_tensors
            return (grad[None, :] * g_od, (grad[None, :] * g_t).sum(dim=1), None, None, None, None, None, None,
                    None, None)
    class _LayerOD(torch.autograd.Function):
        @staticmethod
        def forward(fctx, t_lay, lnw_tot, lnw, cfg):
            lays = _moved_layers(cfg, t_lay, lnw_tot, lnw)
            jr = cfg[0].jrad
            od = _layer_od_run(cfg, lays, [], jr)[0]
            fctx.cfg, fctx.lays, fctx.jrad = cfg, lays, jr
            fctx.devs = (t_lay.device, lnw_tot.device, lnw.device)
            return od

        @staticmethod
        def backward(fctx, grad):
            mols = fctx.cfg[6]
            dirs = [(1.0, 0.0, {}), (0.0, 1.0, {})] + [(0.0, 0.0, {m: 1.0}) for m in mols]
            ch = _layer_od_run(fctx.cfg, fctx.lays, dirs, fctx.jrad)   # recomputed, not stored
            g = grad.to(torch.float64)
            gt = (g * ch[1]).sum(dim=1)
            gwt = (g * ch[2]).sum(dim=1)
            gw = torch.stack([(g * ch[3 + k]).sum(dim=1) for k in range(len(mols))], dim=1) if mols else None
            dv = fctx.devs
            return gt.to(dv[0]), gwt.to(dv[1]), (gw.to(dv[2]) if gw is not None else None), None
except ImportError:  # pragma: no cover - torch is part of the stack
    _Radiance = None


def downwelling_radiance(ctx, od, t_lay, t_lev, v1, dv, lin_in_tau=1, sol_toa=None):
    """Down-welling radiance at the observer of a look-up path (mode3_lookUp_mergeUp,
    src/clblm_src/clblm_noScattRT.f90:843-1108) from float64 cuda tensors: od [L, N]
    (layer 0 at the observer), t_lay [L], t_lev [L+1]; sol_toa [N] (optional) the solar
    radiance at the top.  Returns (radiance [N], path transmittance [N])."""
    import torch
    od = od.contiguous().double()
    if od.dim() != 2 or not od.is_cuda:
        raise ValueError("od must be a [layers, points] cuda tensor")
    L, n = od.shape
    t_lay = t_lay.contiguous().double()
    t_lev = t_lev.contiguous().double()
    sol = sol_toa.contiguous().double() if sol_toa is not None else None
    # the kernel reads t_lay[L], t_lev[L+1] and sol[N] on the device: check them here
    for name, t, want in (("t_lay", t_lay, L), ("t_lev", t_lev, L + 1), ("sol_toa", sol, n)):
        if t is not None and (t.numel() != want or t.device != od.device):
            raise ValueError("%s must hold %d values on %s (has %d on %s)" % (name, want, od.device, t.numel(),
                                                                            t.device))
    rad = torch.empty(n, dtype=torch.float64, device=od.device)
    tx = torch.empty(n, dtype=torch.float64, device=od.device)
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    check(ctx.lib.clblm_rt_down(ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(dv), p(od), p(t_lay), p(t_lev),
                                int(lin_in_tau), p(sol), p(rad), p(tx),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return rad, tx
