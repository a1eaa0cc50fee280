"""Upwelling radiance and its Jacobians through the HIP library.

Mirrors noScattJacob (src/clblm_src/clblm_noScattJac.f90:20-336) for a nadir,
down-looking, thermal case: layer ODs on one uniform grid (DVout, the RT_Jac
grid rule of clblm_setDV.f90:203-208), the merge-up radiance, and per spectral
point the analytic derivative of the radiance with respect to each layer's OD
and source temperature.  Layer Jacobians are assembled as the reference does:

    dR/dT_l      = dR/dtau_l * dtau_l/dT_l + dR/dB_l * dB_l/dT_l
    dR/dln q_l,m = dR/dtau_l * tau_l,m      (tau_l,m: OD of molecule m alone,
                                             clblm_noScattJac.f90:1159-1171)

dtau_l/dT_l is the central difference of two full layer-OD passes at T_l +- dT
(the reference uses a one-sided +1 K difference, clblm_noScattJac.f90:1005-1014).
Layer ODs are recomputed per spectral block rather than stored for the whole range.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import workloads as W
from ._lib import FP32C, FP64, check
from .odlay import Context


class UpwellingJacobian:
    def __init__(self, ctx: Context, handle: int, layers, t_lev, dvs, *, emis=0.98, tskin=295.0, lin_in_tau=1,
                 sec_dn=1.0, precision=FP32C, dt=0.1, continuum=None, dvout=None):
        self.ctx, self.h, self.layers = ctx, handle, list(layers)
        self.t_lev = np.asarray(t_lev, np.float64)
        self.dvs = list(dvs)
        self.emis, self.tskin, self.lin, self.sec_dn = emis, tskin, lin_in_tau, sec_dn
        self.prec, self.dt = precision, dt
        self.continuum = continuum   # callable(layer, v1, v2) -> (absrb, v1abs, npt) or None
        self.dvout = min(self.dvs) if dvout is None else dvout
        self.mols = sorted({m for l in self.layers for m in l.w})

    def _jobs(self, v1, v2):
        """Layer-OD jobs of one spectral block: base, T+dT, T-dT, then each molecule alone."""
        import copy
        jobs, pars = [], []
        for k, lay in enumerate(self.layers):
            variants = [(lay, True)]
            for sgn in (+1, -1):
                l2 = copy.copy(lay)
                l2.tave = lay.tave + sgn * self.dt
                variants.append((l2, True))
            for m in self.mols:
                l2 = copy.copy(lay)
                l2.w = {mm: (v if mm == m else 0.0) for mm, v in lay.w.items()}
                variants.append((l2, m == 1))
            for lv, cont in variants:
                c = None
                if cont and self.continuum is not None:
                    c = self.continuum(lv, v1, v2)[0]
                jobs.append(lv)
                pars.append(Context.params(v1, v2, self.dvs[k], dvout=self.dvout, precision=self.prec,
                                           continuum=c))
        return jobs, pars

    def block(self, v1, v2):
        """Radiance and Jacobians on [v1, v2] (uniform grid DVout).  Returns torch
        tensors: rad [N], jac_t [L, N], jac_lnq [L, M, N], plus tskin/emis derivatives."""
        import torch
        jobs, pars = self._jobs(v1, v2)
        self.ctx.jrad = 1
        od, off = self.ctx.od_batch(self.h, jobs, pars)
        n = int(off[1] - off[0])
        assert np.all(np.diff(off) == n), "all jobs of a block share the DVout grid"
        L, M = len(self.layers), len(self.mols)
        nv = 3 + M
        od = od.view(L, nv, n)
        base = od[:, 0, :].contiguous()
        dtau_dt = (od[:, 1, :] - od[:, 2, :]) / (2.0 * self.dt)
        tau_m = od[:, 3:, :]
        dev = base.device
        tlay = torch.tensor([l.tave for l in self.layers], dtype=torch.float64, device=dev)
        tlev = torch.tensor(self.t_lev, dtype=torch.float64, device=dev)
        rad = torch.empty(n, dtype=torch.float64, device=dev)
        dr_dod = torch.empty(L, n, dtype=torch.float64, device=dev)
        dr_dt = torch.empty(L, n, dtype=torch.float64, device=dev)
        dr_dts = torch.empty(n, dtype=torch.float64, device=dev)
        dr_de = torch.empty(n, dtype=torch.float64, device=dev)
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())
        check(self.ctx.lib.clblm_rt_up(self.ctx.ptr, L, n, ctypes.c_double(v1), ctypes.c_double(self.dvout),
                                       ptr(base), ptr(tlay), ptr(tlev), ctypes.c_double(self.emis),
                                       ctypes.c_double(self.tskin), self.lin, ctypes.c_double(self.sec_dn),
                                       ptr(rad), ptr(dr_dod), ptr(dr_dt), ptr(dr_dts), ptr(dr_de),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        jac_t = dr_dod * dtau_dt + dr_dt
        jac_q = dr_dod[:, None, :] * tau_m
        return dict(rad=rad, jac_t=jac_t, jac_lnq=jac_q, dr_dtskin=dr_dts, dr_demis=dr_de, od=base,
                    dr_dod=dr_dod, n=n)


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
except ImportError:  # pragma: no cover - torch is part of the stack
    _Radiance = None
