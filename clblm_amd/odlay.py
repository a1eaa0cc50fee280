"""Python driver of the layer optical-depth path.

``Context.odlay`` mirrors the reference entry point ODLAY_nonStruct
(src/clblm_src/clblm_ODLAY.f90:129-176): one layer, its molecules and columns,
V1/V2, DVnormal/DVnarrow, SAMPLE and an optional DVout, returning the OD
spectrum and the DV actually used.  JRAD is carried across calls exactly like
the reference's SAVE variable (clblm_ODLAY.f90:182-183, 349-352).
``Context.od_batch`` runs many layers in one set of launches and returns a
ROCm tensor.  Everything goes through libclblm_hip.so; nothing is computed on
the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import FP32C, FP64, ODParams, check


def _layer_struct(layer) -> _lib.Layer:
    s = _lib.Layer()
    s.pave, s.tave, s.wtot = float(layer.pave), float(layer.tave), float(layer.wtot)
    s.zbot, s.ztop = float(getattr(layer, "zbot", 0.0)), float(getattr(layer, "ztop", 0.0))
    for m, v in layer.w.items():
        s.w[int(m)] = float(v)
        s.sel[int(m)] = 1
    return s


class ODGraph:
    """A batch captured by ``Context.od_graph``: ``launch()`` replays it on ``stream``
    (default: torch's current stream), ``check()`` waits for the last replay and
    raises a device error, ``close()`` frees it.  ``out`` and ``offsets`` are those
    of ``Context.od_batch``."""

    def __init__(self, ctx, g, out, offsets):
        self.ctx, self.g, self.out, self.offsets = ctx, g, out, offsets

    def launch(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(self.ctx.lib.clblm_od_graph_launch(self.ctx.ptr, self.g, ctypes.c_void_p(stream)))
        return self.out

    def check(self):
        check(self.ctx.lib.clblm_od_graph_check(self.ctx.ptr, self.g))

    def close(self):
        if self.g:
            check(self.ctx.lib.clblm_od_graph_destroy(self.ctx.ptr, self.g))
            self.g = None


class Context:
    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.ptr = ctypes.c_void_p()
        check(self.lib.clblm_ctx_create(int(device), ctypes.byref(self.ptr)))
        self.jrad = 1
        self._keep = []

    def close(self):
        if self.ptr:
            self.lib.clblm_ctx_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- line list ------------------------------------------------------------
    def upload_lines(self, lines) -> int:
        arrs = [np.ascontiguousarray(a) for a in (lines.nu, lines.s, lines.a_air, lines.a_self, lines.e_lo,
                                                  lines.tmpalf, lines.shift, lines.mol, lines.iflg)]
        assert arrs[0].dtype == np.float64 and arrs[7].dtype == np.int32
        h = ctypes.c_int()
        check(self.lib.clblm_linelist_upload(self.ptr, int(lines.n), *[a.ctypes.data for a in arrs],
                                             ctypes.byref(h)))
        if getattr(lines, "brd_flg", None) is not None:
            b = [np.ascontiguousarray(getattr(lines, f), dt).reshape(lines.n, 7)
                 for f, dt in (("brd_flg", np.int32), ("brd_hw", np.float32), ("brd_tmp", np.float32),
                               ("brd_shft", np.float32))]
            check(self.lib.clblm_linelist_broadening(self.ptr, h.value, *[a.ctypes.data for a in b]))
        return h.value

    def load_tape3(self, path, vlo=1.0, vhi=0.0) -> int:
        """Reads a TAPE3 line file (clblm_linelist_load_tape3) and uploads its panels
        that overlap [vlo, vhi] (every panel when vlo > vhi): give the range of the
        runs' V1 - 25 - 3 DVR4 to V2 + 25 + 3 DVR4, the reach of LINF4Q
        (clblm_ODLAY.f90:913-924).  Returns the handle."""
        h, n = ctypes.c_int(), ctypes.c_int64()
        check(self.lib.clblm_linelist_load_tape3(self.ptr, str(path).encode(), float(vlo), float(vhi),
                                                 ctypes.byref(h), ctypes.byref(n)))
        return h.value

    # -- MT_CKD water-vapour continuum ---------------------------------------------
    def mtckd_load(self, table, xself=1.0, xfrgn=1.0):
        """clblm_mtckd_load from a table of clblm_amd.continuum.load_mtckd."""
        arr = [np.ascontiguousarray(table[k], np.float64) for k in
               ("wavenumbers", "self_absco_ref", "for_absco_ref", "self_texp")]
        check(self.lib.clblm_mtckd_load(self.ptr, int(arr[0].size), *[a.ctypes.data for a in arr],
                                        float(table["ref_temp"]), float(table["ref_press"]), float(xself),
                                        float(xfrgn)))

    def contnm_scales(self, xco2c=1.0, xo3cn=1.0, xo2cn=1.0, xn2cn=1.0, xrayl=1.0):
        """clblm_contnm_scales: CONTNM's scale factors for the layers that ask for its
        other continua (Context.params(..., other_continua=True))."""
        check(self.lib.clblm_contnm_scales(self.ptr, float(xco2c), float(xo3cn), float(xo2cn), float(xn2cn),
                                           float(xrayl)))

    def mtckd_absrb(self, layers, v1, v2, jrad, stream=None):
        """clblm_mtckd_absrb: the water-vapour continuum of each layer on its ABSRB grid;
        returns a list of float64 cuda tensors."""
        return self.contnm_absrb(layers, v1, v2, jrad, _lib.OD_MTCKD, stream)

    def contnm_absrb(self, layers, v1, v2, jrad, flags=_lib.OD_MTCKD | _lib.OD_CONTNM, stream=None):
        """clblm_contnm_absrb: CONTNM's ABSRB of each layer for the continua `flags`
        selects (OD_MTCKD, OD_CONTNM); returns a list of float64 cuda tensors."""
        import torch
        n = len(layers)
        ls = (_lib.Layer * n)(*[_layer_struct(lay) for lay in layers])
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v1, np.float64), (n,)))
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(v2, np.float64), (n,)))
        j = np.ascontiguousarray(np.broadcast_to(np.asarray(jrad, np.int32), (n,)))
        npt = np.zeros(n, np.int32)
        check(self.lib.clblm_contnm_absrb(self.ptr, n, ls, a.ctypes.data, b.ctypes.data, j.ctypes.data, int(flags),
                                          None, 0, npt.ctypes.data, None))
        ld = int(npt.max())
        out = torch.empty((n, ld), dtype=torch.float64, device="cuda")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(self.lib.clblm_contnm_absrb(self.ptr, n, ls, a.ctypes.data, b.ctypes.data, j.ctypes.data, int(flags),
                                          ctypes.c_void_p(out.data_ptr()), ld, npt.ctypes.data,
                                          ctypes.c_void_p(stream)))
        return [out[l, :npt[l]] for l in range(n)]

    def free_lines(self, handle):
        check(self.lib.clblm_linelist_free(self.ptr, int(handle)))

    # -- one layer (host buffers) ----------------------------------------------
    @staticmethod
    def params(v1, v2, dvnormal, dvnarrow=None, sample=4.0, dvout=0.0, precision=FP64, continuum=None,
               species_broad=False, mtckd=False, other_continua=False):
        p = ODParams()
        p.flags = ((_lib.OD_SPECIES_BROAD if species_broad else 0) | (_lib.OD_MTCKD if mtckd else 0) |
                   (_lib.OD_CONTNM if other_continua else 0))
        p.v1, p.v2, p.dvnormal = float(v1), float(v2), float(dvnormal)
        p.dvnarrow = float(dvnormal if dvnarrow is None else dvnarrow)
        p.sample, p.dvout, p.precision = float(sample), float(dvout), int(precision)
        if continuum is not None:
            c = np.ascontiguousarray(continuum, np.float64)
            p.contnm, p.continuum, p.nptabs = 1, c.ctypes.data, int(c.size)
            p._keep = c
        return p

    def odlay(self, handle, layer, v1, v2, dvnormal, dvnarrow=None, sample=4.0, dvout=0.0,
              precision=FP64, continuum=None, mtckd=False, species_broad=False, other_continua=False):
        lay = _layer_struct(layer)
        par = self.params(v1, v2, dvnormal, dvnarrow, sample, dvout, precision, continuum,
                          species_broad=species_broad, mtckd=mtckd, other_continua=other_continua)
        jr = ctypes.c_int32(self.jrad)
        n = ctypes.c_int64()
        dvu = ctypes.c_double()
        check(self.lib.clblm_odlay(self.ptr, handle, ctypes.byref(lay), ctypes.byref(par), ctypes.byref(jr),
                                   ctypes.byref(n), ctypes.byref(dvu), None))
        out = np.zeros(n.value)
        jr = ctypes.c_int32(self.jrad)
        check(self.lib.clblm_odlay(self.ptr, handle, ctypes.byref(lay), ctypes.byref(par), ctypes.byref(jr),
                                   ctypes.byref(n), ctypes.byref(dvu), out.ctypes.data_as(ctypes.c_void_p)))
        self.jrad = jr.value
        return out, dvu.value

    # -- many layers (device buffers) -------------------------------------------
    def od_batch_plan(self, handle, layers, params):
        nl = len(layers)
        lay = (_lib.Layer * nl)(*[_layer_struct(l) for l in layers])
        par = (ODParams * nl)(*params)
        off = np.zeros(nl + 1, np.int64)
        jr = ctypes.c_int32(self.jrad)
        check(self.lib.clblm_od_batch(self.ptr, handle, nl, ctypes.cast(lay, ctypes.c_void_p),
                                      ctypes.cast(par, ctypes.c_void_p), ctypes.byref(jr),
                                      off.ctypes.data_as(ctypes.c_void_p), None, None))
        return lay, par, off

    def od_batch(self, handle, layers, params, out=None, stream=None):
        """Returns (od tensor on the GPU, offsets).  ``out`` may be a preallocated
        float64 cuda tensor of offsets[-1] elements.  The work is enqueued on
        ``stream`` (default: torch's current stream), so torch operations on the
        result are ordered after it; errors found on the device are raised by the
        next call or by ``sync()``."""
        import torch
        lay, par, off = self.od_batch_plan(handle, layers, params)
        if out is None:
            out = torch.empty(int(off[-1]), dtype=torch.float64, device="cuda")
        assert out.dtype == torch.float64 and out.is_cuda and out.numel() >= off[-1]
        jr = ctypes.c_int32(self.jrad)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        st = ctypes.c_void_p(stream)
        check(self.lib.clblm_od_batch(self.ptr, handle, len(layers), ctypes.cast(lay, ctypes.c_void_p),
                                      ctypes.cast(par, ctypes.c_void_p), ctypes.byref(jr),
                                      off.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(out.data_ptr()), st))
        self.jrad = jr.value
        return out, off

    def od_tangent(self, handle, layers, params, dirs, dcontinuum=None, stream=None):
        """Layer ODs and their derivatives (clblm_od_tangent).  dirs: sequence of
        (dt, dlnwtot, {m: dlnw}); dcontinuum: per layer with a continuum, an array
        [len(dirs), nptabs] of the continuum's derivatives (None: zero).  Returns
        (tensor [1 + len(dirs), offsets[-1]] on the GPU, offsets): row 0 the OD of the
        smooth model (closed-form line shapes), row 1 + k its derivative along
        dirs[k].  Synchronous."""
        import torch
        nl = len(layers)
        lay = (_lib.Layer * nl)(*[_layer_struct(l) for l in layers])
        par = (ODParams * nl)(*params)
        nd = len(dirs)
        D = (_lib.ODDir * max(1, nd))()
        for k, (dt, dlw, dq) in enumerate(dirs):
            D[k].dt, D[k].dlnwtot = float(dt), float(dlw)
            for m, v in (dq or {}).items():
                D[k].dlnw[m] = float(v)
        dc = None
        if dcontinuum is not None:
            parts = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in dcontinuum if a is not None]
            dc = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(1))
        off = np.zeros(nl + 1, np.int64)
        jr = ctypes.c_int32(self.jrad)
        args = [self.ptr, handle, nl, ctypes.cast(lay, ctypes.c_void_p), ctypes.cast(par, ctypes.c_void_p),
                ctypes.byref(jr), nd, ctypes.cast(D, ctypes.c_void_p),
                ctypes.c_void_p(dc.ctypes.data) if dc is not None else None, off.ctypes.data_as(ctypes.c_void_p)]
        check(self.lib.clblm_od_tangent(*args, None, None))
        out = torch.empty((nd + 1, int(off[-1])), dtype=torch.float64, device="cuda")
        jr = ctypes.c_int32(self.jrad)
        args[5] = ctypes.byref(jr)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(self.lib.clblm_od_tangent(*args, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
        self.jrad = jr.value
        return out, off

    def od_graph(self, handle, layers, params, out=None):
        """Captures the batch of ``od_batch`` in one hipGraph (nothing runs yet):
        returns an ODGraph whose ``launch()`` fills ``out`` (the graph's float64 cuda
        tensor) anew on each replay."""
        import torch
        lay, par, off = self.od_batch_plan(handle, layers, params)
        if out is None:
            out = torch.empty(int(off[-1]), dtype=torch.float64, device="cuda")
        assert out.dtype == torch.float64 and out.is_cuda and out.numel() >= off[-1]
        jr = ctypes.c_int32(self.jrad)
        g = ctypes.c_void_p()
        check(self.lib.clblm_od_graph_create(self.ptr, handle, len(layers), ctypes.cast(lay, ctypes.c_void_p),
                                             ctypes.cast(par, ctypes.c_void_p), ctypes.byref(jr),
                                             off.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(out.data_ptr()),
                                             ctypes.byref(g)))
        self.jrad = jr.value
        return ODGraph(self, g, out, off)

    # -- instrument scan ---------------------------------------------------------------
    def fft_scan(self, x, v1, dv, v1out, v2out, dvout=0.0, funct_id=5, hwhm=0.6034, fill=0.0, precision=FP64,
                 stream=None, pre_box_hw=-1.0, deconv_pre_box=False):
        """clblm_FFTSCN_array (src/clblm_src/clblm_FFT.f90:308-349) on a float64 cuda
        tensor x on v1 + i*dv: returns (scanned tensor, its v1, its dv).  pre_box_hw > 0
        pre-scans with a boxcar of that half width (deconvolved if deconv_pre_box)."""
        import torch
        assert x.dtype == torch.float64 and x.is_cuda
        x = x.contiguous()
        n, a, b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        args = (int(x.numel()), float(v1), float(dv), float(v1out), float(v2out), float(dvout), int(funct_id),
                float(hwhm), float(fill), float(pre_box_hw), int(bool(deconv_pre_box)), int(precision))
        check(self.lib.clblm_fft_scan(self.ptr, None, *args, None, ctypes.byref(n), ctypes.byref(a),
                                      ctypes.byref(b), None))
        out = torch.empty(n.value, dtype=torch.float64, device=x.device)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(self.lib.clblm_fft_scan(self.ptr, ctypes.c_void_p(x.data_ptr()), *args, ctypes.c_void_p(out.data_ptr()),
                                      ctypes.byref(n), ctypes.byref(a), ctypes.byref(b), ctypes.c_void_p(stream)))
        return out, a.value, b.value

    def scanfn(self, x, v1, dv, v1out, v2out, dvout, funct_id=1, hwhm=0.1, stream=None):
        """clblm_SCANFN_array (src/clblm_src/clblm_ScanFilter.f90:33-67) on a float64 cuda
        tensor x of spectra on v1 + i*dv along its last axis: funct_id 1 rectangle, 2
        triangle, 3 Gaussian, 4 sinc**2, 5 sinc (-4, -5: hwhm is the first zero
        crossing); dvout 0 takes the function's default sampling.  Returns (scanned
        tensor, its v1, its dv)."""
        import torch
        assert x.dtype == torch.float64 and x.is_cuda
        shape = x.shape
        x2 = x.contiguous().reshape(-1, shape[-1])
        rows, n = x2.shape
        nout, a, b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        args = (int(n), int(rows), int(n), float(v1), float(dv), float(v1out), float(v2out), float(dvout),
                int(funct_id), float(hwhm))
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        # the count can depend on the data (functions 2..5), so the query reads x too
        check(self.lib.clblm_scanfn(self.ptr, ctypes.c_void_p(x2.data_ptr()), *args, None, 0, ctypes.byref(nout),
                                    ctypes.byref(a), ctypes.byref(b), ctypes.c_void_p(stream)))
        out = torch.empty((rows, max(1, nout.value)), dtype=torch.float64, device=x.device)
        check(self.lib.clblm_scanfn(self.ptr, ctypes.c_void_p(x2.data_ptr()), *args, ctypes.c_void_p(out.data_ptr()),
                                    int(out.shape[1]), ctypes.byref(nout), ctypes.byref(a), ctypes.byref(b),
                                    ctypes.c_void_p(stream)))
        return out[:, :nout.value].reshape(*shape[:-1], nout.value), a.value, b.value

    def filter(self, x, v1, dv, channels, stream=None):
        """postProc_filter_singleStruct (src/clblm_src/clblm_PostProc.f90:289-336) with
        clblm_FLTRFN: x a float64 cuda tensor (one spectrum on v1 + i*dv); channels a list
        of (v1f, dvf, filter values), dvf < 0 meaning v1f is the centre.  Returns a
        float64 cuda tensor of one value per channel."""
        import torch
        assert x.dtype == torch.float64 and x.is_cuda and x.dim() == 1
        n = len(channels)
        v1f = np.array([c[0] for c in channels], np.float64)
        dvf = np.array([c[1] for c in channels], np.float64)
        vals = [np.ascontiguousarray(c[2], np.float64) for c in channels]
        nptf = np.array([v.size for v in vals], np.int32)
        off = np.concatenate([[0], np.cumsum(nptf)[:-1]]).astype(np.int64)
        xf = np.ascontiguousarray(np.concatenate(vals))
        out = torch.empty(n, dtype=torch.float64, device=x.device)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        x = x.contiguous()
        check(self.lib.clblm_filter(self.ptr, ctypes.c_void_p(x.data_ptr()), int(x.numel()), float(v1), float(dv),
                                    n, v1f.ctypes.data, dvf.ctypes.data, nptf.ctypes.data, off.ctypes.data,
                                    xf.ctypes.data, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
        return out

    def prebox(self, x, v1, dv, boxcar_hw):
        """output_JacArray's boxcar pre-average (clblm_noScattJac.f90:1341-1409) of a
        float64 cuda tensor of spectra on v1 + i*dv (last axis): the rectangle scan with M = int(2 boxcarHW/dv)
        points per box onto v1 + (M-1)dv/2 .. v2 - (M-1)dv/2 at (M-1)dv; the input
        itself when boxcarHW <= dv or M <= 1.  Returns (tensor, v1, dv)."""
        m = int(2.0 * boxcar_hw / dv)
        if not (boxcar_hw > 0.0 and boxcar_hw > dv) or m <= 1:
            return x, v1, dv
        v2 = v1 + dv * (x.shape[-1] - 1)
        return self.scanfn(x, v1, dv, v1 + (m - 1) * dv / 2.0, v2 - (m - 1) * dv / 2.0, (m - 1) * dv, 1, boxcar_hw)

    def interp_spectrum(self, x, v1, dvin, dvout, nout, stream=None):
        """interpSpectrum (src/clblm_src/clblm_Spectrum.f90:326-590) of the float64 cuda
        tensor x on v1 + i*dvin onto nout points on v1 + i*dvout."""
        import torch
        assert x.dtype == torch.float64 and x.is_cuda
        x = x.contiguous()
        out = torch.empty(int(nout), dtype=torch.float64, device=x.device)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(self.lib.clblm_interp_spectrum(self.ptr, ctypes.c_void_p(x.data_ptr()), int(x.numel()), float(v1),
                                             float(dvin), ctypes.c_void_p(out.data_ptr()), int(nout), float(dvout),
                                             ctypes.c_void_p(stream)))
        return out

    def interp_to_finest(self, spectra, dvs, v1, stream=None):
        """interpToFinestGrid (clblm_Spectrum.f90:733-795) of float64 cuda tensors from v1
        on dvs: returns (the spectra on the finest grid, its DV)."""
        import torch
        n = len(spectra)
        xs = [t.contiguous() for t in spectra]
        dv = np.ascontiguousarray(dvs, np.float64)
        nl = np.array([t.numel() for t in xs], np.int64)
        dvo, no = ctypes.c_double(), ctypes.c_int64()
        check(self.lib.clblm_interp_to_finest(self.ptr, n, float(v1), dv.ctypes.data, nl.ctypes.data, None, None,
                                              ctypes.byref(dvo), ctypes.byref(no), None))
        outs = [torch.empty(no.value, dtype=torch.float64, device=x.device) for x in xs]
        pin = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        pout = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(self.lib.clblm_interp_to_finest(self.ptr, n, float(v1), dv.ctypes.data, nl.ctypes.data,
                                              ctypes.cast(pin, ctypes.c_void_p), ctypes.cast(pout, ctypes.c_void_p),
                                              ctypes.byref(dvo), ctypes.byref(no), ctypes.c_void_p(stream)))
        return outs, dvo.value

    def tuning(self, key, value=0):
        """Sets launch-shape parameter ``key`` (``_lib.TUNE_*``) to ``value`` (0: query
        only); returns the previous setting."""
        old = ctypes.c_int()
        check(self.lib.clblm_tuning(self.ptr, int(key), int(value), ctypes.byref(old)))
        return old.value

    def autotune(self, handle, layers, params, keys=None, reps=2):
        """Runs the library's autotuner on this batch (clblm_autotune) and returns the
        chosen settings {key: value} and the time per batch in ms."""
        nl = len(layers)
        lay = (_lib.Layer * nl)(*[_layer_struct(l) for l in layers])
        par = (ODParams * nl)(*params)
        karr = None
        nk = 0
        if keys:
            karr = (ctypes.c_int * len(keys))(*keys)
            nk = len(keys)
        ms = ctypes.c_double()
        check(self.lib.clblm_autotune(self.ptr, handle, nl, ctypes.cast(lay, ctypes.c_void_p),
                                      ctypes.cast(par, ctypes.c_void_p),
                                      ctypes.cast(karr, ctypes.c_void_p) if karr is not None else None, nk,
                                      int(reps), ctypes.byref(ms)))
        chosen = {k: self.tuning(k) for k in (_lib.TUNE_F4_POINTS_FP32C, _lib.TUNE_F4_POINTS_FP64,
                                              _lib.TUNE_CONV_UNROLL, _lib.TUNE_CONV_SPLIT, _lib.TUNE_F4_XCD_CHUNK,
                                              _lib.TUNE_F4_STREAM, _lib.TUNE_F4_FAR_FIELD)}
        return chosen, ms.value

    def sync(self):
        """Wait for the enqueued batches; raises the first error found on the device."""
        check(self.lib.clblm_sync(self.ptr))

    def work(self, reset=False):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        c, d = ctypes.c_double(), ctypes.c_double()
        check(self.lib.clblm_work(self.ptr, int(reset), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                  ctypes.byref(d)))
        return dict(conv_evals=a.value, f4_evals=b.value, ms_conv=c.value, ms_f4=d.value)

    def timing(self, reset=False):
        vals = [ctypes.c_double() for _ in range(4)]
        n = ctypes.c_int64()
        check(self.lib.clblm_timing(self.ptr, int(reset), *[ctypes.byref(v) for v in vals], ctypes.byref(n)))
        return dict(prep=vals[0].value, conv=vals[1].value, f4=vals[2].value, panel=vals[3].value,
                    launches=n.value)
