"""ctypes front end of the C oracle (liboracle.so).  Test infrastructure only."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class _Lines(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(k, ctypes.c_void_p) for k in
                ("nu", "s", "a_air", "e_lo", "mol", "a_self", "tmpalf", "shift", "iflg",
                 "brd_flg", "brd_hw", "brd_tmp", "brd_shft")]


class _Layer(ctypes.Structure):
    _fields_ = [("pave", ctypes.c_double), ("tave", ctypes.c_double), ("wtot", ctypes.c_double),
                ("w", ctypes.c_double * 48), ("sel", ctypes.c_int * 48)]


class _Params(ctypes.Structure):
    _fields_ = [("v1", ctypes.c_double), ("v2", ctypes.c_double), ("dvnormal", ctypes.c_double),
                ("dvnarrow", ctypes.c_double), ("sample", ctypes.c_double), ("dvout", ctypes.c_double),
                ("jrad", ctypes.c_int), ("contnm", ctypes.c_int), ("absrb", ctypes.c_void_p),
                ("nptabs", ctypes.c_int), ("species_broad", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(path)
        _lib.orc_error.restype = ctypes.c_char_p
        _lib.orc_radfn.restype = ctypes.c_double
        _lib.orc_radfn.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_calculate_dv.restype = ctypes.c_double
        _lib.orc_calculate_dv.argtypes = [ctypes.c_double] * 6
        _lib.orc_avrat.restype = ctypes.c_double
        _lib.orc_planck.restype = ctypes.c_double
        _lib.orc_planck.argtypes = [ctypes.c_double, ctypes.c_double]
    return _lib


def tables():
    """F1, F2, F3, FG, XVER (1-based arrays with 3002 slots)."""
    arrs = [np.zeros(3002) for _ in range(5)]
    lib().orc_tables(*[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])
    return arrs


def tips_scor(mol, iso, t):
    out = ctypes.c_double()
    if lib().orc_tips_scor(int(mol), int(iso), ctypes.c_double(t), ctypes.byref(out)):
        raise RuntimeError(lib().orc_error().decode())
    return out.value


def radfni(vi, dvi, xkt, vinew, rdlast):
    vn, rd, rl = ctypes.c_double(vinew), ctypes.c_double(), ctypes.c_double(rdlast)
    lib().orc_radfni.restype = ctypes.c_double
    r = lib().orc_radfni(ctypes.c_double(vi), ctypes.c_double(dvi), ctypes.c_double(xkt),
                         ctypes.byref(vn), ctypes.byref(rd), ctypes.byref(rl))
    return r, vn.value, rd.value, rl.value


def xint(v1a, v2a, dva, a, vft, dvr3, r3, n1r3, n2r3):
    """a, r3: 1-based numpy arrays (element 0 unused); r3 is updated in place."""
    lib().orc_xint(ctypes.c_double(v1a), ctypes.c_double(v2a), ctypes.c_double(dva),
                   a.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(1.0), ctypes.c_double(vft),
                   ctypes.c_double(dvr3), r3.ctypes.data_as(ctypes.c_void_p), int(n1r3), int(n2r3))


def interp(spin, dvin, v1in, dvout, v1out, nout):
    spin = np.ascontiguousarray(spin, np.float64)
    out = np.zeros(nout)
    lib().orc_interp(spin.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(dvin), ctypes.c_double(v1in),
                     int(spin.size), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(dvout),
                     ctypes.c_double(v1out), int(nout))
    return out


class Oracle:
    """Sequential restatement of ODLAY for one line list."""

    def __init__(self, lines):
        self.lines = lines
        arrs = [lines.nu, lines.s, lines.a_air, lines.e_lo, lines.mol, lines.a_self, lines.tmpalf,
                lines.shift, lines.iflg]
        self._keep = [np.ascontiguousarray(a) for a in arrs]
        brd = [None] * 4
        if getattr(lines, "brd_flg", None) is not None:
            self._brd = [np.ascontiguousarray(getattr(lines, f), dt).reshape(lines.n, 7)
                         for f, dt in (("brd_flg", np.int32), ("brd_hw", np.float32), ("brd_tmp", np.float32),
                                       ("brd_shft", np.float32))]
            brd = [a.ctypes.data for a in self._brd]
        self._L = _Lines(lines.n, *[a.ctypes.data for a in self._keep], *brd)
        self.jrad = 1

    def odlay(self, layer, v1, v2, dvnormal, dvnarrow=None, sample=4.0, dvout=0.0, continuum=None,
              species_broad=False, mtckd=None, contnm=None):
        """Returns (od, dv_used).  JRAD persists across calls like the reference's SAVE.
        mtckd: a continuum table (clblm_amd.continuum.load_mtckd); its water-vapour
        continuum is added to ABSRB after `continuum`, with OPDPTH's JRAD after the flip.
        contnm: CONTNM's other thermal-infrared continua (CO2, O2, N2, Rayleigh), the
        scale factors (XCO2C, XO3CN, XO2CN, XN2CN, XRAYL), added after those."""
        lay = _Layer()
        lay.pave, lay.tave, lay.wtot = layer.pave, layer.tave, layer.wtot
        for m, v in layer.w.items():
            lay.w[m] = v
            lay.sel[m] = 1
        absrb = None
        npt = 0
        if mtckd is not None:
            jr = self.jrad
            if v1 / (layer.tave / 1.4387752) < 5. and jr != -1:   # clblm_ODLAY.f90:877-878
                jr = 0
            continuum = mtckd_h2o(mtckd, layer.pave, layer.tave, layer.w.get(1, 0.0), layer.wtot, v1, v2, jr,
                                  absrb=continuum)
        if contnm is not None:
            jr = self.jrad
            if v1 / (layer.tave / 1.4387752) < 5. and jr != -1:
                jr = 0
            continuum, rc = contnm_ir(layer.pave, layer.tave, layer.wtot, layer.w, v1, v2, jr, contnm,
                                      absrb=continuum)
            if rc:
                raise RuntimeError("contnm: ABSRB longer than 5050 points")
        if continuum is not None:
            absrb = np.ascontiguousarray(continuum, np.float64)
            npt = absrb.size
        par = _Params(v1, v2, dvnormal, dvnormal if dvnarrow is None else dvnarrow, sample, dvout,
                      self.jrad, 1 if absrb is not None else 0,
                      absrb.ctypes.data if absrb is not None else None, npt, 1 if species_broad else 0)
        ns, dvu = ctypes.c_int64(), ctypes.c_double()
        L = lib()
        if L.orc_odlay(ctypes.byref(self._L), ctypes.byref(lay), ctypes.byref(par), None,
                       ctypes.c_int64(0), ctypes.byref(ns), ctypes.byref(dvu)):
            raise RuntimeError(L.orc_error().decode())
        out = np.zeros(ns.value)
        par.jrad = self.jrad
        if L.orc_odlay(ctypes.byref(self._L), ctypes.byref(lay), ctypes.byref(par),
                       out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(out.size),
                       ctypes.byref(ns), ctypes.byref(dvu)):
            raise RuntimeError(L.orc_error().decode())
        self.jrad = par.jrad
        return out[: ns.value], dvu.value


# ---- instrument scan (clblm_FFT.f90) ------------------------------------------------
class _Dir(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("dlnwtot", ctypes.c_double), ("dlnw", ctypes.c_double * 48)]


def make_dirs(dirs):
    """dirs: sequence of (dt, dlnwtot, {m: dlnw}) -> ctypes array of orc_dir."""
    arr = (_Dir * max(1, len(dirs)))()
    for k, (dt, dlw, dq) in enumerate(dirs):
        arr[k].dt, arr[k].dlnwtot = float(dt), float(dlw)
        for m, v in (dq or {}).items():
            arr[k].dlnw[m] = float(v)
    return arr


def _odlay_tan(self, layer, v1, v2, dvnormal, dirs, dvnarrow=None, sample=4.0, dvout=0.0, continuum=None,
               dcontinuum=None):
    """The smooth-model OD and its derivatives along dirs ((dt, dlnwtot, {m: dlnw}) each).
    Returns (array [1 + ndir, npts], dv_used); JRAD persists as in odlay.  dcontinuum:
    [ndir, nptabs] derivatives of the continuum along each direction."""
    lay = _Layer()
    lay.pave, lay.tave, lay.wtot = layer.pave, layer.tave, layer.wtot
    for m, v in layer.w.items():
        lay.w[m] = v
        lay.sel[m] = 1
    absrb = None if continuum is None else np.ascontiguousarray(continuum, np.float64)
    dab = None if dcontinuum is None else np.ascontiguousarray(dcontinuum, np.float64)
    npt = 0 if absrb is None else absrb.size
    par = _Params(v1, v2, dvnormal, dvnormal if dvnarrow is None else dvnarrow, sample, dvout, self.jrad,
                  1 if absrb is not None else 0, absrb.ctypes.data if absrb is not None else None, npt, 0)
    D = make_dirs(dirs)
    nd = len(dirs)
    ns, dvu = ctypes.c_int64(), ctypes.c_double()
    L = lib()
    args = (ctypes.byref(self._L), ctypes.byref(lay), ctypes.byref(par), nd, D,
            ctypes.c_void_p(dab.ctypes.data) if dab is not None else None)
    jr0 = self.jrad
    if L.orc_odlay_tan(*args, None, ctypes.c_int64(0), ctypes.byref(ns), ctypes.byref(dvu)):
        raise RuntimeError(L.orc_error().decode())
    cap = ns.value
    out = np.zeros((nd + 1, cap))
    par.jrad = jr0
    if L.orc_odlay_tan(*args, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap), ctypes.byref(ns),
                       ctypes.byref(dvu)):
        raise RuntimeError(L.orc_error().decode())
    self.jrad = par.jrad
    return out[:, : ns.value], dvu.value


Oracle.odlay_tan = _odlay_tan


class ScanPlan(ctypes.Structure):
    _fields_ = [("jfn", ctypes.c_int), ("hwhm", ctypes.c_double), ("a", ctypes.c_double),
                ("pre_interp", ctypes.c_int), ("dvn", ctypes.c_double), ("v1z", ctypes.c_double),
                ("v2z", ctypes.c_double), ("dv", ctypes.c_double), ("n1load", ctypes.c_int64),
                ("ltotal", ctypes.c_int64), ("lrec", ctypes.c_int64), ("nfft", ctypes.c_int64),
                ("n1", ctypes.c_int64), ("ntotal", ctypes.c_int64), ("v1s", ctypes.c_double),
                ("ivx", ctypes.c_int), ("fft", ctypes.c_int), ("prebox", ctypes.c_int), ("m", ctypes.c_int),
                ("ltotal0", ctypes.c_int64), ("lrec0", ctypes.c_int64), ("v1z0", ctypes.c_double),
                ("dv0", ctypes.c_double), ("deconv", ctypes.c_int), ("deconv_a", ctypes.c_double),
                ("deconv_only", ctypes.c_int)]


def scan_interp(x, v1, dv, vv1, vv2, dvo):
    """Intpdr/clblm_INTERP (4-point): returns (values, first frequency)."""
    L = lib()
    L.orc_scan_interp.restype = ctypes.c_int64
    L.orc_scan_interp.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_double] * 5 + \
        [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    x = np.ascontiguousarray(x, np.float64)
    v1r = ctypes.c_double()
    n = L.orc_scan_interp(x.ctypes.data, x.size, v1, dv, vv1, vv2, dvo, None, 0, ctypes.byref(v1r))
    out = np.zeros(max(n, 1))
    L.orc_scan_interp(x.ctypes.data, x.size, v1, dv, vv1, vv2, dvo, out.ctypes.data, n, ctypes.byref(v1r))
    return out[:n], v1r.value


def fft_scan(x, v1, dv, v1out, v2out, dvout, funct_id=5, hwhm=0.6034, fill=0.0, pre_box_hw=-1.0,
             deconv_pre_box=False):
    """clblm_FFTSCN_struct: returns (values, v1, dv).  The transforms are numpy's;
    everything around them is the C restatement: the boxcar pre-scan, the analytic
    apodization (IVX = 1) or the transform of the scanning function (IVX = -1), the
    boxcar deconvolution, the window and the output re-gridding."""
    L = lib()
    L.orc_fftscn_plan2.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.POINTER(ScanPlan)]
    L.orc_fftscn_finish.argtypes = [ctypes.POINTER(ScanPlan), ctypes.c_double, ctypes.c_int64, ctypes.c_double,
                                    ctypes.c_double]
    for fn in (L.orc_fft_apod, L.orc_fft_deconv, L.orc_scan_sfn):
        fn.argtypes = [ctypes.POINTER(ScanPlan), ctypes.c_void_p]
    L.orc_scan_boxcar.argtypes = [ctypes.POINTER(ScanPlan), ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    x = np.ascontiguousarray(x, np.float64)
    p = ScanPlan()
    if L.orc_fftscn_plan2(v1, dv, x.size, v1out, v2out, int(funct_id), hwhm, fill, pre_box_hw,
                          int(bool(deconv_pre_box)), ctypes.byref(p)):
        raise RuntimeError(L.orc_error().decode())
    if p.pre_interp:
        xs, v1p = scan_interp(x, v1, dv, p.v1z, p.v2z, p.dvn)
    else:
        xs, v1p = x, v1
    if L.orc_fftscn_finish(ctypes.byref(p), v1p, xs.size, v1out, v2out):
        raise RuntimeError(L.orc_error().decode())
    loaded = np.full(p.lrec0 * 65536, fill, np.float64)
    loaded[:p.ltotal0] = xs[p.n1load - 1:p.n1load - 1 + p.ltotal0]
    if p.prebox:
        buf = np.zeros(p.nfft)
        L.orc_scan_boxcar(ctypes.byref(p), loaded.ctypes.data, fill, buf.ctypes.data)
    else:
        buf = loaded
    if p.fft:
        X = np.fft.rfft(buf)
        if p.ivx == 1:
            h = np.zeros(p.nfft // 2 + 1)
            L.orc_fft_apod(ctypes.byref(p), h.ctypes.data)
            h[-1] = 0.0                              # Nyquist coefficient
            X = X * h
        elif p.ivx == -1:
            f = np.zeros(p.nfft)
            L.orc_scan_sfn(ctypes.byref(p), f.ctypes.data)
            H = np.fft.rfft(f)
            H = H / H[0].real                        # unit area
            X0, XN = X[0].real * H[0].real, X[-1].real * H[-1].real
            X = X * H
            X[0], X[-1] = X0, XN
        if p.deconv:
            d = np.zeros(p.nfft // 2 + 1)
            L.orc_fft_deconv(ctypes.byref(p), d.ctypes.data)
            d[-1] = 0.0
            X = X * d
        y = np.fft.irfft(X, n=p.nfft)
    else:
        y = buf
    seg = y[p.n1 - 1:p.n1 - 1 + p.ntotal]
    if dvout > 0:
        out, v1o = scan_interp(seg, p.v1s, p.dv, v1out, v2out, dvout)
        return out, v1o, dvout
    return seg, p.v1s, p.dv


# ---- level Jacobians (clblm_noScattJac.f90) ---------------------------------------
def _p(a):
    return a.ctypes.data if a is not None else None


def level2layer(p_lev, t_lev, t_lay, q_lev):
    """level2layer: q_lev [nmol, nlay+1] level densities.  Returns dT_dTup, dT_dTlo [nlay],
    dlnW_dlnQup, dlnW_dlnQlo [nmol, nlay]."""
    L = lib()
    t_lay = np.ascontiguousarray(t_lay, np.float64)
    nlay = t_lay.size
    q = np.ascontiguousarray(np.atleast_2d(q_lev), np.float64) if q_lev is not None else np.zeros((0, nlay + 1))
    nmol = q.shape[0]
    outs = [np.zeros(nlay), np.zeros(nlay), np.zeros((nmol, nlay)), np.zeros((nmol, nlay))]
    p_lev = np.ascontiguousarray(p_lev, np.float64)
    t_lev = np.ascontiguousarray(t_lev, np.float64)
    L.orc_level2layer.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 8
    L.orc_level2layer(nlay, nmol, _p(p_lev), _p(t_lev), _p(t_lay), _p(q), *[_p(o) for o in outs])
    return tuple(outs)


def jacobians(od, t_lay, t_lev, v1, dv, emis=0.98, tskin=295.0, lin=1, sec_dn=1.0, dt_dtup=None, dt_dtlo=None):
    """Per-point Jacobians (down-looking, thermal): returns rad [n], dR/dtau [nlay, n],
    dR/dT_lev [nlay+1, n], dR/dTskin [n], dR/demis [n]."""
    L = lib()
    od = np.ascontiguousarray(od, np.float64)
    nlay, n = od.shape
    up = np.zeros(nlay) if dt_dtup is None else np.ascontiguousarray(dt_dtup, np.float64)
    lo = np.zeros(nlay) if dt_dtlo is None else np.ascontiguousarray(dt_dtlo, np.float64)
    rad, dtau, dtlev, dts, de = np.zeros(n), np.zeros((nlay, n)), np.zeros((nlay + 1, n)), np.zeros(n), np.zeros(n)
    D = ctypes.c_double
    L.orc_jacobians.argtypes = [ctypes.c_int, ctypes.c_int64, D, D, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                D, D, ctypes.c_int, D] + [ctypes.c_void_p] * 7
    t_lay = np.ascontiguousarray(t_lay, np.float64)
    t_lev = np.ascontiguousarray(t_lev, np.float64)
    L.orc_jacobians(nlay, n, v1, dv, _p(od), _p(t_lay), _p(t_lev), emis, tskin, lin, sec_dn, _p(up), _p(lo), _p(rad),
                    _p(dtau), _p(dtlev), _p(dts), _p(de))
    return rad, dtau, dtlev, dts, de


def jacobians_paths(od, t_lay, t_lev, v1, dv, amf_view=None, amf_down=None, amf_sun=None, emis=0.98, tskin=295.0,
                    lin=1, solar=False, sol_toa=None, rsfc_sun=0.0, obs=None, uplook=False, dt_dtup=None,
                    dt_dtlo=None):
    """orc_jacobians_paths: Jacobians with every branch (air-mass factors per layer,
    observer level, up- or down-looking, solar).  Returns a dict: rad [n], dtau
    [nlay, n], dtlev [nlay+1, n], dtskin, demis, drsfc [n]."""
    L = lib()
    od = np.ascontiguousarray(od, np.float64)
    nlay, n = od.shape
    f = lambda a, d: np.full(nlay, d) if a is None else np.ascontiguousarray(a, np.float64)
    av, ad, asun = f(amf_view, 1.0), f(amf_down, 1.0), f(amf_sun, 1.0)
    up, lo = f(dt_dtup, 0.0), f(dt_dtlo, 0.0)
    sol = np.zeros(n) if sol_toa is None else np.ascontiguousarray(sol_toa, np.float64)
    t_lay = np.ascontiguousarray(t_lay, np.float64)
    t_lev = np.ascontiguousarray(t_lev, np.float64)
    out = dict(rad=np.zeros(n), dtau=np.zeros((nlay, n)), dtlev=np.zeros((nlay + 1, n)), dtskin=np.zeros(n),
               demis=np.zeros(n), drsfc=np.zeros(n))
    D, P, I = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    L.orc_jacobians_paths.argtypes = [I, ctypes.c_int64, D, D, P, P, P, P, P, P, D, D, I, P, D, I, I, I, P, P] + [P] * 6
    L.orc_jacobians_paths(nlay, n, v1, dv, _p(od), _p(av), _p(ad), _p(asun), _p(t_lay), _p(t_lev), emis, tskin,
                          1 if solar else 0, _p(sol), rsfc_sun, nlay if obs is None else int(obs), 1 if uplook else 0,
                          lin, _p(up), _p(lo), *[_p(out[k]) for k in ("rad", "dtau", "dtlev", "dtskin", "demis",
                                                                     "drsfc")])
    return out


def level_jac(od_var, dr_dtau, dr_dtlev, wtot, tave, dt, dw_frac, dt_dtup, dt_dtlo, dw_dqup, dw_dqlo):
    """drv_LevelTemp + drv_LevelConc: od_var [3+nmol, nlay, n].  Returns jac_t [nlay+1, n]
    and jac_q [nmol, nlay+1, n]."""
    L = lib()
    od_var = np.ascontiguousarray(od_var, np.float64)
    nv, nlay, n = od_var.shape
    nmol = nv - 3
    jt, jq = np.zeros((nlay + 1, n)), np.zeros((nmol, nlay + 1, n))
    D = ctypes.c_double
    L.orc_level_jac.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [D, D] + \
        [ctypes.c_void_p] * 6
    keep = [np.ascontiguousarray(a, np.float64) for a in
            (dr_dtau, dr_dtlev, wtot, tave, dt_dtup, dt_dtlo, np.reshape(dw_dqup, (nmol, nlay)),
             np.reshape(dw_dqlo, (nmol, nlay)))]
    a = [_p(k) for k in keep]
    L.orc_level_jac(nlay, nmol, n, _p(od_var), a[0], a[1], a[2], a[3], dt, dw_frac, a[4], a[5], a[6], a[7], _p(jt),
                    _p(jq))
    return jt, jq


def scanfn_rect(x, v1, dv, v1out, v2out, dvout, hwhm):
    """orc_scanfn_rect: the rectangle (boxcar) spectral-domain scan; returns (values, v1, dv)."""
    L = lib()
    L.orc_scanfn_rect.restype = ctypes.c_int64
    L.orc_scanfn_rect.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_double] * 6 + \
        [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    x = np.ascontiguousarray(x, np.float64)
    a, b = ctypes.c_double(), ctypes.c_double()
    n = L.orc_scanfn_rect(x.ctypes.data, x.size, v1, dv, v1out, v2out, dvout, hwhm, None, 0, ctypes.byref(a),
                          ctypes.byref(b))
    if n < 0:
        raise RuntimeError(L.orc_error().decode())
    out = np.zeros(n)
    L.orc_scanfn_rect(x.ctypes.data, x.size, v1, dv, v1out, v2out, dvout, hwhm, out.ctypes.data, n, ctypes.byref(a),
                      ctypes.byref(b))
    return out, a.value, b.value


def scanfn(x, v1, dv, v1out, v2out, dvout, funct_id, hwhm):
    """orc_scanfn_rect (functID 1) or orc_scanfn_conv (2..5, -4, -5): the spectral-domain
    scan clblm_SCANFN_array; returns (values, v1, dv)."""
    if funct_id == 1:
        return scanfn_rect(x, v1, dv, v1out, v2out, dvout, hwhm)
    L = lib()
    L.orc_scanfn_conv.restype = ctypes.c_int64
    L.orc_scanfn_conv.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_double] * 5 + \
        [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
         ctypes.POINTER(ctypes.c_double)]
    x = np.ascontiguousarray(x, np.float64)
    a, b = ctypes.c_double(), ctypes.c_double()
    n = L.orc_scanfn_conv(x.ctypes.data, x.size, v1, dv, v1out, v2out, dvout, funct_id, hwhm, None, 0,
                          ctypes.byref(a), ctypes.byref(b))
    if n < 0:
        raise RuntimeError(L.orc_error().decode())
    out = np.zeros(max(n, 1))
    L.orc_scanfn_conv(x.ctypes.data, x.size, v1, dv, v1out, v2out, dvout, funct_id, hwhm, out.ctypes.data, n,
                      ctypes.byref(a), ctypes.byref(b))
    return out[:n], a.value, b.value


def rt_mode1(ods, dvs, v1, v2, t_lay, t_lev, od_dn=None, od_sun=None, fac_up=None, fac_dn=None, fac_sun=None,
             obs_level=None, lin=1, thermal=True, solar=False, emis=0.98, tskin=295.0, refl=0.0, sol=None,
             sol_dv=None):
    """orc_rt_mode1: the merge-down radiance at an observer above the layers (ods: list of
    per-layer view-path OD arrays on v1 + i dvs[l], layer 0 at the surface).  Returns
    (radiance, total transmittance, DV)."""
    L = lib()
    nlay = len(ods)
    nl = np.array([o.size for o in ods], np.int64)
    off = np.concatenate([[0], np.cumsum(nl)[:-1]]).astype(np.int64)
    cat = lambda a: np.ascontiguousarray(np.concatenate(a), np.float64)
    up = cat(ods)
    dn = up if od_dn is None else cat(od_dn)
    sn = up if od_sun is None else cat(od_sun)
    one = np.ones(nlay)
    fu, fd, fs = [np.ascontiguousarray(one if f is None else f, np.float64) for f in (fac_up, fac_dn, fac_sun)]
    dv = np.ascontiguousarray(dvs, np.float64)
    s0 = np.zeros(2) if sol is None else np.ascontiguousarray(sol, np.float64)
    cap = int(max(nl.max(), s0.size, 2))
    rad, tx = np.zeros(cap), np.zeros(cap)
    dvo = ctypes.c_double()
    D, P, I = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    L.orc_rt_mode1.restype = ctypes.c_int64
    L.orc_rt_mode1.argtypes = [I, D, D, P, P, P, P, P, P, P, P, P, P, P, I, I, I, I, D, D, D, P, ctypes.c_int64, D,
                               P, P, ctypes.c_int64, ctypes.POINTER(D)]
    tl = np.ascontiguousarray(t_lay, np.float64)
    tv = np.ascontiguousarray(t_lev, np.float64)
    n = L.orc_rt_mode1(nlay, v1, v2, dv.ctypes.data, nl.ctypes.data, off.ctypes.data, up.ctypes.data, dn.ctypes.data,
                       sn.ctypes.data, fu.ctypes.data, fd.ctypes.data, fs.ctypes.data, tl.ctypes.data, tv.ctypes.data,
                       nlay if obs_level is None else obs_level, lin, int(thermal), int(solar), emis, tskin, refl,
                       s0.ctypes.data, s0.size, float(dv[-1] if sol_dv is None else sol_dv), rad.ctypes.data,
                       tx.ctypes.data, cap, ctypes.byref(dvo))
    if n < 0:
        raise RuntimeError(L.orc_error().decode())
    return rad[:n], tx[:n], dvo.value


def mtckd_h2o(table, pave, tave, wk1, wtot, v1, v2, jrad, xself=1.0, xfrgn=1.0, absrb=None):
    """orc_mtckd_h2o: the MT_CKD water-vapour continuum of one layer on OPDPTH's ABSRB
    grid (V1ABS = INT(V1) - 3 at 1 cm-1), added to absrb (zeros by default)."""
    L = lib()
    D, P = ctypes.c_double, ctypes.c_void_p
    L.orc_mtckd_h2o.restype = ctypes.c_int
    L.orc_mtckd_h2o.argtypes = [P, P, P, P, ctypes.c_int] + [D] * 8 + [ctypes.c_int, D, D, P, ctypes.c_int]
    t = {k: np.ascontiguousarray(table[k], np.float64) for k in
         ("wavenumbers", "self_absco_ref", "for_absco_ref", "self_texp")}
    out = np.zeros(5050) if absrb is None else np.concatenate([np.asarray(absrb, np.float64),
                                                               np.zeros(5050 - len(absrb))])
    n = L.orc_mtckd_h2o(t["wavenumbers"].ctypes.data, t["self_absco_ref"].ctypes.data,
                        t["for_absco_ref"].ctypes.data, t["self_texp"].ctypes.data, t["wavenumbers"].size,
                        float(table["ref_temp"]), float(table["ref_press"]), pave, tave, wk1, wtot, v1, v2, int(jrad),
                        xself, xfrgn, out.ctypes.data, out.size)
    if n < 0:
        raise RuntimeError(L.orc_error().decode())
    return out[:n]


def rt_flux(od, t_lay, t_lev, p_lev, v1, v2, dv, dv_flux, nang=3, lin=1, emis=0.98, tskin=295.0, lambertian=True):
    """orc_rt_flux: fluxes [L+1, nbin] (up, dn, net, heating) from layer ODs [L, N] on one
    grid (layer 0 at the surface)."""
    L = lib()
    D, P, I = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    L.orc_rt_flux.restype = I
    L.orc_rt_flux.argtypes = [I, ctypes.c_int64, D, D, D, P, P, P, P, I, D, D, I, D, I, P, P, P, P]
    od = np.ascontiguousarray(od, np.float64)
    nl, n = od.shape
    tl, tv, pl = (np.ascontiguousarray(x, np.float64) for x in (t_lay, t_lev, p_lev))
    nout = int((v2 - v1) / dv_flux + 1e-4)
    out = {k: np.zeros((nl + 1, max(nout, 1))) for k in ("up", "dn", "net", "heating")}
    r = L.orc_rt_flux(nl, n, v1, v2, dv, od.ctypes.data, tl.ctypes.data, tv.ctypes.data, pl.ctypes.data, lin,
                      emis, tskin, nang, dv_flux, 1 if lambertian else 0,
                      *[out[k].ctypes.data for k in ("up", "dn", "net", "heating")])
    if r < 0:
        raise RuntimeError(L.orc_error().decode())
    return out


def absrb_grid(v1, v2):
    """OPDPTH's ABSRB grid (clblm_ODLAY.f90:880-887): (V1ABS, NPTABS) at 1 cm-1."""
    a = float(int(v1))
    if v1 < 0.:
        a -= 1.
    a -= 3.
    b = float(int(v2 + 3. + 0.5))
    return a, int((b - a) / 1. + 1.5)


def contnm_ir(pave, tave, wtot, wk, v1, v2, jrad, scale=(1.0, 1.0, 1.0, 1.0, 1.0), absrb=None):
    """orc_contnm_ir: CONTNM's continua other than water vapour (CO2; O3 Chappuis,
    Hartley-Huggins and UV; the O2 fundamental, 1.27 um, 1.06 um, A-band, visible,
    Herzberg and far-UV continua; the N2 rotation, fundamental and overtone; Rayleigh) of
    one layer on OPDPTH's ABSRB grid, added to absrb (zeros by default).  wk: {HITRAN
    molecule: column}.  Returns (absrb, rc): rc is -1 when ABSRB is longer than the
    reference's 5050 points, else 0."""
    L = lib()
    D, P = ctypes.c_double, ctypes.c_void_p
    L.orc_contnm_ir.restype = ctypes.c_int
    L.orc_contnm_ir.argtypes = [D, D, D, P, D, D, ctypes.c_int, P, D, ctypes.c_int, P]
    v1abs, n = absrb_grid(v1, v2)
    w = np.zeros(48)
    for m, v in wk.items():
        w[m] = v
    sc = np.ascontiguousarray(scale, np.float64)
    out = np.zeros(n) if absrb is None else np.array(absrb, np.float64)
    assert out.size == n
    beyond = L.orc_contnm_ir(pave, tave, wtot, w.ctypes.data, v1, v2, int(jrad), sc.ctypes.data, v1abs, n,
                             out.ctypes.data)
    return out, beyond


def fltrfn(x, v1, dv, v1f, dvf, xf):
    """orc_fltrfn: the filtered value of the spectrum x on v1 + i dv through the filter
    xf on v1f + i dvf (dvf < 0: v1f is the centre)."""
    L = lib()
    D, P = ctypes.c_double, ctypes.c_void_p
    L.orc_fltrfn.restype = ctypes.c_int
    L.orc_fltrfn.argtypes = [P, ctypes.c_int64, D, D, D, D, ctypes.c_int, P, ctypes.POINTER(D)]
    x = np.ascontiguousarray(x, np.float64)
    f = np.ascontiguousarray(xf, np.float64)
    out = ctypes.c_double()
    if L.orc_fltrfn(x.ctypes.data, x.size, v1, dv, v1f, dvf, f.size, f.ctypes.data, ctypes.byref(out)):
        raise RuntimeError(L.orc_error().decode())
    return out.value


def scan_shape(jfn):
    """orc_scan_shape: the tabulated scanning function XF(1..NJMX) of JFN 1..4."""
    L = lib()
    L.orc_scan_shape.argtypes = [ctypes.c_int, ctypes.c_void_p]
    xf = np.zeros(6019)
    L.orc_scan_shape(jfn, xf.ctypes.data)
    return xf[1:{1: 252, 2: 252, 3: 2761, 4: 6019}[jfn]]


def prebox(x, v1, dv, boxcar_hw):
    """output_JacArray's boxcar pre-average (clblm_noScattJac.f90:1341-1409): the
    rectangle scan with M = int(2 boxcarHW / dv) points per box onto
    v1 + (M-1) dv/2 .. v2 - (M-1) dv/2 at (M-1) dv; the input unchanged when M <= 1
    or boxcarHW <= dv."""
    m = int(2.0 * boxcar_hw / dv)
    if not (boxcar_hw > 0.0 and boxcar_hw > dv) or m <= 1:
        return np.asarray(x, np.float64).copy(), v1, dv
    v2 = v1 + dv * (len(x) - 1)
    return scanfn_rect(x, v1, dv, v1 + (m - 1) * dv / 2.0, v2 - (m - 1) * dv / 2.0, (m - 1) * dv, boxcar_hw)


def interp_to_finest(spectra, dvs, v1):
    """orc_interp_to_finest: spectra (list of 1-D arrays) from v1 on dvs; returns (the
    spectra on the finest grid, index of the finest)."""
    L = lib()
    n = len(spectra)
    ins = [np.ascontiguousarray(s, np.float64) for s in spectra]
    nl = np.array([s.size for s in ins], np.int64)
    dv = np.ascontiguousarray(dvs, np.float64)
    loc = int(np.argmin(dv))
    outs = [np.zeros(max(int(nl[loc]), int(nl[k]))) for k in range(n)]
    P = ctypes.c_void_p
    L.orc_interp_to_finest.restype = ctypes.c_int
    L.orc_interp_to_finest.argtypes = [ctypes.c_int, ctypes.c_double, P, P, P, P]
    pin = (P * n)(*[s.ctypes.data for s in ins])
    pout = (P * n)(*[o.ctypes.data for o in outs])
    loc = L.orc_interp_to_finest(n, v1, dv.ctypes.data, nl.ctypes.data, pin, pout)
    # every spectrum ends on the finest grid, so every one has its length
    return [outs[k][:nl[loc]] for k in range(n)], loc


def rt_up_grids(ods, dvs, v1, t_lay, t_lev, emis=0.98, tskin=295.0, surf_dv=None, surf_nlim=None, lin=1,
                sec_dn=1.0):
    """orc_rt_up_grids: ods (list of per-layer OD arrays on v1 + i dvs[l], layer 0 at the
    surface); the surface grid defaults to layer 0's.  Returns (radiance, its DV)."""
    L = lib()
    nlay = len(ods)
    nl = np.array([o.size for o in ods], np.int64)
    off = np.concatenate([[0], np.cumsum(nl)[:-1]]).astype(np.int64)
    od = np.ascontiguousarray(np.concatenate(ods), np.float64)
    dv = np.ascontiguousarray(dvs, np.float64)
    sdv = float(dv[0] if surf_dv is None else surf_dv)
    sn = int(nl[0] if surf_nlim is None else surf_nlim)
    cap = int(max(nl.max(), sn))
    rad = np.zeros(cap)
    dvo = ctypes.c_double()
    D, P = ctypes.c_double, ctypes.c_void_p
    L.orc_rt_up_grids.restype = ctypes.c_int64
    L.orc_rt_up_grids.argtypes = [ctypes.c_int, D, P, P, P, P, P, P, D, D, D, ctypes.c_int64, ctypes.c_int, D, P,
                                  ctypes.c_int64, ctypes.POINTER(D)]
    tl = np.ascontiguousarray(t_lay, np.float64)
    tv = np.ascontiguousarray(t_lev, np.float64)
    n = L.orc_rt_up_grids(nlay, v1, dv.ctypes.data, nl.ctypes.data, off.ctypes.data, od.ctypes.data, tl.ctypes.data,
                          tv.ctypes.data, emis, tskin, sdv, sn, lin, sec_dn, rad.ctypes.data, cap, ctypes.byref(dvo))
    return rad[:n], dvo.value
