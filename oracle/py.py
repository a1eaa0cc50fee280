"""ctypes front end of the C oracle (liboracle.so).  Test infrastructure only."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class _Lines(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(k, ctypes.c_void_p) for k in
                ("nu", "s", "a_air", "e_lo", "mol", "a_self", "tmpalf", "shift", "iflg")]


class _Layer(ctypes.Structure):
    _fields_ = [("pave", ctypes.c_double), ("tave", ctypes.c_double), ("wtot", ctypes.c_double),
                ("w", ctypes.c_double * 48), ("sel", ctypes.c_int * 48)]


class _Params(ctypes.Structure):
    _fields_ = [("v1", ctypes.c_double), ("v2", ctypes.c_double), ("dvnormal", ctypes.c_double),
                ("dvnarrow", ctypes.c_double), ("sample", ctypes.c_double), ("dvout", ctypes.c_double),
                ("jrad", ctypes.c_int), ("contnm", ctypes.c_int), ("absrb", ctypes.c_void_p),
                ("nptabs", ctypes.c_int)]


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
        self._L = _Lines(lines.n, *[a.ctypes.data for a in self._keep])
        self.jrad = 1

    def odlay(self, layer, v1, v2, dvnormal, dvnarrow=None, sample=4.0, dvout=0.0, continuum=None):
        """Returns (od, dv_used).  JRAD persists across calls like the reference's SAVE."""
        lay = _Layer()
        lay.pave, lay.tave, lay.wtot = layer.pave, layer.tave, layer.wtot
        for m, v in layer.w.items():
            lay.w[m] = v
            lay.sel[m] = 1
        absrb = None
        npt = 0
        if continuum is not None:
            absrb = np.ascontiguousarray(continuum, np.float64)
            npt = absrb.size
        par = _Params(v1, v2, dvnormal, dvnormal if dvnarrow is None else dvnarrow, sample, dvout,
                      self.jrad, 1 if absrb is not None else 0,
                      absrb.ctypes.data if absrb is not None else None, npt)
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
