"""The oracle's single-grid merge-up radiance (orc_rt_up) for the CPU tests."""
import ctypes

import numpy as np

from oracle.py import lib


def rt_up(od, t_lay, t_lev, v1, dv, emis=0.98, tskin=295.0, lin=1, sec_dn=1.0):
    L = lib()
    D = ctypes.c_double
    L.orc_rt_up.argtypes = [ctypes.c_int, ctypes.c_int64, D, D, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            D, D, ctypes.c_int, D, ctypes.c_void_p]
    od = np.ascontiguousarray(od, np.float64)
    rad = np.zeros(od.shape[1])
    L.orc_rt_up(od.shape[0], od.shape[1], v1, dv, od.ctypes.data, np.ascontiguousarray(t_lay, np.float64).ctypes.data,
                np.ascontiguousarray(t_lev, np.float64).ctypes.data, emis, tskin, lin, sec_dn, rad.ctypes.data)
    return rad
