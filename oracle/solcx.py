"""The reference's analytic solCx solution (Underworld/SysTest/AnalyticPlugins/Velic_solCx/solCx.c, compiled by
oracle/Makefile into oracle/_ref/libsolcx.so): free-slip unit box, density -sigma sin(pi z) cos(n pi x), viscosity
eta_A for x < x_c and eta_B beyond.  Test infrastructure: the closed-form checker of the discretisation and the
solve, used only by tests."""
import ctypes as C
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libsolcx.so")
_LIB = None


def available():
    return os.path.exists(PATH)


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(PATH)
        d = C.POINTER(C.c_double)
        _LIB._Velic_solCx.argtypes = [d, C.c_double, C.c_double, C.c_double, C.c_int, d, d, d, d]
        _LIB._Velic_solCx.restype = None
    return _LIB


def solcx(x, z, eta_a, eta_b, x_c, n):
    """Velocity (u_x, u_z) and pressure at the points (x, z)."""
    L = lib()
    x = np.asarray(x, dtype=np.float64).ravel()
    z = np.asarray(z, dtype=np.float64).ravel()
    vel = np.zeros((len(x), 2))
    p = np.zeros(len(x))
    pos, v, pr = (C.c_double * 2)(), (C.c_double * 2)(), C.c_double()
    st, sr = (C.c_double * 3)(), (C.c_double * 3)()
    for k in range(len(x)):
        pos[0], pos[1] = x[k], z[k]
        L._Velic_solCx(pos, eta_a, eta_b, x_c, int(n), v, C.byref(pr), st, sr)
        vel[k] = (v[0], v[1])
        p[k] = pr.value
    return vel, p
