"""ctypes wrapper of the CPU oracle (``oracle/libzefr_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, and only there as the checker or as the timed CPU
baseline.  The product never imports this module.

All arrays come back in the reference's CPU layouts (row-major, element index last):
``U_spts[nSpts][nVars][nEles]``, ``divF[nStages][nSpts][nVars][nEles]``,
``fpt2gfpt[nFpts][nEles]`` and so on (src/elements.cpp:87-121).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# arithmetic type of the oracle build -> its library (oracle/Makefile builds both)
_LIB_PATHS = {"double": os.path.join(_HERE, "libzefr_oracle.so"),
              "long double": os.path.join(_HERE, "libzefr_oracle_ld.so")}
_LIB_PATH = _LIB_PATHS["double"]


class ZoPhys(C.Structure):
    _fields_ = [("viscous", C.c_int), ("gamma", C.c_double), ("prandtl", C.c_double),
                ("mu", C.c_double), ("rt", C.c_double), ("c_sth", C.c_double),
                ("fix_vis", C.c_int), ("rus_k", C.c_double), ("ldg_b", C.c_double),
                ("ldg_tau", C.c_double), ("dt", C.c_double), ("nStages", C.c_int),
                ("res_type", C.c_int), ("fconv", C.c_int)]


def build():
    """Compile the oracle with its Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def lib(precision="double"):
    """The oracle library: "double" (the FP64 restatement, the checker) or "long double"
    (the same code on x87 extended precision, the exact answer FP64 results are measured
    against)."""
    if precision not in _libs:
        path = _LIB_PATHS[precision]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.zo_last_error.restype = C.c_char_p
        L.zo_create.restype = C.c_void_p
        L.zo_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(ZoPhys)]
        L.zo_destroy.argtypes = [C.c_void_p]
        L.zo_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.zo_get_maps.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_byte)]
        L.zo_get_array.restype = C.c_long
        L.zo_get_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_long]
        L.zo_set_U.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.zo_compute_residual.argtypes = [C.c_void_p, C.c_int]
        L.zo_step_rk.argtypes = [C.c_void_p, C.c_int]
        L.zo_report_residuals.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.zo_gl_points.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.zo_lagrange.restype = C.c_double
        L.zo_lagrange.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double]
        L.zo_lagrange_d1.restype = C.c_double
        L.zo_lagrange_d1.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double]
        L.zo_point_flux.argtypes = [C.c_int, C.POINTER(ZoPhys), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.zo_common_flux.argtypes = [C.c_int, C.POINTER(ZoPhys), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.zo_pmg_create.restype = C.c_void_p
        L.zo_pmg_create.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.c_double, C.c_char_p, C.c_int]
        L.zo_pmg_destroy.argtypes = [C.c_void_p]
        L.zo_pmg_cycle.argtypes = [C.c_void_p, C.c_int]
        L.zo_pmg_array.restype = C.c_long
        L.zo_pmg_array.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_long, C.c_int]
        L.zo_pmg_operator.restype = C.c_long
        L.zo_pmg_operator.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_long]
        _libs[precision] = L
    return _libs[precision]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc, L=None):
    if rc != 0:
        raise RuntimeError((L or lib()).zo_last_error().decode())


def make_phys(viscous=False, gamma=1.4, prandtl=0.72, mu=0.0, rt=0.0, c_sth=0.0, fix_vis=True,
              rus_k=0.0, ldg_b=0.5, ldg_tau=0.1, dt=1e-3, nStages=4, res_type=2, fconv=0):
    """fconv: 0 Rusanov (the reference's flux), 2 Roe (textbook, Harten entropy fix)."""
    return ZoPhys(int(viscous), gamma, prandtl, mu, rt, c_sth, int(fix_vis), rus_k, ldg_b,
                  ldg_tau, dt, nStages, res_type, fconv)


class Oracle:
    """One oracle solver instance on a Gmsh file (single element type, periodic)."""

    def __init__(self, meshfile, nDims, order, phys: ZoPhys, precision="double"):
        self.precision = precision
        self._L = lib(precision)
        self._h = self._L.zo_create(meshfile.encode(), nDims, order, C.byref(phys))
        if not self._h:
            raise RuntimeError(self._L.zo_last_error().decode())
        d = (C.c_int * 10)()
        self._L.zo_dims(self._h, d)
        (self.nEles, self.nSpts, self.nFpts, self.nVars, self.nDims, self.nGfpts,
         self.nGfpts_int, self.nStages, self.nFaces, self.nIntFaces) = list(d)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.zo_destroy(h)
            self._h = None

    def maps(self):
        g = np.zeros((self.nFpts, self.nEles), np.int32)
        s = np.zeros((self.nFpts, self.nEles), np.int8)
        self._L.zo_get_maps(self._h, g.ctypes.data_as(C.POINTER(C.c_int)),
                          s.ctypes.data_as(C.POINTER(C.c_byte)))
        return g, s

    def array(self, name, shape=None):
        buf = np.zeros(self._size_hint(name), np.float64)
        cnt = self._L.zo_get_array(self._h, name.encode(), _dp(buf), buf.size)
        if cnt < 0:
            raise RuntimeError(self._L.zo_last_error().decode())
        out = buf[:cnt].copy()
        return out.reshape(shape) if shape is not None else out

    def _size_hint(self, name):
        E, nS, nF, V, D, G = self.nEles, self.nSpts, self.nFpts, self.nVars, self.nDims, self.nGfpts
        sizes = {"oppE": nF * nS, "oppD": D * nS * nS, "oppD_fpts": D * nS * nF,
                 "oppDiv": nS * D * nS, "oppDiv_fpts": nS * nF, "coord_spts": nS * D * E,
                 "coord_fpts": nF * D * E, "jaco_det_spts": nS * E, "inv_jaco_spts": D * nS * D * E,
                 "inv_jaco_fpts": D * nF * D * E, "norm": D * G, "dA": 2 * G, "weights_spts": nS,
                 "loc_spts": nS * D, "loc_fpts": nF * D, "U_spts": nS * V * E, "U_fpts": nF * V * E,
                 "Ucomm": nF * V * E, "F_spts": D * nS * V * E, "Fcomm": nF * V * E,
                 "dU_spts": D * nS * V * E, "dU_fpts": D * nF * V * E,
                 "divF": self.nStages * nS * V * E, "coord_nodes": 1 << 24}
        return sizes[name]

    def coord_spts(self):
        return self.array("coord_spts", (self.nSpts, self.nDims, self.nEles))

    def jaco_det(self):
        return self.array("jaco_det_spts", (self.nSpts, self.nEles))

    def set_U(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.nSpts, self.nVars, self.nEles)
        self._L.zo_set_U(self._h, _dp(U))

    def U(self):
        return self.array("U_spts", (self.nSpts, self.nVars, self.nEles))

    def compute_residual(self, stage=0):
        _check(self._L.zo_compute_residual(self._h, stage), self._L)

    def divF(self, stage=0):
        d = self.array("divF", (self.nStages, self.nSpts, self.nVars, self.nEles))
        return d[stage]

    def step_rk(self, nsteps=1):
        _check(self._L.zo_step_rk(self._h, nsteps), self._L)

    def report_residuals(self):
        r = np.zeros(self.nVars)
        _check(self._L.zo_report_residuals(self._h, _dp(r)), self._L)
        return r


class OraclePMG:
    """p-multigrid over an Oracle (multigrid.cpp:18-518): coarse oracle solvers at the
    orders ``levels`` (finest first, equal to the fine order) on the same mesh."""

    _WHICH = {"U": 0, "divF": 2, "source": 3, "solution": 4, "correction": 5}

    def __init__(self, fine: Oracle, meshfile, levels, steps, rel_fac=1.0, cycle="V", fmg_vcycles=1):
        self.fine = fine  # keeps the fine solver alive as long as the cycle object
        self._L = fine._L
        self.levels = list(levels)
        n = len(self.levels)
        lv = (C.c_int * n)(*self.levels)
        st = (C.c_int * n)(*steps)
        self._h = self._L.zo_pmg_create(fine._h, meshfile.encode(), n, lv, st, rel_fac, cycle.encode(), fmg_vcycles)
        if not self._h:
            raise RuntimeError(self._L.zo_last_error().decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.zo_pmg_destroy(h)
            self._h = None

    def cycle(self, n=1):
        _check(self._L.zo_pmg_cycle(self._h, n), self._L)

    def _shape(self, level):
        f = self.fine
        return ((self.levels[level] + 1) ** f.nDims, f.nVars, f.nEles)

    def array(self, level, which="U"):
        a = np.zeros(self._shape(level))
        if self._L.zo_pmg_array(self._h, level, self._WHICH[which], _dp(a), a.size, 0) < 0:
            raise RuntimeError(self._L.zo_last_error().decode())
        return a

    def set_U(self, level, U):
        a = np.ascontiguousarray(U, np.float64)
        assert a.shape == self._shape(level)
        if self._L.zo_pmg_array(self._h, level, 0, _dp(a), a.size, 1) < 0:
            raise RuntimeError(self._L.zo_last_error().decode())

    def operator(self, level, kind):
        """Dense operator of a level: kind 'res' (to level+1) or 'pro' (to level-1),
        [nSpts_to][nSpts_from]."""
        k = {"res": 0, "pro": 1}[kind]
        cnt = self._L.zo_pmg_operator(self._h, level, k, None, 0)
        if cnt < 0:
            raise RuntimeError(self._L.zo_last_error().decode())
        a = np.zeros(cnt)
        self._L.zo_pmg_operator(self._h, level, k, _dp(a), cnt)
        D = self.fine.nDims
        to = self.levels[level + 1] if k == 0 else self.levels[level - 1]
        return a.reshape((to + 1) ** D, (self.levels[level] + 1) ** D)


def gl_points(n):
    p = np.zeros(n); w = np.zeros(n)
    _check(lib().zo_gl_points(n, _dp(p), _dp(w)))
    return p, w


def lagrange(grid, mode, x, deriv=False):
    g = np.ascontiguousarray(grid, np.float64)
    f = lib().zo_lagrange_d1 if deriv else lib().zo_lagrange
    return f(_dp(g), g.size, mode, float(x))


def roe_flux(nDims, UL, UR, n, gamma=1.4):
    """The oracle's Roe common normal flux for the unit normal n."""
    phys = make_phys(gamma=gamma, fconv=2)
    F = np.zeros(nDims + 2)
    UL = np.ascontiguousarray(UL, np.float64)
    UR = np.ascontiguousarray(UR, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    _check(lib().zo_common_flux(nDims, C.byref(phys), _dp(UL), _dp(UR), _dp(n), _dp(F)))
    return F


def point_flux(nDims, phys, U, dU=None):
    U = np.ascontiguousarray(U, np.float64)
    V = nDims + 2
    F = np.zeros((V, nDims))
    dptr = None
    if dU is not None:
        dU = np.ascontiguousarray(dU, np.float64).reshape(V, nDims)
        dptr = _dp(dU)
    _check(lib().zo_point_flux(nDims, C.byref(phys), _dp(U), dptr, _dp(F)))
    return F
