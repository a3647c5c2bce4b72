"""ORACLE — test infrastructure only.

ctypes binding of ``liboracle.so``, the plain-C FP64 restatement of the
reference pseudo-time step (see ``orc_eos.c`` / ``orc_ops.c`` /
``orc_step.c`` for the per-function reference citations).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module, and only as the checker: the product path never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Cfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "npv", "ntv", "ndv", "ngrd", "iturb", "ides", "nscheme", "nmuscl", "nlim",
        "timemethod", "local", "prec", "ncav", "nsteady", "nprint")] + \
        [(n, C.c_double) for n in (
            "cfl", "pref", "uref", "str", "dt_phy", "c_v", "c_c", "dp_ref", "t_ref", "t_crit")] + \
        [("bc_pv", C.c_double * 9), ("bc_pressure", C.c_double), ("emutref", C.c_double),
         ("cav_clip", C.c_int)]


class Eos2(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho", "h", "drdp", "drdt", "dhdp", "dhdt")]


class Prop2(C.Structure):
    _fields_ = [("vis", C.c_double), ("cond", C.c_double)]


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class GBlock(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("imax", "jmax", "kmax", "nbc", "ncon")] + \
        [("x", _D), ("bc", _I), ("con", _I), ("cx", _D), ("ex", _D), ("tx", _D), ("grd", _D)]


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.orc_powi.restype = C.c_double
        L.orc_powi.argtypes = [C.c_double, C.c_int]
        L.orc_h2o_pww.restype = C.c_double
        L.orc_h2o_pww.argtypes = [C.c_double]
        L.orc_h2o_tww.restype = C.c_double
        L.orc_h2o_tww.argtypes = [C.c_double]
        for f in ("orc_iapws97_l", "orc_iapws97_v", "orc_ideal"):
            getattr(L, f).argtypes = [C.c_double, C.c_double, C.POINTER(Eos2)]
        for f in ("orc_iapws97_l_prop", "orc_iapws97_v_prop", "orc_ideal_prop"):
            getattr(L, f).argtypes = [C.c_double, C.c_double, C.POINTER(Eos2), C.POINTER(Prop2)]
        L.orc_deteos_simple.argtypes = [C.c_double] * 4 + [_D]
        L.orc_deteos_prop.argtypes = [C.c_double] * 4 + [_D, _D]
        L.orc_new.restype = C.c_void_p
        L.orc_new.argtypes = [C.POINTER(Cfg), C.c_int, _I]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_set_grid.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
        L.orc_set_bcs.argtypes = [C.c_void_p, C.c_int, C.c_int, _I]
        L.orc_set_conns.argtypes = [C.c_void_p, C.c_int, C.c_int, _I]
        L.orc_finalize.argtypes = [C.c_void_p]
        L.orc_put_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
        L.orc_get_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
        L.orc_timestep.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _D]
        L.orc_timeinteg.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _D]
        L.orc_get_bcinfo.argtypes = [C.c_void_p, C.c_int, _I]
        L.orc_setbc.argtypes = [C.c_void_p]
        L.orc_calrhs.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
        L.orc_get_dt.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_setqres.argtypes = [C.c_void_p]
        L.orc_residual.argtypes = [C.c_void_p, _D, _I, _D]
        L.orc_unsteady_shift.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_grid_metrics.argtypes = [C.c_int, C.POINTER(GBlock), C.c_int]
        L.orc_integrals.argtypes = [C.c_void_p, C.POINTER(_D), C.c_double, C.c_double, C.c_double, C.c_double, _D]
        L.orc_alt_snd2_count_ptr = C.c_int.in_dll(L, "orc_alt_snd2_count")
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(t))


# ---------------------------------------------------------------- EOS
def iapws97_l(p, t):
    e = Eos2(); lib().orc_iapws97_l(p, t, C.byref(e)); return e


def iapws97_v(p, t):
    e = Eos2(); lib().orc_iapws97_v(p, t, C.byref(e)); return e


def iapws97_l_prop(p, t):
    e, q = Eos2(), Prop2(); lib().orc_iapws97_l_prop(p, t, C.byref(e), C.byref(q)); return e, q


def iapws97_v_prop(p, t):
    e, q = Eos2(), Prop2(); lib().orc_iapws97_v_prop(p, t, C.byref(e), C.byref(q)); return e, q


def ideal_prop(p, t):
    e, q = Eos2(), Prop2(); lib().orc_ideal_prop(p, t, C.byref(e), C.byref(q)); return e, q


def h2o_pww(t):
    return lib().orc_h2o_pww(t)


def h2o_tww(p):
    return lib().orc_h2o_tww(p)


def powi(x, m):
    return lib().orc_powi(x, m)


def deteos_simple(p, t, y1, y2):
    dv = np.zeros(18)
    lib().orc_deteos_simple(p, t, y1, y2, _p(dv))
    return dv


def deteos_prop(p, t, y1, y2):
    dv = np.zeros(18)
    tv = np.zeros(3)
    lib().orc_deteos_prop(p, t, y1, y2, _p(dv), _p(tv))
    return dv, tv[:2]


def grid_metrics(blocks, ngrd):
    """cx, ex, tx, grd of every block (flat, Fortran layouts) from the vertex
    arrays X (imax, jmax, kmax, 3), bc_desc and conn_desc of `blocks`."""
    keep, out = [], []
    arr = (GBlock * len(blocks))()
    for q, b in enumerate(blocks):
        I, J, K = b.imax, b.jmax, b.kmax
        x = np.ascontiguousarray(np.asarray(b.X, dtype=np.float64).transpose(2, 1, 0, 3)).ravel()
        bc = np.ascontiguousarray(np.asarray(b.bc_desc, dtype=np.int32).ravel()) if len(b.bc_desc) else np.zeros(7, np.int32)
        con = np.ascontiguousarray(np.asarray(b.conn_desc, dtype=np.int32).ravel()) if len(b.conn_desc) else np.zeros(22, np.int32)
        o = (np.zeros(3 * I * (J + 1) * (K + 1)), np.zeros(3 * (I + 1) * J * (K + 1)),
             np.zeros(3 * (I + 1) * (J + 1) * K), np.zeros(ngrd * (I + 1) * (J + 1) * (K + 1)))
        keep += [x, bc, con]
        out.append(o)
        arr[q] = GBlock(I, J, K, len(b.bc_desc), len(b.conn_desc), _p(x), _p(bc, C.c_int), _p(con, C.c_int),
                        *(_p(a) for a in o))
    rc = lib().orc_grid_metrics(len(blocks), arr, ngrd)
    if rc != 0:
        raise RuntimeError(f"orc_grid_metrics failed ({rc})")
    return out


def set_threads(n: int):
    lib().orc_set_threads(int(n))


# ---------------------------------------------------------------- whole step
class Oracle:
    """Multi-block context holding the reference-layout arrays of a case."""

    def __init__(self, case):
        L = lib()
        self.case = case
        cf = Cfg()
        for k, v in case.cfg_dict().items():
            if k == "bc_pv":
                arr = (C.c_double * 9)(*list(v) + [0.0] * (9 - len(v)))
                cf.bc_pv = arr
            else:
                setattr(cf, k, v)
        self._cfg = cf
        dims = np.array([d for b in case.blocks for d in (b.imax, b.jmax, b.kmax)], dtype=np.int32)
        self.h = L.orc_new(C.byref(cf), len(case.blocks), _p(dims, C.c_int))
        for ib, b in enumerate(case.blocks):
            L.orc_set_grid(self.h, ib, _p(b.cx), _p(b.ex), _p(b.tx), _p(b.grd))
            bcs = np.ascontiguousarray(np.array(b.bc_desc, dtype=np.int32).reshape(-1))
            if L.orc_set_bcs(self.h, ib, len(b.bc_desc), _p(bcs, C.c_int)) != 0:
                raise ValueError("bad bc description")
            cons = np.ascontiguousarray(np.array(b.conn_desc, dtype=np.int32).reshape(-1)) if b.conn_desc else np.zeros(22, np.int32)
            L.orc_set_conns(self.h, ib, len(b.conn_desc), _p(cons, C.c_int))
        if L.orc_finalize(self.h) != 0:
            raise ValueError("unsupported configuration")

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put_state(self, ib, pv, dv, tv, qq=None):
        b = self.case.blocks[ib]
        ng = (b.imax + 5) * (b.jmax + 5) * (b.kmax + 5)
        ni = (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)
        for a, size in ((pv, self.case.npv * ng), (dv, 18 * ng), (tv, self.case.ntv * ng), (qq, 2 * self.case.npv * ni)):
            if a is not None and a.size != size:
                raise ValueError(f"state array has {a.size} values, expected {size}")
        lib().orc_put_state(self.h, ib, _p(pv), _p(dv), _p(tv), _p(qq))

    def get_state(self, ib):
        b = self.case.blocks[ib]
        npv, ntv = self.case.npv, self.case.ntv
        ng = (b.imax + 5) * (b.jmax + 5) * (b.kmax + 5)
        ni = (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)
        pv = np.zeros(npv * ng); dv = np.zeros(18 * ng); tv = np.zeros(max(ntv, 1) * ng)
        qq = np.zeros(2 * npv * ni)
        lib().orc_get_state(self.h, ib, _p(pv), _p(dv), _p(tv), _p(qq))
        return pv, dv, (tv if ntv else None), qq

    def timestep(self, nt_phy=1, nt=1, timeprev=0.0):
        t = C.c_double(0.0)
        lib().orc_timestep(self.h, nt_phy, nt, timeprev, C.byref(t))
        return t.value

    def timeinteg(self, nt_phy=1, nt=1, timeprev=0.0):
        t = C.c_double(0.0)
        lib().orc_timeinteg(self.h, nt_phy, nt, timeprev, C.byref(t))
        return t.value

    def setbc(self):
        lib().orc_setbc(self.h)

    def calrhs(self, ib):
        b = self.case.blocks[ib]
        ni = (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)
        res = np.zeros(self.case.npv * ni); icav = np.zeros(4 * ni); itt = np.zeros(4 * ni); oc = np.zeros(ni)
        lib().orc_calrhs(self.h, ib, _p(res), _p(icav), _p(itt), _p(oc))
        return res, icav, itt, oc

    def get_dt(self, ib):
        b = self.case.blocks[ib]
        dt = np.zeros((b.imax - 1) * (b.jmax - 1) * (b.kmax - 1))
        lib().orc_get_dt(self.h, ib, _p(dt))
        return dt

    def bcinfo(self, ib):
        out = np.zeros(20 * 32, np.int32)
        lib().orc_get_bcinfo(self.h, ib, _p(out, C.c_int))
        return out.reshape(20, 32)

    def setqres(self):
        lib().orc_setqres(self.h)

    def residual(self):
        l2 = np.zeros(self.case.npv)
        ijk = np.zeros(4, np.int32)
        rm = C.c_double(0.0)
        lib().orc_residual(self.h, _p(l2), _p(ijk, C.c_int), C.byref(rm))
        return l2, ijk, rm.value

    def integrals(self, rhoref, uref, area, aoa=0.0):
        """vapour volume, gas volume, pressure force (3), viscous force (3), cl, cd."""
        xs = [np.ascontiguousarray(np.asarray(b.X, dtype=np.float64).transpose(2, 1, 0, 3)).ravel()
              for b in self.case.blocks]
        ptrs = (_D * len(xs))(*[_p(x) for x in xs])
        out = np.zeros(10)
        lib().orc_integrals(self.h, ptrs, rhoref, uref, area, aoa, _p(out))
        return out

    def unsteady_shift(self):
        lib().orc_unsteady_shift(self.h)
