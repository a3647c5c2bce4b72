"""ctypes binding of libcav3d.so (include/cav3d.h) — the product path.

Mirrors the reference's operator interface for the pseudo-time step:
``Solver.timeinteg`` is ``t_lusgs%timeinteg`` (src/update_module.f90:455),
``calrhs`` is ``t_rhs%calrhs`` (src/rhs_module.f90:260), ``residual`` is
``t_resi%setqres`` + ``t_resi%residual`` (src/residual_module.f90:94-186).
Every call goes to the HIP library; there is no CPU fallback and a missing or
failing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcav3d.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

BC = {
    "wall": 1, "inflow_subsonic": 2, "outflow_subsonic": 3, "farfield": 4, "symmetry": 5,
    "inflow_supersonic": 6, "outflow_supersonic": 7, "degenerate": 8, "inflow": 9, "outflow": 10,
}

SURFACE = 0x100   # CAV3D_BC_SURFACE: OR'ed into a kind, marks the force-integration surface

EXPORTS = [
    "cav3d_create", "cav3d_set_metrics", "cav3d_set_grid", "cav3d_build_metrics", "cav3d_get_metrics",
    "cav3d_integrals", "cav3d_integrals_noalloc", "cav3d_export", "cav3d_import", "cav3d_rhs_vjp",
    "cav3d_rhs_vjp_params", "cav3d_objective", "cav3d_rhs_vjp_eddy", "cav3d_eddy",
    "cav3d_put_state", "cav3d_get_state", "cav3d_timeinteg",
    "cav3d_setbc", "cav3d_calrhs", "cav3d_timestep", "cav3d_residual", "cav3d_unsteady_shift",
    "cav3d_eos", "cav3d_eos_host", "cav3d_profile", "cav3d_synchronize", "cav3d_last_error",
    "cav3d_destroy",
]


class Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("iturb", "nscheme", "nmuscl", "nlim", "local", "prec", "ncav",
                                       "nsteady", "precision", "cav_clip")] + \
        [("ides", C.c_int), ("reserved", C.c_int)] + \
        [(n, C.c_double) for n in ("cfl", "pref", "uref", "str", "dt_phy", "c_v", "c_c", "dp_ref",
                                   "t_ref", "t_crit")] + \
        [("bc_pv", C.c_double * 9), ("bc_pressure", C.c_double), ("emutref", C.c_double)]


class BlockDesc(C.Structure):
    _fields_ = [("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int), ("nbc", C.c_int),
                ("bc", _I), ("ncon", C.c_int), ("con", _I)]


class Cav3dError(RuntimeError):
    pass


def lib(build_if_missing: bool = True):
    """Load libcav3d.so (building it with hipcc only if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        if not build_if_missing:
            raise Cav3dError(f"{LIB_PATH} is missing; run __graft_entry__.build()")
        import importlib
        importlib.import_module(__package__ + ".build").build()
    L = C.CDLL(LIB_PATH)
    L.cav3d_last_error.restype = C.c_char_p
    L.cav3d_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(BlockDesc), C.POINTER(C.c_void_p)]
    L.cav3d_set_metrics.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
    L.cav3d_set_grid.argtypes = [C.c_void_p, C.c_int, _D]
    L.cav3d_build_metrics.argtypes = [C.c_void_p]
    L.cav3d_get_metrics.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
    L.cav3d_integrals.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _D]
    L.cav3d_integrals_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _D]
    L.cav3d_export.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_double]
    L.cav3d_import.argtypes = [C.c_void_p, C.c_char_p, _I, _I, _D]
    L.cav3d_rhs_vjp.argtypes = [C.c_void_p, _D, _D, _D]
    L.cav3d_rhs_vjp_params.argtypes = [C.c_void_p, _D, _D]
    L.cav3d_rhs_vjp_eddy.argtypes = [C.c_void_p, _D, _D, _D]
    L.cav3d_eddy.argtypes = [C.c_void_p]
    L.cav3d_objective.argtypes = [C.c_void_p, C.c_double, C.c_double, _D, _D, _D]
    L.cav3d_put_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
    L.cav3d_get_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
    L.cav3d_timeinteg.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _D]
    L.cav3d_setbc.argtypes = [C.c_void_p]
    L.cav3d_calrhs.argtypes = [C.c_void_p, C.c_double, C.c_int, _D, _D, _D, _D]
    L.cav3d_timestep.argtypes = [C.c_void_p, C.c_int, _D]
    L.cav3d_residual.argtypes = [C.c_void_p, _D, _I, _D]
    L.cav3d_unsteady_shift.argtypes = [C.c_void_p]
    L.cav3d_eos.argtypes = [C.c_void_p, C.c_int] + [_D] * 6
    L.cav3d_eos_host.argtypes = [C.c_int] + [_D] * 6
    L.cav3d_profile.argtypes = [C.c_void_p, C.c_int, _D]
    L.cav3d_synchronize.argtypes = [C.c_void_p]
    L.cav3d_destroy.argtypes = [C.c_void_p]
    L.cav3d_destroy.restype = None
    _lib = L
    return L


def _p(a, t=C.c_double):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.POINTER(t))


def _check(rc: int):
    if rc != 0:
        msg = lib().cav3d_last_error().decode()
        raise Cav3dError(f"cav3d error {rc}: {msg}")


def eos_host(p, t, y1, y2):
    """dv(18,n), tv(2,n) of absolute (p, T, y1, y2) on the host (set-up use)."""
    p, t, y1, y2 = (np.ascontiguousarray(np.broadcast_to(a, np.broadcast(p, t, y1, y2).shape), dtype=np.float64).ravel()
                    for a in (p, t, y1, y2))
    n = p.size
    dv = np.empty((n, 18))
    tv = np.empty((n, 2))
    _check(lib().cav3d_eos_host(n, _p(p), _p(t), _p(y1), _p(y2), _p(dv), _p(tv)))
    return dv, tv


def params_struct(case) -> Params:
    P = Params()
    for k, v in case.params.items():
        if k == "bc_pv":
            P.bc_pv = (C.c_double * 9)(*list(v))
        elif hasattr(P, k):
            setattr(P, k, v)
    P.precision = case.params.get("precision", 0)
    return P


class Solver:
    """One GPU context holding every block of a case."""

    def __init__(self, case, precision: int = 0, metrics: str = "host"):
        """metrics "host": the case's metric arrays are copied in
        (cav3d_set_metrics); "device": they are computed on the GPU from the
        case's vertex coordinates (cav3d_set_grid + cav3d_build_metrics)."""
        if metrics not in ("host", "device"):
            raise ValueError("metrics must be 'host' or 'device'")
        L = lib()
        self.case = case
        self.npv, self.ntv = case.npv, case.ntv
        P = params_struct(case)
        P.precision = precision
        self._keep = []
        descs = (BlockDesc * len(case.blocks))()
        for ib, b in enumerate(case.blocks):
            bc = np.ascontiguousarray(np.array(b.bc_desc, dtype=np.int32).reshape(-1))
            con = np.ascontiguousarray(np.array(b.conn_desc, dtype=np.int32).reshape(-1)) if b.conn_desc \
                else np.zeros(22, np.int32)
            self._keep += [bc, con]
            descs[ib] = BlockDesc(b.imax, b.jmax, b.kmax, len(b.bc_desc), _p(bc, C.c_int),
                                  len(b.conn_desc), _p(con, C.c_int))
        h = C.c_void_p()
        _check(L.cav3d_create(C.byref(P), len(case.blocks), descs, C.byref(h)))
        self.h = h
        if metrics == "host":
            for ib, b in enumerate(case.blocks):
                _check(L.cav3d_set_metrics(self.h, ib, _p(b.cx), _p(b.ex), _p(b.tx), _p(b.grd)))
        else:
            for ib, b in enumerate(case.blocks):
                x = np.ascontiguousarray(np.asarray(b.X, dtype=np.float64).transpose(2, 1, 0, 3)).ravel()
                _check(L.cav3d_set_grid(self.h, ib, _p(x)))
            _check(L.cav3d_build_metrics(self.h))

    def get_metrics(self, ib):
        """cx, ex, tx, grd of block ib as the context holds them (Fortran layouts, flat)."""
        b = self.case.blocks[ib]
        I, J, K = b.imax, b.jmax, b.kmax
        cx = np.zeros(3 * I * (J + 1) * (K + 1))
        ex = np.zeros(3 * (I + 1) * J * (K + 1))
        tx = np.zeros(3 * (I + 1) * (J + 1) * K)
        grd = np.zeros(self.case.ngrd * (I + 1) * (J + 1) * (K + 1))
        _check(lib().cav3d_get_metrics(self.h, ib, _p(cx), _p(ex), _p(tx), _p(grd)))
        return cx, ex, tx, grd

    def close(self):
        if getattr(self, "h", None):
            lib().cav3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state
    def put_state(self, ib, pv, dv=None, tv=None, qq=None):
        b = self.case.blocks[ib]
        ng = (b.imax + 5) * (b.jmax + 5) * (b.kmax + 5)
        ni = (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)
        for name, a, size in (("pv", pv, self.npv * ng), ("dv", dv, 18 * ng), ("tv", tv, self.ntv * ng),
                              ("qq", qq, 2 * self.npv * ni)):
            if a is not None and (a.dtype != np.float64 or a.size != size):
                raise ValueError(f"{name}: expected {size} float64 values, got {a.size} {a.dtype}")
        _check(lib().cav3d_put_state(self.h, ib, _p(pv), _p(dv), _p(tv), _p(qq)))

    def get_state(self, ib):
        b = self.case.blocks[ib]
        ng = (b.imax + 5) * (b.jmax + 5) * (b.kmax + 5)
        ni = (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)
        pv = np.zeros(self.npv * ng)
        dv = np.zeros(18 * ng)
        tv = np.zeros(max(self.ntv, 1) * ng)
        qq = np.zeros(2 * self.npv * ni)
        _check(lib().cav3d_get_state(self.h, ib, _p(pv), _p(dv), _p(tv) if self.ntv else None, _p(qq)))
        return pv, dv, (tv if self.ntv else None), qq

    # ---- operators
    def timeinteg(self, nt_phy=1, nt=1, timeprev=0.0, want_time=True):
        t = C.c_double(0.0)
        _check(lib().cav3d_timeinteg(self.h, nt_phy, nt, timeprev, C.byref(t) if want_time else None))
        return t.value

    def setbc(self):
        _check(lib().cav3d_setbc(self.h))

    def calrhs(self, ib, time=0.0):
        b = self.case.blocks[ib]
        ni = (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)
        res = np.zeros(self.npv * ni)
        icav = np.zeros(4 * ni)
        itt = np.zeros(4 * ni)
        oc = np.zeros(ni)
        _check(lib().cav3d_calrhs(self.h, time, ib, _p(res), _p(icav), _p(itt), _p(oc)))
        return res, icav, itt, oc

    def timestep(self, ib):
        b = self.case.blocks[ib]
        dt = np.zeros((b.imax - 1) * (b.jmax - 1) * (b.kmax - 1))
        _check(lib().cav3d_timestep(self.h, ib, _p(dt)))
        return dt

    def residual(self):
        l2 = np.zeros(self.npv)
        ijk = np.zeros(4, np.int32)
        rm = C.c_double(0.0)
        _check(lib().cav3d_residual(self.h, _p(l2), _p(ijk, C.c_int), C.byref(rm)))
        return l2, ijk, rm.value

    def unsteady_shift(self):
        _check(lib().cav3d_unsteady_shift(self.h))

    def eos(self, p, t, y1, y2):
        p, t, y1, y2 = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (p, t, y1, y2))
        n = p.size
        dv = np.empty((n, 18))
        tv = np.empty((n, 2))
        _check(lib().cav3d_eos(self.h, n, _p(p), _p(t), _p(y1), _p(y2), _p(dv), _p(tv)))
        return dv, tv

    def set_grid(self):
        """Hand the case's vertex coordinates to the context (surface forces
        need them; cav3d_build_metrics is not run)."""
        for ib, b in enumerate(self.case.blocks):
            x = np.ascontiguousarray(np.asarray(b.X, dtype=np.float64).transpose(2, 1, 0, 3)).ravel()
            _check(lib().cav3d_set_grid(self.h, ib, _p(x)))

    def integrals(self, rhoref, uref, area, aoa=0.0):
        """vapour volume, gas volume, pressure force (3), viscous force (3), cl, cd."""
        out = np.zeros(10)
        _check(lib().cav3d_integrals(self.h, rhoref, uref, area, aoa, _p(out)))
        return out

    def integrals_noalloc(self, rhoref, uref, area, aoa=0.0):
        """The same integrals through cav3d_integrals_noalloc (no device allocation)."""
        out = np.zeros(10)
        _check(lib().cav3d_integrals_noalloc(self.h, rhoref, uref, area, aoa, _p(out)))
        return out

    def export(self, path, nt_phy, nt, time):
        _check(lib().cav3d_export(self.h, os.fsencode(path), nt_phy, nt, time))

    def import_(self, path):
        """Read a restart file; returns (nps, nts, timeprev) as the reader advances them."""
        nps, nts, tp = C.c_int(0), C.c_int(0), C.c_double(0.0)
        _check(lib().cav3d_import(self.h, os.fsencode(path), C.byref(nps), C.byref(nts), C.byref(tp)))
        return nps.value, nts.value, tp.value

    # ---- adjoint
    def rhs_vjp(self, res_bars, eddy=False):
        """(d res/d pv)^T res_bar over every block (cav3d_rhs_vjp; with
        eddy=True cav3d_rhs_vjp_eddy, the interior eddy viscosity a function
        of the state).

        res_bars: one array per block in calrhs's res layout.  Returns the
        per-block pv_bar arrays (pv layout, ghosted) and [d/d p_ref, d/d u_ref]
        of res_bar . res."""
        blocks = self.case.blocks
        rb = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.float64).ravel() for r in res_bars]))
        nc = [self.npv * (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1) for b in blocks]
        if rb.size != sum(nc):
            raise ValueError(f"res_bar: expected {sum(nc)} values, got {rb.size}")
        ng = [self.npv * (b.imax + 5) * (b.jmax + 5) * (b.kmax + 5) for b in blocks]
        pb = np.zeros(sum(ng))
        sc = np.zeros(2)
        fn = lib().cav3d_rhs_vjp_eddy if eddy else lib().cav3d_rhs_vjp
        _check(fn(self.h, _p(rb), _p(pb), _p(sc)))
        return np.split(pb, np.cumsum(ng)[:-1]), sc

    def param_vjp(self, res_bars):
        """d(res_bar . res)/d(run parameters) with the ghost layers refilled by
        the boundary fill (cav3d_rhs_vjp_params), in the order of
        cases.PARAM_KEYS: p_ref, u_ref, dp_ref, t_ref, bc_pv[0..8],
        bc_pressure.  The current ghost layers must be the boundary fill of
        the interior (call setbc first)."""
        rb = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.float64).ravel() for r in res_bars]))
        out = np.zeros(14)
        _check(lib().cav3d_rhs_vjp_params(self.h, _p(rb), _p(out)))
        return out

    def objective(self, w_vap=1.0, w_fx=1.0):
        """J = w_vap (vapour volume) + w_fx (x-force on the surface ranges, N),
        dJ/dpv per block (pv layout, ghosted) and dJ/d(p_ref) (cav3d_objective)."""
        ng = [self.npv * (b.imax + 5) * (b.jmax + 5) * (b.kmax + 5) for b in self.case.blocks]
        pb = np.zeros(sum(ng))
        J, dp = C.c_double(), C.c_double()
        _check(lib().cav3d_objective(self.h, w_vap, w_fx, C.byref(J), _p(pb), C.byref(dp)))
        return J.value, np.split(pb, np.cumsum(ng)[:-1]), dp.value

    def put_with_eos(self, pvs, muts=None, qqs=None):
        """Load ghosted states pvs with the cell properties from the device EOS
        and the eddy viscosity muts / BDF history qqs as given."""
        pref = self.case.params["pref"]
        for ib, pv in enumerate(pvs):
            a = np.asarray(pv, dtype=np.float64).reshape(-1, self.npv)
            dv, t2 = self.eos(a[:, 0] + pref, a[:, 4], a[:, 5], a[:, 6])
            tv = None
            if self.ntv:
                tv = np.zeros((a.shape[0], self.ntv))
                tv[:, :2] = t2
                if self.ntv == 3:
                    tv[:, 2] = muts[ib]
            qq = None if qqs is None else qqs[ib]
            self.put_state(ib, np.ascontiguousarray(a.ravel()), np.ascontiguousarray(dv.ravel()),
                           None if tv is None else np.ascontiguousarray(tv.ravel()), qq)

    def eddy(self):
        """The interior cells' SST eddy viscosity from the current state (cav3d_eddy)."""
        _check(lib().cav3d_eddy(self.h))

    def residual_of_eddy(self, pvs, muts, qqs=None):
        """The map cav3d_rhs_vjp_eddy differentiates: residual_of with the
        interior cells' eddy viscosity recomputed from the state (the ghost
        cells' taken from muts)."""
        pref = self.case.params["pref"]
        for ib, pv in enumerate(pvs):
            a = np.asarray(pv, dtype=np.float64).reshape(-1, self.npv)
            dv, t2 = self.eos(a[:, 0] + pref, a[:, 4], a[:, 5], a[:, 6])
            tv = np.zeros((a.shape[0], self.ntv))
            tv[:, :2] = t2
            tv[:, 2] = muts[ib]
            self.put_state(ib, np.ascontiguousarray(a.ravel()), np.ascontiguousarray(dv.ravel()),
                           np.ascontiguousarray(tv.ravel()), None if qqs is None else qqs[ib])
        self.eddy()
        return [self.calrhs(ib)[0] for ib in range(len(pvs))]

    def residual_after_setbc(self, pvs, muts=None, qqs=None):
        """Every block's residual for the states pvs with the ghost layers
        refilled by the boundary fill (the map d/d parameter differentiates),
        cell properties from the EOS, eddy viscosity and BDF history as given."""
        self.residual_of(pvs, muts, qqs)
        self.setbc()
        return [self.calrhs(ib)[0] for ib in range(len(pvs))]

    def frozen(self):
        """The parts of the current state the adjoint's map holds fixed: the
        eddy viscosity (ghosted, per block) and the BDF history (per block)."""
        muts, qqs = [], []
        for ib in range(len(self.case.blocks)):
            _, _, tv, qq = self.get_state(ib)
            muts.append(tv.reshape(-1, self.ntv)[:, 2].copy() if self.ntv == 3 else None)
            qqs.append(qq.copy() if self.case.params["nsteady"] >= 1 else None)
        return {"mut": muts, "qq": qqs}

    def residual_of(self, pvs, muts=None, qqs=None):
        """The map the adjoint differentiates: every block's residual for the
        ghosted states pvs, with the cell properties re-derived by the
        device EOS and the eddy viscosity muts (per block, ghosted) and BDF
        history qqs held as given.  Loads the states into the context."""
        pref = self.case.params["pref"]
        out = []
        for ib, pv in enumerate(pvs):
            a = np.asarray(pv, dtype=np.float64).reshape(-1, self.npv)
            dv, t2 = self.eos(a[:, 0] + pref, a[:, 4], a[:, 5], a[:, 6])
            tv = None
            if self.ntv:
                tv = np.zeros((a.shape[0], self.ntv))
                tv[:, :2] = t2
                if self.ntv == 3:
                    tv[:, 2] = muts[ib]
            qq = None if qqs is None else qqs[ib]
            self.put_state(ib, np.ascontiguousarray(a.ravel()), np.ascontiguousarray(dv.ravel()),
                           None if tv is None else np.ascontiguousarray(tv.ravel()), qq)
        for ib in range(len(pvs)):
            out.append(self.calrhs(ib)[0])
        return out

    def profile(self, enable=True):
        out = np.zeros(10)
        _check(lib().cav3d_profile(self.h, 1 if enable else 0, _p(out)))
        return out

    def synchronize(self):
        _check(lib().cav3d_synchronize(self.h))
