"""The library's pointwise physics (3dcode_amd/csrc/phys.hpp, built for the
host with contraction off) against the oracle on identical inputs.

Every routine is required to agree bit for bit: the kernels follow the
reference's operation order, so the only device-side differences left are
the device math library's exp/pow/tanh roundings."""
import ctypes as C
import math
import importlib

import numpy as np
import pytest

import physhost as H

D = C.POINTER(C.c_double)
PP = C.POINTER(D)


def dp(a):
    return a.ctypes.data_as(D)


def ptrs(arrs):
    return (D * len(arrs))(*[a.ctypes.data_as(D) for a in arrs])


@pytest.fixture(scope="module")
def L(tmp_path_factory):
    return H.build(str(tmp_path_factory.mktemp("physcheck")))


@pytest.fixture(scope="module")
def OL(orc):
    L = orc.lib()
    L.orc_tvd.argtypes = [C.c_void_p, D, D, D]
    L.orc_mlp.argtypes = [C.c_void_p, D, D, D, C.c_void_p]
    L.orc_roe.argtypes = [C.c_void_p, D, D, D, D, D, D]
    L.orc_ausmpup.argtypes = [C.c_void_p, D, D, D, D, D, D, C.c_void_p]
    L.orc_roem.argtypes = [C.c_void_p, D, D, D, D, D, D, D]
    L.orc_ausmpwp.argtypes = [C.c_void_p, D, D, D, D, D, D, D]
    L.orc_vsflux.argtypes = [C.c_void_p, D, PP, PP, D, D, D, D, D, D, D, D]
    L.orc_localdt.argtypes = [C.c_void_p, PP, D, D, D, D]
    L.orc_localdt.restype = C.c_double
    L.orc_getx.argtypes = [C.c_void_p, PP, D, D, D, D, C.c_double, D, D, D, D]
    L.orc_caljac.argtypes = [C.c_void_p, D, D, D, D, D, C.c_int, D]
    L.orc_merkle.argtypes = [C.c_void_p, D, D, D, D, D, D]
    L.orc_cavsource.argtypes = [C.c_void_p, D, D, D, D, D, D]
    L.orc_h2o_sigma.argtypes = [C.c_double]
    L.orc_h2o_sigma.restype = C.c_double
    L.orc_kwsst_source.argtypes = [PP, D, D, D, D, D, D, D]
    L.orc_kwsst_source_des.argtypes = [C.c_int, PP, D, D, D, D, D, D, D]
    L.orc_kwsst_eddy.argtypes = [PP, D, D, D, D]
    L.orc_kwsst_eddy.restype = C.c_double
    return L


def cfg_struct(orc, cfg):
    cf = orc.Cfg()
    for k, v in cfg.items():
        if k == "bc_pv":
            cf.bc_pv = (C.c_double * 9)(*v)
        else:
            setattr(cf, k, v)
    return cf


def test_phase_eos_bit_exact(L, orc):
    rng = np.random.default_rng(1)
    for _ in range(1000):
        p, t = rng.uniform(2e3, 5e6), rng.uniform(275.0, 500.0)
        for fn, ref in ((L.pc_liquid, orc.iapws97_l), (L.pc_vapour, orc.iapws97_v)):
            o = np.zeros(6)
            fn(p, t, dp(o))
            e = ref(p, t)
            assert np.array_equal(o, [e.rho, e.h, e.drdp, e.drdt, e.dhdp, e.dhdt]), (fn, p, t)


def test_mixture_eos_bit_exact(L, orc):
    rng = np.random.default_rng(2)
    for _ in range(1000):
        p, t = rng.uniform(2e3, 5e6), rng.uniform(275.0, 500.0)
        y1 = rng.choice([0.0, rng.uniform(0, 1e-3), rng.uniform(0, 1)])
        y2 = rng.choice([0.0, rng.uniform(0, 1 - y1)])
        full, tv = np.zeros(18), np.zeros(2)
        L.pc_mixture(p, t, y1, y2, dp(full), dp(tv))
        dref, tref = orc.deteos_prop(p, t, y1, y2)
        assert np.array_equal(full, dref) and np.array_equal(tv, tref)


def _stencil(rng, P4, npv, amp):
    k, j, i = rng.integers(3, P4.shape[0] - 3), rng.integers(3, P4.shape[1] - 3), rng.integers(4, P4.shape[2] - 4)
    rows = [P4[k + kk, j + jj, i + ii] for kk in (-1, 0, 1) for jj in (-1, 0, 1) for ii in (0, 1)]
    rows += [P4[k + kk, j + jj, i - 1] for kk in (-1, 0, 1) for jj in (-1, 0, 1)]
    rows += [P4[k + kk, j + jj, i + 2] for kk in (-1, 0, 1) for jj in (-1, 0, 1)]
    rows += [P4[k, j, i - 2], P4[k, j, i + 3]]
    x = np.zeros((38, 9))
    x[:, :npv] = np.array(rows) * (1 + amp * rng.standard_normal((38, npv)))
    return x


SCHEMES = [(1, 0, 0), (1, 1, 1), (1, 2, 8), (4, 1, 2), (4, 2, 8), (4, 1, 5), (1, 2, 3),
           (2, 0, 0), (2, 1, 1), (2, 2, 8), (3, 0, 0), (3, 1, 2), (3, 2, 8)]


@pytest.mark.parametrize("cfgname", ["config1", "config2"])
@pytest.mark.parametrize("nscheme,nmuscl,nlim", SCHEMES)
def test_inviscid_face_bit_exact(L, OL, orc, pkg, cfgname, nscheme, nmuscl, nlim):
    """Reconstruction + clamps + face EOS + Roe / Roe-M / AUSMPW+ / AUSM+-up
    (src/rhs_module.f90:314-337)."""
    case = getattr(pkg.cases, cfgname)(10, 8, 8)
    case.params.update(nscheme=nscheme, nmuscl=nmuscl, nlim=nlim)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    cf = cfg_struct(orc, cfg)
    b = case.blocks[0]
    P4 = case.init[0][0].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)
    rng = np.random.default_rng(nscheme * 100 + nmuscl * 10 + nlim)
    for trial in range(150):
        x = _stencil(rng, P4, case.npv, 1e-2 if trial % 2 else 1e-6)
        n3 = rng.standard_normal(3) * 1e-3
        d9 = orc.deteos_simple(x[8, 0] + cfg["pref"], x[8, 4], x[8, 5], x[8, 6])
        d10 = orc.deteos_simple(x[9, 0] + cfg["pref"], x[9, 4], x[9, 5], x[9, 6])
        fx, xl, xr, g3 = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(18)
        L.pc_face(H.p(iv, C.c_int), dp(dvp), H.p(x), dp(n3), H.p(d9[[0, 1, 5]]), H.p(d10[[0, 1, 5]]), dp(fx),
                  dp(xl), dp(xr), dp(g3))
        ol, orr = np.zeros(9), np.zeros(9)
        if nmuscl == 1:
            OL.orc_tvd(C.byref(cf), dp(x), dp(ol), dp(orr))
        elif nmuscl == 2:
            OL.orc_mlp(C.byref(cf), dp(x), dp(ol), dp(orr), None)
        else:
            ol[:], orr[:] = x[8], x[9]
        if nmuscl:
            dl = orc.deteos_simple(ol[0] + cfg["pref"], ol[4], ol[5], ol[6])
            dr = orc.deteos_simple(orr[0] + cfg["pref"], orr[4], orr[5], orr[6])
        else:
            dl, dr = d9, d10
        fo = np.zeros(9)
        sd = np.ascontiguousarray(x[:18, 0])
        if nscheme == 1:
            OL.orc_roe(C.byref(cf), dp(n3), dp(ol), dp(orr), dp(dl), dp(dr), dp(fo))
        elif nscheme == 2:
            OL.orc_roem(C.byref(cf), dp(n3), dp(ol), dp(orr), dp(dl), dp(dr), dp(sd), dp(fo))
        elif nscheme == 3:
            OL.orc_ausmpwp(C.byref(cf), dp(n3), dp(ol), dp(orr), dp(dl), dp(dr), dp(sd), dp(fo))
        else:
            OL.orc_ausmpup(C.byref(cf), dp(n3), dp(ol), dp(orr), dp(dl), dp(dr), dp(fo), None)
        n = case.npv
        assert np.array_equal(xl[:n], ol[:n]) and np.array_equal(xr[:n], orr[:n]), trial
        assert np.array_equal(fx[:n], fo[:n]), (trial, fx[:n], fo[:n])


@pytest.mark.parametrize("turb", [True, False])
def test_viscous_face_bit_exact(L, OL, orc, pkg, turb):
    case = pkg.cases.config2(8, 6, 6) if turb else pkg.cases.config1(8, 6, 6)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    cf = cfg_struct(orc, cfg)
    rng = np.random.default_rng(5)
    for _ in range(200):
        base = np.array([rng.uniform(-3e4, 3e4), rng.uniform(-15, 15), rng.uniform(-3, 3), rng.uniform(-3, 3),
                         rng.uniform(290, 310), 0.0, 0.0, 0.01, 1000.0])
        x = np.ascontiguousarray(np.tile(base, (38, 1)) * (1 + 1e-2 * rng.standard_normal((38, 9))))
        x[:, 5] = rng.choice([0.0, 1e-4])
        x[:, 6] = 0.0
        n3 = rng.standard_normal(3) * 1e-4
        e4 = [rng.standard_normal(3) * 1e-4 for _ in range(4)]
        t4 = [rng.standard_normal(3) * 1e-4 for _ in range(4)]
        grdl = np.array([rng.uniform(1e-9, 1e-6), 0, 0, 0, rng.uniform(1e-4, 1e-2)])
        grdr = np.array([rng.uniform(1e-9, 1e-6), 0, 0, 0, rng.uniform(1e-4, 1e-2)])
        dl, tl2 = orc.deteos_prop(x[8, 0] + cfg["pref"], x[8, 4], x[8, 5], x[8, 6])
        dr, tr2 = orc.deteos_prop(x[9, 0] + cfg["pref"], x[9, 4], x[9, 5], x[9, 6])
        ml, mr = rng.uniform(1e-5, 1e-2, 2)
        tvl, tvr = np.array([tl2[0], tl2[1], ml]), np.array([tr2[0], tr2[1], mr])
        fo, fg = np.zeros(9), np.zeros(9)
        OL.orc_vsflux(C.byref(cf), dp(n3), ptrs(e4), ptrs(t4), dp(grdl), dp(grdr), dp(x), dp(dl), dp(dr),
                      dp(tvl), dp(tvr), dp(fo))
        L.pc_visc(H.p(iv, C.c_int), dp(dvp), dp(x), dp(n3), H.p(np.concatenate(e4)), H.p(np.concatenate(t4)),
                  H.p(np.array([grdl[0], grdl[4]])), H.p(np.array([grdr[0], grdr[4]])),
                  H.p(H.compact(dl, tl2)), H.p(H.compact(dr, tr2)), ml if turb else 0.0, mr if turb else 0.0, dp(fg))
        assert np.array_equal(fg[:case.npv], fo[:case.npv])


@pytest.mark.parametrize("clip,nsteady", [(0, 0), (1, 0), (1, 2), (0, 1)])
def test_cell_routines_bit_exact(L, OL, orc, pkg, clip, nsteady):
    """local dt, split Jacobian product, getx, Merkle, SST source and eddy viscosity."""
    case = pkg.cases.config2(8, 6, 6)
    case.params.update(cav_clip=clip, nsteady=nsteady)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    cf = cfg_struct(orc, cfg)
    ivp = H.p(iv, C.c_int)
    rng = np.random.default_rng(10 + clip + 3 * nsteady)
    for _ in range(200):
        pv = np.array([rng.uniform(-3e4, 3e4), rng.uniform(-15, 15), rng.uniform(-3, 3), rng.uniform(-3, 3),
                       rng.uniform(290, 310), rng.choice([0.0, rng.uniform(0, 1e-3), rng.uniform(0, 0.5)]), 0.0,
                       rng.uniform(1e-4, 0.1), rng.uniform(10, 1e4)])
        dvv, tv2 = orc.deteos_prop(pv[0] + cfg["pref"], pv[4], pv[5], pv[6])
        mut = rng.uniform(1e-5, 1e-2)
        tv = np.array([tv2[0], tv2[1], mut])
        d13 = H.compact(dvv, tv2)
        vol = rng.uniform(1e-9, 1e-6)
        grd = np.array([vol, 0.1, 0.05, 0.02, rng.uniform(1e-4, 1e-2)])
        f6 = [rng.standard_normal(3) * 1e-4 for _ in range(6)]
        f6a = np.concatenate(f6)
        dt = rng.uniform(1e-5, 1e-2)
        assert L.pc_localdt(ivp, dp(dvp), dp(pv), dp(d13), mut, vol, dp(f6a)) == \
            OL.orc_localdt(C.byref(cf), ptrs(f6), dp(grd), dp(pv), dp(dvv), dp(tv))
        cc = rng.standard_normal(4) * 1e5
        tt = rng.standard_normal(4) * 1e2
        res = rng.standard_normal(9)
        xg, xo = np.zeros(9), np.zeros(9)
        L.pc_getx(ivp, dp(dvp), dp(pv), dp(d13), mut, vol, dp(f6a), dt, dp(cc), dp(tt), dp(res), dp(xg))
        OL.orc_getx(C.byref(cf), ptrs(f6), dp(grd), dp(pv), dp(dvv), dp(tv), dt, dp(cc), dp(tt), dp(res), dp(xo))
        assert np.array_equal(xg, xo)
        xs = np.zeros(9)
        L.pc_getx_sel(ivp, dp(dvp), dp(pv), dp(d13), mut, vol, dp(f6a), dt, dp(cc), dp(tt), dp(res), dp(xs))
        assert np.array_equal(xs, xo)
        n3 = rng.standard_normal(3) * 1e-4
        for sg, sgn in ((1, 1.0), (1, -1.0), (-1, 1.0), (-1, -1.0)):
            x = rng.standard_normal(9)
            y0 = rng.standard_normal(9)
            yr, yc = y0.copy(), y0.copy()
            L.pc_jac_rows(ivp, dp(dvp), dp(pv), dp(d13), mut, vol, dp(n3), sg, sgn, dp(x), dp(yr))
            L.pc_jac_coef(ivp, dp(dvp), dp(pv), dp(d13), mut, vol, dp(n3), sg, sgn, dp(x), dp(yc))
            assert np.array_equal(yr[:case.npv], yc[:case.npv])
        for sg in (1, -1):
            x = rng.standard_normal(9)
            y = np.zeros(9)
            L.pc_jac(ivp, dp(dvp), dp(pv), dp(d13), mut, vol, dp(n3), sg, dp(x), dp(y))
            A = np.zeros(81)
            OL.orc_caljac(C.byref(cf), dp(n3), dp(grd), dp(pv), dp(dvv), dp(tv), sg, dp(A))
            Am = A.reshape(9, 9).T
            yo = np.zeros(9)
            for n in range(9):            # the sweep's sequential row sums
                for l in range(9):
                    yo[n] = yo[n] + Am[n, l] * x[l]
            assert np.array_equal(y, yo)
        x7 = np.ascontiguousarray(np.tile(pv, (7, 1)) * (1 + 1e-2 * rng.standard_normal((7, 9))))
        mo = np.zeros(6)
        L.pc_merkle(ivp, dp(dvp), H.p(x7[1]), dp(d13), vol, dp(mo))
        rc, rv, ic = C.c_double(), C.c_double(), np.zeros(4)
        OL.orc_merkle(C.byref(cf), dp(x7), dp(dvv), dp(grd), C.byref(rc), C.byref(rv), dp(ic))
        assert np.array_equal(mo, [rc.value, rv.value, *ic])
        so = np.zeros(7)
        L.pc_sst_source(dp(x7), dp(f6a), vol, grd[4], dp(d13), mut, dp(so))
        src, itt, oc = np.zeros(2), np.zeros(4), C.c_double()
        OL.orc_kwsst_source(ptrs(f6), dp(grd), dp(x7), dp(dvv), dp(tv), dp(src), dp(itt), C.byref(oc))
        assert np.array_equal(so, [*src, *itt, oc.value])
        assert L.pc_sst_eddy(dp(x7), dp(f6a), vol, grd[4], dp(d13)) == \
            OL.orc_kwsst_eddy(ptrs(f6), dp(grd), dp(x7), dp(dvv), dp(tv))


def test_sweep_row_forms_laminar(L, pkg):
    """The branch-free Jacobian rows and the row-selecting getx of the sweeps
    against the row-by-row forms, for the 7-variable laminar state."""
    case = pkg.cases.config1(8, 6, 6)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    ivp = H.p(iv, C.c_int)
    rng = np.random.default_rng(77)
    for _ in range(200):
        pv = np.array([rng.uniform(-3e4, 3e4), rng.uniform(-15, 15), rng.uniform(-3, 3), rng.uniform(-3, 3),
                       rng.uniform(290, 310), 0.0, 0.0, 0.0, 0.0])
        full = np.zeros(18)
        tv2 = np.zeros(2)
        L.pc_mixture(pv[0] + cfg["pref"], pv[4], 0.0, 0.0, dp(full), dp(tv2))
        d13 = H.compact(full, tv2)
        vol = rng.uniform(1e-9, 1e-6)
        f6a = rng.standard_normal(18) * 1e-4
        xg, xs = np.zeros(9), np.zeros(9)
        cc = rng.standard_normal(4) * 1e5
        tt = rng.standard_normal(4) * 1e2
        res = rng.standard_normal(9)
        dt = rng.uniform(1e-5, 1e-2)
        L.pc_getx(ivp, dp(dvp), dp(pv), dp(d13), 0.0, vol, dp(f6a), dt, dp(cc), dp(tt), dp(res), dp(xg))
        L.pc_getx_sel(ivp, dp(dvp), dp(pv), dp(d13), 0.0, vol, dp(f6a), dt, dp(cc), dp(tt), dp(res), dp(xs))
        assert np.array_equal(xg[:7], xs[:7])
        n3 = rng.standard_normal(3) * 1e-4
        for sg, sgn in ((1, 1.0), (-1, -1.0)):
            x = rng.standard_normal(9)
            y0 = rng.standard_normal(9)
            yr, yc = y0.copy(), y0.copy()
            L.pc_jac_rows(ivp, dp(dvp), dp(pv), dp(d13), 0.0, vol, dp(n3), sg, sgn, dp(x), dp(yr))
            L.pc_jac_coef(ivp, dp(dvp), dp(pv), dp(d13), 0.0, vol, dp(n3), sg, sgn, dp(x), dp(yc))
            assert np.array_equal(yr[:7], yc[:7])


def _cav_state(rng, pww, pref, branch, dp=(50.0, 3e3), y1=(1e-4, 0.05)):
    """A cell state on the condensing (p > p_sat) or vaporising side."""
    d = rng.uniform(*dp) * (1 if branch == "cond" else -1)
    return np.array([pww + d - pref, rng.uniform(-10, 10), rng.uniform(-2, 2), rng.uniform(-2, 2),
                     0.0, rng.uniform(*y1), rng.choice([0.0, 1e-7, rng.uniform(0, 1e-6)]),
                     rng.uniform(1e-4, 0.1), rng.uniform(10, 1e4)])


@pytest.mark.parametrize("ncav", [2, 3, 4, 5])
@pytest.mark.parametrize("clip", [0, 1])
def test_cavitation_models_bit_exact(L, OL, orc, pkg, ncav, clip):
    """Kunz, Singhal, Schnerr-Sauer and Lee: the device source compiled for the
    host against the oracle, on both sides of saturation."""
    case = pkg.cases.config2(8, 6, 6)
    case.params.update(ncav=ncav, cav_clip=clip, c_c=3.0, c_v=5.0)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    cf = cfg_struct(orc, cfg)
    ivp = H.p(iv, C.c_int)
    rng = np.random.default_rng(100 + ncav + 10 * clip)
    for n in range(300):
        t = rng.uniform(280.0, 330.0)
        pv = _cav_state(rng, orc.h2o_pww(t), cfg["pref"], "cond" if n % 2 else "vap")
        pv[4] = t
        dvv, _ = orc.deteos_prop(pv[0] + cfg["pref"], t, pv[5], pv[6])
        vol = rng.uniform(1e-9, 1e-6)
        x7 = np.ascontiguousarray(np.tile(pv, (7, 1)))
        grd = np.array([vol, 0.0, 0.0, 0.0, 1e-3])
        mo = np.zeros(6)
        L.pc_cav_model(ivp, dp(dvp), H.p(pv), dp(dvv), vol, dp(mo))
        rc, rv, ic = C.c_double(), C.c_double(), np.zeros(4)
        OL.orc_cavsource(C.byref(cf), dp(x7), dp(dvv), dp(grd), C.byref(rc), C.byref(rv), dp(ic))
        # the rate clip as written (on T) makes r_v negative everywhere, so the
        # vaporising linearisation runs on the condensing side too; where that
        # takes a negative number to the power -0.5 both sides give NaN
        assert np.array_equal(mo, [rc.value, rv.value, *ic], equal_nan=True), (n, mo, rc.value, rv.value, ic)
    for t in (273.15, 300.0, 373.15, 500.0):
        assert L.pc_h2o_sigma(t) == OL.orc_h2o_sigma(t)
    for p in (2e3, 1.01325e5, 1e6, 1e7):
        assert L.pc_sat_t(p) == orc.h2o_tww(p)


# (model, branch) -> the state components whose coupling entry is the exact
# derivative of r_v - r_c (the others drop the saturation-line or
# surface-tension slope, src/cav_module.f90:157-168, 196-207, 240-257, 296-313, 338-354)
_EXACT = {(1, "cond"): (0, 2, 3), (1, "vap"): (0, 2, 3),
          (2, "cond"): (0, 1, 2, 3),
          (3, "cond"): (0, 2, 3), (3, "vap"): (0, 2, 3),
          (4, "cond"): (0, 2, 3), (4, "vap"): (0, 2, 3),
          (5, "cond"): (1, 2, 3), (5, "vap"): (1, 2, 3)}


@pytest.mark.parametrize("ncav,branch", sorted(_EXACT))
def test_cavitation_linearisation_is_the_rate_derivative(OL, orc, pkg, ncav, branch):
    """icav = d(r_v - r_c)/d(p, T, y1, y2) per unit volume, flipped by the
    sign of its y1 entry: central differences with the EOS re-evaluated,
    for every entry the model linearises exactly (Kunz with y2 = 0 and only
    where condensation alone acts)."""
    case = pkg.cases.config2(8, 6, 6)
    k = {1: 0.05, 2: 1e-3, 3: 0.05, 4: 1e-6, 5: 0.05}[ncav]     # rates well inside the clip
    case.params.update(ncav=ncav, cav_clip=1, c_c=k, c_v=k)
    cfg = case.cfg_dict()
    cf = cfg_struct(orc, cfg)
    rng = np.random.default_rng(7 * ncav + (branch == "vap"))

    def rates(pv):
        dvv, _ = orc.deteos_prop(pv[0] + cfg["pref"], pv[4], pv[5], pv[6])
        x7 = np.ascontiguousarray(np.tile(pv, (7, 1)))
        rc, rv, ic = C.c_double(), C.c_double(), np.zeros(4)
        OL.orc_cavsource(C.byref(cf), dp(x7), dp(dvv), dp(np.array([1.0, 0, 0, 0, 1e-3])), C.byref(rc),
                         C.byref(rv), dp(ic))
        return rv.value - rc.value, ic, dvv

    for _ in range(20):
        t = rng.uniform(285.0, 320.0)
        pv = _cav_state(rng, orc.h2o_pww(t), cfg["pref"], branch, dp=(20.0, 300.0), y1=(1e-6, 1e-5))
        pv[4] = t
        if ncav == 2:
            pv[6] = 0.0
        f0, ic, dvv = rates(pv)
        assert f0 != 0.0
        assert (-f0 if branch == "cond" else f0) < dvv[0] * (pv[5] if branch == "cond" else 1 - pv[5] - pv[6])
        deriv = np.zeros(4)
        for q, idx in enumerate((0, 4, 5, 6)):
            h = 1e-6 * max(abs(pv[idx]), (1e3, 1.0, 1e-3, 1e-3)[q])
            a, b = pv.copy(), pv.copy()
            a[idx] += h
            b[idx] -= h
            deriv[q] = (rates(a)[0] - rates(b)[0]) / (2 * h)
        s = 1.0 if deriv[2] >= 0 else -1.0
        for q in _EXACT[(ncav, branch)]:
            assert ic[q] == pytest.approx(s * deriv[q], rel=2e-5, abs=1e-9 * np.abs(deriv).max()), (q, ic, deriv)


@pytest.mark.parametrize("ides", [1, 2, 3, 4])
def test_des_sources_bit_exact(L, OL, orc, pkg, ides):
    """SST source with the DES / DDES / IDDES / S-IDDES length scale
    (src/turbsource_module.f90:416-576): device code compiled for the host
    against the oracle, near and far from the wall."""
    case = pkg.cases.config2(8, 6, 6)
    case.params.update(ides=ides)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    ivp = H.p(iv, C.c_int)
    rng = np.random.default_rng(300 + ides)
    for n in range(300):
        pv = np.array([rng.uniform(-3e4, 3e4), rng.uniform(-15, 15), rng.uniform(-3, 3), rng.uniform(-3, 3),
                       rng.uniform(290, 310), 0.0, 0.0, rng.uniform(1e-4, 0.1), rng.uniform(10, 1e4)])
        dvv, tv2 = orc.deteos_prop(pv[0] + cfg["pref"], pv[4], pv[5], pv[6])
        mut = rng.uniform(1e-6, 1e-2)
        tv = np.array([tv2[0], tv2[1], mut])
        d13 = H.compact(dvv, tv2)
        h = rng.uniform(1e-4, 1e-2)
        vol = h ** 3 * rng.uniform(0.5, 2.0)
        grd = np.array([vol, 0.1, 0.05, 0.02, h * rng.choice([0.05, 0.5, 5.0, 50.0])])
        f6 = [np.eye(3)[q // 2] * h * h * rng.uniform(0.7, 1.3) + rng.standard_normal(3) * 1e-2 * h * h for q in range(6)]
        f6a = np.concatenate(f6)
        x7 = np.ascontiguousarray(np.tile(pv, (7, 1)) * (1 + 1e-2 * rng.standard_normal((7, 9))))
        so = np.zeros(7)
        L.pc_sst_source_des(ivp, dp(dvp), dp(x7), dp(f6a), vol, grd[4], dp(d13), mut, dp(so))
        src, itt, oc = np.zeros(2), np.zeros(4), C.c_double()
        OL.orc_kwsst_source_des(ides, ptrs(f6), dp(grd), dp(x7), dp(dvv), dp(tv), dp(src), dp(itt), C.byref(oc))
        assert np.array_equal(so, [*src, *itt, oc.value]), (n, so, src, itt)


def test_des_length_closed_forms(OL, orc, pkg):
    """A cube of side h: DES destroys k with rho k^1.5 / min(l_RANS, 0.61 h);
    with l_LES above l_RANS, DES and DDES fall back to RANS; far from the
    wall (d > 0.53 h) IDDES and S-IDDES coincide."""
    cfg = pkg.cases.config2(8, 6, 6).cfg_dict()
    rng = np.random.default_rng(5)

    def source(ides, pv, h, d, mut):
        dvv, tv2 = orc.deteos_prop(pv[0] + cfg["pref"], pv[4], pv[5], pv[6])
        tv = np.array([tv2[0], tv2[1], mut])
        grd = np.array([h ** 3, 0, 0, 0, d])
        f6 = [np.eye(3)[q // 2] * h * h for q in range(6)]
        x7 = np.ascontiguousarray(np.tile(pv, (7, 1)) * (1 + 1e-3 * noise))
        out = []
        for i in (0, ides):
            src, itt, oc = np.zeros(2), np.zeros(4), C.c_double()
            OL.orc_kwsst_source_des(i, ptrs(f6), dp(grd), dp(x7), dp(dvv), dp(tv), dp(src), dp(itt), C.byref(oc))
            out.append(src[0])
        return out, dvv[0]

    for _ in range(20):
        noise = rng.standard_normal((7, 9))
        noise[1] = 0.0                      # the cell itself keeps pv (its k, omega set l_RANS)
        pv = np.array([0.0, rng.uniform(-5, 5), 0.5, 0.2, 300.0, 0.0, 0.0, rng.uniform(1e-3, 0.1), rng.uniform(10, 1e3)])
        k, om = pv[7], pv[8]
        l_rans = math.sqrt(k) / 9.0 / om * 100.0
        h = rng.uniform(0.2, 5.0) * l_rans / 0.61          # l_LES on either side of l_RANS
        (s0, s1), rho = source(1, pv, h, 0.1 * h, 1e-4)
        vol = h ** 3
        want = s0 + rho * k ** 1.5 * (1.0 / l_rans - 1.0 / min(l_rans, 0.61 * h)) * vol
        assert s1 == pytest.approx(want, rel=1e-10, abs=1e-12 * abs(s0))
        hbig = 50.0 * l_rans
        for ides in (1, 2):
            (a, b), _ = source(ides, pv, hbig, 5.0 * hbig, 1e-4)
            assert b == pytest.approx(a, rel=1e-12), ides
        (_, c3), _ = source(3, pv, h, 0.6 * h, 1e-4)
        (_, c4), _ = source(4, pv, h, 0.6 * h, 1e-4)
        assert c3 == c4


def _shielded_source(OL, orc, cfg, ides, k, om, grad, h, d, mut):
    """k source (ides 0 and ides) of a cube cell of side h at wall distance d,
    uniform k and omega, and the linear velocity field u = u_c + grad (x - x_c)
    sampled at the six neighbours (the 7-point Green-Gauss gradient of a
    linear field on a uniform cube is exact); returns the two sources, rho
    and mu"""
    pv = np.array([0.0, 0.4, 0.2, -0.1, 300.0, 0.0, 0.0, k, om])
    dvv, tv2 = orc.deteos_prop(pv[0] + cfg["pref"], pv[4], pv[5], pv[6])
    tv = np.array([tv2[0], tv2[1], mut])
    grd = np.array([h ** 3, 0, 0, 0, d])
    f6 = [np.eye(3)[q // 2] * h * h for q in range(6)]
    x7 = np.tile(pv, (7, 1))                  # rows: i-1, cell, i+1, j-1, j+1, k-1, k+1
    for row, axis, sgn in [(0, 0, -1), (2, 0, 1), (3, 1, -1), (4, 1, 1), (5, 2, -1), (6, 2, 1)]:
        x7[row, 1:4] = pv[1:4] + sgn * h * grad[:, axis]
    out = []
    for i in (0, ides):
        src, itt, oc = np.zeros(2), np.zeros(4), C.c_double()
        OL.orc_kwsst_source_des(i, ptrs(f6), dp(grd), dp(x7), dp(dvv), dp(tv), dp(src), dp(itt), C.byref(oc))
        out.append(src[0])
    return out, dvv[0], tv2[0]


def _menter_f1(rho, mu, k, om, d):
    """Menter's blending F1 = tanh(arg1^4) with uniform k and omega (CD_kw at
    its floor 1e-10): arg1 = min(max(sqrt(k)/(beta* omega d), 500 nu/(d^2
    omega)), 4 rho sigma_w2 k/(CD_kw d^2)), beta* = 0.09, sigma_w2 = 0.856"""
    arg1 = min(max(math.sqrt(k) / (0.09 * om * d), 500.0 * mu / rho / (d * d * om)),
               4.0 * rho * 0.856 * k / (1e-10 * d * d))
    return math.tanh(arg1 ** 4)


def _check_length(OL, orc, cfg, ides, k, om, grad, h, d, mut, l_want):
    (s0, s1), rho, _ = _shielded_source(OL, orc, cfg, ides, k, om, grad, h, d, mut)
    l_rans = math.sqrt(k) / (0.09 * om)
    vol = h ** 3
    want = s0 + rho * k ** 1.5 * (1.0 / l_rans - 1.0 / l_want) * vol
    assert s1 == pytest.approx(want, rel=1e-9, abs=1e-11 * rho * k ** 1.5 / min(l_rans, l_want) * vol), (ides, d, h)


def _grad_with_norm(rng, norm):
    g = rng.standard_normal((3, 3))
    return g * (norm / np.linalg.norm(g))


def test_ddes_shielding_closed_form(OL, orc, pkg):
    """DDES (Spalart et al. 2006) with the SST blend of C_DES (0.78 F1 +
    0.61 (1 - F1)) away from its limits: l = l_RANS - f_d max(0, l_RANS -
    C_DES h_max), f_d = 1 - tanh((20 r_d)^3), r_d = (mu + mu_t) / (kappa^2
    d^2 |grad u|) with kappa = 0.41 and |grad u| the Frobenius norm of the
    velocity gradient (sqrt(U_ij U_ij)).  The reference's r_d takes the
    dynamic viscosities mu + mu_t where the paper has nu + nu_t; the pin
    follows the reference.  r_d is drawn so f_d runs over (0, 1)."""
    cfg = pkg.cases.config2(8, 6, 6).cfg_dict()
    rng = np.random.default_rng(41)
    k, om = 1e-2, 100.0
    l_rans = math.sqrt(k) / (0.09 * om)
    seen = []
    for trial in range(24):
        d = l_rans * math.exp(rng.uniform(math.log(0.2), math.log(3.0)))
        h = l_rans * rng.uniform(0.2, 1.4)
        _, rho, mu = _shielded_source(OL, orc, cfg, 0, k, om, np.eye(3), h, d, 0.0)
        mut = mu * math.exp(rng.uniform(math.log(0.5), math.log(50.0)))
        r_d = rng.uniform(0.01, 0.08)
        grad = _grad_with_norm(rng, (mu + mut) / (0.41 ** 2 * d * d * r_d))
        f1 = _menter_f1(rho, mu, k, om, d)
        fd = 1.0 - math.tanh((20.0 * r_d) ** 3)
        l_les = (0.78 * f1 + 0.61 * (1.0 - f1)) * h
        _check_length(OL, orc, cfg, 2, k, om, grad, h, d, mut, l_rans - fd * max(0.0, l_rans - l_les))
        seen.append((fd, f1))
    fds, f1s = np.array(seen).T
    assert fds.min() < 0.05 and fds.max() > 0.95 and np.any((fds > 0.2) & (fds < 0.8))
    assert f1s.min() < 0.2 and f1s.max() > 0.99


def test_iddes_shielding_closed_form(OL, orc, pkg):
    """SST-IDDES (Gritskevich et al. 2012; Shur et al. 2008), ides 3, and its
    form without the elevating function f_e, ides 4:
      l = f~_d (1 + f_e) l_RANS + (1 - f~_d) C_DES Delta,
      Delta = min(C_w max(d, h_max), h_max), C_w = 0.15,
      alpha = 0.25 - d / h_max, f_B = min(2 exp(-9 alpha^2), 1),
      f~_d = max(1 - f_dt, f_B), f_dt = 1 - tanh((20 r_dt)^3),
      f_e = f_e2 max(f_e1 - 1, 0), f_e1 = 2 exp(-11.09 alpha^2) (alpha >= 0)
      or 2 exp(-9 alpha^2), f_e2 = 1 - max(f_t, f_l),
      f_t = tanh((c_t^2 r_dt)^3), f_l = tanh((c_l^2 r_dl)^10), c_t = 1.87, c_l = 5,
      r_dt = mu_t / (kappa^2 d^2 |grad u|), r_dl = mu / (kappa^2 d^2 |grad u|)
    (dynamic viscosities, as in the reference's DDES).  The draws cover both
    signs of alpha, both branches of f~_d and f_l, f_t away from 0 and 1."""
    cfg = pkg.cases.config2(8, 6, 6).cfg_dict()
    rng = np.random.default_rng(42)
    k, om = 1e-2, 100.0
    l_rans = math.sqrt(k) / (0.09 * om)
    seen = []
    for trial in range(32):
        d = l_rans * math.exp(rng.uniform(math.log(0.2), math.log(3.0)))
        h = d / rng.uniform(0.05, 1.2)
        _, rho, mu = _shielded_source(OL, orc, cfg, 0, k, om, np.eye(3), h, d, 0.0)
        r_dt = math.exp(rng.uniform(math.log(0.01), math.log(0.6)))
        r_dl = rng.uniform(0.015, 0.06)
        mut = mu * r_dt / r_dl
        grad = _grad_with_norm(rng, mut / (0.41 ** 2 * d * d * r_dt))
        f1 = _menter_f1(rho, mu, k, om, d)
        cdes = 0.78 * f1 + 0.61 * (1.0 - f1)
        delta = min(0.15 * max(d, h), h)
        alpha = 0.25 - d / h
        fb = min(2.0 * math.exp(-9.0 * alpha * alpha), 1.0)
        fdt = 1.0 - math.tanh((20.0 * r_dt) ** 3)
        fdd = max(1.0 - fdt, fb)
        fe1 = 2.0 * math.exp((-11.09 if alpha >= 0.0 else -9.0) * alpha * alpha)
        ft = math.tanh((1.87 ** 2 * r_dt) ** 3)
        fl = math.tanh((5.0 ** 2 * r_dl) ** 10)
        fe = (1.0 - max(ft, fl)) * max(fe1 - 1.0, 0.0)
        _check_length(OL, orc, cfg, 3, k, om, grad, h, d, mut, fdd * (1.0 + fe) * l_rans + (1.0 - fdd) * cdes * delta)
        _check_length(OL, orc, cfg, 4, k, om, grad, h, d, mut, fdd * l_rans + (1.0 - fdd) * cdes * delta)
        seen.append((alpha, fb > 1.0 - fdt, fl, ft, fe))
    a, fbwin, fl, ft, fe = np.array(seen).T
    assert a.min() < 0.0 < a.max()
    assert 0 < fbwin.sum() < len(fbwin)
    assert fl.min() < 0.1 and fl.max() > 0.9 and ft.max() > 0.3 and fe.max() > 0.05


@pytest.mark.parametrize("nscheme,nmuscl,nlim", [(1, 1, 1), (1, 0, 0), (4, 2, 8), (2, 2, 8)])
def test_fp32_face_within_single_precision(L, OL, orc, pkg, nscheme, nmuscl, nlim):
    """The mixed-precision face (reconstruction, face EOS, flux in FP32 on
    gauge pressure) against the FP64 face: each flux component within 2e-5
    of that component's largest magnitude over the sample.  (AUSMPW+ keeps
    FP64 in the mixed mode: its pressure weight cancels to ~4e-4 in FP32.)"""
    case = pkg.cases.config2(10, 8, 8)
    case.params.update(nscheme=nscheme, nmuscl=nmuscl, nlim=nlim)
    cfg = case.cfg_dict()
    iv, dvp = H.par_arrays(cfg)
    cf = cfg_struct(orc, cfg)
    b = case.blocks[0]
    P4 = case.init[0][0].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)
    rng = np.random.default_rng(900 + nscheme)
    got, want = [], []
    for trial in range(150):
        x = _stencil(rng, P4, case.npv, 1e-2 if trial % 2 else 1e-6)
        n3 = rng.standard_normal(3) * 1e-3
        d9 = orc.deteos_simple(x[8, 0] + cfg["pref"], x[8, 4], x[8, 5], x[8, 6])
        d10 = orc.deteos_simple(x[9, 0] + cfg["pref"], x[9, 4], x[9, 5], x[9, 6])
        f32 = np.zeros(9)
        L.pc_face32(H.p(iv, C.c_int), dp(dvp), H.p(x), dp(n3), H.p(d9[[0, 1, 5]]), H.p(d10[[0, 1, 5]]), dp(f32))
        f64, xl, xr, g3 = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(18)
        L.pc_face(H.p(iv, C.c_int), dp(dvp), H.p(x), dp(n3), H.p(d9[[0, 1, 5]]), H.p(d10[[0, 1, 5]]), dp(f64),
                  dp(xl), dp(xr), dp(g3))
        got.append(f32[:case.npv])
        want.append(f64[:case.npv])
    got, want = np.array(got), np.array(want)
    scale = np.abs(want).max(axis=0)
    zero = scale == 0.0                      # a species absent everywhere
    assert np.all(got[:, zero] == 0.0)
    err = np.abs(got[:, ~zero] - want[:, ~zero]).max(axis=0) / scale[~zero]
    assert np.all(err < 2e-5), err
    assert np.any(got != want)
