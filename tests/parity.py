"""Shared helpers of the GPU parity tests: build a case, run the HIP
library and the oracle side by side from the same state, and compare."""
import numpy as np


def small_cases(pk):
    """Small bump-channel cases covering the scheme switches of the path."""
    bc = pk.cases.bump_case
    return {
        "lam_roe_tvd": lambda: pk.cases.config1(12, 8, 6),
        "lam_roe_mlp": lambda: bc(12, 8, 6, False, False, 1, 2, 8, cfl=5.0, seed=3),
        "lam_ausm_tvd_vanleer": lambda: bc(12, 8, 6, False, False, 4, 1, 2, cfl=5.0, seed=4),
        "sst_cav_ausm_mlp": lambda: _clip(pk.cases.config2(12, 8, 6), 1),
        "sst_cav_roe_first": lambda: _clip(bc(12, 8, 6, True, True, 1, 0, 0, cfl=5.0, seed=5), 1),
        "sst_cav_as_written": lambda: _clip(pk.cases.config2(12, 8, 6), 0),
        "lam_roe_2block": lambda: pk.cases.split_i(pk.cases.config1(12, 8, 6), [7]),
        "lam_roe_bdf2": lambda: bc(12, 8, 6, False, False, 1, 1, 1, cfl=5.0, nsteady=2, dt_phy=2e-4, prec=2, seed=6),
        "sst_cav_bdf1": lambda: _clip(bc(12, 8, 6, True, True, 4, 2, 8, cfl=5.0, nsteady=1, dt_phy=2e-4, prec=2,
                                         seed=8), 1),
        "sst_cav_3block": lambda: _clip(pk.cases.split_i(pk.cases.config2(14, 8, 6), [5, 10]), 1),
        "sst_cav_roem_mlp": lambda: _clip(bc(12, 8, 6, True, True, 2, 2, 8, cfl=3.0, seed=11), 1),
        "lam_ausmpwp_tvd": lambda: bc(12, 8, 6, False, False, 3, 1, 2, cfl=5.0, seed=12),
        "sst_cav_ausmpwp_first": lambda: _clip(bc(12, 8, 6, True, True, 3, 0, 0, cfl=3.0, seed=13), 1),
        "sst_des": lambda: _des(pk.cases.config2(12, 8, 6), 1),
        "sst_ddes": lambda: _des(pk.cases.config2(12, 8, 6), 2),
        "sst_iddes": lambda: _des(pk.cases.config2(12, 8, 6), 3),
        "sst_siddes": lambda: _des(pk.cases.config2(12, 8, 6), 4),
        "sst_kunz": lambda: _model(pk.cases.config2(12, 8, 6), 2),
        "sst_singhal": lambda: _model(pk.cases.config2(12, 8, 6), 3),
        "sst_schnerr_sauer": lambda: _model(pk.cases.config2(12, 8, 6), 4),
        "sst_lee": lambda: _model(pk.cases.config2(12, 8, 6), 5),
        # global-minimum (local = 0) and fixed (local = -1, dt = cfl seconds)
        # pseudo-time steps, src/timestep_module.f90:214-316
        "lam_roe_mintime": lambda: _local(bc(12, 8, 6, False, False, 1, 1, 1, cfl=5.0, seed=14), 0, 5.0),
        "sst_cav_fixedtime": lambda: _local(_clip(pk.cases.config2(12, 8, 6), 1), -1, 2.0e-3),
        # blocks three to five cells thick in i (the fewest a block may have): the sweeps' stage transfers
        # move cells in pairs, clamped at both ends of such short pencils
        "lam_roe_thin_blocks": lambda: pk.cases.split_i(pk.cases.config1(12, 8, 6), [4, 8]),
        "sst_cav_thin_blocks": lambda: _clip(pk.cases.split_i(pk.cases.config2(12, 8, 6), [4, 8]), 1),
    }


def _local(case, local, cfl):
    case.params["local"] = local
    case.params["cfl"] = cfl
    return case


def _des(case, ides):
    """the config-2 state with a DES length scale (species clip), made to
    reach the LES branches: omega lowered 100-fold (the RANS length
    sqrt(k)/(0.09 omega) exceeds the cell size), a shear u += 2 m/s per cell
    across the channel and mu_t = 1e-6 mu (the shielding functions open)"""
    case.params["ides"] = ides
    case.params["cav_clip"] = 1
    for q, (pv, dv, tv, qq) in enumerate(case.init):
        b = case.blocks[q]
        pv = pv.copy().reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)
        pv[..., 8] *= 0.01
        pv[..., 1] += 2.0 * np.arange(b.jmax + 5)[None, :, None]
        if tv is not None:
            tv = tv.copy().reshape(-1, case.ntv)
            tv[:, 2] = 1e-6 * tv[:, 0]
            tv = tv.ravel()
        case.init[q] = (pv.ravel(), dv, tv, qq)
    return case


def _model(case, ncav):
    """the config-2 state with another mass-transfer model (species clip)"""
    case.params["ncav"] = ncav
    case.params["cav_clip"] = 1
    return case


def _clip(case, mode):
    case.params["cav_clip"] = mode
    return case


def pair(pk, orc, case):
    s = pk.Solver(case)
    o = orc.Oracle(case)
    for ib, (pv, dv, tv, qq) in enumerate(case.init):
        s.put_state(ib, pv, dv, tv, qq)
        o.put_state(ib, pv, dv, tv, qq)
    return s, o


def interior(case, ib, a, nv):
    b = case.blocks[ib]
    A = a.reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, nv)
    return A[3:b.kmax + 2, 3:b.jmax + 2, 3:b.imax + 2].reshape(-1, nv)


def rel_err(a, b, floor=1e-8):
    """Largest per-element relative difference of a against b per column,
    with elements below floor * (column max) compared against that floor.
    A non-finite element on either side counts as an infinite difference
    (a cell that goes NaN on both sides is a failure, not an agreement)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
        b = b[:, None]
    bad = ~(np.isfinite(a) & np.isfinite(b))
    bf = np.where(bad, 0.0, b)
    scale = np.max(np.abs(bf), axis=0)
    den = np.maximum(np.abs(bf), floor * np.maximum(scale, 1e-300))
    with np.errstate(invalid="ignore"):
        e = np.abs(np.where(bad, 0.0, a) - bf) / den
    e[bad] = np.inf
    return np.max(e, axis=0)
