"""HIP library against the oracle on the same inputs (FP64 path).

Bar: 1e-11 relative per cell and equation (elements below 1e-8 of their
field's largest magnitude are compared against that floor); the argmax
cell of the residual is bit-exact."""
import numpy as np
import pytest

from parity import interior, pair, rel_err, small_cases

pytestmark = pytest.mark.gpu

TOL = 1e-11
NAMES = ["lam_roe_tvd", "lam_roe_mlp", "lam_ausm_tvd_vanleer", "sst_cav_ausm_mlp", "sst_cav_roe_first",
         "sst_cav_as_written", "lam_roe_2block", "sst_cav_3block", "lam_roe_bdf2", "sst_cav_bdf1",
         "sst_kunz", "sst_singhal", "sst_schnerr_sauer", "sst_lee", "sst_cav_roem_mlp", "lam_ausmpwp_tvd",
         "sst_cav_ausmpwp_first", "sst_des", "sst_ddes", "sst_iddes", "sst_siddes", "lam_roe_mintime",
         "sst_cav_fixedtime", "lam_roe_thin_blocks", "sst_cav_thin_blocks"]


def _case(pkg, name):
    print(f"case {name}", flush=True)
    return small_cases(pkg)[name]()


def test_device_eos_matches_oracle(pkg, orc):
    case = pkg.cases.config1(8, 6, 6)
    s = pkg.Solver(case)
    rng = np.random.default_rng(7)
    n = 4000
    p = rng.uniform(2e3, 5e6, n)
    t = rng.uniform(275.0, 500.0, n)
    y1 = np.where(rng.random(n) < 0.5, 0.0, rng.uniform(0.0, 1.0, n))
    y2 = np.where(rng.random(n) < 0.7, 0.0, rng.uniform(0.0, 1.0, n) * (1.0 - y1))
    dv, tv = s.eos(p, t, y1, y2)
    ref = np.array([orc.deteos_prop(*q) for q in zip(p, t, y1, y2)], dtype=object)
    dref = np.stack([r[0] for r in ref])
    tref = np.stack([r[1] for r in ref])
    assert np.all(rel_err(dv, dref, floor=1e-14) < TOL)
    assert np.all(rel_err(tv, tref, floor=1e-14) < TOL)
    s.close()


@pytest.mark.parametrize("name", NAMES)
def test_setbc_fills_ghosts_like_oracle(pkg, orc, name):
    case = _case(pkg, name)
    s, o = pair(pkg, orc, case)
    s.setbc()
    o.setbc()
    npv = case.npv
    for ib in range(len(case.blocks)):
        pg, dg, tg, _ = s.get_state(ib)
        po, do, to, _ = o.get_state(ib)
        P_g, P_o = pg.reshape(-1, npv), po.reshape(-1, npv)
        D_o = do.reshape(-1, 18)
        set_cells = D_o[:, 0] != 0.0          # cells the reference fills (interior + ghosts)
        assert set_cells.sum() > case.blocks[ib].ncell
        assert np.all(rel_err(P_g[set_cells], P_o[set_cells]) < TOL), rel_err(P_g[set_cells], P_o[set_cells])
        assert np.all(P_g[~set_cells] == 0.0)
        D_g = dg.reshape(-1, 18)
        e = rel_err(D_g[set_cells], D_o[set_cells])
        assert np.all(e < TOL), e
        if case.ntv:
            T_g, T_o = tg.reshape(-1, case.ntv), to.reshape(-1, case.ntv)
            e = rel_err(T_g[set_cells], T_o[set_cells])
            assert np.all(e < TOL), e
    s.close()


@pytest.mark.parametrize("name", NAMES)
def test_calrhs_and_timestep_match_oracle(pkg, orc, name):
    case = _case(pkg, name)
    s, o = pair(pkg, orc, case)
    s.setbc()
    o.setbc()
    for ib in range(len(case.blocks)):
        rg, icg, itg, ocg = s.calrhs(ib)
        ro, ico, ito, oco = o.calrhs(ib)
        npv = case.npv
        e = rel_err(rg.reshape(-1, npv), ro.reshape(-1, npv), floor=1e-6)
        assert np.all(e < TOL), e
        if case.params["ncav"]:
            e = rel_err(icg.reshape(-1, 4), ico.reshape(-1, 4))
            assert np.all(e < TOL), e
        if case.npv == 9:
            e = rel_err(itg.reshape(-1, 4), ito.reshape(-1, 4))
            assert np.all(e < TOL), e
            assert np.all(rel_err(ocg, oco) < TOL)
        dg = s.timestep(ib)
        o.timestep(1, 1, 0.0)
        do = o.get_dt(ib)
        assert np.all(rel_err(dg, do) < TOL), rel_err(dg, do)
    s.close()


@pytest.mark.parametrize("name", NAMES)
def test_timeinteg_matches_oracle(pkg, orc, name):
    case = _case(pkg, name)
    s, o = pair(pkg, orc, case)
    npv = case.npv
    nit = 1 if name == "sst_cav_as_written" else 3
    for it in range(nit):
        print(f"  iteration {it}", flush=True)
        o.setqres()
        o.timeinteg(1, it + 1, 0.0)
        s.timeinteg(1, it + 1, 0.0)
        for ib in range(len(case.blocks)):
            pg, dg, tg, _ = s.get_state(ib)
            po, do, to, _ = o.get_state(ib)
            e = rel_err(interior(case, ib, pg, npv), interior(case, ib, po, npv))
            assert np.all(e < TOL), (it, e)
            e = rel_err(interior(case, ib, dg, 18), interior(case, ib, do, 18))
            assert np.all(e < TOL), (it, e)
            if case.ntv:
                e = rel_err(interior(case, ib, tg, case.ntv), interior(case, ib, to, case.ntv))
                assert np.all(e < TOL), (it, e)
        l2g, ijkg, rmg = s.residual()
        l2o, ijko, rmo = o.residual()
        assert np.all(rel_err(l2g, l2o) < TOL)
        assert list(ijkg) == list(ijko)
        assert abs(rmg - rmo) <= TOL * abs(rmo)
    s.close()


@pytest.mark.parametrize("name", ["lam_roe_bdf2", "sst_cav_bdf1"])
def test_dual_time_shift_matches_oracle(pkg, orc, name):
    """Two physical steps of dual time: inner iterations, the qq shift
    (src/solve_module.f90:155-183), more inner iterations."""
    case = _case(pkg, name)
    s, o = pair(pkg, orc, case)
    npv = case.npv
    for step in (1, 2):
        for it in range(2):
            tg = s.timeinteg(step, it + 1, 0.0)
            o.setqres()
            to = o.timeinteg(step, it + 1, 0.0)
            assert tg == to
        for ib in range(len(case.blocks)):
            pg, _, _, qg = s.get_state(ib)
            po, _, _, qo = o.get_state(ib)
            assert np.all(rel_err(interior(case, ib, pg, npv), interior(case, ib, po, npv)) < TOL)
            assert np.all(rel_err(qg.reshape(-1, npv), qo.reshape(-1, npv)) < TOL)
        s.unsteady_shift()
        o.unsteady_shift()
    s.close()


@pytest.mark.parametrize("name,nj", [("lam_roe_tvd", 44), ("sst_cav_ausm_mlp", 44), ("lam_roe_tvd", 72)])
def test_timeinteg_many_tiles_matches_oracle(pkg, orc, name, nj):
    """A cross-section of 43 x 43 cells: 121 sweep tiles in flight at once,
    most of them waiting on two upstream tiles; 71 x 71 cells: 324 tiles, more
    than one per workgroup of the sweep's grid on a 256-CU device."""
    case = pkg.cases.config1(20, nj, nj) if name == "lam_roe_tvd" else pkg.cases.config2(20, nj, nj)
    if name != "lam_roe_tvd":
        case.params["cav_clip"] = 1
    s, o = pair(pkg, orc, case)
    npv = case.npv
    for it in range(2):
        print(f"  iteration {it}", flush=True)
        o.setqres()
        o.timeinteg(1, it + 1, 0.0)
        s.timeinteg(1, it + 1, 0.0)
        pg = interior(case, 0, s.get_state(0)[0], npv)
        po = interior(case, 0, o.get_state(0)[0], npv)
        e = rel_err(pg, po)
        if not np.all(e < TOL):
            bad = np.argmax(np.max(np.abs(pg - po) / np.maximum(np.abs(po), 1e-300), axis=1))
            b = case.blocks[0]
            ni, nj = b.imax - 1, b.jmax - 1
            raise AssertionError(f"iteration {it}: rel err {e}; worst cell i={2 + bad % ni} "
                                 f"j={2 + (bad // ni) % nj} k={2 + bad // (ni * nj)}")
    s.close()
