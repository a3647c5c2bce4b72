"""The 4-block NACA 0015 C-grid (the config-3 topology at 4 x 16 x 12 x 6
cells) on the HIP library against the oracle: device metrics from the
vertex coordinates (with the aligned i-face interfaces and the wake cut, a
j-face interface with i reversed), the boundary fill (BCFarfield outside,
symmetry at the span ends, the k-omega wall on the section, the wake-cut
halo under its non-identity transform), the residual, three LU-SGS steps,
and the force and vapour integrals.

Bar: metrics bit for bit; the state 1e-11 relative per cell and equation
(elements below 1e-8 of their field's largest magnitude against that
floor); the residual against a floor of 1e-6 of its field; integrals 1e-12."""
import numpy as np
import pytest

from parity import interior, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _pair(pkg, orc, ncav=1):
    case = pkg.cases.config3(16, 12, 6)
    case.params["ncav"] = ncav
    ref = orc.grid_metrics(case.blocks, case.ngrd)
    for b, m in zip(case.blocks, ref):
        b.cx, b.ex, b.tx, b.grd = m
    s = pkg.Solver(case, metrics="device")
    o = orc.Oracle(case)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
        o.put_state(ib, *st)
    return case, s, o, ref


def test_cgrid_device_metrics_match_oracle(pkg, orc):
    print("case config3 4x16x12x6 metrics", flush=True)
    case, s, o, ref = _pair(pkg, orc)
    for ib in range(4):
        for what, a, r in zip(("cx", "ex", "tx", "grd"), s.get_metrics(ib), ref[ib]):
            assert np.array_equal(a, r), (ib, what, np.abs(a - r).max())
    s.close()


def test_cgrid_boundary_fill_matches_oracle(pkg, orc):
    print("case config3 4x16x12x6 boundary fill", flush=True)
    case, s, o, _ = _pair(pkg, orc)
    s.setbc()
    o.setbc()
    for ib in range(4):
        pg, dg, tg, _ = s.get_state(ib)
        po, do, to, _ = o.get_state(ib)
        D_o = do.reshape(-1, 18)
        live = D_o[:, 0] != 0.0
        assert live.sum() > case.blocks[ib].ncell
        e = rel_err(pg.reshape(-1, 9)[live], po.reshape(-1, 9)[live])
        assert np.all(e < TOL), (ib, e)
        e = rel_err(dg.reshape(-1, 18)[live], D_o[live])
        assert np.all(e < TOL), (ib, e)
        e = rel_err(tg.reshape(-1, 3)[live], to.reshape(-1, 3)[live])
        assert np.all(e < TOL), (ib, e)
    # the wake cut: block 0's first ghost layer below the cut holds block 3's
    # first layer above it with i reversed (cell i <-> I + 1 - i + 1)
    b = case.blocks[0]
    I, J, K = b.imax, b.jmax, b.kmax
    P0 = s.get_state(0)[0].reshape(K + 5, J + 5, I + 5, 9)
    P3 = s.get_state(3)[0].reshape(K + 5, J + 5, I + 5, 9)
    for i in range(2, I + 1):
        np.testing.assert_array_equal(P0[3:K + 2, 1 + 1, i + 1], P3[3:K + 2, 2 + 1, (I + 2 - i) + 1])
    s.close()


@pytest.mark.parametrize("ncav", [1, 0])
def test_cgrid_residual_and_steps_match_oracle(pkg, orc, ncav):
    print(f"case config3 4x16x12x6 ncav {ncav}: residual and three steps", flush=True)
    case, s, o, _ = _pair(pkg, orc, ncav)
    s.setbc()
    o.setbc()
    for ib in range(4):
        e = rel_err(s.calrhs(ib)[0].reshape(-1, 9), o.calrhs(ib)[0].reshape(-1, 9), floor=1e-6)
        assert np.all(e < TOL), (ib, e)
    for it in range(3):
        print(f"  iteration {it}", flush=True)
        o.setqres()
        to = o.timeinteg(1, it + 1, 0.0)
        tg = s.timeinteg(1, it + 1, 0.0)
        assert abs(tg - to) <= TOL * abs(to)
        for ib in range(4):
            pg, dg, tvg, _ = s.get_state(ib)
            po, do, tvo, _ = o.get_state(ib)
            for what, g, r, m in (("pv", pg, po, 9), ("dv", dg, do, 18), ("tv", tvg, tvo, 3)):
                e = rel_err(interior(case, ib, g, m), interior(case, ib, r, m))
                assert np.all(e < TOL), (it, ib, what, e)
        l2g, ijkg, rmg = s.residual()
        l2o, ijko, rmo = o.residual()
        assert np.all(rel_err(l2g, l2o) < TOL)
        assert list(ijkg) == list(ijko)
    p = case.params
    got = s.integrals(p["rhoref"], p["uref"], p["area"], np.radians(p["aoa"]))
    ref = o.integrals(p["rhoref"], p["uref"], p["area"], np.radians(p["aoa"]))
    scale = np.abs(ref[2:]).max()
    assert np.all(np.abs(got[2:] - ref[2:]) <= 1e-12 * scale), (got, ref)
    if ncav:
        assert abs(got[0] - ref[0]) <= 1e-12 * max(abs(ref[0]), 1e-300)
    s.close()
