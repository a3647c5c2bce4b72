"""Block interfaces in the oracle (src/bc_module.f90:355-488, 1300-1341):
the interface ghost layers hold the neighbour block's cells under the
connection's index map, and an inviscid residual does not see a cut."""
import numpy as np
import pytest


def ghosted(case, ib, arr, nv):
    b = case.blocks[ib]
    return arr.reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, nv)


def test_connection_descriptor_matches_reference_rules(pkg):
    c = pkg.cases.config1(10, 6, 5)
    s = pkg.cases.split_i(c, [6])
    (d0,), (d1,) = s.blocks[0].conn_desc, s.blocks[1].conn_desc
    # block 0 sends its last three cells i = 4..6 (cells j = 2..7, k = 2..6)
    # to block 1's ghosts -1..1
    assert d0[:13] == [1, 4, 2, 2, 6, 7, 6, -1, 2, 2, 1, 7, 6]
    # block 1 sends i = 2..4 to block 0's ghosts 7..9
    assert d1[:13] == [0, 2, 2, 2, 4, 7, 6, 7, 2, 2, 9, 7, 6]
    assert d0[13:] == d1[13:] == [1, 0, 0, 0, 1, 0, 0, 0, 1]


@pytest.mark.parametrize("cut", [4, 6, 8])
def test_halo_copies_neighbour_cells(pkg, orc, cut):
    c = pkg.cases.config2(10, 6, 5)
    s = pkg.cases.split_i(c, [cut])
    o = orc.Oracle(s)
    for ib, st in enumerate(s.init):
        o.put_state(ib, *st)
    o.setbc()
    b0, b1 = s.blocks
    P0 = ghosted(s, 0, o.get_state(0)[0], s.npv)
    P1 = ghosted(s, 1, o.get_state(1)[0], s.npv)
    I0 = b0.imax
    for ii in (-1, 0, 1):                    # block 1 ghost ii <- block 0 cell I0 - 1 + ii
        assert np.array_equal(P1[3:-3, 3:-3, ii + 1], P0[3:-3, 3:-3, I0 - 1 + ii + 1])
    for ii in (1, 2, 3):                     # block 0 ghost I0 + ii <- block 1 cell 1 + ii
        assert np.array_equal(P0[3:-3, 3:-3, I0 + ii + 1], P1[3:-3, 3:-3, 1 + ii + 1])


def test_inviscid_residual_does_not_see_the_cut(pkg, orc):
    """Euler residual of the split channel equals the single block's, cell by
    cell, away from the wall/symmetry edges."""
    c = pkg.cases.bump_case(10, 6, 5, False, False, 1, 2, 8, seed=7)
    c.params["iturb"] = -3
    c.npv, c.ntv, c.ngrd = 7, 0, 4
    for q, (pv, dv, tv, qq) in enumerate(c.init):
        c.init[q] = (pv, dv, None, qq)
    s = pkg.cases.split_i(c, [6])
    o1 = orc.Oracle(c)
    o1.put_state(0, *c.init[0])
    o1.setbc()
    r1 = o1.calrhs(0)[0].reshape(c.blocks[0].kmax - 1, c.blocks[0].jmax - 1, c.blocks[0].imax - 1, 7)
    o2 = orc.Oracle(s)
    for ib, st in enumerate(s.init):
        o2.put_state(ib, *st)
    o2.setbc()
    parts = [o2.calrhs(ib)[0].reshape(b.kmax - 1, b.jmax - 1, b.imax - 1, 7) for ib, b in enumerate(s.blocks)]
    r2 = np.concatenate(parts, axis=2)
    assert r2.shape == r1.shape
    inner = (slice(1, -1), slice(1, -1), slice(None))
    assert np.array_equal(r2[inner], r1[inner])


@pytest.mark.parametrize("which", ["config1", "config2"])
def test_viscous_residual_does_not_see_the_cut(pkg, orc, which):
    """With the interface exchange of the metrics (ghost volumes, wall
    distances and tangential face vectors taken from the neighbour block), the
    laminar and the SST + Merkle residuals of the cut channel equal the uncut
    block's bit for bit away from the j/k boundary layers, where the
    neighbour's edge ghosts come from the edge fill instead."""
    c = getattr(pkg.cases, which)(14, 7, 5)
    c.params["cav_clip"] = 1
    s = pkg.cases.split_i(c, [6])
    b, npv = c.blocks[0], c.npv
    o1 = orc.Oracle(c)
    o1.put_state(0, *c.init[0])
    o1.setbc()
    r1 = o1.calrhs(0)[0].reshape(b.kmax - 1, b.jmax - 1, b.imax - 1, npv)
    o2 = orc.Oracle(s)
    for ib, st in enumerate(s.init):
        o2.put_state(ib, *st)
    o2.setbc()
    r2 = np.concatenate([o2.calrhs(ib)[0].reshape(q.kmax - 1, q.jmax - 1, q.imax - 1, npv)
                         for ib, q in enumerate(s.blocks)], axis=2)
    inner = (slice(1, -1), slice(1, -1), slice(None))
    assert np.array_equal(r2[inner], r1[inner])
