"""The NACA 0015 C-grid generator and the oracle on it (CPU): positive,
closed cells in every block; the oracle's metrics equal the host generator's
face vectors and volumes away from the ghost layers; the boundary fill copies
the wake cut under its reversed index map and the aligned i interfaces
unchanged, bit for bit."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def cg(pkg, orc):
    case = pkg.cases.config3(16, 12, 6)
    ms = orc.grid_metrics(case.blocks, case.ngrd)
    for b, m in zip(case.blocks, ms):
        b.cx, b.ex, b.tx, b.grd = m
    return case


def test_cells_positive_and_closed(pkg, cg):
    for b in cg.blocks:
        I, J, K = b.imax, b.jmax, b.kmax
        v, _ = pkg.cases.cell_geometry(b.X)
        assert np.all(v > 0.0)
        cx, ex, tx = pkg.cases.face_vectors(b.X)
        out = (cx[1:I, 1:J, 1:K] - cx[:I - 1, 1:J, 1:K] + ex[1:I, 1:J, 1:K] - ex[1:I, :J - 1, 1:K]
               + tx[1:I, 1:J, 1:K] - tx[1:I, 1:J, :K - 1])
        area = np.abs(cx[1:I, 1:J, 1:K]).sum(-1) + np.abs(ex[1:I, 1:J, 1:K]).sum(-1)
        assert np.all(np.abs(out).max(-1) <= 1e-13 * area)


def test_oracle_metrics_equal_generator(pkg, cg):
    for b in cg.blocks:
        I, J, K = b.imax, b.jmax, b.kmax
        cx, ex, tx = pkg.cases.face_vectors(b.X)
        ocx = b.cx.reshape(K + 1, J + 1, I, 3).transpose(2, 1, 0, 3)
        np.testing.assert_allclose(ocx[:, 1:J, 1:K], cx[:, 1:J, 1:K], rtol=0, atol=1e-15 * np.abs(cx).max())
        v, _ = pkg.cases.cell_geometry(b.X)
        g = b.grd.reshape(K + 1, J + 1, I + 1, 5).transpose(2, 1, 0, 3)
        np.testing.assert_allclose(g[1:I, 1:J, 1:K, 0], v, rtol=1e-13)
        assert np.all(g[1:I, 1:J, 1:K, 4] > 0.0)


def test_boundary_fill_interfaces(pkg, orc, cg):
    o = orc.Oracle(cg)
    for ib, st in enumerate(cg.init):
        o.put_state(ib, *st)
    o.setbc()
    b = cg.blocks[0]
    I, J, K = b.imax, b.jmax, b.kmax
    P = [o.get_state(ib)[0].reshape(K + 5, J + 5, I + 5, 9) for ib in range(4)]
    at = lambda A, i, j: A[3:K + 2, j + 1, i + 1]          # interior k, Fortran (i, j)
    for i in range(2, I + 1):
        for d in range(3):                                  # three ghost layers across the wake cut
            np.testing.assert_array_equal(at(P[0], i, 1 - d), at(P[3], I + 2 - i, 2 + d))
            np.testing.assert_array_equal(at(P[3], i, 1 - d), at(P[0], I + 2 - i, 2 + d))
    for q in range(3):                                      # aligned i interfaces q | q+1
        for j in range(2, J + 1):
            for d in range(3):
                np.testing.assert_array_equal(at(P[q], I + 1 + d, j), at(P[q + 1], 2 + d, j))
                np.testing.assert_array_equal(at(P[q + 1], 1 - d, j), at(P[q], I - d, j))
