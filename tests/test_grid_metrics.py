"""Grid metrics from vertex coordinates (t_grid calnormal / calvolume /
calydns, src/grid_module.f90:350-1145): the oracle's C restatement against
the host generator, closed forms, the closed-surface identity, brute-force
wall distance, and the interface exchange against the uncut block."""
import numpy as np
import pytest


def metrics(orc, blocks, ngrd):
    return orc.grid_metrics(blocks, ngrd)


def shaped(b, cx, ex, tx, grd, ngrd):
    I, J, K = b.imax, b.jmax, b.kmax
    return (cx.reshape(K + 1, J + 1, I, 3), ex.reshape(K + 1, J, I + 1, 3), tx.reshape(K, J + 1, I + 1, 3),
            grd.reshape(K + 1, J + 1, I + 1, ngrd))


@pytest.mark.parametrize("which", ["config1", "config2"])
def test_oracle_matches_host_generator(pkg, orc, which):
    """Two independent implementations: vectorised numpy (k-d tree wall
    distance) and the loop-by-loop C restatement, bit for bit."""
    c = getattr(pkg.cases, which)(12, 8, 6)
    (b,) = c.blocks
    got = metrics(orc, c.blocks, c.ngrd)[0]
    for name, a, ref in zip(("cx", "ex", "tx", "grd"), got, (b.cx, b.ex, b.tx, b.grd)):
        assert np.array_equal(a, ref), name


def affine_block(pkg, A, x0, n=(5, 4, 6)):
    I, J, K = (m + 1 for m in n)
    idx = np.stack(np.meshgrid(np.arange(I), np.arange(J), np.arange(K), indexing="ij"), axis=-1).astype(float)
    X = idx @ np.asarray(A, float).T + np.asarray(x0, float)
    bcs = [pkg.cases.face_bc("wall", f, I, J, K) for f in ("imin", "imax", "jmin", "jmax", "kmin", "kmax")]
    return pkg.cases.Block(I, J, K, X, bcs, [])


def test_closed_forms_on_an_affine_grid(pkg, orc):
    """Cells of x = A (i,j,k) + x0 are congruent parallelepipeds: face
    vectors are the cross products of the edge vectors, volumes |det A|,
    centroids the images of the index centres."""
    A = np.array([[0.3, 0.05, -0.02], [0.01, 0.2, 0.04], [-0.03, 0.02, 0.25]])
    b = affine_block(pkg, A, [1.0, -2.0, 0.5])
    cx, ex, tx, grd = shaped(b, *metrics(orc, [b], 5)[0], 5)
    a0, a1, a2 = A[:, 0], A[:, 1], A[:, 2]
    tol = dict(rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(cx[1:-1, 1:-1, :], np.broadcast_to(np.cross(a1, a2), cx[1:-1, 1:-1, :].shape), **tol)
    np.testing.assert_allclose(ex[1:-1, :, 1:-1], np.broadcast_to(np.cross(a2, a0), ex[1:-1, :, 1:-1].shape), **tol)
    np.testing.assert_allclose(tx[:, 1:-1, 1:-1], np.broadcast_to(np.cross(a0, a1), tx[:, 1:-1, 1:-1].shape), **tol)
    vol = grd[..., 0]
    np.testing.assert_allclose(vol, abs(np.linalg.det(A)), rtol=1e-13)
    I, J, K = b.imax, b.jmax, b.kmax
    ii, jj, kk = np.meshgrid(np.arange(1, I + 2), np.arange(1, J + 2), np.arange(1, K + 2), indexing="ij")
    # cell (i,j,k) of the reference spans CGNS vertices i-1..i: centre at index i - 1.5
    centre = np.stack([ii - 1.5, jj - 1.5, kk - 1.5], axis=-1) @ A.T + [1.0, -2.0, 0.5]
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(grd[..., 1:4][inner], centre.transpose(2, 1, 0, 3)[inner], rtol=1e-13, atol=1e-14)


def test_face_vectors_close_every_cell(pkg, orc):
    """Outward face vectors of a hexahedron sum to zero: cx(i)-cx(i-1)+ex(j)-
    ex(j-1)+tx(k)-tx(k-1) = 0 for every interior cell of the curved channel
    (an index or sign slip in any of the three loops breaks it)."""
    c = pkg.cases.config2(16, 10, 6)
    (b,) = c.blocks
    cx, ex, tx, grd = shaped(b, *metrics(orc, c.blocks, c.ngrd)[0], c.ngrd)
    I, J, K = b.imax, b.jmax, b.kmax
    # interior cells i,j,k = 2..I etc.: python index i-1
    s = (cx[1:K, 1:J, 1:I] - cx[1:K, 1:J, 0:I - 1]
         + ex[1:K, 1:J, 1:I] - ex[1:K, 0:J - 1, 1:I]
         + tx[1:K, 1:J, 1:I] - tx[0:K - 1, 1:J, 1:I])
    scale = np.abs(cx).max()
    assert np.abs(s).max() < 1e-14 * scale
    assert (grd[1:K, 1:J, 1:I, 0] > 0).all()


def test_wall_distance_is_the_brute_force_minimum(pkg, orc):
    c = pkg.cases.config2(12, 8, 6)
    (b,) = c.blocks
    *_, grd = shaped(b, *metrics(orc, c.blocks, c.ngrd)[0], c.ngrd)
    w = pkg.cases.wall_face_centres(b.X, b.bc_desc)
    I, J, K = b.imax, b.jmax, b.kmax
    xc = grd[1:K, 1:J, 1:I, 1:4].reshape(-1, 3)
    d = np.sqrt(((xc[:, None, 0] - w[None, :, 0]) ** 2 + (xc[:, None, 1] - w[None, :, 1]) ** 2)
                + (xc[:, None, 2] - w[None, :, 2]) ** 2)
    assert np.array_equal(grd[1:K, 1:J, 1:I, 4].ravel(), np.minimum(d.min(axis=1), 1.0e6))
    # no wall anywhere: the reference's initial value
    b2 = pkg.cases.Block(b.imax, b.jmax, b.kmax, b.X, [d_ for d_ in b.bc_desc if d_[0] != pkg.cases.BC["wall"]], [])
    *_, g2 = shaped(b2, *metrics(orc, [b2], 5)[0], 5)
    assert (g2[..., 4] == 1.0e6).all()


@pytest.mark.parametrize("cuts", [[5], [4, 9]])
def test_cut_blocks_see_the_uncut_metrics(pkg, orc, cuts):
    """After the exchanges, each block of an i-cut channel holds the uncut
    block's values in its interior and, where the reference exchanges them,
    in its interface ghost layer: grd(1), grd(5) over cells j, k = 2..J, 2..K;
    ex, tx over j = 1..J, k = 1..K."""
    c = pkg.cases.config2(14, 7, 5)
    (b,) = c.blocks
    full = shaped(b, *metrics(orc, c.blocks, c.ngrd)[0], c.ngrd)
    s = pkg.cases.split_i(c, cuts)
    parts = metrics(orc, s.blocks, s.ngrd)
    bounds = [1] + cuts + [b.imax]
    J, K = b.jmax, b.kmax
    for q, sb in enumerate(s.blocks):
        cx, ex, tx, grd = shaped(sb, *parts[q], s.ngrd)
        lo = bounds[q]            # block vertex 1 = uncut vertex lo
        Ib = sb.imax
        # interior cells i = 2..Ib  <->  uncut cells lo+1..lo+Ib-1 (python index +1 offset)
        np.testing.assert_array_equal(grd[1:K, 1:J, 1:Ib], full[3][1:K, 1:J, lo:lo + Ib - 1])
        np.testing.assert_array_equal(ex[1:K, :, 1:Ib], full[1][1:K, :, lo:lo + Ib - 1])
        for side, gi, ui in (("lo", 0, lo - 1), ("hi", Ib, lo + Ib - 1)):
            if (side == "lo" and q == 0) or (side == "hi" and q == len(s.blocks) - 1):
                continue
            for comp in (0, 4):
                np.testing.assert_array_equal(grd[1:K, 1:J, gi, comp], full[3][1:K, 1:J, ui, comp])
            np.testing.assert_array_equal(ex[:K, :J, gi], full[1][:K, :J, ui])
            np.testing.assert_array_equal(tx[:K, :J, gi], full[2][:K, :J, ui])


def flipped_pair(pkg, c, cut):
    """The channel cut at vertex `cut`, the second block with its j and k
    index directions reversed (CGNS transform [1, -2, -3])."""
    (b,) = c.blocks
    I, J, K = b.imax, b.jmax, b.kmax
    kinds = {}
    for d in b.bc_desc:
        s, e = d[1:4], d[4:7]
        for f, (ax, v) in {"imin": (0, 1), "imax": (0, I + 1), "jmin": (1, 1), "jmax": (1, J + 1),
                           "kmin": (2, 1), "kmax": (2, K + 1)}.items():
            if s[ax] == e[ax] == v:
                kinds[f] = d[0]
    inv = {v: k for k, v in pkg.cases.BC.items()}
    X0 = b.X[:cut]
    X1 = b.X[cut - 1:, ::-1, ::-1]
    I0, I1 = cut, I - cut + 1
    swap = {"jmin": "jmax", "jmax": "jmin", "kmin": "kmax", "kmax": "kmin"}
    b0 = pkg.cases.Block(I0, J, K, X0, [pkg.cases.face_bc(inv[k], f, I0, J, K) for f, k in kinds.items() if f != "imax"], [])
    b1 = pkg.cases.Block(I1, J, K, np.ascontiguousarray(X1),
                         [pkg.cases.face_bc(inv[k], swap.get(f, f), I1, J, K) for f, k in kinds.items() if f != "imin"], [])
    b0.conn_desc.append(pkg.cases.connection([I0, 1, 1, I0, J, K], [1, J, K, 1, 1, 1], [1, -2, -3], 1))
    b1.conn_desc.append(pkg.cases.connection([1, 1, 1, 1, J, K], [I0, J, K, I0, 1, 1], [1, -2, -3], 0))
    return b0, b1


def test_reversed_interface_orientation(pkg, orc):
    """A connection with reversed tangential directions: cell quantities
    land on the mirrored cells, ex on the mirrored vertex lines with the sign
    of the reversed j direction; tx follows the reference's vertex-range
    mapping, one cell off the mirror image in j."""
    c = pkg.cases.config2(12, 7, 5)
    (b,) = c.blocks
    J, K = b.jmax, b.kmax
    full = shaped(b, *metrics(orc, c.blocks, c.ngrd)[0], c.ngrd)
    cut = 6
    b0, b1 = flipped_pair(pkg, c, cut)
    m1 = shaped(b1, *metrics(orc, [b0, b1], c.ngrd)[1], c.ngrd)
    cx, ex, tx, grd = m1
    jj = np.arange(2, J + 1)          # block-1 cells j' <-> uncut cells J+2-j'
    kk = np.arange(2, K + 1)
    for comp in (0, 4):
        got = grd[kk - 1][:, jj - 1, 0, comp]
        want = full[3][(K + 2 - kk) - 1][:, (J + 2 - jj) - 1, cut - 1, comp]
        np.testing.assert_array_equal(got, want)
    # interior volumes mirror too, up to the rounding of the reordered pyramids
    np.testing.assert_allclose(grd[kk - 1][:, jj - 1, 1, 0], full[3][(K + 2 - kk) - 1][:, (J + 2 - jj) - 1, cut, 0],
                               rtol=1e-14)
    jv = np.arange(1, J + 1)          # vertex lines j' <-> J+1-j'
    kv = np.arange(1, K + 1)
    np.testing.assert_array_equal(ex[kv - 1][:, jv - 1, 0], -full[1][(K + 1 - kv) - 1][:, (J + 1 - jv) - 1, cut - 1])
    np.testing.assert_array_equal(tx[kv - 1][:, jv - 1, 0], -full[2][(K + 1 - kv) - 1][:, (J + 1 - jv) - 1, cut - 1])
