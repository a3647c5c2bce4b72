"""The discrete adjoint cav3d_rhs_vjp against central finite differences of
the residual it differentiates: the device residual (cav3d_calrhs with the
cell properties re-derived by the EOS) and the oracle's residual
(src/rhs_module.f90:260-643 restated), along seeded random directions over
every block's ghosted state; the p_ref and u_ref derivatives; determinism;
and the torch.autograd harness."""
import copy
import os
import zlib

import numpy as np
import pytest

from parity import small_cases

pytestmark = pytest.mark.gpu

# d(res_bar . res) along a direction: FD vs adjoint, relative to their size
TOL = 1e-5

# AUSMPW+ (nscheme 3) is piecewise at every face -- minima over nine stencil
# pressure ratios, the larger of two pressure weights, the smaller rho c^2 of
# the two sides -- so at the residual level no step of a random direction is
# reliably free of kinks; its face gradient is checked against differences
# face by face on the host (tests/test_adjoint_host.py, schemes 3-*-*).
AUSMPWP = pytest.mark.skip(reason="AUSMPW+ is piecewise at every face; its face gradient is pinned on the host")

CASES = ["lam_roe_tvd", "lam_roe_mlp", "lam_ausm_tvd_vanleer", "sst_cav_ausm_mlp", "sst_cav_roe_first",
         "lam_roe_2block", "lam_roe_bdf2", "sst_cav_bdf1", "sst_cav_roem_mlp",
         pytest.param("lam_ausmpwp_tvd", marks=AUSMPWP), "sst_ddes", "sst_cav_3block",
         "sst_schnerr_sauer", "sst_lee", "sst_kunz", "sst_singhal"]
# directions (of 8) that must settle per case
NEED = {}


_OPEN = []


@pytest.fixture(autouse=True)
def _close_contexts():
    """Every context a test opens is destroyed when it ends, passed or not
    (a failed test's traceback would otherwise keep its solver alive)."""
    yield
    while _OPEN:
        _OPEN.pop().close()


def prepared(pk, name, steps=2, generic=True, make=None):
    """A case after a few pseudo-time steps, the solver holding it, and its
    frozen parts.  The ghosted state is then made generic: ghost layers that
    copy the interior and the spanwise velocity's exact zeros give limiter
    ratios of zero differences, which the 1e-30 regularisation sets and no
    finite step resolves; every ghosted cell is an input of the map, so any
    values are admissible."""
    case = make() if make is not None else small_cases(pk)[name]()
    s = pk.Solver(case)
    _OPEN.append(s)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(steps):
        s.timeinteg(1, it + 1, 0.0)
    s.setbc()       # ghost layers of the updated state (a step fills them at its start)
    rng = np.random.default_rng(zlib.crc32(("state" + name).encode()))
    pref = case.params["pref"]
    pvs = []
    for ib in range(len(case.blocks)):
        a = s.get_state(ib)[0].reshape(-1, case.npv).copy()
        live = a[:, 4] > 0.0                    # cells the boundary fill has set
        m = live.sum()
        a[live, 0] += 1e-4 * np.abs(a[live, 0] + pref) * rng.standard_normal(m)
        a[live, 1:4] += 1e-3 * rng.standard_normal((m, 3))
        a[live, 4] *= 1.0 + 1e-5 * rng.standard_normal(m)
        a[live, 5] *= 1.0 + 1e-3 * rng.standard_normal(m)
        if case.npv > 7:
            a[live, 8] *= 1.0 + 1e-3 * rng.standard_normal(m)
            # k away from walls only: a wall's ghost layer holds -k (zero face average)
            mk = _k_movable(case, ib, a)
            a[mk, 7] *= 1.0 + 1e-3 * rng.standard_normal(mk.sum())
        pvs.append(a.ravel())
    return case, s, pvs, s.frozen()


def _k_movable(case, ib, a):
    """Cells whose k may move: k > 0 there and in the six neighbours (a wall's
    ghost layer holds -k, so the wall face's average k is zero and any step
    of k beside it leaves the domain of its square root)."""
    b = case.blocks[ib]
    k = a[:, 7].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5)
    ok = k > 0.0
    m = ok.copy()
    for ax in range(3):
        for sh in (1, -1):
            m &= np.roll(ok, sh, axis=ax)
    return m.ravel()


def directions(case, pvs, rng, k):
    """k directions over every ghosted entry, each component scaled to 1e-9
    of its magnitude in that cell (gauge pressure by the absolute one); the
    gas fraction, zero throughout, is not moved, nor k beside a wall.  Under
    AUSMPW+ the pressure is not moved either: its sensor takes minima over
    nine stencil pressure ratios at every face, so nearly every step of the
    pressure crosses a kink somewhere (its pressure bars are checked face by
    face on the host, tests/test_adjoint_host.py).  Under Lee's model (ncav
    5) the temperature is not moved: its differences drift at every step
    down to 1.6e-11 of T (kinks at shorter distances in many cells), while
    every other variable's settle; the model's T derivatives are checked on
    the host against differences across saturation
    (tests/test_adjoint_host.py::test_mass_transfer_gradient)."""
    out = []
    pref = case.params["pref"]
    for _ in range(k):
        d = []
        for ib, pv in enumerate(pvs):
            a = pv.reshape(-1, case.npv)
            sc = 1e-9 * np.abs(a)
            sc[:, 0] = 1e-9 * np.abs(a[:, 0] + pref)
            sc[:, 1:4] = 1e-9 * np.maximum(np.abs(a[:, 1:4]), 1.0)
            sc[:, 6] = 0.0
            if case.params["nscheme"] == 3:
                sc[:, 0] = 0.0
            if case.params.get("ncav") == 5:
                sc[:, 4] = 0.0
            if case.npv > 7:
                sc[~_k_movable(case, ib, a), 7] = 0.0
            d.append((rng.standard_normal(a.shape) * sc).ravel())
        out.append(d)
    return out


def seeds(case, res, rng):
    """res_bar: standard normal per entry over the rms of its component."""
    out = []
    rms = np.sqrt(np.mean(np.concatenate([r.reshape(-1, case.npv) for r in res]) ** 2, axis=0))
    rms = np.where(rms > 0.0, rms, 1.0)
    for r in res:
        out.append((rng.standard_normal(r.reshape(-1, case.npv).shape) / rms).ravel())
    return out


def phi(rb, res):
    return sum(float(np.dot(a, b)) for a, b in zip(rb, res))


# the steps (in units of a direction's length) tried in turn, a factor 2
# apart, shortest first
STEPS = (0.0625, 0.125, 0.25, 0.5, 1.0, 2.0, 4.0)


def kink_aware(diff, adj, steps=STEPS, scale=0.0):
    """Compare an adjoint directional derivative with finite differences
    diff(c) at the steps c in turn, shortest first: the first three
    consecutive steps that agree to 1e-4 (no limiter / clamp / |.| kink
    inside the longest, and rounding well below 1e-4) are used, and the
    adjoint is compared with their median, allowing twice their spread.
    Shortest first: once a step straddles a kink at distance x0, the central
    difference tends to the mean of the two one-sided slopes like x0/c, so
    long steps a factor 2 apart can agree on that wrong limit (Lee's
    T = T_sat(p) switch sits close to most cells of the pocket).  Three
    steps, not two: a
    product that cancels to ~1/100 of a typical one is rounding-limited at
    these steps (its differences scatter by a few 1e-4 with no trend as the
    step shrinks), and two of them can agree by chance.  Agreement and error
    are relative to the larger of the product and `scale` (the median
    product over a set of directions): the rounding of a difference is set
    by the sizes of the residual entries it sums, not by how far they
    cancel.  Returns the relative error, or None when no three steps agree."""
    vals = []
    for c in steps:
        vals.append(diff(c))
        if len(vals) < 3:
            continue
        w = sorted(vals[-3:])
        spread = w[2] - w[0]
        if spread <= 1e-4 * max(abs(w[0]), abs(w[2]), scale):
            miss = abs(w[1] - adj) - 2.0 * spread
            return max(miss, 0.0) / max(abs(w[1]), abs(adj), scale, 1e-300)
    return None


def check_directions(f, pvs, pb, rb, dirs, need, steps=STEPS):
    """Every direction whose differences settle (kink_aware) must match the
    adjoint to TOL, and at least `need` directions must settle."""
    worst, counted = 0.0, 0
    adjs = [sum(float(np.dot(a, v)) for a, v in zip(pb, d)) for d in dirs]
    scale = float(np.median(np.abs(adjs)))
    for d, adj in zip(dirs, adjs):
        def diff(c, d=d):
            # residuals differenced entry by entry before the weighting, so small
            # entries are not lost in the rounding of large ones
            rp = f([p + c * v for p, v in zip(pvs, d)])
            rm = f([p - c * v for p, v in zip(pvs, d)])
            return 0.5 * sum(float(np.dot(w, a - b)) for w, a, b in zip(rb, rp, rm)) / c
        err = kink_aware(diff, adj, steps, scale)
        if err is None:
            print(f"  direction skipped: no three steps agree (adjoint {adj:.9e})", flush=True)
            continue
        counted += 1
        worst = max(worst, err)
        assert err < TOL, (adj, err)
    assert counted >= need, (counted, len(dirs))
    return worst


@pytest.mark.parametrize("name", CASES)
def test_vjp_matches_central_differences(pkg, name):
    print(f"case {name} adjoint vs FD", flush=True)
    case, s, pvs, fr = prepared(pkg, name)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    res0 = s.residual_of(pvs, fr["mut"], fr["qq"])
    rb = seeds(case, res0, rng)
    pb, _ = s.rhs_vjp(rb)
    dirs = directions(case, pvs, rng, 8)
    worst = check_directions(lambda p: s.residual_of(p, fr["mut"], fr["qq"]), pvs, pb, rb, dirs, NEED.get(name, 5))
    print(f"case {name}: worst relative difference {worst:.2e} over {len(dirs)} directions", flush=True)
    s.close()


def test_vjp_sixty_four_directions_config2(pkg):
    """The config-2 physics (SST, Merkle, AUSM+-up, MLP) on 64 directions."""
    print("case sst_cav_ausm_mlp 64 directions", flush=True)
    case, s, pvs, fr = prepared(pkg, "sst_cav_ausm_mlp")
    rng = np.random.default_rng(64)
    res0 = s.residual_of(pvs, fr["mut"], fr["qq"])
    rb = seeds(case, res0, rng)
    pb, _ = s.rhs_vjp(rb)
    check_directions(lambda p: s.residual_of(p, fr["mut"], fr["qq"]), pvs, pb, rb, directions(case, pvs, rng, 64), 48)
    s.close()


@pytest.mark.parametrize("name", ["sst_cav_ausm_mlp", "sst_ddes"])
def test_vjp_eddy_matches_central_differences(pkg, name):
    """cav3d_rhs_vjp_eddy: the map recomputes the interior cells' eddy
    viscosity from the state (cav3d_eddy, src/eddy_module.f90:135-187) and
    holds the ghost cells'; 8 directions over every ghosted entry.  The
    eddy viscosity's share must be visible: the same directions' products
    of the fixed-mu_t adjoint differ from it by more than the tolerance."""
    print(f"case {name} eddy-viscosity adjoint vs FD", flush=True)
    case, s, pvs, fr = prepared(pkg, name)
    rng = np.random.default_rng(zlib.crc32(("eddy" + name).encode()))
    f = lambda p: s.residual_of_eddy(p, fr["mut"], fr["qq"])
    rb = seeds(case, f(pvs), rng)
    pb, _ = s.rhs_vjp(rb, eddy=True)
    f(pvs)
    pb0, _ = s.rhs_vjp(rb)
    dirs = directions(case, pvs, rng, 8)
    worst = check_directions(f, pvs, pb, rb, dirs, 5)
    gap = max(abs(sum(float(np.dot(a - b, v)) for a, b, v in zip(pb, pb0, d)))
              / abs(sum(float(np.dot(a, v)) for a, v in zip(pb, d))) for d in dirs)
    print(f"case {name}: worst relative difference {worst:.2e}; eddy share up to {gap:.2e} of a product", flush=True)
    assert gap > 10 * TOL
    s.close()


def test_vjp_full_size_config2(pkg):
    """The config-2 workload at its benchmark size (128x64x64, 524,288
    cells, SST, Merkle, AUSM+-up, MLP) after two steps: 8 seeded directions,
    each supported on its own random 12x12x12 box of ghosted cells (over the
    whole grid every step of a random direction crosses limiter switches in
    some of the half-million cells, so no three steps agree), at least 6
    settled."""
    print("case config2 128x64x64 adjoint vs FD", flush=True)

    def make():
        c = pkg.cases.config2()
        c.params["cav_clip"] = 1
        return c
    case, s, pvs, fr = prepared(pkg, "config2_128x64x64", make=make)
    rng = np.random.default_rng(128)
    res0 = s.residual_of(pvs, fr["mut"], fr["qq"])
    rb = seeds(case, res0, rng)
    pb, _ = s.rhs_vjp(rb)
    b = case.blocks[0]
    shape = (b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)
    dirs = []
    for d in directions(case, pvs, rng, 8):
        a = d[0].reshape(shape)
        k0, j0, i0 = (int(rng.integers(0, n - 12)) for n in shape[:3])
        m = np.zeros(shape[:3], dtype=bool)
        m[k0:k0 + 12, j0:j0 + 12, i0:i0 + 12] = True
        dirs.append([np.where(m[..., None], a, 0.0).ravel()])
    worst = check_directions(lambda p: s.residual_of(p, fr["mut"], fr["qq"]), pvs, pb, rb, dirs, 6)
    print(f"case config2 128x64x64: worst relative difference {worst:.2e}", flush=True)
    s.close()


@pytest.mark.parametrize("name", ["sst_cav_ausm_mlp", "lam_roe_2block", "sst_cav_roem_mlp", "sst_kunz", "sst_singhal"])
def test_vjp_matches_oracle_differences(pkg, orc, name):
    """Central differences of the oracle's residual (its own flux, limiter,
    source code) along the same kind of directions, every variable moved
    (y1 included, against every equation)."""
    print(f"case {name} adjoint vs oracle FD", flush=True)
    case, s, pvs, fr = prepared(pkg, name)
    rng = np.random.default_rng(7)
    o = orc.Oracle(case)
    pref = case.params["pref"]

    def f_orc(ps):
        out = []
        for ib, pv in enumerate(ps):
            a = pv.reshape(-1, case.npv)
            dv, t2 = pkg.eos_host(a[:, 0] + pref, a[:, 4], a[:, 5], a[:, 6])
            tv = None
            if case.ntv:
                tv = np.zeros((a.shape[0], case.ntv))
                tv[:, :2] = t2
                if case.ntv == 3:
                    tv[:, 2] = fr["mut"][ib]
                tv = tv.ravel()
            o.put_state(ib, np.ascontiguousarray(pv), np.ascontiguousarray(dv.ravel()), tv, fr["qq"][ib])
        return [o.calrhs(ib)[0] for ib in range(len(ps))]

    rb = seeds(case, f_orc(pvs), rng)
    s.residual_of(pvs, fr["mut"], fr["qq"])
    pb, _ = s.rhs_vjp(rb)
    check_directions(f_orc, pvs, pb, rb, directions(case, pvs, rng, 6), 4)
    s.close()


@pytest.mark.parametrize("name", ["sst_cav_ausm_mlp", "lam_roe_tvd", "sst_singhal"])
def test_scalar_derivatives(pkg, name):
    """d(res_bar . res)/d p_ref at fixed gauge state and d/d u_ref against
    central differences over contexts built with the shifted scalar."""
    print(f"case {name} scalar derivatives", flush=True)
    case, s, pvs, fr = prepared(pkg, name)
    rng = np.random.default_rng(11)
    rb = seeds(case, s.residual_of(pvs, fr["mut"], fr["qq"]), rng)
    _, sc = s.rhs_vjp(rb)
    def at(key, v):
        c2 = copy.copy(case)
        c2.params = dict(case.params)
        c2.params[key] = v
        s2 = pkg.Solver(c2)
        try:
            return s2.residual_of(pvs, fr["mut"], fr["qq"])
        finally:
            s2.close()

    def diff(key, v0, h):
        rp, rm = at(key, v0 + h), at(key, v0 - h)
        return sum(float(np.dot(w, a - b)) for w, a, b in zip(rb, rp, rm)) / (2.0 * h)

    pb, _ = s.rhs_vjp(rb)
    for q, key in enumerate(("pref", "uref")):
        v0 = case.params[key]
        h = 1e-7 * abs(v0)
        d1, d2 = diff(key, v0, h), diff(key, v0, h / 4)
        fd = d1
        # the p_ref derivative is a sum over every cell's pressure bar: its
        # cancellation sets the scale of what a difference resolves
        floor = 1e-9 * sum(np.abs(x.reshape(-1, case.npv)[:, 0]).sum() for x in pb) if key == "pref" else 0.0
        assert abs(d1 - d2) <= 1e-4 * abs(d1) + floor, (key, d1, d2)      # no kink inside the step
        miss = max(abs(fd - sc[q]) - 2.0 * abs(d1 - d2), 0.0)
        err = miss / max(abs(fd), abs(sc[q]), floor / TOL, 1e-300)
        print(f"  d/d{key}: adjoint {sc[q]:.6e} FD {fd:.6e} rel {err:.1e}", flush=True)
        assert err < TOL, (key, fd, sc[q])
    s.close()


def test_vjp_is_deterministic_and_linear(pkg):
    case, s, pvs, fr = prepared(pkg, "sst_cav_ausm_mlp")
    rng = np.random.default_rng(3)
    res0 = s.residual_of(pvs, fr["mut"], fr["qq"])
    r1, r2 = seeds(case, res0, rng), seeds(case, res0, rng)
    a1, _ = s.rhs_vjp(r1)
    a1b, _ = s.rhs_vjp(r1)
    assert all(np.array_equal(x, y) for x, y in zip(a1, a1b))
    a2, _ = s.rhs_vjp(r2)
    a12, _ = s.rhs_vjp([x + 2.0 * y for x, y in zip(r1, r2)])
    for x, y, z in zip(a1, a2, a12):
        scale = np.abs(z).max()
        assert np.abs(x + 2.0 * y - z).max() <= 1e-12 * scale
    s.close()


def test_autograd_harness():
    """torch reverse mode through ResidualFunction gives the library's VJP,
    and its gradient of a quadratic objective matches central differences.
    Runs in a child process that imports torch before the library (torch
    carries its own HIP runtime; loading it after the library would put two
    runtimes in this process)."""
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "adjoint_torch_check.py")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0


def _first_order_model(pk, ncav):
    """config-2 physics with first-order reconstruction and mass-transfer
    model ncav (species clip): the faces read no limiter, so the residual is
    differentiable in y1 at cells where y1 = 0 (a limiter over flat data has
    a one-sided derivative that is not linear in the direction)."""
    case = pk.cases.bump_case(12, 8, 6, True, True, 4, 0, 0, cfl=3.0, seed=21 + ncav)
    case.params["ncav"] = ncav
    case.params["cav_clip"] = 1
    return case


@pytest.mark.parametrize("ncav", [1, 2, 3, 5])
def test_vjp_zero_vapour_one_sided(pkg, ncav):
    """Cells the update's clamp left at y1 = 0: there the mass-transfer rate
    and its clip rho*y1 tie, and only y1 >= 0 is reachable.  Directions that
    raise y1 in those cells only, against one-sided forward differences of
    the device residual.  The one-sided difference D(c) carries an O(c) term
    that is large here (the mixture's compressibility is strongly curved in
    y1 near 0: the vapour's share of d(rho)/dp equals the liquid's near
    y1 = 6e-9), so the check uses its Richardson extrapolation
    2 D(c/2) - D(c) at c = 32, 16, 8 times 2e-12 (on the oracle's residual the
    extrapolations agree to a few 1e-6, where D itself moves by 1e-3 over
    the same steps)."""
    print(f"case first-order ncav {ncav}: y1 = 0 cells, one-sided", flush=True)
    case = _first_order_model(pkg, ncav)
    s = pkg.Solver(case)
    _OPEN.append(s)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(3):
        s.timeinteg(1, it + 1, 0.0)
    s.setbc()
    fr = s.frozen()
    pvs = [s.get_state(ib)[0].copy() for ib in range(len(case.blocks))]
    f = lambda p: s.residual_of(p, fr["mut"], fr["qq"])
    rng = np.random.default_rng(90 + ncav)
    r0 = f(pvs)
    rb = seeds(case, r0, rng)
    pb, _ = s.rhs_vjp(rb)
    zero = [(pv.reshape(-1, case.npv)[:, 5] == 0.0) & (pv.reshape(-1, case.npv)[:, 4] > 0.0) for pv in pvs]
    nz = sum(int(z.sum()) for z in zero)
    print(f"  {nz} ghosted cells at y1 = 0", flush=True)
    assert nz > 50
    dirs = []
    for trial in range(4):
        d = []
        for pv, z in zip(pvs, zero):
            a = np.zeros((pv.size // case.npv, case.npv))
            a[z, 5] = 2e-12 * rng.uniform(0.5, 1.5, int(z.sum()))
            d.append(a.ravel())
        dirs.append(d)
    adjs = [sum(float(np.dot(a, v)) for a, v in zip(pb, d)) for d in dirs]
    scale = float(np.median(np.abs(adjs)))
    counted = 0
    for trial, (d, adj) in enumerate(zip(dirs, adjs)):
        memo = {}

        def one_sided(c, d=d, memo=memo):
            if c not in memo:
                rp = f([p + c * v for p, v in zip(pvs, d)])
                memo[c] = sum(float(np.dot(w, a - b)) for w, a, b in zip(rb, rp, r0)) / c
            return memo[c]
        err = kink_aware(lambda c: 2.0 * one_sided(0.5 * c) - one_sided(c), adj, (8.0, 16.0, 32.0), scale)
        print(f"  direction {trial}: adjoint {adj:.9e} rel {err}", flush=True)
        if err is None:
            continue
        counted += 1
        assert err < TOL, (adj, err)
    assert counted >= 3
    s.close()


@pytest.mark.parametrize("name", ["sst_cav_ausm_mlp", "sst_cav_roe_first"])
def test_sigma_and_inflow_speed_gradients(pkg, name):
    """The config-4 scalars: d(res_bar . res)/d sigma and d/d U_in, with the
    ghost layers refilled by the boundary fill, through every run parameter
    they move (p_ref; u_ref; the mass-transfer scales dp_ref, t_ref; the
    boundary state: pressure, velocity, k, omega; the back pressure), as
    par_bar . d(parameters)/d(scalar), against central differences over
    contexts built at the shifted scalar from the same interior state.
    The interior velocity is scaled by 0.95 first: the preconditioning speed
    is max(|u|, u_ref) with u_ref = U_in, and the stepped state has |u|
    within 1e-3 of U_in in most cells, so a step of U_in would cross that
    switch in many of them (d/dU_in is not defined there)."""
    print(f"case {name}: sigma and U_in gradients", flush=True)
    case, s, pvs, fr = prepared(pkg, name)
    nb = len(case.blocks)
    pvs = [s.get_state(ib)[0].copy() for ib in range(nb)]
    for pv in pvs:
        pv.reshape(-1, case.npv)[:, 1:4] *= 0.95     # the ghosts are refilled from it
    rng = np.random.default_rng(zlib.crc32(("scalars" + name).encode()))
    rb = seeds(case, s.residual_after_setbc(pvs, fr["mut"], fr["qq"]), rng)
    par = s.param_vjp(rb)
    for which in ("sigma", "uref"):
        dth = pkg.cases.param_derivatives(case, which)
        adj = float(par @ dth)
        v0 = case.params["bump"][which]

        def diff(c, which=which, v0=v0):
            h = 1e-6 * abs(v0) * c
            out = []
            for sg in (1.0, -1.0):
                s2 = pkg.Solver(pkg.cases.with_param(case, which, v0 + sg * h))
                try:
                    out.append(s2.residual_after_setbc(pvs, fr["mut"], fr["qq"]))
                finally:
                    s2.close()
            return sum(float(np.dot(w, a - b)) for w, a, b in zip(rb, *out)) / (2.0 * h)
        err = kink_aware(diff, adj)
        print(f"  d/d{which}: adjoint {adj:.9e} rel err {err}", flush=True)
        assert err is not None and err < TOL, (which, adj, err)
    s.close()


def _with_key(case, key, value):
    """A copy of the case (same grid and state) with one run parameter of
    cases.PARAM_KEYS set to value; "pref" moves bc_pv[0] with it (the
    boundary state's reference pressure, equal to p_ref in these cases)."""
    c2 = copy.copy(case)
    c2.params = dict(case.params)
    bc = list(c2.params["bc_pv"])
    if key.startswith("bc_pv"):
        bc[int(key[5:])] = value
    else:
        c2.params[key] = value
        if key == "pref":
            bc[0] = value
    c2.params["bc_pv"] = bc
    return c2


@pytest.mark.parametrize("name", ["sst_cav_ausm_mlp", "sst_cav_roe_first"])
def test_parameter_partials(pkg, name):
    """Each run parameter's partial d(res_bar . res)/d theta from
    cav3d_rhs_vjp_params (ghost layers refilled by the boundary fill)
    against central differences over contexts built with that parameter
    shifted, on the state of test_sigma_and_inflow_speed_gradients.  p_ref
    is moved together with bc_pv[0] (their partials added): the fills
    evaluate the ghost EOS at bc_pv[0], so only that combination is the
    derivative of the filled map.  u_ref and the boundary velocity bc_pv[1..3]
    are not checked one by one: the characteristic fill reflects the
    interior velocity about the boundary velocity, so at the inflow faces the
    averaged speed sits at U_in = u_ref, on the switch max(|u|, u_ref) of the
    preconditioning, where each partial alone is one-sided; moved together,
    as U_in moves them, both branches have the same slope
    (test_sigma_and_inflow_speed_gradients)."""
    print(f"case {name}: run-parameter partials", flush=True)
    case, s, _, fr = prepared(pkg, name)
    nb = len(case.blocks)
    pvs = [s.get_state(ib)[0].copy() for ib in range(nb)]
    for pv in pvs:
        pv.reshape(-1, case.npv)[:, 1:4] *= 0.95
    rng = np.random.default_rng(zlib.crc32(("partials" + name).encode()))
    rb = seeds(case, s.residual_after_setbc(pvs, fr["mut"], fr["qq"]), rng)
    par = s.param_vjp(rb)
    theta = pkg.cases.param_vector(case.params)
    bad, settled = [], 0
    for q, key in enumerate(pkg.cases.PARAM_KEYS):
        if key in ("uref", "bc_pv0", "bc_pv1", "bc_pv2", "bc_pv3", "bc_pv5", "bc_pv6") or \
                (key == "t_ref" and case.params["ncav"] == 1):
            continue
        adj_q = float(par[q] + par[4]) if key == "pref" else float(par[q])
        v0 = float(theta[q])
        scale = abs(v0) if v0 != 0.0 else 1.0

        def diff(c, key=key, v0=v0, scale=scale):
            h = 1e-6 * scale * c
            out = []
            for sg in (1.0, -1.0):
                s2 = pkg.Solver(_with_key(case, key, v0 + sg * h))
                try:
                    out.append(s2.residual_after_setbc(pvs, fr["mut"], fr["qq"]))
                finally:
                    s2.close()
            return sum(float(np.dot(w, a - b)) for w, a, b in zip(rb, *out)) / (2.0 * h)
        # floor: 1e-3 of the largest parameter sensitivity |par theta|, in this
        # parameter's units (a partial far below that is rounding-limited)
        err = kink_aware(diff, adj_q, scale=1e-3 * float(np.abs(par * np.abs(theta)).max()) / scale)
        print(f"  d/d{key}: adjoint {adj_q:.9e} rel err {err}", flush=True)
        if err is None:
            continue
        settled += 1
        if err >= TOL:
            bad.append((key, adj_q, err))
    assert not bad, bad
    assert settled >= 4, settled
    s.close()


def test_objective_and_its_gradient(pkg):
    """The config-4 objective, vapour volume + x-force on the bump wall:
    its value against cav3d_integrals, dJ/dpv against central differences
    along directions over the interior cells (p, u, v, w, T, y1), and
    dJ/d(p_ref) against a context built at a shifted p_ref."""
    print("case sst_cav_ausm_mlp objective", flush=True)
    from test_post import flag_surface
    case = small_cases(pkg)["sst_cav_ausm_mlp"]()
    flag_surface(pkg, case, "jmin")
    p = case.params
    s = pkg.Solver(case)
    _OPEN.append(s)
    s.set_grid()
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(2):
        s.timeinteg(1, it + 1, 0.0)
    fr = s.frozen()
    pvs = [s.get_state(0)[0].copy()]
    s.put_with_eos(pvs, fr["mut"], fr["qq"])
    J, pb, dJp = s.objective(1.0, 1.0)
    area = 0.01
    out = s.integrals(p["rhoref"], p["uref"], area, 0.0)
    ref = 2.0 / p["rhoref"] / p["uref"] ** 2 / area
    want = out[0] + (out[2] + out[5]) / ref
    assert abs(J - want) <= 1e-12 * abs(want), (J, want)
    Jv = s.objective(1.0, 0.0)[0]
    assert Jv > 0.0 and abs(Jv - out[0]) <= 1e-12 * out[0]
    b = case.blocks[0]
    live = np.zeros((b.kmax + 5, b.jmax + 5, b.imax + 5), bool)
    live[3:b.kmax + 2, 3:b.jmax + 2, 3:b.imax + 2] = True
    live = live.ravel()
    rng = np.random.default_rng(404)
    a0 = pvs[0].reshape(-1, case.npv)
    counted = 0
    for trial in range(6):
        d = np.zeros_like(a0)
        d[live, 0] = 1e-9 * np.abs(a0[live, 0] + p["pref"]) * rng.standard_normal(live.sum())
        d[live, 1:4] = 1e-9 * np.maximum(np.abs(a0[live, 1:4]), 1.0) * rng.standard_normal((live.sum(), 3))
        d[live, 4] = 1e-9 * a0[live, 4] * rng.standard_normal(live.sum())
        d[live, 5] = 1e-9 * a0[live, 5] * rng.standard_normal(live.sum())
        d = d.ravel()
        adj = float(np.dot(pb[0], d))

        def diff(c, d=d):
            vals = []
            for sg in (1.0, -1.0):
                s.put_with_eos([pvs[0] + sg * c * d], fr["mut"], fr["qq"])
                vals.append(s.objective(1.0, 1.0)[0])
            return (vals[0] - vals[1]) / (2.0 * c)
        err = kink_aware(diff, adj)
        print(f"  direction {trial}: adjoint {adj:.9e} rel {err}", flush=True)
        if err is None:
            continue
        counted += 1
        assert err < TOL, (adj, err)
    assert counted >= 4
    # p_ref at fixed gauge state
    h = 1e-6 * p["pref"]
    vals = []
    for sg in (1.0, -1.0):
        c2 = pkg.cases.with_param(case, "sigma", case.params["bump"]["sigma"])
        c2.params["pref"] = p["pref"] + sg * h
        s2 = pkg.Solver(c2)
        try:
            s2.set_grid()
            s2.put_with_eos(pvs, fr["mut"], fr["qq"])
            vals.append(s2.objective(1.0, 1.0)[0])
        finally:
            s2.close()
    fd = (vals[0] - vals[1]) / (2.0 * h)
    assert abs(fd - dJp) <= 1e-5 * max(abs(fd), abs(dJp)), (fd, dJp)
    s.close()
