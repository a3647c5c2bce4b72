"""Post integrals restated in the oracle (vof fields times volume,
src/postvariable_module.f90:282-295; clcd_writing :507-633) against closed
forms: a uniform vapour fraction, a uniform gauge pressure on the bump wall,
and laminar shear of a uniform stream over a flat wall."""
import math

import numpy as np
import pytest


def uniform_state(orc, case, p_gauge, u, t, y1, y2=0.0):
    """Ghosted pv/dv/tv of one block holding one state everywhere."""
    (b,) = case.blocks
    p = case.params
    pv = np.zeros(case.npv)
    pv[:7] = [p_gauge, u[0], u[1], u[2], t, y1, y2]
    if case.npv == 9:
        pv[7:] = [p["kref"], p["oref"]]
    dv, tv = orc.deteos_prop(p_gauge + p["pref"], t, y1, y2)
    tvf = np.zeros(case.ntv)
    tvf[:2] = tv[:2]
    if case.ntv == 3:
        tvf[2] = p["emutref"]
    ng = b.nghost
    return (np.ascontiguousarray(np.tile(pv, ng)), np.ascontiguousarray(np.tile(dv, ng)),
            np.ascontiguousarray(np.tile(tvf, ng)) if case.ntv else None, dv, tvf)


def flag_surface(pkg, case, face="jmin"):
    (b,) = case.blocks
    I, J, K = b.imax, b.jmax, b.kmax
    ax, v = {"imin": (0, 1), "imax": (0, I + 1), "jmin": (1, 1), "jmax": (1, J + 1),
             "kmin": (2, 1), "kmax": (2, K + 1)}[face]
    for d in b.bc_desc:
        if d[1 + ax] == d[4 + ax] == v:
            d[0] |= pkg.solver.SURFACE


def test_vapour_volume_of_a_uniform_fraction(pkg, orc):
    """rho*y1/rho_v is the vapour volume fraction of the mixture: with y1 set
    from alpha, the integral is alpha times the block volume."""
    c = pkg.cases.config2(8, 6, 5)
    p = c.params
    alpha = 0.3
    dv0, _ = orc.deteos_prop(p["pref"], 300.0, 0.0, 0.0)
    rl, rv = dv0[2], dv0[3]
    y1 = alpha * rv / (alpha * rv + (1.0 - alpha) * rl)
    pv, dv, tv, *_ = uniform_state(orc, c, 0.0, (0.0, 0.0, 0.0), 300.0, y1)
    o = orc.Oracle(c)
    o.put_state(0, pv, dv, tv)
    out = o.integrals(p["rhoref"], p["uref"], 1.0)
    (b,) = c.blocks
    vol = b.grd.reshape(b.kmax + 1, b.jmax + 1, b.imax + 1, c.ngrd)[1:-1, 1:-1, 1:-1, 0].sum()
    assert out[0] == pytest.approx(alpha * vol, rel=1e-12)
    assert out[1] == 0.0
    assert np.all(out[2:] == 0.0)          # no surface flagged


def test_uniform_pressure_on_the_bump_wall(pkg, orc):
    """-p0 times the summed face vectors of the floor: the y component is the
    projected area L*W, the bump's x components cancel."""
    c = pkg.cases.config2(12, 6, 5)
    p = c.params
    flag_surface(pkg, c, "jmin")
    p0 = 750.0
    pv, dv, tv, *_ = uniform_state(orc, c, p0, (0.0, 0.0, 0.0), 300.0, 0.0)
    o = orc.Oracle(c)
    o.put_state(0, pv, dv, tv)
    area = 0.01
    out = o.integrals(p["rhoref"], p["uref"], area)
    ref = 2.0 / p["rhoref"] / p["uref"] ** 2 / area
    H = pkg.cases.H
    assert out[3] == pytest.approx(-p0 * (4 * H) * H * ref, rel=1e-12)
    assert abs(out[2]) < 1e-12 * abs(out[3]) and out[4] == 0.0
    assert np.all(out[5:8] == 0.0)          # fluid at rest: no shear
    assert out[8] == out[3] and out[9] == out[2]
    out90 = o.integrals(p["rhoref"], p["uref"], area, math.pi / 2)
    assert out90[8] == pytest.approx(-out[2], abs=1e-12 * abs(out[3]))
    assert out90[9] == pytest.approx(out[3], rel=1e-12)


def test_laminar_shear_of_a_uniform_stream(pkg, orc):
    """Flat channel, u = (U,0,0): each wall face contributes mu U / (dy1/2)
    times its area, dy1 the first cell height."""
    c1 = pkg.cases.config1(8, 6, 5)
    b = pkg.cases.bump_channel(8, 6, 5, bump=0.0)
    pkg.cases.finish_block(b, c1.ngrd)
    c = pkg.cases.Case("flat", dict(c1.params), [b], c1.npv, c1.ntv, c1.ngrd)
    flag_surface(pkg, c, "jmin")
    p = c.params
    U = 2.5
    pv, dv, tv, dv1, tv1 = uniform_state(orc, c, 0.0, (U, 0.0, 0.0), 300.0, 0.0)
    o = orc.Oracle(c)
    o.put_state(0, pv, dv, tv)
    out = o.integrals(p["rhoref"], p["uref"], 1.0)
    ref = 2.0 / p["rhoref"] / p["uref"] ** 2
    H = pkg.cases.H
    dy1 = b.X[0, 1, 0, 1] - b.X[0, 0, 0, 1]
    assert out[5] == pytest.approx(tv1[0] * U / (0.5 * dy1) * (4 * H) * H * ref, rel=1e-12)
    assert abs(out[6]) < 1e-12 * out[5] and abs(out[7]) < 1e-12 * out[5]
    assert np.all(out[2:5] == 0.0)
