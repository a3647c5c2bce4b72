"""Device post integrals against the oracle, and the restart file: its byte
layout (src/variable_module.f90:108-182) and a restart that continues bit
for bit (src/initial_module.f90:245-389)."""
import struct

import numpy as np
import pytest

from parity import small_cases
from test_post import flag_surface

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["sst_cav_ausm_mlp", "lam_roe_tvd"])
def test_device_integrals_match_oracle(pkg, orc, name):
    print(f"case {name} integrals", flush=True)
    case = small_cases(pkg)[name]()
    flag_surface(pkg, case, "jmin")
    p = case.params
    s = pkg.Solver(case)
    s.set_grid()
    o = orc.Oracle(case)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
        o.put_state(ib, *st)
    for it in range(2):
        s.timeinteg(1, it + 1, 0.0)
        o.timeinteg(1, it + 1, 0.0)
    for aoa in (0.0, 0.14):
        got = s.integrals(p["rhoref"], p["uref"], 0.01, aoa)
        ref = o.integrals(p["rhoref"], p["uref"], 0.01, aoa)
        assert got[0] == pytest.approx(ref[0], rel=1e-12, abs=1e-300)
        assert got[1] == pytest.approx(ref[1], rel=1e-12, abs=1e-300)
        scale = np.abs(ref[2:]).max()
        assert scale > 0
        assert np.all(np.abs(got[2:] - ref[2:]) <= 1e-12 * scale), (got, ref)
    s.close()


def _records(case, b):
    ng, nc = b.nghost, b.ncell
    p = case.params
    return [("pv", case.npv * ng), ("emut", ng if p["iturb"] >= -1 else 0),
            ("qq", 2 * case.npv * nc if p["nsteady"] >= 1 else 0), ("rcav", 2 * nc if p["ncav"] else 0)]


@pytest.mark.parametrize("name", ["sst_cav_3block", "lam_roe_bdf2"])
def test_export_layout(pkg, name, tmp_path):
    print(f"case {name} export", flush=True)
    case = small_cases(pkg)[name]()
    s = pkg.Solver(case)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(2):
        s.timeinteg(3, it + 1, 0.0)
    f = tmp_path / "out_00000003.dat"
    s.export(str(f), 3, 2, 1.25e-3)
    raw = f.read_bytes()
    at = 0
    for ib, b in enumerate(case.blocks):
        nt_phy, nt, t = struct.unpack_from("=iid", raw, at)
        assert (nt_phy, nt, t) == (3, 2, 1.25e-3)
        at += 16
        pv, _, tv, qq = s.get_state(ib)
        for what, n in _records(case, b):
            a = np.frombuffer(raw, dtype=np.float64, count=n, offset=at) if n else None
            at += 8 * n
            if what == "pv":
                assert np.array_equal(a, pv)
            elif what == "emut" and n:
                assert np.array_equal(a, tv.reshape(-1, case.ntv)[:, 2])
            elif what == "qq" and n:
                assert np.array_equal(a, qq)
            elif what == "rcav" and n:
                assert np.all(np.isfinite(a))
    assert at == len(raw)
    s.close()


@pytest.mark.parametrize("name", ["sst_cav_3block", "lam_roe_bdf2"])
def test_restart_continues_bit_for_bit(pkg, name, tmp_path):
    """Run A steps straight through; run B stops halfway, writes the restart
    file, and a fresh context reads it and finishes: same state bits."""
    print(f"case {name} restart", flush=True)
    case = small_cases(pkg)[name]()
    unsteady = case.params["nsteady"] >= 1

    def fresh():
        s = pkg.Solver(case)
        for ib, st in enumerate(case.init):
            s.put_state(ib, *st)
        return s

    def advance(s, phase):
        if unsteady:
            if phase == 0:
                for it in range(2):
                    s.timeinteg(1, it + 1, 0.0)
                s.unsteady_shift()
            else:
                for it in range(2):
                    s.timeinteg(2, it + 1, 0.0)
        else:
            for it in range(3):
                s.timeinteg(1, 3 * phase + it + 1, 0.0)

    a = fresh()
    advance(a, 0)
    advance(a, 1)
    want = [a.get_state(ib) for ib in range(len(case.blocks))]
    a.close()

    b = fresh()
    advance(b, 0)
    f = tmp_path / "out_restart.dat"
    b.export(str(f), 1, 3, 0.0)
    b.close()
    c = fresh()
    nps, nts, tp = c.import_(str(f))
    assert (nps, nts, tp) == ((2, 1, 0.0) if unsteady else (1, 4, 0.0))
    advance(c, 1)
    for ib in range(len(case.blocks)):
        got = c.get_state(ib)
        assert np.array_equal(got[0], want[ib][0])
        if unsteady:
            assert np.array_equal(got[3], want[ib][3])
    c.close()
