"""cav3d_integrals_noalloc (the integrals with no device allocation, its
sums in one fixed device array per GPU) against cav3d_integrals on the same
states: bit for bit, on two calls in a row, the second on a smaller case
(the array still holds the first call's larger set of partial sums), and
both against the oracle."""
import numpy as np
import pytest

from parity import small_cases
from test_post import flag_surface

pytestmark = pytest.mark.gpu


def _stepped(pkg, case, steps=2):
    flag_surface(pkg, case, "jmin")
    s = pkg.Solver(case)
    s.set_grid()
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(steps):
        s.timeinteg(1, it + 1, 0.0)
    return s


def test_noalloc_integrals_match_allocating_entry_point(pkg, orc):
    big = pkg.cases.config2(40, 24, 16)
    big.params["cav_clip"] = 1
    small = small_cases(pkg)["sst_cav_ausm_mlp"]()
    for name, case in (("config2 40x24x16", big), ("config2 12x8x6", small)):
        print(f"case {name}: {case.ncell} cells", flush=True)
        p = case.params
        s = _stepped(pkg, case)
        o = orc.Oracle(case)
        for ib, st in enumerate(case.init):
            o.put_state(ib, *st)
        for it in range(2):
            o.timeinteg(1, it + 1, 0.0)
        for aoa in (0.0, 0.14):
            ref = s.integrals(p["rhoref"], p["uref"], 0.01, aoa)
            got = s.integrals_noalloc(p["rhoref"], p["uref"], 0.01, aoa)
            again = s.integrals_noalloc(p["rhoref"], p["uref"], 0.01, aoa)
            assert np.array_equal(got, ref), (got, ref)
            assert np.array_equal(again, ref)
            orc_ref = o.integrals(p["rhoref"], p["uref"], 0.01, aoa)
            assert abs(got[0] - orc_ref[0]) <= 1e-12 * abs(orc_ref[0])
            scale = np.abs(orc_ref[2:]).max()
            assert np.all(np.abs(got[2:] - orc_ref[2:]) <= 1e-12 * scale)
        s.close()
