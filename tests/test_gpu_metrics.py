"""Device grid metrics (cav3d_set_grid + cav3d_build_metrics) against the
oracle's restatement of calnormal / calvolume / calydns, bit for bit, and the
pseudo-time step run on them."""
import numpy as np
import pytest

from test_grid_metrics import flipped_pair

pytestmark = pytest.mark.gpu


def _cases(pkg):
    c2 = pkg.cases.config2(14, 7, 5)
    flip = pkg.cases.Case("config2_flipped_pair", dict(c2.params), list(flipped_pair(pkg, c2, 6)), c2.npv, c2.ntv,
                          c2.ngrd)
    return {
        "config2_single": lambda: pkg.cases.config2(12, 8, 6),
        "config1_laminar": lambda: pkg.cases.config1(10, 6, 5),
        "config2_split2": lambda: pkg.cases.split_i(pkg.cases.config2(14, 7, 5), [5]),
        "config2_split3": lambda: pkg.cases.split_i(pkg.cases.config2(14, 7, 5), [4, 9]),
        "config2_flipped_pair": lambda: flip,
    }


@pytest.mark.parametrize("name", ["config2_single", "config1_laminar", "config2_split2", "config2_split3",
                                  "config2_flipped_pair"])
def test_device_metrics_match_oracle(pkg, orc, name):
    print(f"case {name}", flush=True)
    case = _cases(pkg)[name]()
    s = pkg.Solver(case, metrics="device")
    ref = orc.grid_metrics(case.blocks, case.ngrd)
    for ib in range(len(case.blocks)):
        got = s.get_metrics(ib)
        for what, a, r in zip(("cx", "ex", "tx", "grd"), got, ref[ib]):
            assert np.array_equal(a, r), (name, ib, what, np.abs(a - r).max())
    s.close()


def test_device_metrics_full_size_match_host_generator(pkg):
    """128x64x64 channel: 524,288 cells against 16,384 wall faces, against
    the host generator's k-d tree distances, bit for bit."""
    print("case config2 128x64x64 metrics", flush=True)
    case = pkg.cases.config2()
    (b,) = case.blocks
    s = pkg.Solver(case, metrics="device")
    for what, a, r in zip(("cx", "ex", "tx", "grd"), s.get_metrics(0), (b.cx, b.ex, b.tx, b.grd)):
        assert np.array_equal(a, r), (what, np.abs(a - r).max())
    s.close()


def test_step_on_device_metrics_equals_step_on_host_metrics(pkg):
    print("case config2 split3 two steps", flush=True)
    case = pkg.cases.split_i(pkg.cases.config2(14, 7, 5), [4, 9])
    case.params["cav_clip"] = 1      # the species clip keeps the run finite
    out = []
    for m in ("host", "device"):
        s = pkg.Solver(case, metrics=m)
        for ib, st in enumerate(case.init):
            s.put_state(ib, *st)
        for it in range(2):
            s.timeinteg(1, it + 1, 0.0)
        out.append([s.get_state(ib)[0] for ib in range(len(case.blocks))])
        s.close()
    for a, b in zip(*out):
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)
