"""The mixed precision mode (CAV3D_MIXED: reconstruction, face EOS and the
inviscid and viscous face fluxes in FP32; sources, Jacobians, D^-1, sweeps,
update, EOS of the cells and norms in FP64) against the FP64 path on the
same inputs.  Its error is stated, not bit-exact: after 20 pseudo-time
iterations every field stays within 2e-3 of its largest magnitude, and the
residual norms within 1e-2."""
import numpy as np
import pytest

from parity import small_cases

pytestmark = pytest.mark.gpu


def _run(pkg, case, precision, n):
    s = pkg.Solver(case, precision=precision)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(n):
        s.timeinteg(1, it + 1, 0.0)
    out = [s.get_state(ib)[0] for ib in range(len(case.blocks))], s.residual()[0]
    s.close()
    return out


@pytest.mark.parametrize("name", ["lam_roe_tvd", "sst_cav_ausm_mlp", "sst_cav_roem_mlp", "sst_cav_3block"])
def test_mixed_tracks_fp64(pkg, name):
    print(f"case {name} mixed", flush=True)
    case = small_cases(pkg)[name]()
    (p64, l64), (p32, l32) = _run(pkg, case, 0, 20), _run(pkg, case, 1, 20)
    differs = False
    for a, b in zip(p64, p32):
        A, B = a.reshape(-1, case.npv), b.reshape(-1, case.npv)
        assert np.all(np.isfinite(B))
        scale = np.maximum(np.abs(A).max(axis=0), 1e-300)
        err = np.abs(A - B).max(axis=0) / scale
        assert np.all(err < 2e-3), err
        differs |= bool(np.any(A != B))
    assert differs                      # the FP32 kernels ran
    nz = np.abs(l64) > 0
    assert np.all(np.abs(l32[nz] - l64[nz]) / np.abs(l64[nz]) < 1e-2), (l32, l64)


def test_ausmpwp_stays_fp64_in_mixed_mode(pkg):
    """Inviscid (no viscous faces), AUSMPW+: the mixed mode leaves its face
    kernel in FP64, so both modes give the same bits."""
    print("case euler ausmpwp mixed", flush=True)
    case = small_cases(pkg)["lam_ausmpwp_tvd"]()
    case.params["iturb"] = -3
    case.ntv = 0
    case.init = [(pv, dv, None, qq) for (pv, dv, tv, qq) in case.init]
    (p64, _), (p32, _) = _run(pkg, case, 0, 3), _run(pkg, case, 1, 3)
    for a, b in zip(p64, p32):
        assert np.array_equal(a, b)
