"""Child process of test_gpu_adjoint.test_autograd_harness: the
torch.autograd harness over the residual (3dcode_amd/autograd.py) on a small
first-order SST + Merkle case.  torch is imported before the library."""
import os
import sys

import torch  # noqa: F401  (first: its HIP runtime is the process's)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import importlib  # noqa: E402

import numpy as np  # noqa: E402

import test_gpu_adjoint as T  # noqa: E402


def main():
    pk = importlib.import_module("3dcode_amd")
    ag = importlib.import_module("3dcode_amd.autograd")
    # a generic state without pseudo-time steps: the parent test process holds
    # the device's cooperative-launch queue (the LU-SGS sweeps)
    case, s, pvs, fr = T.prepared(pk, "sst_cav_roe_first", steps=0)
    try:
        pv = torch.from_numpy(np.concatenate(pvs)).requires_grad_(True)
        res = ag.residual(s, pv, fr)
        w = torch.from_numpy(np.random.default_rng(5).standard_normal(res.shape[0]))
        loss = 0.5 * ((res * w) ** 2).sum()
        loss.backward()
        g = pv.grad.numpy().copy()
        rb = (res.detach() * w * w).numpy()
        pb, _ = s.rhs_vjp(ag._split(s, rb, "res"))
        assert np.allclose(g, np.concatenate(pb), rtol=1e-12, atol=0.0), "autograd gradient != library VJP"
        d = np.concatenate(T.directions(case, pvs, np.random.default_rng(6), 1)[0])

        def loss_at(c):
            with torch.no_grad():
                r = ag.residual(s, torch.from_numpy(np.concatenate(pvs) + c * d), fr)
                return 0.5 * ((r * w) ** 2).sum().item()
        adj = float(np.dot(g, d))
        err = T.kink_aware(lambda c: 0.5 * (loss_at(c) - loss_at(-c)) / c, adj)
        print(f"autograd harness: directional derivative adjoint {adj:.9e} rel err {err}", flush=True)
        assert err is not None and err < T.TOL
    finally:
        s.close()


if __name__ == "__main__":
    main()
