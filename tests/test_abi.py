"""The C-ABI library loads on a machine without a GPU, exports every entry
point include/cav3d.h declares, and its host-side equation of state (the
same templates the kernels use) agrees with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "cav3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cav3d_\w+)\s*\(", src)))


def test_header_and_binding_agree(pkg):
    assert declared() == sorted(pkg.solver.EXPORTS)


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    for name in declared():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_struct_layout_matches_header(pkg, tmp_path):
    """ctypes structures have the C compiler's sizes and offsets."""
    import subprocess
    prog = tmp_path / "layout.c"
    prog.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "cav3d.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(cav3d_params),"
        " offsetof(cav3d_params, cav_clip), offsetof(cav3d_params, cfl), offsetof(cav3d_params, bc_pv),"
        " offsetof(cav3d_params, emutref), sizeof(cav3d_block), offsetof(cav3d_block, con)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    P, B = pkg.solver.Params, pkg.solver.BlockDesc
    assert got == [ctypes.sizeof(P), P.cav_clip.offset, P.cfl.offset, P.bc_pv.offset, P.emutref.offset,
                   ctypes.sizeof(B), B.con.offset]


def test_create_without_device_fails_cleanly(pkg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    case = pkg.cases.config1(8, 6, 6)
    with pytest.raises(pkg.Cav3dError, match="device"):
        pkg.Solver(case)


def test_create_rejects_unsupported_options(pkg):
    case = pkg.cases.config1(8, 6, 6)
    case.params["nscheme"] = 5
    with pytest.raises(pkg.Cav3dError, match="nscheme"):
        pkg.Solver(case)
    case.params["nscheme"] = 1
    case.params["ncav"] = 6
    with pytest.raises(pkg.Cav3dError, match="ncav"):
        pkg.Solver(case)
    case.params["ncav"] = 0
    case.params["dt_phy"] = 0.0
    with pytest.raises(pkg.Cav3dError, match="dt_phy"):
        pkg.Solver(case)


def test_host_eos_matches_oracle(pkg, orc):
    rng = np.random.default_rng(11)
    n = 3000
    p = rng.uniform(2e3, 5e6, n)
    t = rng.uniform(275.0, 500.0, n)
    y1 = np.where(rng.random(n) < 0.5, 0.0, rng.uniform(0.0, 1.0, n))
    y2 = np.where(rng.random(n) < 0.7, 0.0, rng.uniform(0.0, 1.0, n) * (1.0 - y1))
    dv, tv = pkg.eos_host(p, t, y1, y2)
    for q in range(n):
        dref, tref = orc.deteos_prop(p[q], t[q], y1[q], y2[q])
        np.testing.assert_allclose(dv[q], dref, rtol=1e-11, atol=0)
        np.testing.assert_allclose(tv[q], tref, rtol=1e-11, atol=0)
