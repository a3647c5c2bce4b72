"""The Fortran bind(c) module (include/cav3d_iface.f90) a reference host
would use: it compiles with the image's Fortran compiler, its derived types
have the C structs' sizes, every function the C header declares has an
interface in it, and a Fortran program evaluates the library's host EOS
through it (compared with the oracle on the same states).  CPU only."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

PROGRAM = r"""
program check
  use cav3d_iface
  implicit none
  type(cav3d_params) :: p
  type(cav3d_block) :: b
  real(c_double) :: pr(3), t(3), y1(3), y2(3), dv(18, 3), tv(2, 3)
  integer :: rc, i
  write(*, '(a, i0, 1x, i0)') 'sizes ', c_sizeof(p), c_sizeof(b)
  pr = [1.0d5, 3.0d6, 2.0d3]
  t = [300.0d0, 350.0d0, 290.0d0]
  y1 = [0.0d0, 0.0d0, 0.5d0]
  y2 = [0.0d0, 0.0d0, 0.0d0]
  rc = cav3d_eos_host(3_c_int, pr, t, y1, y2, dv, tv)
  write(*, '(a, i0)') 'rc ', rc
  do i = 1, 3
    write(*, '(a, 3es25.17)') 'eos ', dv(1, i), dv(2, i), tv(1, i)
  end do
end program check
"""


def test_fortran_module_compiles_and_calls_the_library(pkg, orc, tmp_path):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler in this image")
    src = tmp_path / "check.f90"
    src.write_text(PROGRAM)
    exe = tmp_path / "check"
    lib = os.path.join(ROOT, "3dcode_amd")
    r = subprocess.run([FC, "-O0", "-J", str(tmp_path), os.path.join(ROOT, "include", "cav3d_iface.f90"), str(src),
                        "-L", lib, "-lcav3d", f"-Wl,-rpath,{lib}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in out[0].split()[1:]]
    assert sizes == [ctypes.sizeof(pkg.solver.Params), ctypes.sizeof(pkg.solver.BlockDesc)]
    assert out[1].split()[1] == "0"
    for line, q in zip(out[2:5], [(1.0e5, 300.0, 0.0, 0.0), (3.0e6, 350.0, 0.0, 0.0), (2.0e3, 290.0, 0.5, 0.0)]):
        rho, h, mu = (float(x) for x in line.split()[1:])
        dv, tv = orc.deteos_prop(*q)
        assert abs(rho - dv[0]) <= 1e-13 * dv[0] and abs(h - dv[1]) <= 1e-13 * abs(dv[1])
        assert abs(mu - tv[0]) <= 1e-13 * tv[0]


def test_every_c_function_has_a_fortran_interface():
    h = open(os.path.join(ROOT, "include", "cav3d.h")).read()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    c_funcs = set(re.findall(r"\b(cav3d_\w+)\s*\(", h))
    f = open(os.path.join(ROOT, "include", "cav3d_iface.f90")).read()
    f_funcs = set(re.findall(r"(?:function|subroutine)\s+(cav3d_\w+)\s*\(", f))
    assert c_funcs == f_funcs, (c_funcs ^ f_funcs)
