"""Helpers of the CPU tests that compare the library's pointwise physics,
compiled for the host from 3dcode_amd/csrc/phys.hpp (tests/native), with
the oracle's routines on identical inputs."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "physcheck.cpp")
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build(outdir):
    so = os.path.join(outdir, "libphyscheck.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-ffp-contract=off",
                    "-std=c++17", "-fPIC", "-shared", "-w", SRC, "-o", so], check=True)
    L = C.CDLL(so)
    for f in ("pc_localdt",):
        getattr(L, f).restype = C.c_double
    L.pc_sst_eddy.restype = C.c_double
    L.pc_localdt.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D]
    L.pc_getx.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D, C.c_double, _D, _D, _D, _D]
    L.pc_jac.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D, C.c_int, _D, _D]
    L.pc_jac_coef.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D, C.c_int, C.c_double, _D, _D]
    L.pc_jac_rows.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D, C.c_int, C.c_double, _D, _D]
    L.pc_getx_sel.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D, C.c_double, _D, _D, _D, _D]
    L.pc_merkle.argtypes = [_I, _D, _D, _D, C.c_double, _D]
    L.pc_cav_model.argtypes = [_I, _D, _D, _D, C.c_double, _D]
    L.pc_sst_source_des.argtypes = [_I, _D, _D, _D, C.c_double, C.c_double, _D, C.c_double, _D]
    L.pc_sat_t.argtypes = [C.c_double]
    L.pc_sat_t.restype = C.c_double
    L.pc_h2o_sigma.argtypes = [C.c_double]
    L.pc_h2o_sigma.restype = C.c_double
    L.pc_sst_source.argtypes = [_D, _D, C.c_double, C.c_double, _D, C.c_double, _D]
    L.pc_sst_eddy.argtypes = [_D, _D, C.c_double, C.c_double, _D]
    L.pc_liquid.argtypes = [C.c_double, C.c_double, _D]
    L.pc_vapour.argtypes = [C.c_double, C.c_double, _D]
    L.pc_mixture.argtypes = [C.c_double] * 4 + [_D, _D]
    L.pc_face.argtypes = [_I, _D, _D, _D, _D, _D, _D, _D, _D, _D]
    L.pc_face32.argtypes = [_I, _D, _D, _D, _D, _D, _D]
    L.pc_face_adj.argtypes = [_I, _D, _D, _D, _D, _D, _D, _D, _D, _D]
    L.pc_eos_adj.argtypes = [C.c_int, _D, _D, _D]
    L.pc_cav_adj.argtypes = [_I, _D, _D, _D, _D]
    L.pc_visc.argtypes = [_I, _D, _D, _D, _D, _D, _D, _D, _D, _D, C.c_double, C.c_double, _D]
    return L


def p(a, t=C.c_double):
    a = np.ascontiguousarray(a)
    return a.ctypes.data_as(C.POINTER(t))


def par_arrays(cfg):
    """int / double parameter vectors of pc_* from an oracle-style cfg dict."""
    iv = np.array([cfg["npv"], cfg["ntv"], cfg["iturb"], cfg["nscheme"], cfg["nmuscl"], cfg["nlim"], cfg["local"],
                   cfg["prec"], cfg["ncav"], cfg["nsteady"], cfg.get("cav_clip", 0), cfg.get("ides", 0)],
                  dtype=np.int32)
    bdf = 1.0 if cfg["nsteady"] == 1 else (0.75 if cfg["nsteady"] == 2 else 0.0)
    dv = np.array([cfg["cfl"], cfg["pref"], cfg["uref"], cfg["str"], cfg["dt_phy"], cfg["c_v"], cfg["c_c"],
                   cfg["dp_ref"], cfg["t_ref"], cfg["t_crit"], *list(cfg["bc_pv"]), cfg["bc_pressure"],
                   cfg["emutref"], bdf])
    return iv, dv


# compact per-cell properties from dv(18), tv(2)
COMPACT = [0, 1, 5, 6, 7, 8, 9, 10, 11, 12, 13]


def compact(dv, tv):
    return np.concatenate([np.asarray(dv)[COMPACT], np.asarray(tv)[:2]])


def ulps(a, b):
    """Distance in units of the last place of b (0 for identical values)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    sp = np.spacing(np.maximum(np.abs(b), 1e-300))
    return np.where(a == b, 0.0, np.abs(a - b) / sp)
