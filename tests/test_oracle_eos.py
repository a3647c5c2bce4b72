"""The oracle's equation of state against the IAPWS-IF97 verification
tables, region-4 closed forms and the mixture rules
(src/eos_module.f90:1064-1234, 1331-1659, 2266-2354)."""
import math
from fractions import Fraction

import numpy as np
import pytest

# IAPWS-IF97 (2007) verification values: (T [K], p [MPa], v [m3/kg], h [kJ/kg], cp [kJ/kg/K], w [m/s])
TABLE5_R1 = [
    (300.0, 3.0, 0.100215168e-2, 0.115331273e3, 0.417301218e1, 0.150773921e4),
    (300.0, 80.0, 0.971180894e-3, 0.184142828e3, 0.401008987e1, 0.163469054e4),
    (500.0, 3.0, 0.120241800e-2, 0.975542239e3, 0.465580682e1, 0.124071337e4),
]
TABLE15_R2 = [
    (300.0, 0.0035, 0.394913866e2, 0.254991145e4, 0.191300162e1, 0.427920172e3),
    (700.0, 0.0035, 0.923015898e2, 0.333568375e4, 0.208141274e1, 0.644289068e3),
]
TABLE18_META = [
    (450.0, 1.0, 0.192516540, 0.276881115e4, 0.276349265e1, 0.498408101e3),
    (440.0, 1.0, 0.186212297, 0.274015123e4, 0.298166443e1, 0.489363295e3),
    (450.0, 1.5, 0.121685206, 0.272134539e4, 0.362795578e1, 0.481941819e3),
]
TABLE35_PSAT = [(300.0, 0.353658941e-2), (500.0, 0.263889776e1), (600.0, 0.123443146e2)]
TABLE36_TSAT = [(0.1, 0.372755919e3), (1.0, 0.453035632e3), (10.0, 0.584149488e3)]

# region-1 density prefactor p*/(R T*) = 16.53 MPa / (0.461526 kJ/kg/K * 1386 K); the
# reference writes it as a default-real literal (src/eos_module.f90:1367)
C_RHO1 = 16.53e3 / (0.461526 * 1386.0)
C_RHO1_F = float(np.float32(25.841246054191803727))


def sound2(e):
    """c^2 of a single phase from rho(p,T), h(p,T) (isentropic dh = dp/rho)."""
    return e.dhdt / (e.drdp * e.dhdt + e.drdt * (1.0 / e.rho - e.dhdp))


def _r1_coeffs():
    """Region-1 coefficients as the oracle stores them (parsed from its source)."""
    import os
    src = open(os.path.join(os.path.dirname(__file__), "..", "oracle", "orc_eos.c")).read()
    def arr(name):
        blk = src[src.index(name):]
        blk = blk[blk.index("{") + 1:blk.index("}")]
        return [x.strip() for x in blk.replace("\n", " ").split(",")]
    return [int(x) for x in arr("R1_I[34]")], [int(x) for x in arr("R1_J[34]")], [Fraction(x) for x in arr("R1_N[34]")]


def r1_exact(p, T, pi_per_pa):
    """IF97 region 1 in exact rational arithmetic: v, h, cp, dh/dp (SI), with the
    reduced pressure pi = p * pi_per_pa inside the Gibbs function."""
    I, J, N = _r1_coeffs()
    R = Fraction("461.526")
    pi = Fraction(p) * pi_per_pa
    tau = Fraction(1386) / Fraction(T)
    a, b = Fraction(71, 10) - pi, tau - Fraction(1222, 1000)
    gp = sum(-n * i * a ** (i - 1) * b ** j for n, i, j in zip(N, I, J))
    gt = sum(n * a ** i * j * b ** (j - 1) for n, i, j in zip(N, I, J))
    gtt = sum(n * a ** i * j * (j - 1) * b ** (j - 2) for n, i, j in zip(N, I, J))
    gpt = sum(-n * i * a ** (i - 1) * j * b ** (j - 1) for n, i, j in zip(N, I, J))
    pstar = Fraction(16530000)
    v = gp * R * Fraction(T) / pstar
    h = tau * gt * R * Fraction(T)
    cp = -tau * tau * gtt * R
    dhdp = R * Fraction(1386) * gpt / pstar
    return float(v), float(h), float(cp), float(dhdp)


@pytest.mark.parametrize("row", TABLE5_R1)
def test_region1_coefficients_table5(row):
    """The oracle's coefficient table, evaluated exactly with 1/16.53 MPa, gives table 5."""
    T, p, v, h, cp, w = row
    ve, he, cpe, _ = r1_exact(p * 1e6, T, Fraction(1, 16530000))
    assert abs(ve / v - 1.0) < 5e-9
    assert abs(he / 1e3 / h - 1.0) < 5e-9
    assert abs(cpe / 1e3 / cp - 1.0) < 5e-9


@pytest.mark.parametrize("T,p", [(300.0, 3e6), (300.0, 80e6), (500.0, 3e6), (280.0, 4e4), (350.0, 1e5)])
def test_region1_reference_constants(orc, T, p):
    """Oracle = exact IF97 with the reference's constants: pi = p*6.049607e-8
    inside the Gibbs function (src/eos_module.f90:1342) and the
    single-precision rho prefactor (:1367)."""
    e = orc.iapws97_l(p, T)
    ve, he, cpe, dhdpe = r1_exact(p, T, Fraction("0.00000006049607"))
    shift = C_RHO1_F / C_RHO1
    assert abs(shift - 1.0) > 1e-8
    assert abs(1.0 / (e.rho / shift) / ve - 1.0) < 1e-12
    assert abs(e.h / he - 1.0) < 1e-12
    assert abs(e.dhdt / cpe - 1.0) < 1e-12
    # dh/dp uses a single-precision 1/rho* literal (:1371): agreement to float rounding
    assert abs(e.dhdp / dhdpe - 1.0) < 1e-7


@pytest.mark.parametrize("row", TABLE5_R1)
def test_region1_sound_speed(orc, row):
    T, p, v, h, cp, w = row
    e = orc.iapws97_l(p * 1e6, T)
    assert abs(math.sqrt(sound2(e)) / w - 1.0) < 5e-8


@pytest.mark.parametrize("row", TABLE15_R2 + TABLE18_META)
def test_region2_and_metastable(orc, row):
    T, p, v, h, cp, w = row
    e = orc.iapws97_v(p * 1e6, T)
    assert abs(1.0 / e.rho / v - 1.0) < 5e-9
    assert abs(e.h / 1e3 / h - 1.0) < 5e-9
    assert abs(e.dhdt / 1e3 / cp - 1.0) < 5e-9
    # drho/dT carries the single-precision 1/540 literal (src/eos_module.f90:1457)
    assert abs(math.sqrt(sound2(e)) / w - 1.0) < 5e-8


@pytest.mark.parametrize("T,ps", TABLE35_PSAT)
def test_saturation_pressure(orc, T, ps):
    assert abs(orc.h2o_pww(T) / (ps * 1e6) - 1.0) < 5e-9


@pytest.mark.parametrize("p,Ts", TABLE36_TSAT)
def test_saturation_temperature(orc, p, Ts):
    assert abs(orc.h2o_tww(p * 1e6) / Ts - 1.0) < 5e-9


def test_saturation_clamps(orc):
    # below the triple point and above the critical point the temperature is clamped
    assert orc.h2o_pww(200.0) == orc.h2o_pww(273.15)
    assert orc.h2o_pww(800.0) == orc.h2o_pww(647.096)
    assert abs(orc.h2o_pww(647.096) / 22.064e6 - 1.0) < 1e-6


def test_powi_binary_exponentiation(orc):
    assert orc.powi(2.0, 10) == 1024.0
    assert orc.powi(3.0, 0) == 1.0
    x = 1.0000001
    # binary exponentiation: x^5 = x * (x^2)^2
    assert orc.powi(x, 5) == x * ((x * x) * (x * x))
    assert orc.powi(0.5, -3) == 8.0
    for m in range(-41, 33):
        assert abs(orc.powi(1.2345, m) / 1.2345 ** m - 1.0) < 1e-13


def _mix(orc, p, T, y1, y2):
    dv, tv = orc.deteos_prop(p, T, y1, y2)
    return dv, tv


def test_pure_phases(orc):
    p, T = 2.0e5, 300.0
    dl, tl = _mix(orc, p, T, 0.0, 0.0)
    el, ql = orc.iapws97_l_prop(p, T)
    # the mixture rule 1/(y/rho) reproduces the phase to an ulp
    assert abs(dl[0] / el.rho - 1.0) < 5e-16 and dl[1] == el.h and abs(dl[2] / el.rho - 1.0) < 5e-16
    assert abs(tl[0] / ql.vis - 1.0) < 1e-15 and abs(tl[1] / ql.cond - 1.0) < 1e-15
    assert abs(dl[5] / sound2(el) - 1.0) < 1e-12
    # vapour at 700 K, 3.5 kPa (stable region 2)
    dv_, tv_ = _mix(orc, 3500.0, 700.0, 1.0, 0.0)
    ev, qv = orc.iapws97_v_prop(3500.0, 700.0)
    assert abs(dv_[0] / ev.rho - 1.0) < 5e-16 and dv_[1] == ev.h and abs(dv_[3] / ev.rho - 1.0) < 5e-16
    assert abs(dv_[5] / sound2(ev) - 1.0) < 1e-12


@pytest.mark.parametrize("y1,y2", [(1e-4, 0.0), (0.3, 0.0), (0.2, 0.3), (0.0, 0.5)])
def test_mixture_rules(orc, y1, y2):
    p, T = 4.0e4, 300.0
    dv, tv = _mix(orc, p, T, y1, y2)
    el = orc.iapws97_l(p, T)
    ev = orc.iapws97_v(p, T)
    eg, qg = orc.ideal_prop(p, T)
    yl = 1.0 - y1 - y2
    rho = 1.0 / (yl / el.rho + y1 / ev.rho + y2 / eg.rho)
    h = yl * el.h + y1 * ev.h + y2 * eg.h
    assert abs(dv[0] / rho - 1.0) < 1e-14
    assert abs(dv[1] / h - 1.0) < 1e-14
    # drho/dy1 = -rho^2 (1/rho_v - 1/rho_l), drho/dy2 = -rho^2 (1/rho_g - 1/rho_l)
    assert abs(dv[8] / (-rho * rho * (1.0 / ev.rho - 1.0 / el.rho)) - 1.0) < 1e-12
    assert abs(dv[9] / (-rho * rho * (1.0 / eg.rho - 1.0 / el.rho)) - 1.0) < 1e-12
    # dh/dy1 = h_v - h_l, dh/dy2 = h_g - h_l
    assert abs(dv[12] - (ev.h - el.h)) < 1e-9 * abs(ev.h)
    assert abs(dv[13] - (eg.h - el.h)) < 1e-9 * abs(eg.h)
    # volume fractions: rho_phase entries are y*rho/alpha consistent
    drdp = rho * rho * (yl * el.drdp / el.rho ** 2 + y1 * ev.drdp / ev.rho ** 2 + y2 * eg.drdp / eg.rho ** 2)
    assert abs(dv[6] / drdp - 1.0) < 1e-12


def test_transport_iaps85(orc):
    # IAPS-85 at 300 K / 0.1 MPa: mu ~ 853.8e-6 Pa s, kappa ~ 0.6103 W/m/K
    e, q = orc.iapws97_l_prop(1.0e5, 300.0)
    assert abs(q.vis / 853.8e-6 - 1.0) < 5e-3
    assert abs(q.cond / 0.6103 - 1.0) < 2e-2
