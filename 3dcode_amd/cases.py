"""Host-side set-up of the synthetic cases: analytic grids, metrics, boundary
tables, reference values and initial states, in the reference's layouts.

This mirrors what the reference host keeps (config, grid, bc tables, initial
state) and hands to the pseudo-time step:
  grid metrics      src/grid_module.f90:609-942 (calnormal), 944-1241
                    (calvolume), 350-607 (calydns)
  BC point ranges   src/grid_module.f90:149-206
  reference values  src/config_module.f90:178-199 (setref), src/main.f90:31-32
  initial state     src/initial_module.f90:678-768

Arrays are numpy C-order with the Fortran index order reversed, so their
flat memory equals the reference's column-major arrays, e.g.
pv(npv,-1:imax+3,-1:jmax+3,-1:kmax+3) is pv[k+1, j+1, i+1, n].
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .solver import BC, SURFACE, eos_host

H = 0.1  # channel height [m]


@dataclass
class Block:
    imax: int
    jmax: int
    kmax: int
    X: np.ndarray                       # vertices (imax, jmax, kmax, 3)
    bc_desc: list = field(default_factory=list)
    conn_desc: list = field(default_factory=list)
    cx: np.ndarray = None
    ex: np.ndarray = None
    tx: np.ndarray = None
    grd: np.ndarray = None

    @property
    def ncell(self):
        return (self.imax - 1) * (self.jmax - 1) * (self.kmax - 1)

    @property
    def nghost(self):
        return (self.imax + 5) * (self.jmax + 5) * (self.kmax + 5)


@dataclass
class Case:
    name: str
    params: dict
    blocks: list
    npv: int
    ntv: int
    ngrd: int
    ref: dict = field(default_factory=dict)
    init: list = field(default_factory=list)   # per block (pv, dv, tv, qq)

    def cfg_dict(self):
        """Fields of the oracle configuration."""
        p = self.params
        keys = ["iturb", "nscheme", "nmuscl", "nlim", "local", "prec", "ncav", "nsteady", "cfl", "pref",
                "uref", "str", "dt_phy", "c_v", "c_c", "dp_ref", "t_ref", "t_crit", "bc_pv", "bc_pressure",
                "emutref", "cav_clip"]
        d = {k: p[k] for k in keys}
        d.update(npv=self.npv, ntv=self.ntv, ndv=18, ngrd=self.ngrd, ides=p.get("ides", 0), timemethod=3, nprint=1)
        return d

    @property
    def ncell(self):
        return sum(b.ncell for b in self.blocks)


# ------------------------------------------------------------------ metrics
def face_vectors(X):
    """cx, ex, tx (Fortran order reversed) from vertices X (imax,jmax,kmax,3);
    src/grid_module.f90:620-723."""
    I, J, K = X.shape[:3]
    cx = np.zeros((I, J + 1, K + 1, 3))
    e = X[:, 1:, :-1] - X[:, :-1, 1:]
    s = X[:, 1:, 1:] - X[:, :-1, :-1]
    cx[:, 1:J, 1:K] = 0.5 * np.cross(e, s)
    cx[:, 0] = cx[:, 1]
    cx[:, J] = cx[:, J - 1]
    cx[:, :, 0] = cx[:, :, 1]
    cx[:, :, K] = cx[:, :, K - 1]
    ex = np.zeros((I + 1, J, K + 1, 3))
    c = X[:-1, :, :-1] - X[1:, :, 1:]
    s = X[:-1, :, 1:] - X[1:, :, :-1]
    ex[1:I, :, 1:K] = 0.5 * np.cross(c, s)
    ex[0] = ex[1]
    ex[I] = ex[I - 1]
    ex[:, :, 0] = ex[:, :, 1]
    ex[:, :, K] = ex[:, :, K - 1]
    tx = np.zeros((I + 1, J + 1, K, 3))
    c = X[1:, :-1, :] - X[:-1, 1:, :]
    e = X[1:, 1:, :] - X[:-1, :-1, :]
    tx[1:I, 1:J, :] = 0.5 * np.cross(c, e)
    tx[0] = tx[1]
    tx[I] = tx[I - 1]
    tx[:, 0] = tx[:, 1]
    tx[:, J] = tx[:, J - 1]
    return cx, ex, tx


def _pyramid(pa, pb, pc, pd, xc):
    cs = 0.25 * (pa + pb + pc + pd)
    va = pc - pa
    vb = pd - pb
    v = ((xc[..., 0] - cs[..., 0]) * (va[..., 1] * vb[..., 2] - va[..., 2] * vb[..., 1])
         + (xc[..., 1] - cs[..., 1]) * (va[..., 2] * vb[..., 0] - va[..., 0] * vb[..., 2])
         + (xc[..., 2] - cs[..., 2]) * (va[..., 0] * vb[..., 1] - va[..., 1] * vb[..., 0]))
    return np.abs(v)


def cell_geometry(X):
    """volume and centroid of cells 2..imax (src/grid_module.f90:954-1071)."""
    a = X[:-1, :-1, :-1]; b = X[1:, :-1, :-1]; c = X[:-1, 1:, :-1]; d = X[:-1, :-1, 1:]
    e = X[1:, 1:, :-1]; f = X[:-1, 1:, 1:]; g = X[1:, :-1, 1:]; h = X[1:, 1:, 1:]
    xc = 0.125 * (a + b + c + d + e + f + g + h)
    v = (_pyramid(a, b, e, c, xc) + _pyramid(d, g, h, f, xc) + _pyramid(a, b, g, d, xc)
         + _pyramid(c, e, h, f, xc) + _pyramid(a, c, f, d, xc) + _pyramid(b, e, h, g, xc)) / 6.0
    return v, xc


def wall_face_centres(X, bc_desc):
    """centroids of BCWall faces (src/grid_module.f90:364-448)."""
    pts = []
    I, J, K = X.shape[:3]
    for kind, *r in bc_desc:
        if kind != BC["wall"]:
            continue
        s, e = r[:3], r[3:]
        if s[0] == e[0]:
            iv = 0 if s[0] == 1 else I - 1
            q = X[iv, s[1] - 2:e[1], s[2] - 2:e[2]]
            pts.append(0.25 * (q[:-1, :-1] + q[1:, :-1] + q[:-1, 1:] + q[1:, 1:]).reshape(-1, 3))
        elif s[1] == e[1]:
            jv = 0 if s[1] == 1 else J - 1
            q = X[s[0] - 2:e[0], jv, s[2] - 2:e[2]]
            pts.append(0.25 * (q[:-1, :-1] + q[1:, :-1] + q[:-1, 1:] + q[1:, 1:]).reshape(-1, 3))
        else:
            kv = 0 if s[2] == 1 else K - 1
            q = X[s[0] - 2:e[0], s[1] - 2:e[1], kv]
            pts.append(0.25 * (q[:-1, :-1] + q[1:, :-1] + q[:-1, 1:] + q[1:, 1:]).reshape(-1, 3))
    return np.concatenate(pts) if pts else np.zeros((0, 3))


def grd_array(X, ngrd, walls):
    """grd(ngrd,1:imax+1,1:jmax+1,1:kmax+1) reversed: (K+1, J+1, I+1, ngrd)."""
    I, J, K = X.shape[:3]
    v, xc = cell_geometry(X)
    G = np.zeros((I + 1, J + 1, K + 1, ngrd))
    G[1:I, 1:J, 1:K, 0] = v
    G[1:I, 1:J, 1:K, 1:4] = xc
    if ngrd > 4:
        if len(walls):
            from scipy.spatial import cKDTree
            dist, _ = cKDTree(walls).query(xc.reshape(-1, 3))
            G[1:I, 1:J, 1:K, 4] = dist.reshape(xc.shape[:3])
        else:
            G[1:I, 1:J, 1:K, 4] = 1.0e6
    # ghost copies / linear extrapolation (src/grid_module.f90:1073-1113, 488-505)
    for q in range(ngrd):
        lin = 1 <= q <= 3
        Gq = G[..., q]
        Gq[1:I, 0, 1:K] = 2 * Gq[1:I, 1, 1:K] - Gq[1:I, 2, 1:K] if lin else Gq[1:I, 1, 1:K]
        Gq[1:I, J, 1:K] = 2 * Gq[1:I, J - 1, 1:K] - Gq[1:I, J - 2, 1:K] if lin else Gq[1:I, J - 1, 1:K]
        Gq[0, :, 1:K] = 2 * Gq[1, :, 1:K] - Gq[2, :, 1:K] if lin else Gq[1, :, 1:K]
        Gq[I, :, 1:K] = 2 * Gq[I - 1, :, 1:K] - Gq[I - 2, :, 1:K] if lin else Gq[I - 1, :, 1:K]
        if lin:   # kmin uses grd(:,3,j,3) as written at src/grid_module.f90:1104-1106
            Gq[:, :, 0] = 2 * Gq[:, :, 1] - Gq[2, :, 2][None, :]
        else:
            Gq[:, :, 0] = Gq[:, :, 1]
        Gq[:, :, K] = 2 * Gq[:, :, K - 1] - Gq[:, :, K - 2] if lin else Gq[:, :, K - 1]
    return G


def finish_block(b: Block, ngrd: int, walls=None):
    cx, ex, tx = face_vectors(b.X)
    b.cx = np.ascontiguousarray(cx.transpose(2, 1, 0, 3)).ravel()
    b.ex = np.ascontiguousarray(ex.transpose(2, 1, 0, 3)).ravel()
    b.tx = np.ascontiguousarray(tx.transpose(2, 1, 0, 3)).ravel()
    if walls is None:
        walls = wall_face_centres(b.X, b.bc_desc)
    b.grd = np.ascontiguousarray(grd_array(b.X, ngrd, walls).transpose(2, 1, 0, 3)).ravel()


def face_bc(kind, face, I, J, K):
    """BC point range of a whole block face in grid_module convention."""
    lo = [2, 2, 2]
    hi = [I, J, K]
    ax = "ijk".index(face[0])
    v = 1 if face[1:] == "min" else [I, J, K][ax] + 1
    lo[ax] = hi[ax] = v
    return [BC[kind], *lo, *hi]


# ------------------------------------------------------------------ grids
def bump_channel(ni, nj, nk, beta=2.0, bump=0.1):
    """3-D channel, length 4H, height H, span H; cosine bump 0.1H on
    x in [1.5H, 2.5H]; tanh clustering toward both walls."""
    x = np.linspace(0.0, 4.0 * H, ni + 1)
    s = np.linspace(0.0, 1.0, nj + 1)
    z = np.linspace(0.0, H, nk + 1)
    eta = 0.5 * (1.0 + np.tanh(beta * (2.0 * s - 1.0)) / math.tanh(beta))
    yb = np.where((x >= 1.5 * H) & (x <= 2.5 * H), 0.5 * bump * H * (1.0 - np.cos(2.0 * math.pi * (x - 1.5 * H) / H)), 0.0)
    X = np.zeros((ni + 1, nj + 1, nk + 1, 3))
    X[..., 0] = x[:, None, None]
    X[..., 1] = (yb[:, None] + (H - yb[:, None]) * eta[None, :])[:, :, None]
    X[..., 2] = z[None, None, :]
    I, J, K = ni + 1, nj + 1, nk + 1
    bcs = [face_bc("inflow_subsonic", "imin", I, J, K), face_bc("outflow_subsonic", "imax", I, J, K),
           face_bc("wall", "jmin", I, J, K), face_bc("wall", "jmax", I, J, K),
           face_bc("symmetry", "kmin", I, J, K), face_bc("symmetry", "kmax", I, J, K)]
    return Block(I, J, K, X, bcs, [])


# ------------------------------------------------------------------ parameters
def reference_values(p):
    """rho_ref, mu_ref at (pref, tref, y1ref, y2ref) (src/main.f90:31-32)."""
    dv, tv = eos_host(p["pref"], p["tref"], p["y1ref"], p["y2ref"])
    return float(dv[0, 0]), float(tv[0, 0])


def make_params(iturb, nscheme, nmuscl, nlim, prec, ncav, nsteady, cfl, pref, uref, tref=300.0,
                y1ref=0.0, y2ref=0.0, aoa=0.0, aos=0.0, dt_phy=1.0, l_chord=H, l_character=H, l_domain=4 * H,
                c_v=1.0, c_c=80.0, local=1, turb_ref=None):
    p = dict(iturb=iturb, nscheme=nscheme, nmuscl=nmuscl, nlim=nlim, local=local, prec=prec, ncav=ncav,
             nsteady=nsteady, cfl=cfl, pref=pref, uref=uref, tref=tref, y1ref=y1ref, y2ref=y2ref, dt_phy=dt_phy,
             c_v=c_v, c_c=c_c, t_crit=647.096)
    p["str"] = l_character / math.pi / dt_phy
    rho, mu = reference_values(p)
    p["rhoref"], p["visref"] = rho, mu
    p["dp_ref"] = 2.0 / (rho * uref ** 2)
    p["t_ref"] = uref / l_chord
    if iturb == 0:
        if turb_ref is None:   # config_module setref
            oref = 10.0 * uref / l_domain
            emut = 1.0e-5 * mu
            kref = emut / rho * oref
        else:
            kref, oref, emut = turb_ref(rho, mu)
    else:
        kref = oref = emut = 0.0
    p["kref"], p["oref"], p["emutref"] = kref, oref, emut
    aoa_r, aos_r = math.radians(aoa), math.radians(aos)
    p["bc_pv"] = [pref, uref * math.cos(aos_r) * math.cos(aoa_r), uref * math.cos(aos_r) * math.sin(aoa_r),
                  uref * math.sin(aos_r), tref, y1ref, y2ref, kref, oref]
    p["bc_pressure"] = pref
    p["precision"] = 0
    p["cav_clip"] = 0
    return p


# ------------------------------------------------------------------ initial state
def splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def noise(seed, block, ncell, nvar, amp):
    """Counter-based uniform noise in [-amp, amp] keyed by (seed, block, cell, var)."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(40)) ^ (np.uint64(block) << np.uint64(32))
        ctr = np.arange(ncell * nvar, dtype=np.uint64) + key
        u = splitmix64(ctr) >> np.uint64(11)
    return (u.astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0).reshape(ncell, nvar) * amp


def conserved_noise_to_primitive(pv, pref, eps, qdyn):
    """Apply relative noise eps (n, npv) to the conserved state of primitives
    pv (n, npv) and return the new primitives.

    Momentum, total energy, species and turbulence densities are multiplied
    by (1 + eps).  A relative density change of 1e-3 in liquid water is a
    pressure change of about 2 MPa (below vacuum at these reference
    pressures), so the mass equation's noise is applied as a gauge-pressure
    change eps * qdyn instead.  The temperature then follows from the
    perturbed total energy by Newton on rho*h(p,T) - p + |m|^2/(2 rho(p,T)).
    """
    n, npv = pv.shape
    pabs = pv[:, 0] + pref
    T = pv[:, 4].copy()
    dv, _ = eos_host(pabs, T, pv[:, 5], pv[:, 6])
    rho, h = dv[:, 0], dv[:, 1]
    u2 = np.sum(pv[:, 1:4] ** 2, axis=1)
    m = rho[:, None] * pv[:, 1:4] * (1.0 + eps[:, 1:4])
    E = (rho * (h + 0.5 * u2) - pabs) * (1.0 + eps[:, 4])
    ry = rho[:, None] * pv[:, 5:] * (1.0 + eps[:, 5:])
    p = pabs + eps[:, 0] * qdyn
    m2 = np.sum(m ** 2, axis=1)
    for _ in range(40):
        y = ry[:, :2] / rho[:, None]
        dv, _ = eos_host(p, T, y[:, 0], y[:, 1])
        rho = dv[:, 0]
        f = rho * dv[:, 1] - p + 0.5 * m2 / rho - E
        dfdT = dv[:, 7] * dv[:, 1] + rho * dv[:, 11] - 0.5 * m2 * dv[:, 7] / rho ** 2
        dT = f / dfdT
        T -= dT
        if np.max(np.abs(dT) / T) < 1e-15:
            break
    y = ry[:, :2] / rho[:, None]
    dv, _ = eos_host(p, T, y[:, 0], y[:, 1])
    rho = dv[:, 0]
    out = np.empty_like(pv)
    out[:, 0] = p - pref
    out[:, 1:4] = m / rho[:, None]
    out[:, 4] = T
    out[:, 5:] = ry / rho[:, None]
    return out


def init_state(case, seed, pocket=None):
    """Uniform reference state on every interior cell (initial_module
    initial, :678-704), optional vapour pocket, conserved-variable noise of
    amplitude 1e-3, then dv/tv by the EOS (set_others, :707-768)."""
    p = case.params
    npv, ntv = case.npv, case.ntv
    case.init = []
    for ib, b in enumerate(case.blocks):
        I, J, K = b.imax, b.jmax, b.kmax
        nc = b.ncell
        cell = np.zeros((nc, npv))
        cell[:, 0] = 0.0
        cell[:, 1:4] = p["bc_pv"][1:4]
        cell[:, 4] = p["tref"]
        cell[:, 5] = p["y1ref"]
        cell[:, 6] = p["y2ref"]
        if npv > 7:
            cell[:, 7] = p["kref"]
            cell[:, 8] = p["oref"]
        if pocket is not None:
            _, xc = cell_geometry(b.X)
            xc = xc.transpose(2, 1, 0, 3).reshape(-1, 3)
            alpha = pocket(xc)
            dvl, _ = eos_host(p["pref"], p["tref"], 0.0, 0.0)
            dvv = eos_host(p["pref"], p["tref"], 1.0, 0.0)[0]
            rl, rv = dvl[0, 2], dvv[0, 3]
            cell[:, 5] = alpha * rv / (alpha * rv + (1.0 - alpha) * rl)
        eps = noise(seed, ib, nc, npv, 1.0e-3)
        cell = conserved_noise_to_primitive(cell, p["pref"], eps, 0.5 * p["rhoref"] * p["uref"] ** 2)
        pv = np.zeros((K + 5, J + 5, I + 5, npv))
        pv[3:K + 2, 3:J + 2, 3:I + 2, :] = cell.reshape(K - 1, J - 1, I - 1, npv)
        dv = np.zeros((K + 5, J + 5, I + 5, 18))
        tv = np.zeros((K + 5, J + 5, I + 5, max(ntv, 1)))
        d, t = eos_host(cell[:, 0] + p["pref"], cell[:, 4], cell[:, 5], cell[:, 6])
        dv[3:K + 2, 3:J + 2, 3:I + 2, :] = d.reshape(K - 1, J - 1, I - 1, 18)
        if ntv >= 2:
            tv[3:K + 2, 3:J + 2, 3:I + 2, 0:2] = t.reshape(K - 1, J - 1, I - 1, 2)
        if ntv == 3:
            tv[3:K + 2, 3:J + 2, 3:I + 2, 2] = p["emutref"]
        qq = None
        if p["nsteady"] > 0:
            c = cell
            rho, hh = d[:, 0], d[:, 1]
            q = np.empty((nc, npv))
            q[:, 0] = rho
            q[:, 1:4] = rho[:, None] * c[:, 1:4]
            q[:, 4] = rho * (hh + 0.5 * np.sum(c[:, 1:4] ** 2, axis=1)) - c[:, 0] - p["pref"]
            q[:, 5:] = rho[:, None] * c[:, 5:]
            qq = np.ascontiguousarray(np.stack([q, q], axis=1)).ravel()
        case.init.append((pv.ravel(), dv.ravel(), tv.ravel() if ntv else None, qq))
    return case


# ------------------------------------------------------------------ the configs
def bump_params(turbulent, cavitating, nscheme, nmuscl, nlim, cfl=5.0, sigma=0.8, uref=10.0, nsteady=0,
                dt_phy=1.0, prec=1):
    """Run parameters of the bump channel (water at 300 K) as functions of the
    cavitation number sigma and the inflow speed U_in = uref:
    p_ref = p_sat(300 K) + sigma rho_l U^2 / 2 (cavitating) or 101325 Pa,
    the boundary state (p_ref, U_in, 0, 0, 300 K, 0, 0, k, omega) with
    k = 1.5 (0.01 U)^2 and omega = rho k / (10 mu) (SST), and the
    mass-transfer scales dp_ref = 2/(rho_ref U^2), t_ref = U/H
    (src/cav_module.f90:81-82) at rho_ref = rho(p_ref, 300 K)."""
    iturb = 0 if turbulent else -2
    if cavitating:
        rho_l = float(eos_host(1.0e5, 300.0, 0.0, 0.0)[0][0, 0])
        psat = sat_pressure(300.0)
        pref = psat + sigma * 0.5 * rho_l * uref ** 2
        ncav = 1
    else:
        pref = 101325.0
        ncav = 0

    def sst_inflow(rho, mu):
        k = 1.5 * (0.01 * uref) ** 2
        emut = 10.0 * mu
        return k, rho * k / emut, emut

    p = make_params(iturb, nscheme, nmuscl, nlim, prec, ncav, nsteady, cfl, pref, uref, dt_phy=dt_phy,
                    turb_ref=sst_inflow if turbulent else None)
    p["bump"] = dict(turbulent=turbulent, cavitating=cavitating, nscheme=nscheme, nmuscl=nmuscl, nlim=nlim, cfl=cfl,
                     sigma=sigma, uref=uref, nsteady=nsteady, dt_phy=dt_phy, prec=prec)
    return p


# the run parameters the residual reads, in the order of cav3d_rhs_vjp_params
PARAM_KEYS = ["pref", "uref", "dp_ref", "t_ref"] + [f"bc_pv{q}" for q in range(9)] + ["bc_pressure"]


def param_vector(p):
    return np.array([p["pref"], p["uref"], p["dp_ref"], p["t_ref"], *list(p["bc_pv"]), p["bc_pressure"]])


def param_derivatives(case, which, rel=1e-5):
    """d(run parameters)/d(which) for which = "sigma" or "uref" (U_in) of a
    bump-channel case, in PARAM_KEYS order: central differences of the
    host-side parameter map bump_params (smooth closed forms and the EOS at
    the reference state), step rel * |value|."""
    b = dict(case.params["bump"])
    v0 = b[which]
    h = rel * abs(v0)
    b[which] = v0 + h
    hi = param_vector(bump_params(**b))
    b[which] = v0 - h
    lo = param_vector(bump_params(**b))
    return (hi - lo) / (2.0 * h)


def with_param(case, which, value):
    """A copy of a bump-channel case (same grid and state) whose run
    parameters are those at `which` = value."""
    import copy
    b = dict(case.params["bump"])
    b[which] = value
    p = bump_params(**b)
    for k in ("cav_clip", "precision", "ides"):
        if k in case.params:
            p[k] = case.params[k]
    p["ncav"] = case.params["ncav"]
    c2 = copy.copy(case)
    c2.params = p
    return c2


def bump_case(ni, nj, nk, turbulent, cavitating, nscheme, nmuscl, nlim, cfl=5.0, sigma=0.8, uref=10.0,
              nsteady=0, dt_phy=1.0, prec=1, seed=1, name="bump"):
    """Bump channel of height H=0.1 m with water at 300 K."""
    p = bump_params(turbulent, cavitating, nscheme, nmuscl, nlim, cfl=cfl, sigma=sigma, uref=uref, nsteady=nsteady,
                    dt_phy=dt_phy, prec=prec)
    npv = 9 if turbulent else 7
    ntv = 3 if turbulent else 2
    ngrd = 5 if turbulent else 4
    b = bump_channel(ni, nj, nk)
    finish_block(b, ngrd)
    case = Case(name, p, [b], npv, ntv, ngrd)
    pocket = None
    if cavitating:
        def pocket(xc):
            r2 = (xc[:, 0] - 2 * H) ** 2 + (xc[:, 1] - 0.3 * H) ** 2 + (xc[:, 2] - 0.5 * H) ** 2
            return 0.6 * np.exp(-r2 / (0.2 * H) ** 2)
    init_state(case, seed, pocket)
    return case


def sat_pressure(t):
    """IF97 region-4 saturation pressure (same closed form as the EOS)."""
    tt = min(max(t, 273.15), 647.096)
    th = tt - 0.23855557567849e0 / (tt - 0.65017534844798e3)
    a = th * th + 0.11670521452767e4 * th - 0.72421316703206e6
    b = -0.17073846940092e2 * th * th + 0.12020824702470e5 * th - 0.32325550322333e7
    c = 0.14915108613530e2 * th * th - 0.48232657361591e4 * th + 0.40511340542057e6
    x = 2.0 * c / (-b + math.sqrt(b * b - 4.0 * a * c))
    return 1.0e6 * x ** 4


def config1(ni=64, nj=32, nk=32, seed=1):
    """Laminar IF97 liquid, Roe + low-Mach preconditioning, MUSCL minmod."""
    return bump_case(ni, nj, nk, turbulent=False, cavitating=False, nscheme=1, nmuscl=1, nlim=1, cfl=5.0,
                     seed=seed, name=f"config1_{ni}x{nj}x{nk}")


def config2(ni=128, nj=64, nk=64, seed=2):
    """Cavitating bump: AUSM+-up, MLP (5th-order limiter), k-omega SST, Merkle, sigma = 0.8.
    CFL 3: with the species clip the 128x64x64 case runs 500 pseudo-time
    iterations with a falling residual (at CFL 5 it diverges near the bump
    crest after about 35)."""
    return bump_case(ni, nj, nk, turbulent=True, cavitating=True, nscheme=4, nmuscl=2, nlim=8, cfl=3.0,
                     sigma=0.8, seed=seed, name=f"config2_{ni}x{nj}x{nk}")


# ------------------------------------------------------------ multiblock
def connection(ipnts, ipnts_donor, transvec, donor):
    """Interface descriptor (22 ints) from a CGNS 1-to-1 connection: the
    point ranges as grid_module converts them (src/grid_module.f90:222-300)
    and the three-layer extension of bc_module (src/bc_module.f90:355-488)."""
    T = np.zeros((3, 3), dtype=int)
    for c in range(3):
        T[abs(transvec[c]) - 1, c] = int(np.sign(transvec[c]))
    ist, ien = list(ipnts[:3]), list(ipnts[3:])
    isd, ied = list(ipnts_donor[:3]), list(ipnts_donor[3:])
    for m in range(3):
        if ipnts[m] == ipnts[3 + m]:
            if ipnts[m] == 1:
                ist, ien = list(ipnts[:3]), list(ipnts[3:])
                ist[m] += 1
                ien[m] += 1
        if ipnts_donor[m] == ipnts_donor[3 + m] and ipnts_donor[m] != 1:
            isd, ied = list(ipnts_donor[:3]), list(ipnts_donor[3:])
            isd[m] += 1
            ied[m] += 1
    # calnormal's closing re-ordering moves the tangential ranges from vertex
    # to cell indices (src/grid_module.f90:924-938)
    for m in range(3):
        if ist[m] != ien[m]:
            ist[m] += 1
        if isd[m] < ied[m]:
            isd[m] += 1
        elif isd[m] > ied[m]:
            ied[m] += 1
    for m in range(3):                          # the face direction of this block
        if ist[m] == ien[m]:
            # the donor index that this block's normal direction maps onto
            col = T[:, m]
            q = int(np.nonzero(col)[0][0])
            if ist[m] == 2:
                ien[m] += 2
                if col[q] == 1:
                    ied[q] += 2
                else:
                    ied[q] -= 2
            else:
                ist[m] -= 2
                if col[q] == 1:
                    isd[q] -= 2
                else:
                    isd[q] += 2
            break
    return [donor, *ist, *ien, *isd, *ied, *T.reshape(-1).tolist()]


def exchange_i_metrics(L, R, ngrd):
    """The metric exchange across an i-cut with aligned indices, L's imax face
    joined to R's imin face: each block's ghost layer takes the other's first
    interior layer, as calnormal (tangential face vectors ex, tx over vertex
    ranges j = 1..J, k = 1..K; src/grid_module.f90:734-840), calvolume (grd(1)
    over cells 2..J, 2..K; :1127-1190) and calydns (grd(5); :509-560) do."""
    IL, IR, J, K = L.imax, R.imax, L.jmax, L.kmax
    exL, exR = L.ex.reshape(K + 1, J, IL + 1, 3), R.ex.reshape(K + 1, J, IR + 1, 3)
    txL, txR = L.tx.reshape(K, J + 1, IL + 1, 3), R.tx.reshape(K, J + 1, IR + 1, 3)
    gL, gR = L.grd.reshape(K + 1, J + 1, IL + 1, ngrd), R.grd.reshape(K + 1, J + 1, IR + 1, ngrd)
    fromL = (exL[:K, :J, IL - 1].copy(), txL[:K, :J, IL - 1].copy(), gL[1:K, 1:J, IL - 1].copy())
    fromR = (exR[:K, :J, 1].copy(), txR[:K, :J, 1].copy(), gR[1:K, 1:J, 1].copy())
    exR[:K, :J, 0], txR[:K, :J, 0] = fromL[0], fromL[1]
    exL[:K, :J, IL], txL[:K, :J, IL] = fromR[0], fromR[1]
    qs = [0, 4] if ngrd > 4 else [0]
    gR[1:K, 1:J, 0, qs] = fromL[2][..., qs]
    gL[1:K, 1:J, IL, qs] = fromR[2][..., qs]


def split_i(case, cuts):
    """The same case cut into blocks along i at the given vertex indices
    (1-based), with interfaces between consecutive blocks.  Metrics are
    computed per block (wall distance over all walls) and exchanged across
    the cuts; the state is the single block's, cut."""
    (b0,) = case.blocks
    I, J, K = b0.imax, b0.jmax, b0.kmax
    bounds = [1] + list(cuts) + [I]
    walls = wall_face_centres(b0.X, b0.bc_desc)
    kinds = {}
    for d in b0.bc_desc:
        s, e = d[1:4], d[4:7]
        face = next(f for f, (ax, v) in {"imin": (0, 1), "imax": (0, I + 1), "jmin": (1, 1), "jmax": (1, J + 1),
                                          "kmin": (2, 1), "kmax": (2, K + 1)}.items() if s[ax] == e[ax] == v)
        kinds[face] = d[0]
    inv = {v: k for k, v in BC.items()}
    blocks, init = [], []
    n = len(bounds) - 1
    for q in range(n):
        lo, hi = bounds[q], bounds[q + 1]
        X = b0.X[lo - 1:hi]
        Ib = hi - lo + 1
        nb = Block(Ib, J, K, X)
        for face, kind in kinds.items():
            if face == "imin" and q > 0:
                continue
            if face == "imax" and q < n - 1:
                continue
            nb.bc_desc.append(face_bc(inv[kind], face, Ib, J, K))
        if q > 0:      # imin face joins the previous block's imax face
            prevI = bounds[q] - bounds[q - 1] + 1
            nb.conn_desc.append(connection([1, 1, 1, 1, J, K], [prevI, 1, 1, prevI, J, K], [1, 2, 3], q - 1))
        if q < n - 1:  # imax face joins the next block's imin face
            nb.conn_desc.append(connection([Ib, 1, 1, Ib, J, K], [1, 1, 1, 1, J, K], [1, 2, 3], q + 1))
        finish_block(nb, case.ngrd, walls)
        if q > 0:
            exchange_i_metrics(blocks[-1], nb, case.ngrd)
        blocks.append(nb)
        sl = slice(lo - 1, hi + 5)          # ghosted i range of the sub-block: lo-2 .. hi+3 (Fortran)
        pv, dv, tv, qq = case.init[0]
        P = pv.reshape(K + 5, J + 5, I + 5, case.npv)[:, :, sl]
        Dv = dv.reshape(K + 5, J + 5, I + 5, 18)[:, :, sl]
        Tv = tv.reshape(K + 5, J + 5, I + 5, -1)[:, :, sl] if tv is not None else None
        Q = None
        if qq is not None:
            Q = qq.reshape(K - 1, J - 1, I - 1, 2, case.npv)[:, :, lo - 1:hi - 1]
            Q = np.ascontiguousarray(Q).ravel()
        init.append((np.ascontiguousarray(P).ravel(), np.ascontiguousarray(Dv).ravel(),
                     None if Tv is None else np.ascontiguousarray(Tv).ravel(), Q))
    out = Case(case.name + f"_split{n}", dict(case.params), blocks, case.npv, case.ntv, case.ngrd)
    out.init = init
    return out


# ------------------------------------------------------------ NACA0015 C-grid (config 3)
def naca0015_thickness(x, c):
    """Half thickness of the NACA 0015 section with the closed trailing edge
    (last coefficient -0.1036), chord c."""
    s = np.clip(x / c, 0.0, 1.0)
    return 0.75 * c * (0.2969 * np.sqrt(s) - 0.1260 * s - 0.3516 * s ** 2 + 0.2843 * s ** 3 - 0.1036 * s ** 4)


def _geometric(n, first):
    """n+1 points from 0 to 1, the first step `first`, growing by a constant ratio."""
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if first * (r ** n - 1.0) / (r - 1.0) > 1.0:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    s = np.concatenate([[0.0], first * np.cumsum(r ** np.arange(n))])
    return s / s[-1]


def naca_cgrid(ni, nj, nk, chord=0.1, span=None, radius=5.0, wake=5.0, first=2.0e-6):
    """Four blocks of ni x nj x nk cells around a NACA 0015 section extruded
    over `span` (default one chord): block 0 the lower wake (outflow -> trailing
    edge), 1 the lower surface (trailing -> leading edge), 2 the upper surface,
    3 the upper wake.  i runs along the C line, j from the section / wake line
    out to the far field at about `radius` chords, k along the span.  The j
    lines leave the section along its outward normal where that normal leans
    upstream, straight down (up) elsewhere, with geometric clustering from a
    first spacing `first` [m] (y+ about 1 at Re 1.2e6 for the default).
    Boundaries: the section is BCWall (the force-integration surface), the
    outer boundary and both outflow planes BCFarfield, the span ends
    BCSymmetryPlane; blocks 0-1, 1-2, 2-3 join at aligned i faces and the
    wake cut joins block 0's jmin face to block 3's with i reversed (CGNS
    transform (-1, -2, 3))."""
    span = chord if span is None else span
    R, Lw = radius * chord, wake * chord
    # inner C line, vertices along i for each block
    tw = _geometric(ni, 1.0 / (4.0 * ni))[::-1]              # clustered toward the trailing edge
    x_wake = chord + Lw * tw                                 # outflow -> trailing edge
    beta = np.linspace(0.0, math.pi, ni + 1)
    x_foil = chord * 0.5 * (1.0 + np.cos(beta))              # trailing -> leading edge, cosine spacing
    lines = []
    for side in (-1.0, 1.0):
        xf = x_foil if side < 0 else x_foil[::-1]
        yf = side * naca0015_thickness(xf, chord)
        xw = x_wake if side < 0 else x_wake[::-1]
        lines.append((np.stack([xf, yf], 1), np.stack([xw, np.zeros_like(xw)], 1)))
    (lo_foil, lo_wake), (up_foil, up_wake) = lines
    inner = [lo_wake, lo_foil, up_foil, up_wake]
    # j-line directions: the outward surface normal where it leans upstream,
    # else straight out (monotone along the C line, so the lines do not cross)
    dirs = []
    for q, P in enumerate(inner):
        side = -1.0 if q < 2 else 1.0
        d = np.zeros_like(P)
        d[:, 1] = side
        if q in (1, 2):
            t = np.gradient(P, axis=0)
            n = np.stack([-t[:, 1], t[:, 0]], 1)              # outward for the C line's direction
            n /= np.linalg.norm(n, axis=1, keepdims=True)
            lean = n[:, 0] < 0.0                                # pointing upstream
            d[lean] = n[lean]
            d[np.isclose(P[:, 0], 0.0)] = np.array([-1.0, 0.0])   # the leading edge
            d[np.isclose(P[:, 0], chord)] = np.array([0.0, side])  # the trailing edge
        dirs.append(d)
    sj = _geometric(nj, first / R)
    z = np.linspace(0.0, span, nk + 1)
    I, J, K = ni + 1, nj + 1, nk + 1
    blocks = []
    for q in range(4):
        P, d = inner[q], dirs[q]
        L = R - np.abs(P[:, 1])
        X = np.zeros((I, J, K, 3))
        X[..., 0] = (P[:, 0][:, None] + sj[None, :] * (L * d[:, 0])[:, None])[:, :, None]
        X[..., 1] = (P[:, 1][:, None] + sj[None, :] * (L * d[:, 1])[:, None])[:, :, None]
        X[..., 2] = z[None, None, :]
        bcs = [face_bc("farfield", "jmax", I, J, K), face_bc("symmetry", "kmin", I, J, K),
               face_bc("symmetry", "kmax", I, J, K)]
        if q in (1, 2):
            w = face_bc("wall", "jmin", I, J, K)
            w[0] |= SURFACE
            bcs.append(w)
        if q == 0:
            bcs.append(face_bc("farfield", "imin", I, J, K))
        if q == 3:
            bcs.append(face_bc("farfield", "imax", I, J, K))
        b = Block(I, J, K, X, bcs, [])
        if q > 0:
            b.conn_desc.append(connection([1, 1, 1, 1, J, K], [I, 1, 1, I, J, K], [1, 2, 3], q - 1))
        if q < 3:
            b.conn_desc.append(connection([I, 1, 1, I, J, K], [1, 1, 1, 1, J, K], [1, 2, 3], q + 1))
        if q == 0:
            b.conn_desc.append(connection([1, 1, 1, I, 1, K], [I, 1, 1, 1, 1, K], [-1, -2, 3], 3))
        if q == 3:
            b.conn_desc.append(connection([1, 1, 1, I, 1, K], [I, 1, 1, 1, 1, K], [-1, -2, 3], 0))
        blocks.append(b)
    return blocks


def config3(ni=128, nj=160, nk=96, seed=3, cfl=50.0):
    """Cavitating NACA 0015 wing section: 4 blocks of ni x nj x nk cells on a
    C-grid (chord 0.1 m, span one chord), angle of attack 8 deg, U = 12 m/s,
    sigma = 1.0 (p_ref = p_sat(300 K) + sigma rho U^2 / 2), Re = 1.2e6 on the
    chord, k-omega SST (1 % intensity, mu_t/mu = 10 at the inflow), AUSM+-up,
    MLP, Merkle, preconditioned LU-SGS at CFL 50; uniform free stream with
    1e-3 conserved-variable noise.  Metrics come from the vertex coordinates
    (Solver(case, metrics="device")); the blocks carry no host metrics."""
    uref, chord = 12.0, 0.1
    rho_l = float(eos_host(1.0e5, 300.0, 0.0, 0.0)[0][0, 0])
    pref = sat_pressure(300.0) + 1.0 * 0.5 * rho_l * uref ** 2

    def sst_inflow(rho, mu):
        k = 1.5 * (0.01 * uref) ** 2
        emut = 10.0 * mu
        return k, rho * k / emut, emut

    p = make_params(0, 4, 2, 8, 1, 1, 0, cfl, pref, uref, aoa=8.0, l_chord=chord, l_character=chord,
                    l_domain=chord, turb_ref=sst_inflow)
    p["cav_clip"] = 1
    p["area"] = chord * chord
    p["aoa"] = 8.0
    blocks = naca_cgrid(ni, nj, nk, chord=chord)
    case = Case(f"config3_4x{ni}x{nj}x{nk}", p, blocks, 9, 3, 5)
    init_state(case, seed)
    return case
