"""CPU oracle for the Allen-Cahn phase-field LBM step (numpy restatement).

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module;
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
only as the checker (or as the timed CPU baseline), never as the thing measured
or shipped.

It restates, in plain numpy, the behaviour of JunjunQing/pytorch-lbm:
  lattice tables           lbm/lattice.py:36-96
  derived parameters       lbm/fe_config.py:114-163
  geometry precompute      lbm/fe_wetting.py:122-311 (is_boundary, normals,
                           near-wall distance/weight, per-direction masks),
                           lbm/fe_ac_solver.py:193-237 (partial / eps)
  init_fields              lbm/fe_ac_solver.py:239-261
  stencil                  lbm/fe_ac_solver.py:296-320 + fe_wetting.py:418-446
  chemical potential       lbm/fe_ac_solver.py:326-340
  gradient correction      lbm/fe_wetting.py:448-539
  step (S1..S13)           lbm/fe_ac_solver.py:380-639, including the
                           stab_mode variants and body force (:415-439), the
                           Cahn-Hilliard source and mu drive (:457-486, with
                           the gradient of mu, :267-280) and direct wetting
                           (:609-613, fe_wetting.py:313-416)
  inject_droplet           lbm/fe_ac_solver.py:645-732
  diagnostics              lbm/fe_measure.py:11-241

Arithmetic is done in the run dtype (float32 or float64) with the same
rounding of Python scalars that the reference's tensor ops perform (a Python
float meeting a float32 array is rounded to float32 first; numpy >= 2 follows
the same rule).  Sums over the lattice directions and over the vector
components run sequentially in index order.  Values the reference keeps in
float32 even in a float64 run (lattice weights, wall normals, n.e, wall
weight) are kept in float32 here too.

Parity pins (tests/test_oracle_*.py):
  * golden vectors written by tests/golden/make_golden.py from the reference
    on CPU: geometry masks bit-exact, operator outputs, 60-step trajectories
    (every solver option above, and the state right after an injection);
  * closed forms: lattice moment identities, derived constants, MMS gradient
    error numbers published with the reference (N=32: 1.33e-3, N=128: 2.09e-5);
  * np.roll / np.gradient library semantics for streaming and normals;
  * invariants: mass after every step, constant fields, correction no-op where
    the wall normal is zero; for inject_droplet, a drop injected into an empty
    domain (phi = C on the core, equilibrium moments), untouched nodes outside
    the core, idempotence of max(phi, C).
Parity unpinned: none of the functions below; the fp32 summation order inside
torch reductions (Sum over directions, einsum) is not reproducible and is
covered by tolerance, not bit-exactness.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# Lattices (lbm/lattice.py:36-96).  Weights are float32 as in the reference.
# ---------------------------------------------------------------------------
D2Q9_E = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1],
                   [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int64)
D2Q9_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float32)
D2Q9_OPP = np.array([0, 2, 1, 4, 3, 7, 8, 5, 6], dtype=np.int64)

D3Q19_E = np.array([[0, 0, 0],
                    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                    [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0],
                    [1, 0, 1], [-1, 0, 1], [-1, 0, -1], [1, 0, -1],
                    [0, 1, 1], [0, -1, 1], [0, -1, -1], [0, 1, -1]], dtype=np.int64)
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
D3Q19_OPP = np.array([0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12,
                      17, 18, 15, 16], dtype=np.int64)

CS2 = 1.0 / 3.0


def lattice(dim):
    return (D2Q9_E, D2Q9_W, D2Q9_OPP) if dim == 2 else (D3Q19_E, D3Q19_W, D3Q19_OPP)


# ---------------------------------------------------------------------------
# Derived parameters (lbm/fe_config.py:114-150)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    nx: int
    ny: int
    nz: int = 1
    rho_l: float = 828.0
    rho_g: float = 1.0
    sigma: float = 0.01
    xi: float = 5.0
    beta: float = 0.0
    kappa: float = 0.0
    tau_l: float = 0.0
    tau_g: float = 0.6
    tau_h: float = 1.0
    theta_eq: float = 162.0
    phi_c: float = 0.0
    g_force: Tuple[float, ...] = (0.0, 0.0, 0.0)

    def __post_init__(self):
        if self.sigma > 0.0 and self.beta == 0.0:
            self.beta = 12.0 * self.sigma / self.xi
        elif self.sigma == 0.0 and self.beta > 0.0:
            self.sigma = self.beta * self.xi / 12.0
        if self.kappa == 0.0:
            self.kappa = self.beta * self.xi ** 2 / 8.0
        if self.phi_c == 0.0 and self.theta_eq != 90.0:
            self.phi_c = -math.cos(math.radians(self.theta_eq)) * math.sqrt(
                2.0 * self.kappa * self.beta)

    @property
    def dim(self):
        return 2 if self.nz == 1 else 3

    @property
    def shape(self):
        return (self.nx, self.ny) if self.nz == 1 else (self.nx, self.ny, self.nz)


def wall_target(theta):
    """phi target of the boundary relax (lbm/fe_ac_solver.py:165-177)."""
    if theta == 90.0:
        return None
    c = math.cos(math.radians(theta))
    disc = 1.0 - 4.0 * (-c / 3.0)
    if disc >= 0:
        return 0.5 * (1.0 - math.sqrt(disc))
    return 0.5 * (1.0 - math.cos(math.pi - math.radians(theta)))


def _shift(a, e, sign):
    """a(x - sign*e) on the periodic grid: np.roll by +sign*e along each axis."""
    for d, s in enumerate(e):
        if s:
            a = np.roll(a, int(sign * s), axis=d)
    return a


def neighbor(a, e):
    """a(x + e), periodic (torch.roll(shifts=-e) in the reference)."""
    return _shift(a, e, -1)


def upstream(a, e):
    """a(x - e), periodic (torch.roll(shifts=+e))."""
    return _shift(a, e, +1)


# ---------------------------------------------------------------------------
# Geometry precompute (lbm/fe_wetting.py:122-311, lbm/fe_ac_solver.py:205-224)
# ---------------------------------------------------------------------------
@dataclass
class Geometry:
    solid: np.ndarray
    is_boundary: np.ndarray
    wall_normal: np.ndarray          # float32 (D, *shape)
    wall_weight: np.ndarray          # float32
    near_wall: np.ndarray
    dist: np.ndarray                 # float32 distance field (0..11 or 100)
    masks: list                      # per direction: boundary & solid(x+e_i)
    n_dot_e: list                    # per direction float32, None if no mask
    partial: Optional[np.ndarray]
    frac: Optional[np.ndarray]
    mode: Optional[str]              # 'linear' | 'quadratic' | None (no wetting)


def boundary_nodes(solid, e):
    b = np.zeros_like(solid)
    for i in range(1, len(e)):
        b |= upstream(solid, e[i]) & ~solid
    return b


def smooth_normals(frac, boundary):
    """-grad(frac)/|grad(frac)| via np.gradient (non-periodic), fp32 storage."""
    grads = np.gradient(frac)
    if frac.ndim == 1:
        grads = [grads]
    mag = np.sqrt(sum(g ** 2 for g in grads))
    mag[mag < 1e-10] = 1e-10
    return np.stack([(-g / mag) * boundary for g in grads]).astype(np.float32)


def discrete_normals(solid, boundary, e):
    """Neighbour counting with non-periodic bounds (lbm/fe_wetting.py:152-203)."""
    D = solid.ndim
    n = np.zeros((D,) + solid.shape, dtype=np.float32)
    for c in np.argwhere(boundary):
        v = np.zeros(D)
        cnt = 0
        for i in range(1, len(e)):
            p = c + e[i]
            if np.all(p >= 0) and np.all(p < np.array(solid.shape)):
                if solid[tuple(p)]:
                    v -= e[i]
                    cnt += 1
        if cnt > 0:
            v = v / cnt
            nrm = np.sqrt(sum(v[d] ** 2 for d in range(D)))
            if nrm > 1e-10:
                v = v / nrm
        for d in range(D):
            n[(d,) + tuple(c)] = v[d]
    return n


def near_wall_fields(solid, boundary, xi_wall=5.0):
    dist = np.full(solid.shape, 100.0, dtype=np.float32)
    dist[boundary] = 0.0
    for _ in range(int(2 * xi_wall) + 1):
        new = dist.copy()
        for d in range(solid.ndim):
            for s in (-1, 1):
                new = np.minimum(new, np.roll(dist, s, axis=d) + np.float32(1.0))
        dist = new
    dist[solid] = 100.0
    # torch's float32 exp is correctly rounded on these 13 arguments
    # (checked bit-exact against the golden wall_weight); evaluate in float64
    # and round once.
    arg = -dist / np.float32(xi_wall / 2.0)
    w = np.exp(arg.astype(np.float64)).astype(np.float32)
    w[solid] = 0.0
    near = (dist < 2.0 * xi_wall) & ~solid
    return dist, w, near


def make_geometry(p: Params, solid, frac=None):
    e, w, opp = lattice(p.dim)
    solid = np.asarray(solid, dtype=bool)
    partial = None
    if frac is not None:
        partial = (frac > 0.01) & (frac < 0.99) & ~solid
    mode = None
    if p.theta_eq != 90.0 and p.kappa > 0:
        mode = 'linear' if (p.theta_eq > 130.0 or p.theta_eq < 50.0) else 'quadratic'
    b = boundary_nodes(solid, e)
    if frac is not None:
        nrm = smooth_normals(frac, b)
    else:
        nrm = discrete_normals(solid, b, e)
    dist, ww, near = near_wall_fields(solid, b)
    masks, nde = [None], [None]
    for i in range(1, len(e)):
        m = neighbor(solid, e[i]) & b
        masks.append(m)
        if m.any():
            acc = np.zeros(solid.shape, dtype=np.float32)
            for d in range(p.dim):
                acc = acc + nrm[d] * np.float32(e[i, d])
            nde.append(acc)
        else:
            nde.append(None)
    return Geometry(solid, b, nrm, ww, near, dist, masks, nde, partial, frac, mode)


def next_interior(solid, boundary, normal):
    """Per boundary node, the node one rounded wall-normal step into the fluid
    (coordinates clamped to the domain, not wrapped) and whether it is fluid
    (lbm/fe_wetting.py:313-359).  Returns (bnd_coords, nsw_coords, nsw_fluid)."""
    bc = np.argwhere(boundary)
    shift = np.stack([np.rint(normal[d][tuple(bc.T)]) for d in range(solid.ndim)], 1)
    nc = bc + shift.astype(np.int64)
    for d in range(solid.ndim):
        nc[:, d] = np.clip(nc[:, d], 0, solid.shape[d] - 1)
    return bc, nc, ~solid[tuple(nc.T)]


def cubic_boundary_phi(phi, bc, nc, nfluid, a):
    """Boundary phi from the cubic wall energy (lbm/fe_wetting.py:369-416):
    a phi_b^2 + (1 - a) phi_b - phi_nsw = 0 solved for phi_b in the run
    dtype, where the next-interior node is fluid and 0.01 < phi_b,old < 0.99;
    phi elsewhere."""
    if len(bc) == 0 or abs(a) < 1e-10:
        return phi
    dt = phi.dtype.type
    pn = phi[tuple(nc.T)]
    pold = phi[tuple(bc.T)]
    disc = dt((1.0 - a) ** 2) + dt(4.0 * a) * pn
    disc = np.maximum(disc, dt(0.0))
    pb = (dt(-(1.0 - a)) + np.sqrt(disc)) / dt(2.0 * a)
    pb = np.clip(pb, dt(0.0), dt(1.0))
    act = nfluid & (pold > dt(0.01)) & (pold < dt(0.99))
    out = phi.copy()
    sel = bc[act]
    out[tuple(sel.T)] = pb[act]
    return out


# ---------------------------------------------------------------------------
# 16-bit storage model (the build's precision extension; no reference
# counterpart): f* is stored as f*_i(y) - w_i phi(y + e_i), rounded to
# nearest-even in bf16 or fp16, then reconstructed in fp32.  g* is stored as
# is in fp16; in bf16 as g*_i(y) - (3 w_i)(e_i . u(y)) with the writing
# node's new velocity (the velocity-carrying odd term of g_eq, which bf16's
# 8-bit mantissa would otherwise round away), and the same reference is added
# back when the slot is read.
# ---------------------------------------------------------------------------
def round_bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def round16(a, kind):
    if kind == 'fp16':
        return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)
    return round_bf16(a)


# ---------------------------------------------------------------------------
# The solver state machine
# ---------------------------------------------------------------------------
class OracleSolver:
    """numpy restatement of AllenCahnSolver for the default Allen-Cahn path.

    Options mirror lbm/fe_ac_solver.py:54-58 (stab_mode fakhari|liang|none,
    boundary_relax, ghost_scale, sharpen_reduce, mu_wall_scale,
    geometric_wetting, geo_amplification scalar or per-node array, ac_scale).
    """

    def __init__(self, p: Params, dtype=np.float32, stab_mode='fakhari',
                 boundary_relax=0.3, ghost_scale=None, sharpen_reduce=0.8,
                 mu_wall_scale=1.0, geometric_wetting=False,
                 geo_amplification=1.0, ac_scale=1.0, storage=None, direct_wetting=False,
                 interface_mode='allen-cahn', mu_drive=0.0):
        self.p = p
        self.direct_wetting = direct_wetting
        self.interface_mode = interface_mode
        self.mu_drive = mu_drive
        self.storage = storage          # None | 'bf16' | 'fp16' (fp32 compute)
        if storage is not None:
            dtype = np.float32
        self.dt = np.dtype(dtype)
        self.stab_mode = 'fakhari' if stab_mode == 'auto' else stab_mode
        self.boundary_relax = boundary_relax
        self.ghost_scale = ghost_scale
        self.sharpen_reduce = sharpen_reduce
        self.mu_wall_scale = mu_wall_scale
        self.geometric_wetting = geometric_wetting
        self.geo_amp = geo_amplification
        self.ac_scale = ac_scale
        self.e, w32, self.opp = lattice(p.dim)
        self.Q = len(self.e)
        self.D = p.dim
        self.w = w32.astype(self.dt)
        self.ef = self.e.astype(self.dt)
        self.tau_f = p.tau_h if p.tau_h > 0 else p.tau_l
        self.rho_floor = 0.5 * p.rho_l
        self.target = wall_target(p.theta_eq)
        self.geom = None
        self.M0 = None
        self.phi_u_prev = None
        self.step_count = 0
        shape = p.shape
        self.solid = np.zeros(shape, dtype=bool)
        self.wet = None

    # -- geometry ----------------------------------------------------------
    def set_solid(self, solid, frac=None):
        g = make_geometry(self.p, solid, frac)
        self.geom = g
        self.nsw = next_interior(g.solid, g.is_boundary, g.wall_normal)
        self.solid = g.solid
        self.partial_eps = None
        if g.partial is not None and g.partial.any():
            self.partial_eps = np.asarray(frac)[g.partial].astype(self.dt)
        self.wet = g.mode
        if self.wet == 'linear':
            c = math.cos(math.radians(self.p.theta_eq))
            h = self.p.sigma * c
            gs = self.ghost_scale
            if gs is None:
                gs = 3.0 if abs(self.p.theta_eq - 90.0) > 60.0 else 2.0
            self.hk = h / self.p.kappa * gs
            self.mu_wall = -h * 2.0
        elif self.wet == 'quadratic':
            self.pck = self.p.phi_c / self.p.kappa

    # -- operators -----------------------------------------------------------
    def _ghost(self, phi, i, psi):
        g = self.geom
        m = g.masks[i]
        if self.wet is None or m is None or not m.any():
            return psi
        nde = g.n_dot_e[i]
        if self.wet == 'linear':
            val = phi - (nde * np.float32(self.hk))
        else:
            val = phi + (nde * np.float32(self.pck)) * phi * (1.0 - phi)
        return np.where(m, val.astype(self.dt), psi)

    def grad_lap(self, phi):
        """Isotropic 1-ring gradient and Laplacian (+ wetting ghost values)."""
        dt = self.dt
        grad = np.zeros((self.D,) + phi.shape, dtype=dt)
        lap = np.zeros(phi.shape, dtype=dt)
        woc = (self.w / CS2).astype(dt)
        for i in range(1, self.Q):
            psi = neighbor(phi, self.e[i])
            psi = self._ghost(phi, i, psi)
            diff = psi - phi
            wi = woc[i]
            for d in range(self.D):
                grad[d] += (wi * self.ef[i, d]) * psi
            lap += (dt.type(2.0) * wi) * diff
        return grad, lap

    def grad_only(self, a):
        """Stand-alone gradient (lbm/fe_ac_solver.py:267-280) as the step
        applies it to mu: ((w_i e_id) psi) / cs2 accumulated in direction
        order, with the wetting ghost formula of phi evaluated on `a`."""
        grad = np.zeros((self.D,) + a.shape, dtype=self.dt)
        for i in range(1, self.Q):
            psi = self._ghost(a, i, neighbor(a, self.e[i]))
            for d in range(self.D):
                grad[d] += ((self.w[i] * self.ef[i, d]) * psi) / CS2
        return grad

    def lap_only(self, phi):
        """Stand-alone Laplacian used once by init_fields (fe_ac_solver.py:282-294):
        ((2 w_i) (psi - phi)) / cs2, summed in direction order."""
        lap = np.zeros(phi.shape, dtype=self.dt)
        for i in range(1, self.Q):
            psi = self._ghost(phi, i, neighbor(phi, self.e[i]))
            lap += ((self.dt.type(2.0) * self.w[i]) * (psi - phi)) / CS2
        return lap

    def chem_pot(self, phi, lap):
        p = self.p
        mu = (4.0 * p.beta) * phi * (phi - 1.0) * (phi - 0.5) - p.kappa * lap
        mu = mu.astype(self.dt)
        if self.wet is not None:
            b = self.geom.is_boundary
            if self.wet == 'linear':
                mu[b] += self.mu_wall * self.mu_wall_scale
            else:
                c = phi[b]
                mu[b] += p.phi_c * c * (1.0 - c)
        return mu

    def correct_gradient(self, grad):
        g = self.geom
        if not g.near_wall.any():
            return grad
        th = math.radians(self.p.theta_eq)
        cs, sn = math.cos(th), math.sin(th)
        if abs(sn) < 1e-10:
            return grad
        dt = self.dt
        n = g.wall_normal.astype(dt)
        gn = n[0] * grad[0]
        for d in range(1, self.D):
            gn = gn + n[d] * grad[d]
        gt = grad - gn[None] * n
        gtm = gt[0] * gt[0]
        for d in range(1, self.D):
            gtm = gtm + gt[d] * gt[d]
        gtm = np.maximum(np.sqrt(gtm), dt.type(1e-10))
        gd = (-gtm) * cs / sn
        a = self.geo_amp
        if isinstance(a, np.ndarray):
            a = a.astype(dt)
            gd = np.where(a > 1.0, gn + (gd - gn) * a, gd)
        elif a is not None and a > 1.0:
            gd = gn + (gd - gn) * a
        gc = gt + gd[None] * n
        gm = grad[0] * grad[0]
        for d in range(1, self.D):
            gm = gm + grad[d] * grad[d]
        gm = np.sqrt(gm)
        apply = g.near_wall & (gm > 1e-6)
        blend = g.wall_weight.astype(dt)
        corr = grad + blend[None] * (gc - grad)
        return np.where(apply[None], corr, grad).astype(dt)

    def _edot(self, v):
        """(Q, *shape) array of e_i . v, components summed in index order."""
        out = np.empty((self.Q,) + v.shape[1:], dtype=self.dt)
        for i in range(self.Q):
            acc = self.ef[i, 0] * v[0]
            for d in range(1, self.D):
                acc = acc + self.ef[i, d] * v[d]
            out[i] = acc
        return out

    def _usq(self, u):
        s = u[0] * u[0]
        for d in range(1, self.D):
            s = s + u[d] * u[d]
        return s

    def _wv(self):
        return self.w.reshape((-1,) + (1,) * self.D)

    def f_eq(self, phi, u):
        ec = self._edot(u) / CS2
        usq = self._usq(u)
        gam = self._wv() * (((1.0 + ec) + (ec * ec) / 2.0) - usq[None] / (2.0 * CS2))
        return (phi[None] * gam).astype(self.dt)

    def g_eq(self, P, u):
        ec = self._edot(u) / CS2
        usq = self._usq(u)
        s = self._wv() * ((ec + (ec * ec) / 2.0) - usq[None] / (2.0 * CS2))
        return (P[None] * self._wv() + s).astype(self.dt)

    def _gref(self, u):
        """(3 w_i) (e_i . u) in fp32 as the kernels form the bf16 g reference:
        e_i . u summed over the nonzero components in index order."""
        ref = np.zeros((self.Q,) + u.shape[1:], dtype=np.float32)
        for i in range(self.Q):
            acc = None
            for d in range(self.D):
                e = self.e[i][d]
                if e != 0:
                    t = u[d] if e > 0 else -u[d]
                    acc = t if acc is None else acc + t
            if acc is not None:
                ref[i] = np.float32(3.0 * float(self.w[i])) * acc
        return ref

    def _store_g(self, g, u):
        """g as stored (16-bit) and read back; see the storage model above."""
        if self.storage == 'bf16':
            ref = self._gref(u)
            return round16(g - ref, 'bf16') + ref
        return round16(g, self.storage)

    @staticmethod
    def _qsum(a):
        """Sum over the direction axis in the order torch's CPU reduction uses
        (sequential runs of 16 rows, run totals added in order; measured on
        random data: bit-identical to tensor.sum(dim=0) for Q=3, 9 and 19)."""
        tot = None
        for c0 in range(0, a.shape[0], 16):
            s = a[c0].copy()
            for i in range(c0 + 1, min(c0 + 16, a.shape[0])):
                s = s + a[i]
            tot = s if tot is None else tot + s
        return tot

    # -- init -------------------------------------------------------------
    def init_fields(self, phi_init, u_init=None):
        dt = self.dt
        phi = np.clip(np.asarray(phi_init).astype(dt), 0.0, 1.0)
        if u_init is None:
            u = np.zeros((self.D,) + phi.shape, dtype=dt)
        else:
            u = np.asarray(u_init).astype(dt)
        self.phi = phi
        self.rho = (phi * self.p.rho_l + (1.0 - phi) * self.p.rho_g).astype(dt)
        self.u = u
        self.P = np.zeros_like(phi)
        self.M0 = float(np.sum(phi, dtype=np.float64))
        lap = self.lap_only(phi)
        self.mu = self.chem_pot(phi, lap)
        self.f = self.f_eq(phi, u)
        self.g = self.g_eq(self.P, u)
        if self.storage is not None:
            wv = self._wv()
            self.f = round16(self.f - wv * phi[None], self.storage) + wv * phi[None]
            self.g = self._store_g(self.g, u)
        self.phi_u_prev = None

    def inject_droplet(self, center, radius, xi, u_impact=None):
        """Mid-run droplet injection (lbm/fe_ac_solver.py:645-732): where the
        new drop's C > 0.1, phi = max(phi, C); u blends towards u_impact with
        weight C everywhere; f, g re-equilibrate at those core nodes (g with
        the current P); solid nodes keep phi = 0 and u = 0; M0 becomes the new
        sum of phi (fp64 here) and the next step has no d(phi u).  With 16-bit
        storage the whole state is stored again in its deviation form."""
        dt = self.dt
        shape = self.phi.shape
        D = len(shape)
        if u_impact is None:
            u_impact = (0.0,) * D
        mesh = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        dist_sq = 0
        for d, M in enumerate(mesh):
            dist_sq = dist_sq + (M - center[d]) ** 2
        dist = np.sqrt(dist_sq)
        C = np.clip(0.5 + 0.5 * np.tanh(2.0 * (radius - dist) / xi), 0.0, 1.0)
        C = C.astype(np.float32).astype(dt)
        core = C > 0.1
        solid = self.solid
        phi = np.where(core, np.maximum(self.phi, C), self.phi)
        phi = np.clip(phi, 0.0, 1.0).astype(dt)
        phi = np.where(solid, dt.type(0.0), phi)
        u_new = np.zeros((D,) + shape, dtype=dt)
        for d in range(D):
            u_new[d] = u_impact[d]
        Cw = C[None]
        u = ((1.0 - Cw) * self.u + Cw * u_new).astype(dt)
        u = np.where(solid[None], dt.type(0.0), u)
        feq = self.f_eq(phi, u)
        geq = self.g_eq(self.P, u)
        f = np.where(core[None], feq, self.f).astype(dt)
        g = np.where(core[None], geq, self.g).astype(dt)
        if self.storage is not None:
            wv = self._wv()
            f = round16(f - wv * phi[None], self.storage) + wv * phi[None]
            g = self._store_g(g, u)
        self.phi = phi
        self.rho = (phi * self.p.rho_l + (1.0 - phi) * self.p.rho_g).astype(dt)
        self.u = u
        self.f = f
        self.g = g
        self.M0 = float(np.sum(phi, dtype=np.float64))
        self.phi_u_prev = None
        return int(core.sum())

    # -- the step -------------------------------------------------------------
    def step(self, return_stages=False):
        p, dt = self.p, self.dt
        st = {}
        solid = self.solid
        # S1
        phi = np.clip(self._qsum(self.f), 0.0, 1.0).astype(dt)
        phi = np.where(solid, dt.type(0.0), phi)
        # S2
        rho = (phi * p.rho_l + (1.0 - phi) * p.rho_g).astype(dt)
        rho_eff = np.maximum(rho, dt.type(self.rho_floor))
        grad, lap = self.grad_lap(phi)
        mu = self.chem_pot(phi, lap)
        if self.wet is not None and self.geometric_wetting:
            grad = self.correct_gradient(grad)
        ch = self.interface_mode == 'cahn-hilliard'
        gmu = self.grad_only(mu) if (ch or self.mu_drive > 0.0) else None
        st.update(phi=phi, grad=grad, lap=lap, mu=mu, grad_mu=gmu)
        # S3
        P = self._qsum(self.g)
        # S4 (old u)
        gneq = self.g - self.g_eq(P, self.u)
        D = self.D
        sig = np.empty((D, D) + phi.shape, dtype=dt)
        for a in range(D):
            for b in range(D):
                ee = (self.ef[:, a] * self.ef[:, b])
                acc = ee[0] * gneq[0]
                for i in range(1, self.Q):
                    acc = acc + ee[i] * gneq[i]
                sig[a, b] = acc
        # S5
        drho = p.rho_l - p.rho_g
        Fs = mu[None] * grad
        Fp = ((-P[None]) * drho) * CS2 * grad
        vc = (0.5 - p.tau_g) / p.tau_g
        sg = np.empty_like(grad)
        for a in range(D):
            acc = sig[a, 0] * grad[0]
            for b in range(1, D):
                acc = acc + sig[a, b] * grad[b]
            sg[a] = acc
        Fm = (vc * sg) * drho
        gf = tuple(p.g_force[:D])
        if any(x != 0 for x in gf):
            Fb = np.array(gf, dtype=dt).reshape((D,) + (1,) * D) * rho[None]
        else:
            Fb = np.zeros_like(Fs)
        if self.stab_mode == 'fakhari':
            F = ((Fs + Fp) + Fm) + Fb
        elif self.stab_mode == 'liang':
            F = (Fs + Fp) + Fb
        else:
            F = Fs + Fb
        F = F.astype(dt)
        # S6
        ug = np.empty_like(grad)
        for d in range(D):
            acc = self.ef[0, d] * self.g[0]
            for i in range(1, self.Q):
                acc = acc + self.ef[i, d] * self.g[i]
            ug[d] = acc
        u = (ug + (0.5 * F) / rho_eff[None]).astype(dt)
        u = np.where(solid[None], dt.type(0.0), u)
        # S7
        phiu = phi[None] * u
        if self.phi_u_prev is None:
            self.phi_u_prev = phiu.copy()
        dphiu = phiu - self.phi_u_prev
        self.phi_u_prev = phiu.copy()
        mob = CS2 * (self.tau_f - 0.5)
        if ch:
            # lbm/fe_ac_solver.py:457-461
            guo = self.tau_f / (self.tau_f - 0.5)
            src = (dphiu + (mob * (gmu - grad)) * guo).astype(dt)
        else:
            gm = np.sqrt(self._usq(grad))
            gm = np.maximum(gm, dt.type(1e-10))
            nvec = grad / gm[None]
            lam = ((4.0 * phi) * (1.0 - phi)) / p.xi
            sh = ((CS2 * lam[None]) * nvec) * self.ac_scale
            if self.wet is not None:
                fac = 1.0 - self.sharpen_reduce * self.geom.wall_weight.astype(dt)
                sh = sh * fac[None]
            src = (dphiu + sh).astype(dt)
            if self.mu_drive > 0.0:
                # lbm/fe_ac_solver.py:484-486
                src = (src + (self.mu_drive * mob) * gmu).astype(dt)
        acoef = 1.0 - 0.5 / self.tau_f
        Fsrc = ((acoef * self._wv()) * self._edot(src)) / CS2
        # S8
        feq = self.f_eq(phi, u)
        om = 1.0 / self.tau_f
        f = ((self.f - om * (self.f - feq)) + Fsrc).astype(dt)
        if p.tau_g != p.tau_l:
            eta = (phi * p.rho_l * (p.tau_l - 0.5)
                   + (1.0 - phi) * p.rho_g * (p.tau_g - 0.5))
            rd = phi * p.rho_l + (1.0 - phi) * p.rho_g
            teff = (0.5 + eta / rd).astype(dt)
        else:
            teff = np.full(phi.shape, p.tau_l, dtype=dt)
        ns = 1.0 - 0.5 / teff
        G = ((ns[None] * self._wv()) * self._edot(F)) / (rho_eff[None] * CS2)
        geq = self.g_eq(P, u)
        omg = (1.0 / teff).astype(dt)
        g = ((self.g - omg[None] * (self.g - geq)) + G).astype(dt)
        if self.storage is not None:
            # slot (i, y) keeps f*_i(y) - w_i phi(y + e_i); the pulling node x =
            # y + e_i adds back w_i phi(x): the same value, rounded once.
            for i in range(self.Q):
                ref = self.w[i] * neighbor(phi, self.e[i])
                f[i] = round16(f[i] - ref, self.storage) + ref
            g = self._store_g(g, u)
        st.update(P=P, sigma=sig, F=F, u=u, Fsrc=Fsrc, fstar=f.copy(), gstar=g.copy())
        # S9-S10b: stream (pull from x - e_i), bounce-back, volume penalization
        fs, gs = f, g
        fn = np.empty_like(f)
        gn = np.empty_like(g)
        for i in range(self.Q):
            fn[i] = upstream(fs[i], self.e[i])
            gn[i] = upstream(gs[i], self.e[i])
        if solid.any():
            for i in range(self.Q):
                fn[i][solid] = fs[self.opp[i]][solid]
                gn[i][solid] = gs[self.opp[i]][solid]
        if self.partial_eps is not None:
            pm = self.geom.partial
            eps = self.partial_eps
            for i in range(self.Q):
                j = self.opp[i]
                fn[i][pm] = (1.0 - eps) * fn[i][pm] + eps * fs[j][pm]
                gn[i][pm] = (1.0 - eps) * gn[i][pm] + eps * gs[j][pm]
        # S11: mass factor from the S1 phi (sum in float64)
        phi_out = phi
        if self.M0 is not None:
            m = float(np.sum(phi, dtype=np.float64))
            c = self.M0 / max(m, 1e-6) if m > 1e-6 else 1.0
            c = dt.type(c)
            phi_out = np.clip(phi * c, 0.0, 1.0).astype(dt)
            fn = (fn * c).astype(dt)
            st['c'] = c
        # S12: boundary phi relax (target: the wall phi of the contact angle, or
        # with direct_wetting the cubic wall-energy phi of each boundary node)
        if (self.wet is not None and self.boundary_relax > 0
                and (self.direct_wetting
                     or (self.target is not None and abs(self.target - 0.5) > 0.05))):
            b = self.geom.is_boundary
            nw = self.geom.wall_normal.astype(dt)
            nz = np.abs(nw[D - 1])
            nm = np.maximum(np.sqrt(self._usq(nw)), dt.type(1e-10))
            top = (nz / nm) > 0.5
            act = b & top & (phi_out > 0.15) & (phi_out < 0.85)
            if act.any():
                if self.direct_wetting:
                    c = math.cos(math.radians(self.p.theta_eq))
                    a = 6.0 * self.p.sigma * c / self.p.kappa
                    t = cubic_boundary_phi(phi_out, *self.nsw, a)[act]
                else:
                    t = dt.type(self.target)
                pn = np.clip(phi_out[act] + self.boundary_relax * (t - phi_out[act]),
                             0.0, 1.0).astype(dt)
                ssum = self._qsum(fn)[act]
                s = np.clip(pn / np.maximum(ssum, dt.type(1e-10)), 0.5, 2.0).astype(dt)
                phi_out = phi_out.copy()
                phi_out[act] = pn
                for i in range(self.Q):
                    fi = fn[i]
                    fi[act] = fi[act] * s
        self.f, self.g = fn, gn
        self.phi = phi_out
        self.rho = rho
        self.u = u
        self.P = P
        self.mu = mu
        self.step_count += 1
        return st if return_stages else None


# ---------------------------------------------------------------------------
# Diagnostics (lbm/fe_measure.py:11-82; scripts/run_thesis_sweep.py:185-206)
# ---------------------------------------------------------------------------
def spread(phi, solid):
    c = np.where(solid, 0.0, phi)
    inter = c > 0.5
    out = {"volume": float(np.sum(c, dtype=np.float64)), "C_max": float(c.max())}
    if not inter.any():
        out.update(Dx=0.0, Dy=0.0, Dz=0.0, k=0.0, z_com=0.0)
        return out
    co = np.argwhere(inter)
    mn, mx = co.min(0), co.max(0)
    Dx = float(mx[0] - mn[0] + 1)
    Dy = float(mx[1] - mn[1] + 1)
    Dz = float(mx[2] - mn[2] + 1) if phi.ndim == 3 else 0.0
    out.update(Dx=Dx, Dy=Dy, Dz=Dz, k=Dx / Dy if Dy > 0 else 0.0,
               z_com=float(co[:, -1].mean()))
    return out


def contact_radius(phi, solid, cx, cy):
    """Contact-line radius: the largest horizontal distance from (cx, cy) of a
    liquid node (phi > 0.5, not solid) on the first fluid layer above each
    column's topmost solid node (the first-fluid-layer search of
    lbm/fe_measure.py:208-218, taken per column so that curved substrates are
    followed).  0 when no such node exists.  3D only."""
    nx, ny, nz = phi.shape
    has = solid.any(axis=2)
    top = np.where(has, nz - 1 - np.argmax(solid[:, :, ::-1], axis=2), -1)
    z = top + 1
    ok = z < nz
    zc = np.minimum(z, nz - 1)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    p = phi[X, Y, zc]
    s = solid[X, Y, zc]
    liquid = ok & ~s & (p > 0.5)
    if not liquid.any():
        return 0.0
    dx = X.astype(np.float64) - cx
    dy = Y.astype(np.float64) - cy
    r2 = dx * dx + dy * dy
    return float(np.sqrt(r2[liquid].max()))


def base_radius(phi, solid, cx, cy):
    """Base radius of the drop: the largest horizontal distance from (cx, cy)
    of a liquid node (phi > 0.5, not solid) on the lowest layer (last axis)
    that holds liquid; 0 without liquid.  3D only.  The reference's 3D
    contact-angle estimate takes its radius on the first fluid layer above the
    wall (lbm/fe_measure.py:188-241); this follows the drop's own base."""
    liquid = (phi > 0.5) & ~solid
    if not liquid.any():
        return 0.0
    zs = np.nonzero(liquid.any(axis=(0, 1)))[0]
    layer = liquid[:, :, zs[0]]
    xs, ys = np.nonzero(layer)
    dx = xs.astype(np.float64) - cx
    dy = ys.astype(np.float64) - cy
    return float(np.sqrt((dx * dx + dy * dy).max()))
