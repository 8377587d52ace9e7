"""Torch-CPU restatement of the default Allen-Cahn step, for the CPU baseline.

TEST INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg times it and tests/
pins it; nothing in the product package imports it.

The reference's own CPU path is torch with device='cpu' running
AllenCahnSolver.step() (lbm/fe_ac_solver.py:380-639) on every host core
(torch.set_num_threads).  The reference cannot travel to the GPU host, so this
module advances the numpy oracle's state (oracle.ac_oracle.OracleSolver after
set_solid and init_fields) with torch tensors on the CPU instead: the same
operations as OracleSolver.step in the same order, vectorised over the grid,
multi-threaded by torch.  It covers the path the benchmark runs: the
Allen-Cahn interface source without the mu drive, stab_mode 'fakhari' without
a body force, a scalar geo_amplification, linear / quadratic / no wetting, the
boundary relax towards the contact-angle target, bounce-back and volume
penalization.  Pinned against OracleSolver.step by
tests/test_oracle_torch_step.py (the fp64 mass sum runs in torch's order,
so agreement is to rounding, not bit for bit).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from oracle.ac_oracle import CS2


class TorchStep:
    def __init__(self, o):
        if o.interface_mode != 'allen-cahn' or o.mu_drive > 0.0 or o.direct_wetting:
            raise ValueError("TorchStep covers the default Allen-Cahn path only")
        if o.stab_mode != 'fakhari' or any(x != 0 for x in o.p.g_force[:o.D]):
            raise ValueError("TorchStep covers stab_mode 'fakhari' without a body force")
        if isinstance(o.geo_amp, np.ndarray) or o.storage is not None:
            raise ValueError("TorchStep covers a scalar amplification and plain storage")
        self.o = o
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
        self.dt = torch.float64 if o.dt == np.float64 else torch.float32
        self.D, self.Q = o.D, o.Q
        self.e = [tuple(int(v) for v in o.e[i]) for i in range(o.Q)]
        self.opp = [int(v) for v in o.opp]
        self.w = [o.w[i].item() for i in range(o.Q)]          # run-dtype weights
        # w / cs2 rounded to the run dtype, as the stencil takes it
        self.woc = [v.item() for v in (o.w / CS2).astype(o.dt)]
        self.wt = torch.tensor(o.w, dtype=self.dt)
        self.f, self.g, self.u = T(o.f), T(o.g), T(o.u)
        self.phi_u_prev = None if o.phi_u_prev is None else T(o.phi_u_prev)
        self.M0 = o.M0
        self.solid = T(o.solid)
        g = o.geom
        self.wet = o.wet
        self.masks = [None if m is None or not m.any() else T(m) for m in g.masks]
        self.nde = [None if n is None else T(n) for n in g.n_dot_e]
        self.bnd = T(g.is_boundary)
        self.normal = T(g.wall_normal).to(self.dt)
        self.ww = T(g.wall_weight).to(self.dt)
        self.near = T(g.near_wall)
        self.partial = None
        if o.partial_eps is not None:
            self.partial = T(g.partial)
            self.eps = T(o.partial_eps)
        if self.wet == 'linear':
            self.hk32 = np.float32(o.hk).item()
            self.mu_wall = o.mu_wall * o.mu_wall_scale
        elif self.wet == 'quadratic':
            self.pck32 = np.float32(o.pck).item()
        self.top = None
        if (self.wet is not None and o.boundary_relax > 0 and o.target is not None
                and abs(o.target - 0.5) > 0.05):
            nw = self.normal
            nz = nw[self.D - 1].abs()
            nm = torch.clamp(torch.sqrt(self._usq(nw)), min=1e-10)
            self.top = self.bnd & ((nz / nm) > 0.5)

    # -- helpers (same operation order as the numpy oracle) -------------------
    def _roll(self, a, e, sign):
        dims = [d for d in range(self.D) if e[d]]
        if not dims:
            return a
        return torch.roll(a, shifts=[sign * e[d] for d in dims], dims=dims)

    def _usq(self, u):
        s = u[0] * u[0]
        for d in range(1, self.D):
            s = s + u[d] * u[d]
        return s

    def _edot(self, v):
        out = torch.empty((self.Q,) + tuple(v.shape[1:]), dtype=self.dt)
        for i in range(self.Q):
            acc = float(self.e[i][0]) * v[0]
            for d in range(1, self.D):
                acc = acc + float(self.e[i][d]) * v[d]
            out[i] = acc
        return out

    def _qsum(self, a):
        tot = None
        for c0 in range(0, a.shape[0], 16):
            s = a[c0].clone()
            for i in range(c0 + 1, min(c0 + 16, a.shape[0])):
                s = s + a[i]
            tot = s if tot is None else tot + s
        return tot

    def _wv(self):
        return self.wt.reshape((-1,) + (1,) * self.D)

    def _ghost(self, phi, i, psi):
        m = self.masks[i]
        if self.wet is None or m is None:
            return psi
        if self.wet == 'linear':
            val = phi - (self.nde[i] * self.hk32)
        else:
            val = phi + (self.nde[i] * self.pck32) * phi * (1.0 - phi)
        return torch.where(m, val.to(self.dt), psi)

    def step(self):
        o, p, D, Q, dt = self.o, self.o.p, self.D, self.Q, self.dt
        solid = self.solid
        zero = torch.zeros((), dtype=dt)
        # S1-S2
        phi = torch.clamp(self._qsum(self.f), 0.0, 1.0)
        phi = torch.where(solid, zero, phi)
        rho = phi * p.rho_l + (1.0 - phi) * p.rho_g
        rho_eff = torch.clamp(rho, min=o.rho_floor)
        grad = torch.zeros((D,) + tuple(phi.shape), dtype=dt)
        lap = torch.zeros_like(phi)
        for i in range(1, Q):
            psi = self._ghost(phi, i, self._roll(phi, self.e[i], -1))
            wi = self.woc[i]
            for d in range(D):
                if self.e[i][d]:
                    grad[d] += (wi * float(self.e[i][d])) * psi
                else:
                    grad[d] += (wi * 0.0) * psi
            lap += (2.0 * wi) * (psi - phi)
        mu = (4.0 * p.beta) * phi * (phi - 1.0) * (phi - 0.5) - p.kappa * lap
        if self.wet == 'linear':
            mu = torch.where(self.bnd, mu + self.mu_wall, mu)
        elif self.wet == 'quadratic':
            mu = torch.where(self.bnd, mu + p.phi_c * phi * (1.0 - phi), mu)
        if self.wet is not None and o.geometric_wetting:
            grad = self._correct(grad)
        # S3-S4
        P = self._qsum(self.g)
        gneq = self.g - self._g_eq(P, self.u)
        sig = torch.empty((D, D) + tuple(phi.shape), dtype=dt)
        for a in range(D):
            for b in range(D):
                acc = float(self.e[0][a] * self.e[0][b]) * gneq[0]
                for i in range(1, Q):
                    acc = acc + float(self.e[i][a] * self.e[i][b]) * gneq[i]
                sig[a, b] = acc
        # S5 (fakhari, no body force)
        drho = p.rho_l - p.rho_g
        Fs = mu[None] * grad
        Fp = ((-P[None]) * drho) * CS2 * grad
        vc = (0.5 - p.tau_g) / p.tau_g
        sg = torch.empty_like(grad)
        for a in range(D):
            acc = sig[a, 0] * grad[0]
            for b in range(1, D):
                acc = acc + sig[a, b] * grad[b]
            sg[a] = acc
        Fm = (vc * sg) * drho
        F = ((Fs + Fp) + Fm) + torch.zeros_like(Fs)
        # S6
        ug = torch.empty_like(grad)
        for d in range(D):
            acc = float(self.e[0][d]) * self.g[0]
            for i in range(1, Q):
                acc = acc + float(self.e[i][d]) * self.g[i]
            ug[d] = acc
        u = ug + (0.5 * F) / rho_eff[None]
        u = torch.where(solid[None], zero, u)
        # S7
        phiu = phi[None] * u
        if self.phi_u_prev is None:
            self.phi_u_prev = phiu.clone()
        dphiu = phiu - self.phi_u_prev
        self.phi_u_prev = phiu.clone()
        gm = torch.clamp(torch.sqrt(self._usq(grad)), min=1e-10)
        nvec = grad / gm[None]
        lam = ((4.0 * phi) * (1.0 - phi)) / p.xi
        sh = ((CS2 * lam[None]) * nvec) * o.ac_scale
        if self.wet is not None:
            sh = sh * (1.0 - o.sharpen_reduce * self.ww)[None]
        src = dphiu + sh
        acoef = 1.0 - 0.5 / o.tau_f
        Fsrc = ((acoef * self._wv()) * self._edot(src)) / CS2
        # S8
        feq = self._f_eq(phi, u)
        om = 1.0 / o.tau_f
        f = (self.f - om * (self.f - feq)) + Fsrc
        if p.tau_g != p.tau_l:
            eta = (phi * p.rho_l * (p.tau_l - 0.5)
                   + (1.0 - phi) * p.rho_g * (p.tau_g - 0.5))
            rd = phi * p.rho_l + (1.0 - phi) * p.rho_g
            teff = 0.5 + eta / rd
        else:
            teff = torch.full(tuple(phi.shape), p.tau_l, dtype=dt)
        ns = 1.0 - 0.5 / teff
        G = ((ns[None] * self._wv()) * self._edot(F)) / (rho_eff[None] * CS2)
        geq = self._g_eq(P, u)
        g = (self.g - (1.0 / teff)[None] * (self.g - geq)) + G
        # S9-S10b
        fn = torch.empty_like(f)
        gn = torch.empty_like(g)
        for i in range(Q):
            fn[i] = self._roll(f[i], self.e[i], +1)
            gn[i] = self._roll(g[i], self.e[i], +1)
        for i in range(Q):
            j = self.opp[i]
            fn[i] = torch.where(solid, f[j], fn[i])
            gn[i] = torch.where(solid, g[j], gn[i])
        if self.partial is not None:
            pm, eps = self.partial, self.eps
            for i in range(Q):
                j = self.opp[i]
                fn[i][pm] = (1.0 - eps) * fn[i][pm] + eps * f[j][pm]
                gn[i][pm] = (1.0 - eps) * gn[i][pm] + eps * g[j][pm]
        # S11
        m = float(phi.double().sum())
        c = self.M0 / max(m, 1e-6) if m > 1e-6 else 1.0
        c = dt_scalar(dt, c)
        phi_out = torch.clamp(phi * c, 0.0, 1.0)
        fn = fn * c
        # S12
        if self.top is not None:
            act = self.top & (phi_out > 0.15) & (phi_out < 0.85)
            if bool(act.any()):
                t = dt_scalar(dt, o.target)
                pn = torch.clamp(phi_out[act] + o.boundary_relax * (t - phi_out[act]), 0.0, 1.0)
                ssum = self._qsum(fn)[act]
                s = torch.clamp(pn / torch.clamp(ssum, min=1e-10), 0.5, 2.0)
                phi_out = phi_out.clone()
                phi_out[act] = pn
                for i in range(Q):
                    fi = fn[i]
                    fi[act] = fi[act] * s
        self.f, self.g, self.u, self.phi = fn, gn, u, phi_out

    def _correct(self, grad):
        o, D, dt = self.o, self.D, self.dt
        if not bool(self.near.any()):
            return grad
        th = math.radians(o.p.theta_eq)
        cs, sn = math.cos(th), math.sin(th)
        if abs(sn) < 1e-10:
            return grad
        n = self.normal
        gn = n[0] * grad[0]
        for d in range(1, D):
            gn = gn + n[d] * grad[d]
        gt = grad - gn[None] * n
        gtm = torch.clamp(torch.sqrt(self._usq(gt)), min=1e-10)
        gd = (-gtm) * cs / sn
        a = o.geo_amp
        if a is not None and a > 1.0:
            gd = gn + (gd - gn) * a
        gc = gt + gd[None] * n
        gm = torch.sqrt(self._usq(grad))
        apply = self.near & (gm > 1e-6)
        corr = grad + self.ww[None] * (gc - grad)
        return torch.where(apply[None], corr, grad)

    def _f_eq(self, phi, u):
        ec = self._edot(u) / CS2
        usq = self._usq(u)
        gam = self._wv() * (((1.0 + ec) + (ec * ec) / 2.0) - usq[None] / (2.0 * CS2))
        return phi[None] * gam

    def _g_eq(self, P, u):
        ec = self._edot(u) / CS2
        usq = self._usq(u)
        s = self._wv() * ((ec + (ec * ec) / 2.0) - usq[None] / (2.0 * CS2))
        return P[None] * self._wv() + s


def dt_scalar(dt, v):
    """A Python float rounded to the run dtype (as a float32 tensor meeting a
    Python float sees it)."""
    return float(np.float32(v)) if dt == torch.float32 else float(v)
