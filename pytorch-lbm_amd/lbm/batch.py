"""Batched and served sweeps: many droplet-impact cases on one GPU.

The reference runs a sweep case by case (scripts/run_thesis_sweep.py:57-284;
scripts/night_queue.py:28-50 queues whole processes).  Here every case lives on
the device at once and one multi-case launch per pass advances all of them
(pytorch-lbm_amd/csrc/lbm_batch.hpp), so small and large cases share the
machine and no per-case launch or host round trip sits between steps.

``CaseSpec`` describes one case the way the reference's sweep driver builds
it: FEConfig-style physics, a substrate from geometry/substrate.py, a droplet
from lbm/fe_droplet.py placed by the driver's rule (gap 2 above the substrate
top at the centre column, the highest rim for a concave cavity, clamped below
the lid), and a step target.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as nat
from .fe_config import FEConfig

SUBSTRATES = {"flat": 0, "ridge": 1, "convex": 2, "concave": 3}
STORAGE = {"fp32": nat.LBM_F32, "bf16": nat.LBM_BF16, "fp16": nat.LBM_F16}


@dataclass
class CaseSpec:
    grid: Tuple[int, int, int]
    substrate: str = "ridge"
    R_star: float = 1.0
    R_drop: float = 22.5
    theta_eq: float = 162.0
    u_impact: float = -0.05            # along the wall normal (last axis)
    steps: int = 1500
    sigma: float = 0.01424
    xi: float = 4.0
    tau: float = 0.53
    tau_h: float = 1.0
    rho_l: float = 1.0
    rho_g: float = 1.0 / 828.0
    geo_amplification: float = 1.5
    geometric_wetting: bool = True
    boundary_relax: float = 0.3
    center: Optional[Tuple[float, ...]] = None   # None: the sweep driver's placement
    tag: str = ""

    def config(self):
        nx, ny, nz = self.grid
        return FEConfig(nx=nx, ny=ny, nz=nz, rho_l=self.rho_l, rho_g=self.rho_g,
                        sigma=self.sigma, xi=self.xi, tau_l=self.tau, tau_g=self.tau,
                        tau_h=self.tau_h, theta_eq=self.theta_eq)

    @property
    def nodes(self):
        nx, ny, nz = self.grid
        return nx * ny * nz


def _params(spec: CaseSpec, precision):
    cfg = spec.config()
    p = nat.LbmParams()
    nat.lib.lbm_params_default(ctypes.byref(p), cfg.nx, cfg.ny, cfg.nz)
    p.precision = precision
    for k in ("rho_l", "rho_g", "sigma", "xi", "beta", "kappa", "tau_l", "tau_g", "tau_h",
              "theta_eq", "phi_c"):
        setattr(p, k, float(getattr(cfg, k)))
    p.boundary_relax = float(spec.boundary_relax)
    p.geometric_wetting = 1 if spec.geometric_wetting else 0
    p.geo_amplification = float(spec.geo_amplification)
    return p


def footprint(spec: CaseSpec, storage="bf16"):
    """Device bytes one case takes (admission control)."""
    p = _params(spec, STORAGE[storage])
    return int(nat.lib.lbm_case_footprint(ctypes.byref(p)))


class BatchEngine:
    """A set of device-resident cases advanced together."""

    def __init__(self, storage="bf16", budget_bytes=0, device=None):
        self.device = torch.device(device or "cuda")
        self.storage = storage
        self.precision = STORAGE[storage]
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_batch_create(self.precision, int(budget_bytes),
                                               ctypes.byref(h)), "lbm_batch_create")
        self._h = h
        self.cases = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                nat.lib.lbm_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return nat.stream_ptr(self.device)

    def add(self, spec: CaseSpec, inputs="host"):
        """Add a case; returns its id.  Raises LbmError with code -4 when it
        does not fit the HBM budget.

        inputs="host" (default): the substrate and droplet come from the
        reference's generators (geometry/substrate.py, lbm/fe_droplet.py, run
        on the host in numpy) and are handed over as arrays, so the case
        starts from the same bits as an AllenCahnSolver built the same way.
        inputs="device": the library builds them on the device from the spec
        (faster for large grids; the droplet's tanh may then differ from
        numpy's in the last fp32 bit of a few values).  A tuple is taken as
        host_inputs(spec) prepared in advance."""
        if isinstance(inputs, tuple):
            return self._add_host(spec, inputs)
        if inputs == "host":
            return self._add_host(spec, host_inputs(spec))
        if inputs != "device":
            raise ValueError(f"inputs must be 'host', 'device' or the tuple of "
                             f"host_inputs(spec), not {inputs!r}")
        cs = nat.LbmCaseSpec()
        cs.p = _params(spec, self.precision)
        cs.substrate = SUBSTRATES[spec.substrate]
        cs.R_star = float(spec.R_star)
        cs.R_d = float(spec.R_drop)
        nx, ny, nz = spec.grid
        c = spec.center if spec.center is not None else (nx / 2.0, ny / 2.0, -1.0)
        for d in range(3):
            cs.center[d] = float(c[d])
        cs.radius = float(spec.R_drop)
        cs.u_impact[0] = cs.u_impact[1] = 0.0
        cs.u_impact[2] = float(spec.u_impact)
        cs.steps = int(spec.steps)
        cid = ctypes.c_int()
        with torch.cuda.device(self.device):
            rc = nat.lib.lbm_batch_add(self._h, ctypes.byref(cs), self._stream(),
                                       ctypes.byref(cid))
        nat.check(rc, "lbm_batch_add")
        self.cases[cid.value] = spec
        return cid.value

    def _add_host(self, spec: CaseSpec, arrays):
        solid, frac, phi0, u0 = arrays
        dev = self.device
        cs = nat.LbmCaseSpec()
        cs.p = _params(spec, self.precision)
        cs.steps = int(spec.steps)
        solid_t = torch.from_numpy(solid.astype(np.uint8)).to(dev)
        fdt = nat.FRAC_NONE
        frac_t = None
        if frac is not None:
            fdt = nat.FRAC_F64 if frac.dtype == np.float64 else nat.FRAC_F32
            frac_t = torch.from_numpy(np.ascontiguousarray(frac)).to(dev)
        phi_t = torch.from_numpy(phi0).to(dev)
        u_t = torch.from_numpy(u0).to(dev)
        cid = ctypes.c_int()
        with torch.cuda.device(dev):
            rc = nat.lib.lbm_batch_add_fields(self._h, ctypes.byref(cs), nat.ptr(solid_t),
                                              nat.ptr(frac_t), fdt, nat.ptr(phi_t),
                                              nat.ptr(u_t), self._stream(), ctypes.byref(cid))
        nat.check(rc, "lbm_batch_add_fields")
        self.cases[cid.value] = spec
        return cid.value

    def remove(self, cid):
        nat.check(nat.lib.lbm_batch_remove(self._h, int(cid)), "lbm_batch_remove")
        self.cases.pop(cid, None)

    def step(self, nsteps=1):
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_batch_step(self._h, int(nsteps), self._stream()),
                      "lbm_batch_step")

    def autotune(self, nsteps_each=5):
        a, b = ctypes.c_int(), ctypes.c_int()
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_batch_autotune(self._h, int(nsteps_each), self._stream(),
                                                 ctypes.byref(a), ctypes.byref(b)),
                      "lbm_batch_autotune")
        return a.value, b.value

    def set_split(self, split):
        """Form of pass B: 1 (or True, default) plain-fluid / wall chunk-list
        launches, 2 one launch over every case's single-launch list, 0 (or
        False) one launch over all nodes; results are identical."""
        nat.check(nat.lib.lbm_batch_set_split(self._h, int(split)), "lbm_batch_set_split")

    def set_occupancy(self, cap_phi, cap_collide):
        """Cap resident waves per SIMD of the two batch passes (0: no cap)."""
        nat.check(nat.lib.lbm_batch_set_occupancy(self._h, int(cap_phi), int(cap_collide)),
                  "lbm_batch_set_occupancy")

    def launch(self):
        """(block size A, block size B, cap A, cap B, form of pass B)."""
        out = (ctypes.c_int * 5)()
        nat.check(nat.lib.lbm_batch_get_launch(self._h, out), "lbm_batch_get_launch")
        return tuple(out)

    def capture_graph(self, nsteps):
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_batch_graph_capture(self._h, int(nsteps), self._stream()),
                      "lbm_batch_graph_capture")

    def run_graph(self, replays=1):
        with torch.cuda.device(self.device):
            for _ in range(int(replays)):
                nat.check(nat.lib.lbm_batch_graph_launch(self._h, self._stream()),
                          "lbm_batch_graph_launch")

    def status(self, cid):
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        nat.check(nat.lib.lbm_batch_status(self._h, int(cid), ctypes.byref(a), ctypes.byref(b)),
                  "lbm_batch_status")
        return a.value, b.value

    def remaining(self, cid):
        a, b = self.status(cid)
        return b - a

    def bytes_used(self):
        return int(nat.lib.lbm_batch_bytes_used(self._h))

    def _solver(self, cid):
        s = ctypes.c_void_p()
        nat.check(nat.lib.lbm_batch_solver(self._h, int(cid), ctypes.byref(s)),
                  "lbm_batch_solver")
        return s

    def save_case(self, cid):
        """Host copy (uint8 tensor) of case cid's whole state; synchronises."""
        s = self._solver(cid)
        n = int(nat.lib.lbm_state_bytes(s))
        buf = torch.empty(n, dtype=torch.uint8, pin_memory=True)
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_save_state(s, nat.ptr(buf), self._stream()), "lbm_save_state")
        return buf

    def load_case(self, cid, state):
        """Resume case cid from a save_case() image (same spec, same storage)."""
        buf = state if state.is_pinned() or state.is_cuda else state.pin_memory()
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_batch_load_state(self._h, int(cid), nat.ptr(buf),
                                                   self._stream()), "lbm_batch_load_state")

    def field(self, cid, name):
        """A copy of an observable field of case cid (phi, u, f, ...)."""
        spec = self.cases[cid]
        s = self._solver(cid)
        shape = tuple(spec.grid)
        comps = {"u": 3, "f": 19, "g": 19}.get(name)
        if comps:
            shape = (comps,) + shape
        t = torch.empty(shape, dtype=torch.float32, device=self.device)
        nat.check(nat.lib.lbm_get_field(s, nat.FIELD[name], nat.ptr(t), self._stream()),
                  "lbm_get_field")
        return t

    def measure(self, cid):
        out = (ctypes.c_double * 10)()
        nat.check(nat.lib.lbm_measure(self._solver(cid), out, self._stream()), "lbm_measure")
        keys = ("Dx", "Dy", "Dz", "k", "volume", "C_max", "z_com", "n_liquid", "phi_sum", "nan")
        return {k: float(v) for k, v in zip(keys, out)}

    def measure_async(self, cid, out):
        """measure() of case cid written to `out` (10 float64 on the device)
        without synchronising."""
        nat.check(nat.lib.lbm_measure_async(self._solver(cid), nat.ptr(out), self._stream()),
                  "lbm_measure_async")
        return out

    def contact_radius(self, cid):
        spec = self.cases[cid]
        out = ctypes.c_double()
        nat.check(nat.lib.lbm_contact_radius(self._solver(cid), spec.grid[0] / 2.0,
                                             spec.grid[1] / 2.0, ctypes.byref(out),
                                             self._stream()), "lbm_contact_radius")
        return out.value

    def base_radius(self, cid):
        spec = self.cases[cid]
        out = ctypes.c_double()
        nat.check(nat.lib.lbm_base_radius(self._solver(cid), spec.grid[0] / 2.0,
                                          spec.grid[1] / 2.0, ctypes.byref(out),
                                          self._stream()), "lbm_base_radius")
        return out.value


def droplet_center(spec: CaseSpec, solid):
    """The sweep driver's droplet placement (scripts/run_thesis_sweep.py:
    146-157) on the solid mask handed to set_solid: centre column (nx/2, ny/2),
    gap 2 above the substrate top there (the highest rim anywhere for a
    concave cavity), clamped below the lid; an explicit spec.center wins."""
    if spec.center is not None:
        return tuple(float(c) for c in spec.center)
    nx, ny, nz = spec.grid
    has = solid.any(axis=2)
    top = np.where(has, nz - 1 - np.argmax(solid[:, :, ::-1], axis=2), 0)
    ridge_top = int(top.max()) if spec.substrate == "concave" else int(top[nx // 2, ny // 2])
    cz = min(ridge_top + 2.0 + spec.R_drop, nz - spec.R_drop - 2)
    return (nx / 2.0, ny / 2.0, cz)


def host_inputs(spec: CaseSpec):
    """(solid, fraction, phi0, u0) of a case from the reference's generators
    (geometry/substrate.py:162-267, lbm/fe_droplet.py:22-112), placed by
    droplet_center()."""
    from geometry.substrate import create_substrate_with_fraction
    from .fe_droplet import create_fe_droplet_with_impact
    nx, ny, nz = spec.grid
    solid, frac = create_substrate_with_fraction(nx, ny, nz, spec.substrate, spec.R_star,
                                                 spec.R_drop)
    c = droplet_center(spec, solid)
    phi0, _, u0 = create_fe_droplet_with_impact(nx, ny, nz, center=c, radius=spec.R_drop,
                                                xi=spec.xi, u_impact=(0.0, 0.0, spec.u_impact))
    return solid, frac, phi0, u0


def sweep_cases(seed=2024, n=48, steps=1500):
    """The 48-case impact sweep: grids {96^3, 128x128x96, 150x150x107},
    substrates cycling flat/ridge/convex/concave, R* in {0.5, 1, 2}, droplet
    radius 0.15 nx, u_z = -U[0.02, 0.08], theta in {90, 108, 126, 144, 162}.
    Convex R*=2 on the two larger grids puts the drop into the cap under the
    placement rule (the drop is clamped below the lid), so R* is drawn again
    for those cases until it is 0.5 or 1."""
    import numpy as np
    rng = np.random.default_rng(seed)
    grids = [(96, 96, 96), (128, 128, 96), (150, 150, 107)]
    subs = ["flat", "ridge", "convex", "concave"]
    rstars = [0.5, 1.0, 2.0]
    thetas = [90.0, 108.0, 126.0, 144.0, 162.0]
    out = []
    for i in range(n):
        g = grids[int(rng.integers(0, 3))]
        sub = subs[i % 4]
        rs = rstars[int(rng.integers(0, 3))]
        while sub == "convex" and rs == 2.0 and g[0] >= 128:
            rs = rstars[int(rng.integers(0, 3))]
        th = thetas[int(rng.integers(0, 5))]
        u = -float(rng.uniform(0.02, 0.08))
        out.append(CaseSpec(grid=g, substrate=sub, R_star=rs, R_drop=0.15 * g[0], theta_eq=th,
                            u_impact=u, steps=steps, tag=f"c{i:02d}"))
    return out


def case_table(cases):
    """JSON-able rows of a case list (the committed sweep fixture format)."""
    return [{"tag": c.tag, "grid": list(c.grid), "substrate": c.substrate, "R_star": c.R_star,
             "R_drop": c.R_drop, "theta_eq": c.theta_eq, "u_impact": c.u_impact,
             "steps": c.steps} for c in cases]
