"""AllenCahnSolver: the reference's public class (lbm/fe_ac_solver.py:51-732),
backed by the HIP kernels of liblbm_amd.so.

Constructor, ``set_solid``, ``init_fields``, ``step``, ``inject_droplet`` and
the private ``_lattice_gradient_and_laplacian`` keep the reference's
signatures and semantics.  All arithmetic runs on the device (pytorch-lbm_amd/csrc); this
class only moves inputs to the device, calls the C ABI on torch's current
stream and materialises the observable attributes (f, g, phi, rho, u, P, p,
mu_phi) on first access after each step.  There is no CPU path: the solver
requires a HIP device and fails loudly otherwise.

Every constructor option of the reference is built: interface_mode
('allen-cahn' | 'cahn-hilliard'), mu_drive, direct_wetting, stab_mode,
geometric_wetting with scalar or per-node amplification; ch_wall_blend is
accepted and, as in the reference, only stored.
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as nat
from .fe_chemical_potential import compute_density
from .fe_config import FEConfig
from .fe_wetting import ConningtonLeeWetting
from .lattice import D2Q9, D3Q19

_STAB = {"fakhari": 0, "liang": 1, "none": 2}
_INTERFACE = {"allen-cahn": 0, "cahn-hilliard": 1}
_STORAGE = {None: None, "fp32": nat.LBM_F32, "fp64": nat.LBM_F64,
            "bf16": nat.LBM_BF16, "fp16": nat.LBM_F16}


def _to_device(x, dtype, device):
    if torch.is_tensor(x):
        return x.detach().to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device)


class AllenCahnSolver:
    """Conservative Allen-Cahn + pressure-velocity NS LBM on one HIP device."""

    def __init__(self, config: FEConfig, dtype=torch.float64, stab_mode="auto",
                 mu_drive=0.0, boundary_relax=0.3, interface_mode="allen-cahn",
                 ghost_scale=None, sharpen_reduce=0.8, mu_wall_scale=1.0,
                 geometric_wetting=False, direct_wetting=False, ch_wall_blend=0.0,
                 geo_amplification=1.0, ac_scale=1.0, storage=None, verbose=False):
        if interface_mode not in _INTERFACE:
            raise ValueError(f"unknown interface_mode {interface_mode!r}")
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be torch.float32 or torch.float64")
        self.config = cfg = config
        self.device = torch.device(cfg.device) if not isinstance(cfg.device, torch.device) \
            else cfg.device
        if self.device.type != "cuda":
            raise RuntimeError("AllenCahnSolver runs on a HIP device (device='cuda'); "
                               f"got {self.device}")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.step_count = 0
        self.t = 0.0
        self.stab_mode = "fakhari" if stab_mode == "auto" else stab_mode
        if self.stab_mode not in _STAB:
            raise ValueError(f"unknown stab_mode {stab_mode!r}")
        self.mu_drive = mu_drive
        self.boundary_relax = boundary_relax
        self.interface_mode = interface_mode
        self._ghost_scale = ghost_scale
        self._sharpen_reduce = sharpen_reduce
        self._mu_wall_scale = mu_wall_scale
        self.geometric_wetting = geometric_wetting
        self.direct_wetting = direct_wetting
        self.ch_wall_blend = ch_wall_blend
        self._geo_amplification = geo_amplification
        self._ac_scale = ac_scale
        self.shape = cfg.grid_shape
        self.ndim = cfg.dim
        self.lattice = D2Q9() if cfg.dim == 2 else D3Q19()
        self.cs2 = self.lattice.cs2
        self._rho_floor = 0.5 * cfg.rho_l
        self._tau_f = cfg.tau_h if cfg.tau_h > 0 else cfg.tau_l
        self._tau_g = cfg.tau_g
        self._mobility = self.cs2 * (self._tau_f - 0.5)
        self._phi_wall_target = None
        if cfg.theta_eq != 90.0:
            c = math.cos(math.radians(cfg.theta_eq))
            disc = 1.0 - 4.0 * (-c / 3.0)
            self._phi_wall_target = (0.5 * (1.0 - math.sqrt(disc)) if disc >= 0 else
                                     0.5 * (1.0 - math.cos(math.pi - math.radians(cfg.theta_eq))))
        prec = _STORAGE[storage]
        if prec is None:
            prec = nat.LBM_F32 if dtype == torch.float32 else nat.LBM_F64
        self.storage = storage or ("fp32" if prec == nat.LBM_F32 else "fp64")
        self._compute_dtype = torch.float64 if prec == nat.LBM_F64 else torch.float32
        p = nat.LbmParams()
        nat.lib.lbm_params_default(ctypes.byref(p), cfg.nx, cfg.ny, cfg.nz)
        p.precision = prec
        for k in ("rho_l", "rho_g", "sigma", "xi", "beta", "kappa", "tau_l", "tau_g",
                  "tau_h", "theta_eq", "phi_c"):
            setattr(p, k, float(getattr(cfg, k)))
        gf = tuple(cfg.g_force) + (0.0,) * 3
        for d in range(3):
            p.g_force[d] = float(gf[d]) if d < cfg.dim else 0.0
        p.stab_mode = _STAB[self.stab_mode]
        p.boundary_relax = float(boundary_relax)
        p.sharpen_reduce = float(sharpen_reduce)
        p.mu_wall_scale = float(mu_wall_scale)
        p.ac_scale = float(ac_scale)
        p.ghost_scale = float(ghost_scale) if ghost_scale is not None else 0.0
        p.geometric_wetting = 1 if geometric_wetting else 0
        p.direct_wetting = 1 if direct_wetting else 0
        p.interface_mode = _INTERFACE[interface_mode]
        p.mu_drive = float(mu_drive)
        amp_tensor = torch.is_tensor(geo_amplification)
        p.geo_amplification = 1.0 if amp_tensor else float(geo_amplification)
        self._params = p
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_create(ctypes.byref(p), ctypes.byref(h)), "lbm_create")
        self._h = h
        self._amp_tensor = geo_amplification if amp_tensor else None
        self.solid = torch.zeros(self.shape, dtype=torch.bool, device=self.device)
        self.solid_fraction = None
        self.phi_mass_init = None
        self.wetting = None
        self.bounce_back = None
        self._has_partial = False
        self._cache = {}
        self._overrides = {}
        self._geom_set = False
        if verbose:
            print(f"AllenCahn Solver (HIP): grid={self.shape}, {self.device}, "
                  f"storage={self.storage}")

    # ------------------------------------------------------------------
    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
                nat.lib.lbm_destroy(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return nat.stream_ptr(self.device)

    def _read_field(self, which, out):
        nat.check(nat.lib.lbm_get_field(self._h, which, nat.ptr(out), self._stream()),
                  "lbm_get_field")
        return out

    # ------------------------------------------------------------------
    def set_solid(self, solid_mask, solid_fraction=None):
        """Geometry and wetting precompute (lbm/fe_ac_solver.py:193-237)."""
        dev = self.device
        solid_np = solid_mask.detach().cpu().numpy() if torch.is_tensor(solid_mask) \
            else np.asarray(solid_mask)
        if tuple(solid_np.shape) != tuple(self.shape):
            raise ValueError(f"solid mask shape {solid_np.shape} != grid {self.shape}")
        self.solid = torch.as_tensor(solid_np.astype(bool), device=dev)
        solid_u8 = self.solid.to(torch.uint8).contiguous()
        frac_t, fdt = None, nat.FRAC_NONE
        if solid_fraction is not None:
            fr = solid_fraction.detach().cpu().numpy() if torch.is_tensor(solid_fraction) \
                else np.asarray(solid_fraction)
            if fr.dtype == np.float64:
                frac_t, fdt = torch.as_tensor(fr, device=dev), nat.FRAC_F64
            else:
                frac_t = torch.as_tensor(fr.astype(np.float32), device=dev)
                fdt = nat.FRAC_F32
            self.solid_fraction = frac_t.to(self._compute_dtype)
        with torch.cuda.device(dev):
            nat.check(nat.lib.lbm_set_geometry(self._h, nat.ptr(solid_u8), nat.ptr(frac_t),
                                               fdt, self._stream()), "lbm_set_geometry")
            if self._amp_tensor is not None:
                amp = _to_device(self._amp_tensor, self._compute_dtype, dev)
                nat.check(nat.lib.lbm_set_amplification_field(self._h, nat.ptr(amp),
                                                              self._stream()),
                          "lbm_set_amplification_field")
                torch.cuda.current_stream(dev).synchronize()
        self._geom_set = True
        self._keep = (solid_u8, frac_t)
        if solid_fraction is not None:
            part = torch.empty(self.shape, dtype=torch.uint8, device=dev)
            self._read_field(nat.FIELD["partial"], part)
            self._partial_mask = part.bool()
            self._has_partial = bool(self._partial_mask.any())
            self._partial_eps = self.solid_fraction[self._partial_mask]
        else:
            self._has_partial = False
        self._solid_any = bool(self.solid.any())
        cfg = self.config
        self.wetting = None
        if cfg.theta_eq != 90.0 and cfg.kappa > 0:
            mode = "linear" if (cfg.theta_eq > 130.0 or cfg.theta_eq < 50.0) else "quadratic"
            self.wetting = ConningtonLeeWetting(self, cfg.phi_c, cfg.kappa, cfg.sigma,
                                                cfg.theta_eq, mode, self._ghost_scale,
                                                self._geo_amplification)
        self._cache.clear()

    def init_fields(self, phi_init, u_init=None):
        """f = f_eq(phi, u), g = g_eq(0, u), M0 = sum(phi) (lbm/fe_ac_solver.py:239-261)."""
        dt, dev = self._compute_dtype, self.device
        phi_t = _to_device(phi_init, dt, dev)
        if tuple(phi_t.shape) != tuple(self.shape):
            raise ValueError(f"phi shape {tuple(phi_t.shape)} != grid {self.shape}")
        u_t = None
        if u_init is not None:
            u_t = _to_device(u_init, dt, dev)
            if tuple(u_t.shape) != (self.ndim,) + tuple(self.shape):
                raise ValueError("u_init must have shape (dim, *grid)")
        with torch.cuda.device(dev):
            nat.check(nat.lib.lbm_init_fields(self._h, nat.ptr(phi_t), nat.ptr(u_t),
                                              self._stream()), "lbm_init_fields")
        self.step_count = 0
        self.t = 0.0
        self._cache.clear()
        self._overrides.clear()
        m0 = ctypes.c_double()
        nat.check(nat.lib.lbm_get_mass(self._h, ctypes.byref(m0), None, self._stream()),
                  "lbm_get_mass")
        self.phi_mass_init = m0.value
        self._phi_u_prev = None

    def step(self, nsteps=1):
        """Advance nsteps time steps (AllenCahnSolver.step, lbm/fe_ac_solver.py:380-639).

        No host synchronisation; all launches go to torch's current stream."""
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_step(self._h, int(nsteps), self._stream()), "lbm_step")
        self.step_count += int(nsteps)
        self.t += float(nsteps)
        self._cache.clear()
        self._overrides.clear()

    def profile_steps(self, nsteps):
        """Run nsteps with HIP events around each kernel; returns mean ms of
        (pass A, pass B, whole step).  Synchronises."""
        out = (ctypes.c_double * 3)()
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_profile_steps(self._h, int(nsteps), self._stream(), out),
                      "lbm_profile_steps")
        self.step_count += int(nsteps)
        self.t += float(nsteps)
        self._cache.clear()
        self._overrides.clear()
        return tuple(float(v) for v in out)

    def autotune(self, nsteps_each=10):
        """Pick the fastest launch shape (block size, occupancy cap, form of
        pass B) for the two step kernels on this case.  Advances the solver by
        27*(nsteps_each+1) steps; returns the block sizes (launch() gives the
        rest)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        k0 = int(nat.lib.lbm_step_count(self._h))
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_autotune(self._h, int(nsteps_each), self._stream(),
                                           ctypes.byref(a), ctypes.byref(b)), "lbm_autotune")
        n = int(nat.lib.lbm_step_count(self._h)) - k0
        self.step_count += n
        self.t += float(n)
        self._cache.clear()
        self._overrides.clear()
        return a.value, b.value

    def set_block_sizes(self, bs_phi, bs_collide):
        nat.check(nat.lib.lbm_set_block_sizes(self._h, int(bs_phi), int(bs_collide)),
                  "lbm_set_block_sizes")

    def set_occupancy(self, cap_phi, cap_collide):
        """Cap resident waves per SIMD of the two step kernels (0: no cap)."""
        nat.check(nat.lib.lbm_set_occupancy(self._h, int(cap_phi), int(cap_collide)),
                  "lbm_set_occupancy")

    def state_dict(self, device=None):
        """A snapshot of the solver state as a uint8 tensor (on the solver's
        device unless `device` says otherwise): distributions, velocity, phi
        buffers, mass factors, step count.  Restore with load_state_dict on a
        solver with the same grid, precision and geometry."""
        n = int(nat.lib.lbm_state_bytes(self._h))
        t = torch.empty(n, dtype=torch.uint8, device=device or self.device)
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_save_state(self._h, nat.ptr(t), self._stream()),
                      "lbm_save_state")
        return t

    def load_state_dict(self, state):
        """Restore a snapshot taken by state_dict()."""
        t = state.contiguous()
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_load_state(self._h, nat.ptr(t), self._stream()),
                      "lbm_load_state")
        self.step_count = int(nat.lib.lbm_step_count(self._h))
        self.t = float(self.step_count)
        self._cache.clear()
        self._overrides.clear()
        m0 = ctypes.c_double()
        nat.check(nat.lib.lbm_get_mass(self._h, ctypes.byref(m0), None, self._stream()),
                  "lbm_get_mass")
        self.phi_mass_init = m0.value

    def set_split(self, split):
        """Form of pass B: 0 (False) one launch over all nodes; 1 (True) two
        launches over the plain-fluid and the wall chunk lists; 2 one launch
        over both lists, each XCD's range holding its wall chunks last; 3 the
        fused step (both passes in one launch; 3D, fp32/fp64 storage)."""
        nat.check(nat.lib.lbm_set_split(self._h, int(split)), "lbm_set_split")

    def set_fused_tile(self, ty=0, tx=0, zlen=0):
        """Tile of the fused step: ty x tx columns per workgroup, zlen planes
        per segment (0: automatic)."""
        nat.check(nat.lib.lbm_set_fused_tile(self._h, int(ty), int(tx), int(zlen)),
                  "lbm_set_fused_tile")

    def fused_tile(self):
        """(ty, tx, tiles along y, tiles along x, zlen, segments) of the fused
        step in use (zeros when another form is selected)."""
        out = (ctypes.c_int * 6)()
        nat.check(nat.lib.lbm_get_fused_tile(self._h, out), "lbm_get_fused_tile")
        return tuple(out)

    def launch(self):
        """(block size A, block size B, cap A, cap B, form of pass B) of the step
        kernels (see set_split)."""
        out = (ctypes.c_int * 5)()
        nat.check(nat.lib.lbm_get_launch(self._h, out), "lbm_get_launch")
        return tuple(out)

    def capture_graph(self, nsteps=2):
        """Record `nsteps` (even) steps into a HIP graph owned by the solver; replay
        with run_graph().  Needs one eager step since init_fields."""
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_graph_capture(self._h, int(nsteps)), "lbm_graph_capture")
        self._graph_steps = int(nsteps)

    def run_graph(self, replays=1):
        """Replay the captured graph `replays` times on the current stream."""
        with torch.cuda.device(self.device):
            for _ in range(int(replays)):
                nat.check(nat.lib.lbm_graph_launch(self._h, self._stream()),
                          "lbm_graph_launch")
        n = self._graph_steps * int(replays)
        self.step_count += n
        self.t += float(n)
        self._cache.clear()
        self._overrides.clear()

    # ------------------------------------------------------------------
    # Observable state
    # ------------------------------------------------------------------
    def _field(self, name, comps=None):
        if name in self._overrides:
            return self._overrides[name]
        if name not in self._cache:
            shape = tuple(self.shape)
            if comps is not None:
                shape = (comps,) + shape
            t = torch.empty(shape, dtype=self._compute_dtype, device=self.device)
            self._read_field(nat.FIELD[name], t)
            self._cache[name] = t
        return self._cache[name]

    def _set_override(self, name, value):
        self._overrides[name] = value

    phi = property(lambda s: s._field("phi"), lambda s, v: s._set_override("phi", v))
    rho = property(lambda s: s._field("rho"), lambda s, v: s._set_override("rho", v))
    u = property(lambda s: s._field("u", s.ndim), lambda s, v: s._set_override("u", v))
    P = property(lambda s: s._field("P"), lambda s, v: s._set_override("P", v))
    mu_phi = property(lambda s: s._field("mu"), lambda s, v: s._set_override("mu", v))
    f = property(lambda s: s._field("f", s.lattice.q), lambda s, v: s._set_override("f", v))
    g = property(lambda s: s._field("g", s.lattice.q), lambda s, v: s._set_override("g", v))

    @property
    def p(self):
        if "p" in self._overrides:
            return self._overrides["p"]
        if self.step_count == 0:
            return self.rho * self.cs2
        return self.P * self.cs2

    @p.setter
    def p(self, v):
        self._overrides["p"] = v

    @property
    def phi_raw(self):
        """Start-of-step phi of the last step (clamped, 0 at solid)."""
        return self._field("phi_raw")

    # ------------------------------------------------------------------
    def _lattice_gradient_and_laplacian(self, phi):
        """1-ring isotropic gradient and Laplacian with wetting ghosts
        (lbm/fe_ac_solver.py:296-320)."""
        dt = self._compute_dtype
        x = _to_device(phi, dt, self.device)
        grad = torch.empty((self.ndim,) + tuple(self.shape), dtype=dt, device=self.device)
        lap = torch.empty(tuple(self.shape), dtype=dt, device=self.device)
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_grad_lap(self._h, nat.ptr(x), nat.ptr(grad), nat.ptr(lap),
                                           self._stream()), "lbm_grad_lap")
        return grad, lap

    def mass(self):
        """(M0, fp64 sum of the last start-of-step phi); synchronises."""
        a, b = ctypes.c_double(), ctypes.c_double()
        nat.check(nat.lib.lbm_get_mass(self._h, ctypes.byref(a), ctypes.byref(b),
                                       self._stream()), "lbm_get_mass")
        return a.value, b.value

    def measure(self):
        """Device-reduced diagnostics of the current phi (lbm/fe_measure.py:11-82)."""
        out = (ctypes.c_double * 10)()
        nat.check(nat.lib.lbm_measure(self._h, out, self._stream()), "lbm_measure")
        keys = ("Dx", "Dy", "Dz", "k", "volume", "C_max", "z_com", "n_liquid",
                "phi_sum", "nan")
        return {k: float(v) for k, v in zip(keys, out)}

    def measure_async(self, out=None):
        """The diagnostics of measure() written to a device tensor of 10
        float64 (Dx, Dy, Dz, k, volume, C_max, z_com, n_liquid, phi_sum,
        nan_flag) in stream order, without synchronising."""
        if out is None:
            out = torch.empty(10, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_measure_async(self._h, nat.ptr(out), self._stream()),
                      "lbm_measure_async")
        return out

    def contact_radius(self, cx=None, cy=None):
        """Max distance from (cx, cy) of liquid nodes on the first fluid layer
        above the substrate (3D)."""
        cx = self.shape[0] / 2.0 if cx is None else cx
        cy = self.shape[1] / 2.0 if cy is None else cy
        out = ctypes.c_double()
        nat.check(nat.lib.lbm_contact_radius(self._h, float(cx), float(cy), ctypes.byref(out),
                                             self._stream()), "lbm_contact_radius")
        return out.value

    def base_radius(self, cx=None, cy=None):
        """Max distance from (cx, cy) of liquid nodes on the lowest layer that
        holds liquid (3D): the radius of the drop's base, also where it rides
        on a gas cushion above the substrate."""
        cx = self.shape[0] / 2.0 if cx is None else cx
        cy = self.shape[1] / 2.0 if cy is None else cy
        out = ctypes.c_double()
        nat.check(nat.lib.lbm_base_radius(self._h, float(cx), float(cy), ctypes.byref(out),
                                          self._stream()), "lbm_base_radius")
        return out.value

    def inject_droplet(self, center, radius, xi, u_impact=None):
        """Add a droplet between steps (lbm/fe_ac_solver.py:645-732).

        C = clip(0.5 + 0.5 tanh(2 (radius - |x - center|) / xi), 0, 1) is built
        on the host in fp64 and rounded to float32, as the reference does;
        the state edit (phi = max(phi, C) and equilibrium f, g where C > 0.1,
        u blended towards u_impact with weight C, new mass target) is one
        device pass.  P and mu_phi keep their last-step values until the next
        step."""
        ndim = self.ndim
        shape = tuple(self.shape)
        if u_impact is None:
            u_impact = (0.0,) * ndim
        if len(center) != ndim or len(u_impact) != ndim:
            raise ValueError(f"center and u_impact need {ndim} components")
        mesh = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        dist = np.sqrt(sum((M - center[d]) ** 2 for d, M in enumerate(mesh)))
        C = np.clip(0.5 + 0.5 * np.tanh(2.0 * (radius - dist) / xi), 0.0, 1.0)
        C = C.astype(np.float32)
        dt = self._compute_dtype
        c_t = torch.from_numpy(C).to(device=self.device, dtype=dt)
        np_dt = np.float64 if dt == torch.float64 else np.float32
        n_core = int(np.count_nonzero(C.astype(np_dt) > np_dt(0.1)))
        uimp = (ctypes.c_double * 3)(*[float(v) for v in u_impact] + [0.0] * (3 - ndim))
        with torch.cuda.device(self.device):
            nat.check(nat.lib.lbm_inject_droplet(self._h, nat.ptr(c_t), uimp, self._stream()),
                      "lbm_inject_droplet")
        self._cache.clear()
        self._overrides.clear()
        m0 = ctypes.c_double()
        nat.check(nat.lib.lbm_get_mass(self._h, ctypes.byref(m0), None, self._stream()),
                  "lbm_get_mass")
        self.phi_mass_init = m0.value
        self._phi_u_prev = None
        center_str = ",".join(f"{c:.0f}" for c in center)
        vel_str = ",".join(f"{u:.3f}" for u in u_impact)
        print(f"  [inject] Droplet at ({center_str}) R={radius:.1f} u=({vel_str}) "
              f"core_nodes={n_core}")
        return n_core
