"""Wetting boundary view (the observable side of ConningtonLeeWetting,
lbm/fe_wetting.py:31-367 of the reference).

The precompute itself runs on the device inside ``lbm_set_geometry``
(pytorch-lbm_amd/csrc/lbm_geometry.hpp).  This object carries the scalar
constants the reference stores (h, h/kappa*ghost_scale, mu_wall, mode, ...) and
materialises the per-node arrays (is_boundary, wall_normal, wall_weight,
near_wall_mask) from the device on first access.
"""
import math

import torch

from ._native import FIELD


class ConningtonLeeWetting:
    def __init__(self, solver, phi_c, kappa, sigma, theta_deg, mode, ghost_scale,
                 geo_amplification):
        self._solver = solver
        self.phi_c = phi_c
        self.kappa = kappa
        self.sigma = sigma
        self.theta_deg = theta_deg
        self.mode = mode
        self.ndim = solver.ndim
        self._xi = 5.0
        self._geo_amplification = geo_amplification
        if mode == "linear":
            self._h = sigma * math.cos(math.radians(theta_deg))
            if ghost_scale is None:
                ghost_scale = 3.0 if abs(theta_deg - 90.0) > 60.0 else 2.0
            self._h_over_kappa = self._h / kappa * ghost_scale
            self._mu_wall = -self._h * 2.0
            self._ghost_scale = ghost_scale
        else:
            self._phi_c_over_kappa = phi_c / kappa
            self._h = 0.0
            self._h_over_kappa = 0.0
            self._mu_wall = 0.0
        self._cache = {}

    def _get(self, name, dtype, comps=1):
        if name not in self._cache:
            s = self._solver
            shape = ((comps,) if comps > 1 else ()) + tuple(s.shape)
            t = torch.empty(shape, dtype=dtype, device=s.device)
            s._read_field(FIELD[name], t)
            self._cache[name] = t
        return self._cache[name]

    @property
    def is_boundary(self):
        return self._get("is_boundary", torch.uint8).bool()

    @property
    def near_wall_mask(self):
        return self._get("near_wall", torch.uint8).bool()

    @property
    def wall_weight(self):
        return self._get("wall_weight", torch.float32)

    @property
    def wall_normal(self):
        return self._get("wall_normal", torch.float32, comps=self.ndim)

    @property
    def _near_wall_any(self):
        return bool(self.near_wall_mask.any())

    @property
    def _is_boundary_any(self):
        return bool(self.is_boundary.any())

    @staticmethod
    def phi_c_from_contact_angle(theta_deg, kappa, beta):
        """phi_c = -cos(theta) sqrt(2 kappa beta) (lbm/fe_wetting.py:579-588)."""
        return -math.cos(math.radians(theta_deg)) * math.sqrt(2.0 * kappa * beta)
