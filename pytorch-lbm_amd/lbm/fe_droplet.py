"""Droplet initial fields (lbm/fe_droplet.py:22-112 of the reference).

C = 0.5 + 0.5 tanh(2 (R - |x - c|) / xi), computed in float64 and rounded to
float32 once, exactly as the reference does; u = u_impact * C.
"""
import numpy as np


def create_fe_droplet(nx, ny, nz, center=None, radius=None, xi=5.0, rho_l=828.0,
                      rho_g=1.0):
    two_d = nz == 1
    shape = (nx, ny) if two_d else (nx, ny, nz)
    if radius is None:
        radius = 0.15 * (min(nx, ny) if two_d else min(nx, ny, nz))
    if center is None:
        center = (nx / 2.0, ny * 0.6) if two_d else (nx / 2.0, ny / 2.0, nz * 0.6)
    # squared distance summed as the reference sums it over its ij-indexed
    # mesh, ((0 + dx^2) + dy^2) + dz^2, with each axis term broadcast instead
    # of materialised: the same scalar operations in the same order (and the
    # same integer/float promotion), so the same bits, 3-4x faster
    r2 = 0
    for d, (n, c) in enumerate(zip(shape, center)):
        t = (np.arange(n) - c) ** 2
        r2 = r2 + t[tuple(slice(None) if k == d else None for k in range(len(shape)))]
    r = np.sqrt(r2)
    C = np.clip(0.5 + 0.5 * np.tanh(2.0 * (radius - r) / xi), 0.0, 1.0).astype(np.float32)
    rho = (C * rho_l + (1.0 - C) * rho_g).astype(np.float32)
    u = np.zeros((len(shape),) + shape, dtype=np.float32)
    return C, rho, u


def create_fe_droplet_with_impact(nx, ny, nz, center=None, radius=None, xi=5.0,
                                  rho_l=828.0, rho_g=1.0, u_impact=None):
    C, rho, u = create_fe_droplet(nx, ny, nz, center, radius, xi, rho_l, rho_g)
    if u_impact is None:
        u_impact = (0.0,) * u.shape[0]
    for d in range(u.shape[0]):
        u[d] = u_impact[d] * C
    return C, rho, u
