"""Substrate masks and solid fractions (geometry/substrate.py:13-267 of the
reference).

Solid fraction = clip(signed distance + 0.5, 0, 1) with a fully solid floor
(last axis index 0); solid = fraction >= 0.5.  Curved substrates produce a
float64 fraction, the flat one a float32 fraction -- the dtype matters for the
wall normals, which are computed in the fraction's precision.  The boolean
masks of create_substrate use the strict inequalities of the reference.
"""
import numpy as np


def _grid(shape):
    return np.ogrid[tuple(slice(0, n) for n in shape)]


def create_substrate(nx, ny, nz=1, substrate_type="flat", R_star=None, R_d=None, dx=1.0):
    """Boolean solid mask (geometry/substrate.py:13-41, 86-159)."""
    two_d = nz == 1
    shape = (nx, ny) if two_d else (nx, ny, nz)
    solid = np.zeros(shape, dtype=bool)
    solid[..., 0] = True
    if substrate_type == "flat" or R_star is None or R_d is None:
        return solid
    Rg = abs(R_star) * R_d / dx
    if two_d:
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        cx = nx / 2.0
        if substrate_type == "convex":
            solid |= ((i - cx) ** 2 + (j - Rg) ** 2) < Rg ** 2
        elif substrate_type == "concave":
            solid |= (((i - cx) ** 2 + (j - Rg) ** 2) > Rg ** 2) & (j < Rg)
        elif substrate_type == "ridge":
            solid |= (((i - cx) ** 2 + j ** 2) < Rg ** 2) & (j > 0)
        return solid
    i, j, k = _grid(shape)
    cx, cy = nx / 2.0, ny / 2.0
    if substrate_type == "convex":
        solid |= ((i - cx) ** 2 + (j - cy) ** 2 + (k - Rg) ** 2) < Rg ** 2
    elif substrate_type == "concave":
        solid |= (((i - cx) ** 2 + (j - cy) ** 2 + (k - Rg) ** 2) > Rg ** 2) & (k < Rg)
    elif substrate_type == "ridge":
        solid |= (((i - cx) ** 2 + k ** 2) < Rg ** 2) & (k > 0)
    return solid


def _surface_height(solid):
    idx = np.arange(solid.shape[-1])
    return np.where(solid, idx, -1).max(axis=-1)


def substrate_surface_height(nx, ny, nz=1, substrate_type="flat", R_star=None, R_d=None,
                             dx=1.0):
    """Topmost solid index per column (geometry/substrate.py:44-74)."""
    return _surface_height(create_substrate(nx, ny, nz, substrate_type, R_star, R_d, dx))


def create_substrate_with_fraction(nx, ny, nz=1, substrate_type="flat", R_star=None,
                                   R_d=None, dx=1.0):
    """(solid, fraction) for volume penalization (geometry/substrate.py:162-267)."""
    two_d = nz == 1
    shape = (nx, ny) if two_d else (nx, ny, nz)
    frac = np.zeros(shape, dtype=np.float32)
    frac[..., 0] = 1.0
    if substrate_type == "flat" or R_star is None or R_d is None:
        return frac >= 0.5, frac
    Rg = abs(R_star) * R_d / dx
    if two_d:
        I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        cx = nx / 2.0
        X, Z, Y2 = I, J, 0.0
    else:
        I, J, K = _grid(shape)
        cx, cy = nx / 2.0, ny / 2.0
        X, Z, Y2 = I, K, (J - cy) ** 2.0
    if substrate_type == "ridge":
        dist = np.sqrt((X - cx) ** 2.0 + Z.astype(float) ** 2.0)
        part = np.clip(Rg - dist + 0.5, 0.0, 1.0)
        frac = np.where(Z > 0, np.maximum(frac, part), frac)
    elif substrate_type in ("convex", "concave"):
        dist = np.sqrt((X - cx) ** 2.0 + Y2 + (Z - Rg) ** 2.0) if not two_d else \
            np.sqrt((X - cx) ** 2.0 + (Z - Rg) ** 2.0)
        if substrate_type == "convex":
            frac = np.maximum(frac, np.clip(Rg - dist + 0.5, 0.0, 1.0))
        else:
            part = np.clip(dist - Rg + 0.5, 0.0, 1.0)
            frac = np.where(Z < Rg, np.maximum(frac, part), frac)
    return frac >= 0.5, frac
