"""Measurement helpers with the reference's names and results
(lbm/fe_measure.py:11-241): spread factors of the C > 0.5 liquid region and
the contact-angle estimate at the wall.

Like the reference, these take a field (tensor on any device, or an array)
and evaluate it on the host, once per call; a driver that measures often
should use the solver's on-device forms instead (``AllenCahnSolver.measure``
and ``measure_async``, which reduce on the GPU without copying the field).
"""
import math

import numpy as np
import torch


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def measure_spread_factor(C, solid_mask=None) -> dict:
    """Extents Dx, Dy[, Dz] (lattice units, inclusive) of the nodes with
    C > 0.5 outside the solid, k = Dx / Dy, the liquid volume sum(C) and
    max(C) (lbm/fe_measure.py:11-82).  All zero but volume / C_max when no
    node exceeds 0.5."""
    c = _host(C)
    if solid_mask is not None:
        c = np.where(_host(solid_mask), c.dtype.type(0), c)
    liquid = c > 0.5
    volume = float(c.sum())
    c_max = float(c.max())
    if not liquid.any():
        return {"Dx": 0.0, "Dy": 0.0, "Dz": 0.0, "k": 0.0, "volume": volume, "C_max": c_max}
    ext = []
    for axis in range(c.ndim):
        other = tuple(a for a in range(c.ndim) if a != axis)
        occupied = np.nonzero(liquid.any(axis=other))[0]
        ext.append(float(occupied[-1] - occupied[0] + 1))
    dx, dy = ext[0], ext[1]
    dz = ext[2] if c.ndim == 3 else 0.0
    return {"Dx": dx, "Dy": dy, "Dz": dz, "k": dx / dy if dy > 0 else 0.0,
            "volume": volume, "C_max": c_max}


def measure_contact_angle(C, solid_mask, axis: str = "z") -> float:
    """Contact angle (degrees) from the liquid footprint on the wall
    (lbm/fe_measure.py:85-241): the wall is the last axis; the layer one
    above the lowest layer that is not entirely solid is searched for C > 0.5
    nodes; the one farthest from the domain's centre axis (first in index
    order on ties) gives the footprint radius r; along its column the C = 0.5
    crossing above that layer (linear interpolation) gives the height h over
    the lowest non-solid layer (at least 0.5); theta = atan2(r, h).  Returns
    180 when there is no footprint, r < 1 or no crossing."""
    c = _host(C)
    solid = _host(solid_mask).astype(bool)
    shape = c.shape
    nwall = shape[-1]
    # lowest layer along the wall axis that is not entirely solid
    layer_solid = solid.reshape(-1, nwall).all(axis=0)
    open_layers = np.nonzero(~layer_solid)[0]
    z_wall = int(open_layers[0]) if len(open_layers) else 0
    z_fluid = z_wall + 1
    if z_fluid >= nwall:
        return 180.0
    foot = c[..., z_fluid] > 0.5
    if not foot.any():
        return 180.0
    pts = np.argwhere(foot)
    centre = np.array([n / 2.0 for n in shape[:-1]])
    if c.ndim == 2:
        radii = np.abs(pts[:, 0] - centre[0])
    else:
        radii = np.sqrt((pts[:, 0] - centre[0]) ** 2 + (pts[:, 1] - centre[1]) ** 2)
    i = int(np.argmax(radii))
    r = radii[i]
    if r < 1.0:
        return 180.0
    col = c[tuple(pts[i])]
    height = None
    for z in range(z_fluid, nwall - 1):
        if col[z] >= 0.5 and col[z + 1] < 0.5:
            height = z + (col[z] - 0.5) / (col[z] - col[z + 1] + 1e-10)
            break
    if height is None:
        return 180.0
    h = max(height - z_wall, 0.5)
    return float(np.degrees(np.arctan2(r, h)))
