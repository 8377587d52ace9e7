"""The oracle's diagnostics (spread extents, volume, z_com, contact-line
radius) against brute-force loops on small random grids."""
import math

import numpy as np

from oracle.ac_oracle import contact_radius, spread


def brute_spread(phi, solid):
    nx, ny, nz = phi.shape
    vol, cmax, pts = 0.0, -1.0, []
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                c = 0.0 if solid[x, y, z] else float(phi[x, y, z])
                vol += c
                cmax = max(cmax, c)
                if c > 0.5:
                    pts.append((x, y, z))
    if not pts:
        return dict(Dx=0.0, Dy=0.0, Dz=0.0, volume=vol, C_max=cmax, z_com=0.0)
    ext = [max(p[d] for p in pts) - min(p[d] for p in pts) + 1 for d in range(3)]
    return dict(Dx=float(ext[0]), Dy=float(ext[1]), Dz=float(ext[2]), volume=vol, C_max=cmax,
                z_com=sum(p[2] for p in pts) / len(pts))


def brute_contact(phi, solid, cx, cy):
    nx, ny, nz = phi.shape
    best = None
    for x in range(nx):
        for y in range(ny):
            top = -1
            for z in range(nz):
                if solid[x, y, z]:
                    top = z
            z = top + 1
            if z < nz and not solid[x, y, z] and phi[x, y, z] > 0.5:
                r = math.hypot(x - cx, y - cy)
                best = r if best is None else max(best, r)
    return 0.0 if best is None else best


def test_spread_matches_brute_force():
    rng = np.random.default_rng(7)
    for _ in range(5):
        phi = rng.random((7, 6, 5)).astype(np.float32)
        solid = rng.random((7, 6, 5)) < 0.2
        a, b = spread(phi, solid), brute_spread(phi, solid)
        for k in ("Dx", "Dy", "Dz", "C_max", "z_com"):
            assert a[k] == b[k] or abs(a[k] - b[k]) < 1e-12, k
        assert abs(a["volume"] - b["volume"]) < 1e-9


def test_spread_empty_liquid():
    phi = np.full((4, 4, 4), 0.2, np.float32)
    out = spread(phi, np.zeros((4, 4, 4), bool))
    assert out["Dx"] == out["Dy"] == out["Dz"] == 0.0 and out["z_com"] == 0.0


def test_contact_radius_matches_brute_force():
    rng = np.random.default_rng(11)
    for _ in range(8):
        phi = rng.random((9, 8, 6)).astype(np.float32)
        h = rng.integers(-1, 5, size=(9, 8))
        solid = np.arange(6)[None, None, :] <= h[:, :, None]
        # holes: solid below a non-solid node must not change the column's top
        solid[0, 0, 0] = True
        cx, cy = 4.5, 3.0
        assert abs(contact_radius(phi, solid, cx, cy) - brute_contact(phi, solid, cx, cy)) < 1e-12


def test_contact_radius_closed_form():
    # a liquid disc of radius 3 on a flat floor: radius = farthest lattice point
    nx = ny = 15
    phi = np.zeros((nx, ny, 6), np.float32)
    solid = np.zeros_like(phi, dtype=bool)
    solid[:, :, 0] = True
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    disc = (X - 7) ** 2 + (Y - 7) ** 2 <= 9
    phi[:, :, 1][disc] = 1.0
    assert contact_radius(phi, solid, 7.0, 7.0) == 3.0
