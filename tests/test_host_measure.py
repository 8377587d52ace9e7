"""Host measurement helpers with the reference's names (lbm/fe_measure.py,
io_utils/monitor.py): pinned by brute force and closed forms on synthetic
fields."""
import math

import numpy as np
import pytest

from io_utils import Monitor
from lbm.fe_config import FEConfig
from lbm.fe_measure import measure_contact_angle, measure_spread_factor


def brute_spread(c, solid):
    c = np.where(solid, 0, c)
    idx = [tuple(i) for i in np.ndindex(*c.shape) if c[tuple(i)] > 0.5]
    if not idx:
        return None
    ext = [max(p[d] for p in idx) - min(p[d] for p in idx) + 1 for d in range(c.ndim)]
    return ext


@pytest.mark.parametrize("shape", [(13, 9), (9, 11, 7)])
def test_spread_factor_brute_force(shape):
    rng = np.random.default_rng(3)
    for _ in range(5):
        c = (rng.random(shape) ** 3).astype(np.float32)
        solid = rng.random(shape) < 0.2
        m = measure_spread_factor(c, solid)
        ext = brute_spread(c, solid)
        assert [m["Dx"], m["Dy"]] == [float(e) for e in ext[:2]]
        assert m["Dz"] == (float(ext[2]) if len(shape) == 3 else 0.0)
        assert m["k"] == m["Dx"] / m["Dy"]
        assert m["volume"] == float(np.where(solid, 0, c).sum())
        assert m["C_max"] == float(np.where(solid, 0, c).max())
    empty = measure_spread_factor(np.zeros(shape, np.float32))
    assert empty["Dx"] == empty["k"] == 0.0


def test_contact_angle_closed_form():
    """Wall layers z = 0..2 solid; footprint disk on layer 4 (one above the
    lowest open layer 3); its outermost node's column crosses 0.5 between
    z = 7 (0.8) and 8 (0.2): height 7.5 - 3 = 4.5, theta = atan2(r, 4.5)."""
    nx, ny, nz = 20, 20, 16
    solid = np.zeros((nx, ny, nz), bool)
    solid[:, :, :3] = True
    c = np.zeros((nx, ny, nz), np.float32)
    pts = [(x, y) for x in range(nx) for y in range(ny) if (x - 9.5) ** 2 + (y - 10.5) ** 2 <= 16]
    for x, y in pts:
        c[x, y, 4] = 1.0
    radii = [math.hypot(x - 10.0, y - 10.0) for x, y in pts]
    far = pts[int(np.argmax(radii))]
    c[far[0], far[1], 4:8] = [1.0, 1.0, 1.0, 0.8]
    c[far[0], far[1], 8] = 0.2
    theta = measure_contact_angle(c, solid)
    assert abs(theta - math.degrees(math.atan2(max(radii), 4.5))) < 1e-5
    # no footprint, footprint at the axis only, no crossing: 180
    assert measure_contact_angle(np.zeros_like(c), solid) == 180.0
    c2 = np.zeros_like(c)
    c2[10, 10, 4] = 1.0
    assert measure_contact_angle(c2, solid) == 180.0
    c3 = np.zeros_like(c)
    c3[4, 10, 4:] = 1.0
    assert measure_contact_angle(c3, solid) == 180.0


def test_contact_angle_2d_height_floor():
    """2D: the wall is the last axis; a crossing just above the wall gives
    h < 0.5, which is raised to 0.5."""
    nx, ny = 24, 12
    solid = np.zeros((nx, ny), bool)
    solid[:, 0] = True
    c = np.zeros((nx, ny), np.float32)
    c[8:16, 2] = 1.0          # footprint on layer 2 (lowest open layer is 1)
    c[8, 2], c[8, 3] = 0.6, 0.4   # outermost column (x = 8, r = 4) crosses at 2.5
    theta = measure_contact_angle(c, solid, axis="y")
    assert abs(theta - math.degrees(math.atan2(4.0, 1.5))) < 1e-5


def test_monitor_film_and_gradient_angle():
    cfg = FEConfig(nx=16, ny=12, nz=1, rho_l=1.0, rho_g=0.001, sigma=0.01, xi=4.0,
                   tau_l=0.6, tau_g=0.6, theta_eq=90.0, device="cpu")
    mon = Monitor(cfg)
    rho = np.full((16, 12), 0.001)
    solid = np.zeros((16, 12), bool)
    solid[:, :3] = True                       # substrate top at j = 2
    rho[5:11, 7:] = 1.0                       # liquid from j = 7: gap 5
    assert mon.compute_film_thickness(rho, solid) == 5 * cfg.dx
    gmap = mon.compute_film_thickness_map(rho, solid)
    assert np.isnan(gmap[0]) and gmap[5] == 5.0
    # a density tent rho = a (4 - |x - 7.5|) + b y continued into the wall:
    # wet columns 1..14 on the first fluid layer, gradient (+-a, b) at the
    # three columns inside each contact point, angle arccos(|b| / |(a, b)|)
    a, b = 0.03, 0.05
    X, Y = np.meshgrid(np.arange(16), np.arange(12), indexing="ij")
    rho2 = 0.6 + a * (4.0 - np.abs(X - 7.5)) + b * (Y - 3)
    want = math.degrees(math.acos(b / math.hypot(a, b)))
    left, right = mon.compute_contact_angle(rho2, solid)
    assert abs(left - want) < 1e-9 and abs(right - want) < 1e-9


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_base_radius_brute_force(seed):
    """oracle.base_radius against a plain loop over every node: lowest layer
    holding liquid, then the farthest liquid node of that layer."""
    import math
    from oracle.ac_oracle import base_radius
    rng = np.random.default_rng(seed)
    nx, ny, nz = 7, 6, 9
    phi = rng.uniform(0.0, 1.0, (nx, ny, nz)) * (rng.uniform(size=(nx, ny, nz)) < 0.3)
    solid = rng.uniform(size=(nx, ny, nz)) < 0.15
    if seed == 3:
        phi[:] = 0.2          # no liquid at all
    cx, cy = 3.5, 2.0
    zlow = None
    for z in range(nz):
        for x in range(nx):
            for y in range(ny):
                if phi[x, y, z] > 0.5 and not solid[x, y, z]:
                    zlow = z if zlow is None else min(zlow, z)
    want = 0.0
    if zlow is not None:
        for x in range(nx):
            for y in range(ny):
                if phi[x, y, zlow] > 0.5 and not solid[x, y, zlow]:
                    want = max(want, math.sqrt((x - cx) ** 2 + (y - cy) ** 2))
    assert base_radius(phi, solid, cx, cy) == want
