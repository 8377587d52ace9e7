"""Host-side input generators of the package against the reference's outputs
(golden fixtures): substrate masks/fractions and droplet fields, bit-exact."""
import numpy as np
import pytest

from _cases import case_names, load
from geometry.substrate import (create_substrate, create_substrate_with_fraction,
                                substrate_surface_height)
from lbm.fe_droplet import create_fe_droplet_with_impact
from lbm.fe_config import FEConfig


@pytest.mark.parametrize("name", case_names())
def test_substrate_and_droplet_bit_exact(name):
    z, m = load(name)
    nx, ny, nz = m["grid"]
    if "fraction" in z.files:
        solid, frac = create_substrate_with_fraction(nx, ny, nz, m["sub"], m["R_star"], m["R_d"])
        assert frac.dtype == z["fraction"].dtype
        np.testing.assert_array_equal(frac, z["fraction"])
    else:
        solid = create_substrate(nx, ny, nz, m["sub"], m["R_star"], m["R_d"])
    np.testing.assert_array_equal(solid, z["solid"])
    C, _, u = create_fe_droplet_with_impact(nx, ny, nz, center=tuple(m["center"]),
                                            radius=m["radius"], xi=4.0, rho_l=1.0,
                                            rho_g=1.0 / 828.0, u_impact=tuple(m["u"]))
    np.testing.assert_array_equal(C, z["phi_init"])
    np.testing.assert_array_equal(u, z["u_init"])


def test_ridge_cfg2_counts():
    """150x150x107 ridge R*=1, R_d=22.5: node counts restated from the reference's
    closed-form geometry (solid 138,900; partial 3,900)."""
    solid, frac = create_substrate_with_fraction(150, 150, 107, "ridge", 1.0, 22.5)
    assert frac.dtype == np.float64
    assert int(solid.sum()) == 138900
    partial = (frac > 0.01) & (frac < 0.99) & ~solid
    assert int(partial.sum()) == 3900
    h = substrate_surface_height(150, 150, 107, "ridge", 1.0, 22.5)
    assert int(h[75, 75]) == 22


def test_feconfig_closed_forms():
    """Appendix constants for sigma=0.01424, xi=4, theta=162."""
    c = FEConfig(nx=150, ny=150, nz=107, rho_l=1.0, rho_g=1 / 828, sigma=0.01424, xi=4.0,
                 tau_l=0.53, tau_g=0.53, theta_eq=162.0)
    assert abs(c.beta - 0.04272) < 1e-12
    assert abs(c.kappa - 0.08544) < 1e-12
    assert abs(c.M - 0.02 / 0.04272) < 1e-12
    assert abs(c.phi_c - 0.0812583) < 1e-7
    assert c.dim == 3 and c.grid_shape == (150, 150, 107)
    c2 = FEConfig(nx=8, ny=8, nz=1, theta_eq=90.0, sigma=0.01, xi=4.0)
    assert c2.phi_c == 0.0 and c2.dim == 2


def test_sweep_case_table_matches_fixture():
    """The 48-case sweep drawn from seed 2024 is the committed case table."""
    import json
    import os
    from lbm.batch import case_table, sweep_cases
    here = os.path.dirname(os.path.abspath(__file__))
    fx = json.load(open(os.path.join(here, "golden", "cfg4_cases.json")))
    assert case_table(sweep_cases(fx["seed"], 48, 1500)) == fx["cases"]
    rows = fx["cases"]
    assert {tuple(r["grid"]) for r in rows} <= {(96, 96, 96), (128, 128, 96), (150, 150, 107)}
    assert not any(r["substrate"] == "convex" and r["R_star"] == 2.0 and r["grid"][0] >= 128
                   for r in rows)
    assert all(abs(r["R_drop"] - 0.15 * r["grid"][0]) < 1e-12 for r in rows)
    assert all(-0.08 <= r["u_impact"] <= -0.02 for r in rows)
