"""Pins of the oracle's droplet injection (OracleSolver.inject_droplet, after
lbm/fe_ac_solver.py:645-732) that do not restate it: a drop indicator built
node by node with math.tanh, equilibrium moments, untouched nodes, solid
nodes, axis order and the mass target."""
import math

import numpy as np
import pytest

from _cases import config_kwargs, load, solver_kwargs
from oracle.ac_oracle import OracleSolver, Params


def indicator(shape, center, radius, xi):
    """Drop indicator node by node: 0.5 + 0.5 tanh(2 (R - |x - c|) / xi),
    clipped to [0, 1] and rounded to float32."""
    out = np.zeros(shape, np.float32)
    for idx in np.ndindex(*shape):
        d = math.sqrt(sum((i - c) ** 2 for i, c in zip(idx, center)))
        out[idx] = min(1.0, max(0.0, 0.5 + 0.5 * math.tanh(2.0 * (radius - d) / xi)))
    return out


def empty(shape, dtype=np.float64):
    o = OracleSolver(Params(nx=shape[0], ny=shape[1], nz=shape[2] if len(shape) > 2 else 1,
                            tau_l=0.8), dtype=dtype)
    o.set_solid(np.zeros(shape, bool))
    o.init_fields(np.zeros(shape))
    return o


@pytest.mark.parametrize("shape,center", [((9, 7), (2.0, 4.0)), ((7, 8, 9), (3.0, 2.0, 5.0))])
def test_inject_into_empty_domain(shape, center):
    """phi = C on the core (C > 0.1) and 0 elsewhere, u = C u_imp, f and g
    are equilibria: sum f = phi, sum e f = phi u, sum g = P (0 here),
    sum e g = u; the mass target is the fp64 sum of the new phi."""
    D = len(shape)
    o = empty(shape)
    uimp = (0.03, -0.05, 0.02)[:D]
    C = indicator(shape, center, 2.5, 2.0).astype(np.float64)
    n = o.inject_droplet(center, 2.5, 2.0, uimp)
    core = C > 0.1
    assert n == int(core.sum()) > 0
    np.testing.assert_array_equal(o.phi, np.where(core, C, 0.0))
    assert np.unravel_index(np.argmax(o.phi), shape) == tuple(int(c) for c in center)
    for d in range(D):
        np.testing.assert_allclose(o.u[d], C * uimp[d], rtol=0, atol=1e-17)
    # the lattice weights are float32 constants (as in the reference), so the
    # moments hold to ~1e-8 relative; a dropped or wrong term is ~1e-3
    E = o.e.astype(np.float64)
    f, g = o.f, o.g
    np.testing.assert_allclose(f.sum(0), o.phi, atol=1e-9)
    np.testing.assert_allclose(np.einsum("qa,q...->a...", E, f)[:, core],
                               (o.phi[None] * o.u)[:, core], atol=1e-9)
    np.testing.assert_allclose(g.sum(0)[core], 0.0, atol=1e-9)
    np.testing.assert_allclose(np.einsum("qa,q...->a...", E, g)[:, core], o.u[:, core],
                               atol=1e-9)
    # outside the core nothing was there and nothing is added
    assert not f[:, ~core].any() and not g[:, ~core].any()
    assert o.M0 == float(np.sum(o.phi, dtype=np.float64))
    assert o.phi_u_prev is None


def ridge_after(steps=3):
    z, m = load("ridge162_readme_f64")
    o = OracleSolver(Params(**config_kwargs(m)), dtype=np.float64, **solver_kwargs(m))
    o.set_solid(z["solid"], z["fraction"])
    o.init_fields(z["phi_init"], z["u_init"])
    for _ in range(steps):
        o.step()
    return o, z


def test_inject_keeps_the_rest_and_the_solid():
    """Outside the core f, g and phi are untouched; u blends with weight C;
    solid nodes keep phi = u = 0; P and mu are not changed; phi only grows;
    the equilibria at core nodes use the last step's P."""
    o, z = ridge_after()
    solid = z["solid"]
    f0, g0, phi0, u0 = o.f.copy(), o.g.copy(), o.phi.copy(), o.u.copy()
    P0, mu0 = o.P.copy(), o.mu.copy()
    center, R, xi, uimp = (14.0, 6.0, 6.0), 3.0, 4.0, (0.02, 0.0, -0.03)
    C = indicator(solid.shape, center, R, xi).astype(np.float64)
    core = C > 0.1
    assert (core & solid).any() and (core & ~solid).any()
    o.inject_droplet(center, R, xi, uimp)
    keep = ~core
    np.testing.assert_array_equal(o.f[:, keep], f0[:, keep])
    np.testing.assert_array_equal(o.g[:, keep], g0[:, keep])
    np.testing.assert_array_equal(o.phi[keep], phi0[keep])
    assert (o.phi[solid] == 0).all() and (o.u[:, solid] == 0).all()
    fl = ~solid
    assert (o.phi[fl] >= phi0[fl]).all()
    np.testing.assert_array_equal(o.phi[core & fl], np.maximum(phi0, C)[core & fl])
    for d in range(3):
        np.testing.assert_allclose(o.u[d][fl], ((1 - C) * u0[d] + C * uimp[d])[fl],
                                   rtol=0, atol=1e-17)
    np.testing.assert_array_equal(o.P, P0)
    np.testing.assert_array_equal(o.mu, mu0)
    cf = core & fl
    np.testing.assert_allclose(o.g.sum(0)[cf], P0[cf], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(o.f.sum(0)[cf], o.phi[cf], atol=1e-9)


def test_inject_same_drop_twice_keeps_phi():
    """max(phi, C) is idempotent: a second identical injection leaves phi and
    the mass target as they were."""
    o, _ = ridge_after(2)
    o.inject_droplet((6.0, 12.0, 12.0), 3.0, 4.0, (0.0, 0.0, -0.05))
    phi1, m1 = o.phi.copy(), o.M0
    o.inject_droplet((6.0, 12.0, 12.0), 3.0, 4.0, (0.0, 0.0, -0.05))
    np.testing.assert_array_equal(o.phi, phi1)
    assert o.M0 == m1


def test_step_after_inject_conserves_new_mass():
    """The next steps hold the sum of phi at the new target (the mass factor
    uses the injected total), and the first step has no d(phi u) term."""
    o, _ = ridge_after(2)
    o.inject_droplet((6.0, 12.0, 12.0), 3.0, 4.0, (0.0, 0.0, -0.05))
    m = o.M0
    st = o.step(return_stages=True)
    assert abs(st["c"] - 1.0) < 1e-2
    o.step()
    assert o.M0 == m
