"""Pins of the oracle that do not need the reference: closed forms, library
routines and invariants, chosen so that a dropped term, a wrong sign or index
or a transposed operand fails one of them."""
import math
import os

import numpy as np
import pytest

from oracle.ac_oracle import (CS2, D2Q9_E, D2Q9_OPP, D2Q9_W, D3Q19_E, D3Q19_OPP,
                              D3Q19_W, OracleSolver, Params, upstream, wall_target)

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("E,W,OPP", [(D2Q9_E, D2Q9_W, D2Q9_OPP), (D3Q19_E, D3Q19_W, D3Q19_OPP)])
def test_lattice_moments(E, W, OPP):
    w = W.astype(np.float64)
    assert abs(w.sum() - 1.0) < 1e-7
    np.testing.assert_allclose(np.einsum("q,qa->a", w, E), 0.0, atol=1e-8)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, E, E), CS2 * np.eye(E.shape[1]),
                               atol=1e-7)
    np.testing.assert_array_equal(E[OPP], -E)
    # fourth-order isotropy: sum w e_a e_b e_c e_d = cs^4 (d_ab d_cd + d_ac d_bd + d_ad d_bc)
    D = E.shape[1]
    t = np.einsum("q,qa,qb,qc,qd->abcd", w, E, E, E, E)
    I = np.eye(D)
    ref = CS2 ** 2 * (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
                      + np.einsum("ad,bc->abcd", I, I))
    np.testing.assert_allclose(t, ref, atol=1e-7)


def _solver(shape, dtype=np.float64, theta=90.0):
    nx, ny = shape[:2]
    nz = shape[2] if len(shape) == 3 else 1
    p = Params(nx=nx, ny=ny, nz=nz, rho_l=1.0, rho_g=1 / 828, sigma=0.01424, xi=4.0,
               tau_l=0.53, tau_g=0.53, theta_eq=theta)
    o = OracleSolver(p, dtype=dtype)
    o.set_solid(np.zeros(shape, dtype=bool), None)
    return o


@pytest.mark.parametrize("shape", [(12, 10), (10, 8, 6)])
def test_equilibrium_moments(shape):
    """sum f_eq = phi, sum e f_eq = phi u, sum g_eq = P, sum e g_eq = u."""
    o = _solver(shape)
    rng = np.random.default_rng(1)
    phi = rng.uniform(0, 1, shape)
    P = rng.uniform(-0.01, 0.01, shape)
    u = rng.uniform(-0.05, 0.05, (len(shape),) + shape)
    fe, ge = o.f_eq(phi, u), o.g_eq(P, u)
    E = o.e.astype(np.float64)
    np.testing.assert_allclose(fe.sum(0), phi, atol=1e-7)
    np.testing.assert_allclose(ge.sum(0), P, atol=1e-8)
    np.testing.assert_allclose(np.einsum("qa,q...->a...", E, fe), phi * u, atol=1e-8)
    np.testing.assert_allclose(np.einsum("qa,q...->a...", E, ge), u, atol=1e-8)


def test_stencil_exact_on_constant_and_linear_fields():
    o = _solver((16, 16, 16))
    c = np.full((16, 16, 16), 0.37)
    g, l = o.grad_lap(c)
    np.testing.assert_allclose(g, 0.0, atol=1e-15)
    np.testing.assert_allclose(l, 0.0, atol=1e-15)
    # a periodic quadratic in x: lap = 2a exactly in the interior
    x = np.arange(16.0)[:, None, None] * np.ones((16, 16, 16))
    phi = 0.01 * (x - 7.5) ** 2
    g, l = o.grad_lap(phi)
    # exact up to the fp32 rounding of the lattice weights (~1e-8 relative)
    np.testing.assert_allclose(l[2:-2], 0.02, rtol=1e-7)
    np.testing.assert_allclose(g[0][2:-2], 0.02 * (x[2:-2] - 7.5), rtol=1e-7, atol=1e-14)
    np.testing.assert_allclose(g[1:], 0.0, atol=1e-14)


def test_mms_gradient_matches_published_errors():
    """The gradient operator on the periodic sin.cos.cos field: L2 error 1.33e-3
    at N=32 as published with the reference (paper_pof/sections/
    4_experiments.tex:188, harness scripts/mms_spatial.py:26-55), and the
    reference operator's own outputs at N=16, 32 (golden)."""
    z = np.load(os.path.join(GOLD, "mms_operator.npz"))
    for N in (16, 32):
        o = _solver((N, N, N))
        g, l = o.grad_lap(z[f"N{N}_phi"])
        np.testing.assert_array_equal(g, z[f"N{N}_grad"])
        np.testing.assert_array_equal(l, z[f"N{N}_lap"])
    N = 32
    i = np.arange(N)[:, None, None]
    j = np.arange(N)[None, :, None]
    k = np.arange(N)[None, None, :]
    a = 2 * np.pi / N
    phi = np.sin(a * i) * np.cos(a * j) * np.cos(a * k)
    gx = a * np.cos(a * i) * np.cos(a * j) * np.cos(a * k)
    gy = -a * np.sin(a * i) * np.sin(a * j) * np.cos(a * k)
    gz = -a * np.sin(a * i) * np.cos(a * j) * np.sin(a * k)
    g, _ = _solver((N, N, N)).grad_lap(phi)
    err = np.sqrt(np.mean([((g[0] - gx) ** 2).mean(), ((g[1] - gy) ** 2).mean(),
                           ((g[2] - gz) ** 2).mean()]))
    assert abs(err - 1.33e-3) / 1.33e-3 < 0.01


def test_streaming_index_map_is_np_roll():
    """f_i(x) <- f*_i(x - e_i): the pull of an index grid equals np.roll by +e_i,
    the library routine with torch.roll's semantics."""
    shape = (5, 4, 3)
    idx = np.arange(np.prod(shape)).reshape(shape)
    for e in D3Q19_E:
        got = upstream(idx, e)
        exp = idx
        for d in range(3):
            exp = np.roll(exp, int(e[d]), axis=d)
        np.testing.assert_array_equal(got, exp)
        # element check: value at x is the index of x - e (mod shape)
        x = (1, 2, 0)
        src = tuple((x[d] - e[d]) % shape[d] for d in range(3))
        assert got[x] == idx[src]


def test_bounce_back_and_vp_brute_force():
    """One step on a 4^3 box with a solid node and a partial node: the streamed
    f at the solid node is its own reversed post-collision set, and at the
    partial node the eps blend, checked element by element."""
    shape = (4, 4, 4)
    solid = np.zeros(shape, bool)
    solid[1, 1, 1] = True
    frac = np.zeros(shape)
    frac[1, 1, 1] = 1.0
    frac[2, 1, 1] = 0.3
    p = Params(nx=4, ny=4, nz=4, rho_l=1.0, rho_g=1 / 828, sigma=0.01424, xi=4.0,
               tau_l=0.53, tau_g=0.53, theta_eq=90.0)
    o = OracleSolver(p, dtype=np.float64)
    o.set_solid(solid, frac)
    rng = np.random.default_rng(3)
    o.init_fields(rng.uniform(0.2, 0.8, shape), rng.uniform(-0.02, 0.02, (3,) + shape))
    o.M0 = None  # no mass factor: compare the raw stream
    st = o.step(return_stages=True)
    fs = st["fstar"]
    for i in range(19):
        assert o.f[i][1, 1, 1] == fs[D3Q19_OPP[i]][1, 1, 1]
        e = D3Q19_E[i]
        srcx = tuple((np.array((2, 1, 1)) - e) % 4)
        exp = 0.7 * fs[i][srcx] + 0.3 * fs[D3Q19_OPP[i]][2, 1, 1]
        assert abs(o.f[i][2, 1, 1] - exp) < 1e-15
        srcy = tuple((np.array((0, 3, 2)) - e) % 4)
        assert o.f[i][0, 3, 2] == fs[i][srcy]


def test_mass_invariant_and_wall_target():
    o = _solver((16, 16, 16))
    rng = np.random.default_rng(5)
    x = np.arange(16)[:, None, None]
    phi = 0.5 + 0.5 * np.tanh(2 * (5 - np.abs(x - 8.0)) / 4.0) * np.ones((16, 16, 16))
    o.init_fields(phi, None)
    for _ in range(5):
        o.step()
        assert abs(o.phi.sum() - o.M0) / o.M0 < 1e-12
    # boundary-relax targets (lbm/fe_ac_solver.py:165-177)
    assert wall_target(90.0) is None
    assert abs(wall_target(162.0) - 0.0244717) < 1e-6
    assert abs(wall_target(120.0) - 0.211325) < 1e-6
    assert abs(wall_target(144.0) - 0.0954915) < 1e-6


def test_correct_gradient_noop_where_normal_zero_and_closed_form():
    """Zero wall normal: exact no-op.  Nonzero: the normal component becomes
    g_n + alpha (-|g_t| cot(theta) - g_n), tangential part unchanged."""
    z = np.load(os.path.join(GOLD, "ridge162_readme_f64.npz"))
    p = Params(nx=20, ny=20, nz=16, rho_l=1.0, rho_g=1 / 828, sigma=0.01424, xi=4.0,
               tau_l=0.53, tau_g=0.53, theta_eq=162.0)
    o = OracleSolver(p, dtype=np.float64, geometric_wetting=True, geo_amplification=1.5)
    o.set_solid(z["solid"], z["fraction"])
    rng = np.random.default_rng(7)
    grad = rng.normal(size=(3, 20, 20, 16))
    gc = o.correct_gradient(grad)
    n = o.geom.wall_normal.astype(np.float64)
    zero = (n == 0).all(0)
    np.testing.assert_array_equal(gc[:, zero], grad[:, zero])
    b = o.geom.is_boundary & ~zero
    th = math.radians(162.0)
    gn = (n * grad).sum(0)
    gt = grad - gn * n
    want_n = gn + 1.5 * (-np.linalg.norm(gt, axis=0) * math.cos(th) / math.sin(th) - gn)
    got_n = (n * gc).sum(0)
    # normals are fp32-rounded (|n|^2 = 1 +- 1e-7), hence the tolerance
    np.testing.assert_allclose(got_n[b], want_n[b], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose((gc - (n * gc).sum(0) * n)[:, b], gt[:, b], atol=1e-5)
