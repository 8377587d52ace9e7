"""lbm_grad_lap_noalloc (the stencil operator without allocations, its
scratch a static per-GPU array) against lbm_grad_lap through the C ABI: bit
for bit on two calls in a row, the second on a smaller grid, and a grid too
large for its scratch refused."""
import ctypes

import numpy as np
import pytest
import torch

from _cases import config_kwargs, load, solver_kwargs

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def solver(name):
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    z, m = load(name)
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float32, **solver_kwargs(m))
    s.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    s.init_fields(z["phi_init"], z["u_init"])
    s.step(3)
    return s


def both(s):
    from lbm import _native as nat
    phi = s.phi.contiguous()
    shape = tuple(s.shape)
    g1 = torch.empty((s.ndim,) + shape, dtype=phi.dtype, device=phi.device)
    l1 = torch.empty(shape, dtype=phi.dtype, device=phi.device)
    g2, l2 = torch.full_like(g1, np.nan), torch.full_like(l1, np.nan)
    st = nat.stream_ptr()
    nat.check(nat.lib.lbm_grad_lap(s._h, nat.ptr(phi), nat.ptr(g1), nat.ptr(l1), st),
              "lbm_grad_lap")
    nat.check(nat.lib.lbm_grad_lap_noalloc(s._h, nat.ptr(phi), nat.ptr(g2), nat.ptr(l2), st),
              "lbm_grad_lap_noalloc")
    torch.cuda.synchronize()
    return (g1, l1), (g2, l2)


def test_noalloc_grad_lap_matches_allocating_entry_point():
    need_gpu()
    print("[gpu] grad_lap noalloc vs allocating", flush=True)
    big = solver("ridge162_readme_f32")        # 3D, wetting ghosts
    small = solver("flat162_2d_f32")           # smaller grid, second call
    assert np.prod(small.shape) < np.prod(big.shape)
    for s in (big, small, big):
        (g1, l1), (g2, l2) = both(s)
        np.testing.assert_array_equal(g2.cpu().numpy(), g1.cpu().numpy())
        np.testing.assert_array_equal(l2.cpu().numpy(), l1.cpu().numpy())


def test_noalloc_grad_lap_refuses_grids_beyond_its_scratch():
    need_gpu()
    from lbm import _native as nat
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    cfg = FEConfig(nx=256, ny=256, nz=130, rho_l=1.0, rho_g=1.0 / 828.0, sigma=0.01424,
                   xi=4.0, tau_l=0.53, tau_g=0.53, theta_eq=90.0, device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float32)    # 8,519,680 nodes > 8,388,608
    phi = torch.zeros(tuple(s.shape), dtype=torch.float32, device="cuda")
    grad = torch.empty((3,) + tuple(s.shape), dtype=torch.float32, device="cuda")
    lap = torch.empty_like(phi)
    rc = nat.lib.lbm_grad_lap_noalloc(s._h, nat.ptr(phi), nat.ptr(grad), nat.ptr(lap),
                                      nat.stream_ptr())
    assert rc == -1
    assert b"scratch" in nat.lib.lbm_last_error()
