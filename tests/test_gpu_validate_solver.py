"""The reference's own regression gate, scripts/validate_solver.py:15-160:
an 80 x 80 x 65 ridge impact (We = 7.9, theta = 162, tau = 0.555,
tau_h = tau + 0.04, alpha_geo = 1.8, boundary_relax = 0, fp32), 500 steps,
checked against its recorded baseline: k = 1.2963 +- 0.05, z_com = 17.6 +- 1,
Dx = 35 +- 2, Dy = 27 +- 2, relative mass change <= 1e-4, phi in [0, 1].
Run here through the public API on the HIP path with the reference's own
driver arithmetic (fp32 sums of phi, extents of phi > 0.5 and not solid)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_reference_validation_gate():
    need_gpu()
    print("[gpu] validate_solver gate", flush=True)
    from geometry.substrate import create_substrate_with_fraction
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_droplet import create_fe_droplet_with_impact
    D0 = 24.0
    R = D0 / 2.0
    rho_l, rho_g, xi, U0, We, theta, tau = 1.0, 1.0 / 828.0, 4.0, -0.05, 7.9, 162.0, 0.555
    n, N = 80, 500
    sigma = rho_l * U0 ** 2 * D0 / We
    beta = 12.0 * sigma / xi
    kappa = beta * xi ** 2 / 8.0
    nz = min(int(R + 2 + R + 2 * R + 15), 300)
    cfg = FEConfig(nx=n, ny=n, nz=nz, rho_l=rho_l, rho_g=rho_g, sigma=sigma, xi=xi,
                   beta=beta, kappa=kappa, M=0.02 / beta, tau_l=tau, tau_g=tau,
                   tau_h=tau + 0.04, theta_eq=theta, device="cuda", max_steps=N,
                   output_interval=N + 1, g_force=(0.0, 0.0, 0.0))
    s = AllenCahnSolver(cfg, dtype=torch.float32, stab_mode="fakhari", boundary_relax=0.0,
                        geometric_wetting=True, geo_amplification=1.8)
    solid, frac = create_substrate_with_fraction(n, n, nz, substrate_type="ridge",
                                                 R_star=1.0, R_d=R)
    s.set_solid(solid, solid_fraction=frac)
    height = np.array([[np.nonzero(solid[i, j])[0].max() if solid[i, j].any() else 0
                        for j in range(n)] for i in range(n)])
    cz = height[n // 2, n // 2] + 2.0 + R
    C0, _, u0 = create_fe_droplet_with_impact(n, n, nz, center=(n / 2.0, n / 2.0, cz),
                                              radius=R, xi=xi, rho_l=rho_l, rho_g=rho_g,
                                              u_impact=(0.0, 0.0, U0))
    s.init_fields(C0, u0)
    mass0 = s.phi.sum().item()
    worst = 0.0
    for step in range(50, N + 1, 50):
        s.step(50)
        phi = s.phi
        worst = max(worst, abs(phi.sum().item() - mass0) / max(mass0, 1e-10))
        assert float(phi.min()) >= 0.0 and float(phi.max()) <= 1.0
    phi = s.phi.cpu().numpy()
    inter = (phi > 0.5) & ~np.asarray(solid)
    c = np.argwhere(inter)
    Dx = float(c[:, 0].max() - c[:, 0].min() + 1)
    Dy = float(c[:, 1].max() - c[:, 1].min() + 1)
    k = Dx / Dy
    z_com = float(c[:, 2].mean())
    print(f"[gpu] k={k:.4f} Dx={Dx} Dy={Dy} z_com={z_com:.2f} mass={worst:.2e}", flush=True)
    assert abs(k - 1.2963) <= 0.05
    assert abs(z_com - 17.6) <= 1.0
    assert abs(Dx - 35.0) <= 2.0
    assert abs(Dy - 27.0) <= 2.0
    assert worst <= 1e-4
