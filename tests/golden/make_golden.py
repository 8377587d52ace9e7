#!/usr/bin/env python3
"""Write golden vectors for the Allen-Cahn step by running the reference on CPU.

This script is the only place that imports the reference implementation
(JunjunQing/pytorch-lbm, mounted read-only at /root/reference).  It runs it with
``device='cpu'`` on small grids and stores inputs and outputs as compressed
``.npz`` fixtures next to this file.  Nothing here is imported by the product
or by the GPU tests; the fixtures are plain data.

What the fixtures pin (everything else is pinned by closed forms, library
routines and invariants in tests/test_oracle_*.py):
  * the composed multi-step trajectory of ``AllenCahnSolver.step()``
    (lbm/fe_ac_solver.py:380-639): stage order, clamps, lagged mass factor,
    S12 boundary relax, fp32 torch rounding over N steps;
  * M0 and the initial ``mu_phi`` display state of ``init_fields``
    (lbm/fe_ac_solver.py:239-261);
  * the wetting precompute outputs (lbm/fe_wetting.py:122-367) for each
    substrate type, including the discrete-normal path;
  * ``_lattice_gradient_and_laplacian`` + ``correct_gradient`` on the
    step-1 phi (lbm/fe_ac_solver.py:296-320, lbm/fe_wetting.py:448-539);
  * the stab_mode variants and the body force of the force stage
    (lbm/fe_ac_solver.py:415-439);
  * ``inject_droplet`` between two steps (lbm/fe_ac_solver.py:645-732);
  * direct wetting (lbm/fe_ac_solver.py:609-613, lbm/fe_wetting.py:313-416),
    the Cahn-Hilliard source and the mu drive (lbm/fe_ac_solver.py:457-486).

Usage (in the build container only):  python tests/golden/make_golden.py
"""
import os
import sys
import json
import contextlib
import io

import numpy as np
import torch

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REF)

from lbm.fe_config import FEConfig  # noqa: E402
from lbm.fe_ac_solver import AllenCahnSolver  # noqa: E402
from lbm.fe_droplet import create_fe_droplet_with_impact  # noqa: E402
from geometry.substrate import (create_substrate_with_fraction,  # noqa: E402
                                create_substrate)

torch.set_num_threads(1)

# name: grid, substrate, R_star, R_d, theta, solver kwargs, dtype, droplet, frac?
CASES = {
    # README construction (tau_h=1.0, boundary_relax=0.3), linear wetting
    "ridge162_readme_f32": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float32", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(1, 2)),
    "ridge162_readme_f64": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float64", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(1,)),
    # production construction (tau_h=tau+0.04, boundary_relax=0)
    "ridge162_prod_f32": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True,
                boundary_relax=0.0, stab_mode="fakhari"),
        dtype="float32", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=0.57, fg_steps=(1,)),
    # quadratic wetting, S12 active (target 0.2113)
    "ridge120_f32": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=120.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float32", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(1,)),
    "ridge120_f64": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=120.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float64", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=()),
    # no wetting object (theta=90), flat wall, float32 fraction
    "flat90_f32": dict(
        grid=(20, 20, 16), sub="flat", R_star=None, R_d=None, theta=90.0,
        kw=dict(), dtype="float32", center=(10.0, 10.0, 7.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(1,)),
    # discrete-normal path (no solid_fraction), convex, linear 144
    "convex144_nofrac_f32": dict(
        grid=(20, 20, 16), sub="convex", R_star=1.0, R_d=4.0, theta=144.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float32", center=(10.0, 10.0, 11.0), radius=3.5,
        u=(0.0, 0.0, -0.05), frac=False, tau_h=None, fg_steps=()),
    # concave with fraction, quadratic 126, scalar alpha = 1 (no amplification)
    "concave126_f32": dict(
        grid=(20, 20, 16), sub="concave", R_star=1.0, R_d=5.0, theta=126.0,
        kw=dict(geo_amplification=1.0, geometric_wetting=True),
        dtype="float32", center=(10.0, 10.0, 9.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=()),
    # 2D D2Q9: flat wall, linear 162, geometric wetting
    "flat162_2d_f32": dict(
        grid=(32, 32, 1), sub="flat", R_star=None, R_d=None, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float32", center=(16.0, 9.0), radius=6.0,
        u=(0.0, -0.05), frac=True, tau_h=None, fg_steps=(1,)),
    # 2D D2Q9 ridge (semicircle) fp64, quadratic 108
    "ridge108_2d_f64": dict(
        grid=(32, 32, 1), sub="ridge", R_star=1.0, R_d=5.0, theta=108.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float64", center=(16.0, 13.0), radius=5.0,
        u=(0.0, -0.04), frac=True, tau_h=None, fg_steps=(1,)),
    # 2D D2Q9 static drop, theta=90 (cfg1-like, small)
    "flat90_2d_f32": dict(
        grid=(32, 32, 1), sub="flat", R_star=None, R_d=None, theta=90.0,
        kw=dict(), dtype="float32", center=(16.0, 9.0), radius=6.0,
        u=(0.0, 0.0), frac=True, tau_h=None, fg_steps=()),
    # force-stage variants: Liang (no viscous correction), none (surface
    # tension only) with a body force
    "ridge120_liang_2d_f64": dict(
        grid=(32, 32, 1), sub="ridge", R_star=1.0, R_d=5.0, theta=120.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True, stab_mode="liang"),
        dtype="float64", center=(16.0, 13.0), radius=5.0,
        u=(0.0, -0.04), frac=True, tau_h=None, fg_steps=(1,)),
    "flat144_none_grav_2d_f64": dict(
        grid=(32, 32, 1), sub="flat", R_star=None, R_d=None, theta=144.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True, stab_mode="none"),
        dtype="float64", center=(16.0, 12.0), radius=7.0,
        u=(0.0, -0.02), frac=True, tau_h=None, fg_steps=(1,), g_force=(0.0, -1e-5)),
    # direct wetting: S12 relaxes towards the cubic wall-energy boundary phi
    "ridge120_direct_2d_f64": dict(
        grid=(32, 32, 1), sub="ridge", R_star=1.0, R_d=5.0, theta=120.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True, direct_wetting=True),
        dtype="float64", center=(16.0, 11.5), radius=5.0,
        u=(0.0, -0.04), frac=True, tau_h=None, fg_steps=(1, 2)),
    "ridge144_direct_f32": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=144.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True, direct_wetting=True),
        dtype="float32", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(1,)),
    # Cahn-Hilliard interface source and the mu drive (both use grad mu)
    "ridge120_ch_2d_f64": dict(
        grid=(32, 32, 1), sub="ridge", R_star=1.0, R_d=5.0, theta=120.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True,
                interface_mode="cahn-hilliard"),
        dtype="float64", center=(16.0, 11.5), radius=5.0,
        u=(0.0, -0.04), frac=True, tau_h=None, fg_steps=(1,)),
    "flat144_mudrive_2d_f64": dict(
        grid=(32, 32, 1), sub="flat", R_star=None, R_d=None, theta=144.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True, mu_drive=0.5),
        dtype="float64", center=(16.0, 8.0), radius=6.0,
        u=(0.0, -0.02), frac=True, tau_h=None, fg_steps=(1,)),
    "ridge162_ch_f32": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True,
                interface_mode="cahn-hilliard"),
        dtype="float32", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(1,)),
    # per-node amplification field (tensor geo_amplification,
    # lbm/fe_wetting.py:503-512): values in [1, 2], exactly 1 (un-amplified)
    # on some nodes
    "ridge144_alphafield_2d_f64": dict(
        grid=(32, 32, 1), sub="ridge", R_star=1.0, R_d=5.0, theta=144.0,
        kw=dict(geo_amplification="field", geometric_wetting=True),
        dtype="float64", center=(16.0, 11.5), radius=5.0,
        u=(0.0, -0.04), frac=True, tau_h=None, fg_steps=(1,)),
    "ridge120_alphafield_f32": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=120.0,
        kw=dict(geo_amplification="field", geometric_wetting=True),
        dtype="float32", center=(10.0, 10.0, 9.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=()),
    # a second drop injected after step 3 (3D, fp64) and after step 2 (2D, fp32)
    "ridge162_inject_f64": dict(
        grid=(20, 20, 16), sub="ridge", R_star=1.0, R_d=5.0, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float64", center=(10.0, 10.0, 10.0), radius=4.0,
        u=(0.0, 0.0, -0.05), frac=True, tau_h=None, fg_steps=(),
        inject=dict(after=3, center=(5.0, 13.0, 12.0), radius=3.0, xi=4.0,
                    u=(0.01, 0.0, -0.04))),
    "flat162_inject_2d_f32": dict(
        grid=(32, 32, 1), sub="flat", R_star=None, R_d=None, theta=162.0,
        kw=dict(geo_amplification=1.5, geometric_wetting=True),
        dtype="float32", center=(12.0, 9.0), radius=5.0,
        u=(0.0, -0.03), frac=True, tau_h=None, fg_steps=(2,),
        inject=dict(after=2, center=(22.0, 10.0), radius=4.0, xi=4.0,
                    u=(0.0, -0.05))),
}

SAVE_STEPS = (1, 2, 5, 20, 60)


def amp_field(grid):
    """A per-node alpha in [1, 2] (float32, as scripts/adaptive_alpha_study.py
    builds it), exactly 1 where the sine vanishes."""
    nx, ny, nz = grid
    shape = (nx, ny) if nz == 1 else (nx, ny, nz)
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    a = 1.0 + 0.9 * np.abs(np.sin(0.35 * idx[0]) * np.cos(0.21 * idx[1]))
    return np.clip(a, 1.0, 2.0).astype(np.float32)


def run_case(name, c):
    nx, ny, nz = c["grid"]
    dt = torch.float32 if c["dtype"] == "float32" else torch.float64
    cfg_kw = dict(nx=nx, ny=ny, nz=nz, rho_l=1.0, rho_g=1.0 / 828.0,
                  sigma=0.01424, xi=4.0, tau_l=0.53, tau_g=0.53,
                  theta_eq=c["theta"], device="cpu")
    if c["tau_h"] is not None:
        cfg_kw["tau_h"] = c["tau_h"]
    if "g_force" in c:
        cfg_kw["g_force"] = tuple(c["g_force"]) + (0.0,) * (3 - len(c["g_force"]))
    cfg = FEConfig(**cfg_kw)
    out = {}
    kw = dict(c["kw"])
    if kw.get("geo_amplification") == "field":
        out["amp_field"] = amp_field(c["grid"])
        kw["geo_amplification"] = torch.from_numpy(out["amp_field"])
    with contextlib.redirect_stdout(io.StringIO()):
        sol = AllenCahnSolver(cfg, dtype=dt, **kw)
        if c["frac"]:
            solid, frac = create_substrate_with_fraction(
                nx, ny, nz, substrate_type=c["sub"], R_star=c["R_star"],
                R_d=c["R_d"])
            sol.set_solid(solid, solid_fraction=frac)
            out["fraction"] = np.asarray(frac)
        else:
            solid = create_substrate(nx, ny, nz, substrate_type=c["sub"],
                                     R_star=c["R_star"], R_d=c["R_d"])
            sol.set_solid(solid)
        C0, _, u0 = create_fe_droplet_with_impact(
            nx, ny, nz, center=c["center"], radius=c["radius"], xi=4.0,
            rho_l=1.0, rho_g=1.0 / 828.0, u_impact=c["u"])
        sol.init_fields(C0, u0)
    out["solid"] = np.asarray(solid)
    out["phi_init"] = C0
    out["u_init"] = u0
    if sol._has_partial:
        out["partial_mask"] = sol._partial_mask.numpy()
        out["partial_eps"] = sol._partial_eps.numpy()
    w = sol.wetting
    if w is not None:
        out["is_boundary"] = w.is_boundary.numpy()
        out["wall_normal"] = w.wall_normal.numpy()
        out["wall_weight"] = w.wall_weight.numpy()
        out["near_wall_mask"] = w.near_wall_mask.numpy()
        out["wet_mode"] = np.array(w.mode)
    # init state
    out["M0"] = np.array(sol.phi_mass_init, dtype=np.float64)
    out["s0_phi"] = sol.phi.numpy()
    out["s0_mu"] = sol.mu_phi.numpy()
    out["s0_f"] = sol.f.numpy()
    out["s0_g"] = sol.g.numpy()
    # step-1 operator intermediates, through the reference's own methods
    phi1 = torch.clamp(sol.f.sum(dim=0), 0.0, 1.0)
    phi1 = torch.where(sol.solid, torch.zeros_like(phi1), phi1)
    grad, lap = sol._lattice_gradient_and_laplacian(phi1)
    out["op_phi"] = phi1.numpy()
    out["op_grad"] = grad.numpy()
    out["op_lap"] = lap.numpy()
    if w is not None and sol.geometric_wetting:
        out["op_grad_corr"] = w.correct_gradient(grad.clone()).numpy()
    # trajectory
    inj = c.get("inject")
    with contextlib.redirect_stdout(io.StringIO()):
        for s in range(1, max(SAVE_STEPS) + 1):
            sol.step()
            if inj is not None and s == inj["after"]:
                sol.inject_droplet(inj["center"], inj["radius"], inj["xi"], inj["u"])
                out["inj_phi"] = sol.phi.numpy().copy()
                out["inj_u"] = sol.u.numpy().copy()
                out["inj_f"] = sol.f.numpy().copy()
                out["inj_g"] = sol.g.numpy().copy()
                out["inj_M0"] = np.array(sol.phi_mass_init, dtype=np.float64)
            if s in SAVE_STEPS:
                out[f"s{s}_phi"] = sol.phi.numpy().copy()
                out[f"s{s}_u"] = sol.u.numpy().copy()
                out[f"s{s}_P"] = sol.P.numpy().copy()
                out[f"s{s}_rho"] = sol.rho.numpy().copy()
                out[f"s{s}_mu"] = sol.mu_phi.numpy().copy()
                if s in c["fg_steps"]:
                    out[f"s{s}_f"] = sol.f.numpy().copy()
                    out[f"s{s}_g"] = sol.g.numpy().copy()
    meta = dict(c)
    meta["kw"] = {k: v for k, v in c["kw"].items()}
    meta["save_steps"] = list(SAVE_STEPS)
    out["meta"] = np.array(json.dumps(meta))
    return out


def mms_vectors():
    """Reference gradient operator on the MMS field (scripts/mms_spatial.py:26-55)."""
    res = {}
    for N in (16, 32):
        i = np.arange(N)[:, None, None]
        j = np.arange(N)[None, :, None]
        k = np.arange(N)[None, None, :]
        phi = (np.sin(2 * np.pi * i / N) * np.cos(2 * np.pi * j / N)
               * np.cos(2 * np.pi * k / N))
        cfg = FEConfig(nx=N, ny=N, nz=N, rho_l=1.0, rho_g=1.0 / 828.0,
                       sigma=0.01, xi=4.0, beta=1.0, kappa=1.0, M=1.0,
                       tau_l=0.53, tau_g=0.53, tau_h=0.57, theta_eq=162.0,
                       device="cpu")
        with contextlib.redirect_stdout(io.StringIO()):
            sol = AllenCahnSolver(cfg, dtype=torch.float64,
                                  boundary_relax=0.0, geometric_wetting=False)
        g, l = sol._lattice_gradient_and_laplacian(torch.from_numpy(phi))
        res[f"N{N}_phi"] = phi
        res[f"N{N}_grad"] = g.numpy()
        res[f"N{N}_lap"] = l.numpy()
    return res


def main():
    only = sys.argv[1:]
    for name, c in CASES.items():
        if only and name not in only:
            continue
        out = run_case(name, c)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {os.path.getsize(path) / 1e3:.0f} kB", flush=True)
    if not only or "mms" in only:
        path = os.path.join(HERE, "mms_operator.npz")
        np.savez_compressed(path, **mms_vectors())
        print(f"mms: {os.path.getsize(path) / 1e3:.0f} kB", flush=True)


if __name__ == "__main__":
    main()
