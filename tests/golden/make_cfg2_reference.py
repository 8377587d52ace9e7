#!/usr/bin/env python3
"""Run the reference at the full 150x150x107 ridge-impact size on CPU and keep
checkpoint summaries plus its CPU step time.

Configuration: the README construction (README.md:56-81) in fp32 --
FEConfig(nx=150, ny=150, nz=107, rho_l=1, rho_g=1/828, sigma=0.01424, xi=4,
tau_l=tau_g=0.53, theta_eq=162), AllenCahnSolver(cfg, dtype=float32,
geo_amplification=1.5, geometric_wetting=True) with its defaults
(tau_h=1.0, boundary_relax=0.3), ridge R*=1, R_d=22.5 with fraction, droplet
R=22.5 at (75,75,80), u=(0,0,-0.05).

Writes tests/golden/cfg2_ridge162_summary.npz (per checkpoint: sum/sumsq/max of
phi, sum of |u|^2, extents Dx/Dy/Dz/z_com of phi>0.5 & ~solid, mid-plane
slices phi[:,75,:] and u_z[:,75,:]) and tests/golden/reference_cpu_timing.json.
With REF_DTYPE=float64 the same run is made in the reference's fp64 form
(dtype=torch.float64, the same float32 generator inputs) and written to
tests/golden/cfg2_ridge162_f64_summary.npz (slices kept in fp64); the
distance between the two is the reference's own fp32 rounding envelope.

REF_OUT=<dir> writes the summary there instead (copy it into tests/golden/
once the run has reached its last checkpoint).

Usage (build container only):
    python tests/golden/make_cfg2_reference.py [steps]
    REF_DTYPE=float64 python tests/golden/make_cfg2_reference.py [steps]
"""
import os
import sys
import io
import json
import time
import platform
import contextlib

import numpy as np
import torch

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REF)

from lbm.fe_config import FEConfig  # noqa: E402
from lbm.fe_ac_solver import AllenCahnSolver  # noqa: E402
from lbm.fe_droplet import create_fe_droplet_with_impact  # noqa: E402
from geometry.substrate import create_substrate_with_fraction  # noqa: E402

CHECKPOINTS = (1, 10, 100, 500, 1000, 1500, 2000)


def summary(sol):
    phi = sol.phi.numpy().astype(np.float64)
    u = sol.u.numpy().astype(np.float64)
    solid = sol.solid.numpy()
    inter = (phi > 0.5) & ~solid
    d = {"phi_sum": phi.sum(), "phi_sumsq": (phi * phi).sum(),
         "phi_max": phi.max(), "u2_sum": (u * u).sum(),
         "P_sum": sol.P.numpy().astype(np.float64).sum()}
    if inter.any():
        c = np.argwhere(inter)
        d.update(Dx=float(np.ptp(c[:, 0]) + 1), Dy=float(np.ptp(c[:, 1]) + 1),
                 Dz=float(np.ptp(c[:, 2]) + 1), z_com=float(c[:, 2].mean()))
    return d


def main():
    nsteps = int(sys.argv[1]) if len(sys.argv) > 1 else 2000
    nthreads = int(os.environ.get("REF_THREADS", "6"))
    dtype = getattr(torch, os.environ.get("REF_DTYPE", "float32"))
    f64 = dtype == torch.float64
    fname = "cfg2_ridge162_f64_summary.npz" if f64 else "cfg2_ridge162_summary.npz"
    torch.set_num_threads(nthreads)
    cfg = FEConfig(nx=150, ny=150, nz=107, rho_l=1.0, rho_g=1.0 / 828.0,
                   sigma=0.01424, xi=4.0, tau_l=0.53, tau_g=0.53,
                   theta_eq=162.0, device="cpu", max_steps=2000)
    with contextlib.redirect_stdout(io.StringIO()):
        sol = AllenCahnSolver(cfg, dtype=dtype, geo_amplification=1.5,
                              geometric_wetting=True)
        solid, frac = create_substrate_with_fraction(
            150, 150, 107, substrate_type="ridge", R_star=1.0, R_d=22.5)
        sol.set_solid(solid, solid_fraction=frac)
        phi0, _, u0 = create_fe_droplet_with_impact(
            150, 150, 107, center=(75, 75, 80), radius=22.5, xi=4.0,
            u_impact=(0, 0, -0.05))
        sol.init_fields(phi0, u0)
    out = {"M0": np.array(sol.phi_mass_init)}
    times = []
    for s in range(1, nsteps + 1):
        t0 = time.perf_counter()
        with contextlib.redirect_stdout(io.StringIO()):
            sol.step()
        times.append(time.perf_counter() - t0)
        if s in CHECKPOINTS:
            for k, v in summary(sol).items():
                out[f"s{s}_{k}"] = np.array(v)
            out[f"s{s}_phi_xz"] = sol.phi.numpy()[:, 75, :].copy()
            out[f"s{s}_uz_xz"] = sol.u.numpy()[2, :, 75, :].copy()
            # written whole, then renamed: a run cut short keeps the last
            # complete checkpoint set
            dest = os.environ.get("REF_OUT", HERE)
            tmp = os.path.join(dest, fname + ".part.npz")
            np.savez_compressed(tmp, **out)
            os.replace(tmp, os.path.join(dest, fname))
            print(f"step {s}: {np.mean(times[-min(len(times), 20):]):.3f} s/step",
                  flush=True)
        if s == 25 and not f64:
            t = np.array(times[5:25])
            rec = {"workload": "ridge impact D3Q19 150x150x107 fp32, README "
                               "construction (tau_h=1.0, boundary_relax=0.3)",
                   "steps_timed": 20, "warmup": 5,
                   "s_per_step": float(t.mean()),
                   "MLUPS": 150 * 150 * 107 / float(t.mean()) / 1e6,
                   "torch_threads": nthreads, "torch": torch.__version__,
                   "cpu": platform.processor() or "see lscpu",
                   "cpu_model": open("/proc/cpuinfo").read().split(
                       "model name")[1].split("\n")[0].split(":", 1)[-1].strip(),
                   "logical_cpus": os.cpu_count()}
            with open(os.path.join(HERE, "reference_cpu_timing.json"), "w") as f:
                json.dump(rec, f, indent=1)
            print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
