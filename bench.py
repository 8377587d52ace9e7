#!/usr/bin/env python3
"""Benchmark of the Allen-Cahn LBM step on one MI355X, printed as one JSON line.

value: MLUPS of the plain fp32 D3Q19 step on the 150x150x107 ridge impact,
with `roofline` (the collide kernel against the 8 TB/s HBM peak), the step's
% of the HBM roofline at 356 algorithmic B/LU, and beside it
  sweep    the 48-case uneven impact sweep (1500 steps each) batched in fp16
           storage (--sweep-storage) against the same cases run one after
           another in the plain fp32 form: solves/hour and the speed-up;
  serve    a seeded Poisson stream (4 requests/s for 20 s) of impact solves
           served with continuous batching: p50/p99 time to solution and the
           aggregate MLUPS;
  cpu_baseline            the step restated in torch on the CPU
                          (oracle/torch_step.py) on every host core, on the
                          same case;
  cpu_baseline_reference  the reference's own step timed on the CPU in the
                          build container (a committed fixture).

Workload of `value` (the README construction of the reference, fp32):
FEConfig(nx=150, ny=150, nz=107, rho_l=1, rho_g=1/828, sigma=0.01424, xi=4,
tau_l=tau_g=0.53, theta_eq=162), AllenCahnSolver(dtype=float32,
geo_amplification=1.5, geometric_wetting=True) with boundary_relax=0.3 and
tau_h=1.0; ridge R*=1, R_d=22.5 with solid fraction; droplet R=22.5 at
(75,75,80), u=(0,0,-0.05).  Inputs are resident in HBM before the timed
region; a "step" is one solver.step() over all 2,407,500 nodes (solid
included).  Only one GPU is supported.

Timing: after the autotuner and the W warm-up steps, the timed steps are
captured into a graph, which is first replayed untimed for at least
SUSTAIN_STEPS steps (`sustain_steps` in the line): the device reaches its
sustained-load speed only after about 0.1 s of stepping (measured, 20 timed
steps right after capture against after 500 untimed steps: 175.1-176.9 vs
173.5 us/step on one box, 170.7-174.1 vs 166.3-167.3 on another).  The timed
region is then exactly K steps between two synchronisations.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "pytorch-lbm_amd"))

import numpy as np  # noqa: E402

METRIC = ("MLUPS + % HBM roofline, D3Q19 150×150×107; "
          "sweep solves/hr + p99 latency")
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per lattice update (compulsory state traffic once per step):
# f and g read+write (2*19*2*4 = 304), u read+write (24), phi*u_prev read+write
# (24), phi write (4)  ->  356 B/LU for the whole fp32 step.
B_STEP = 356
# The collide kernel alone: f, g read+write (304), u read+write (24), phi read
# (4), previous phi read (4; phi*u_prev is recomputed from it) -> 336 B/LU.
B_COLLIDE = 336
# 16-bit distribution storage: f and g at 2 B -> 152 + 24 + 24 + 4 = 204 B/LU
# for the step, 152 + 24 + 4 + 4 = 184 B/LU for the collide kernel.
B_STEP_16 = 204
B_COLLIDE_16 = 184


def cfg2(storage=None, dtype="float32", theta=162.0):
    import torch
    from geometry.substrate import create_substrate_with_fraction
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_droplet import create_fe_droplet_with_impact
    cfg = FEConfig(nx=150, ny=150, nz=107, rho_l=1.0, rho_g=1.0 / 828.0, sigma=0.01424,
                   xi=4.0, tau_l=0.53, tau_g=0.53, theta_eq=theta, device="cuda:0")
    s = AllenCahnSolver(cfg, dtype=getattr(torch, dtype), geo_amplification=1.5,
                        geometric_wetting=True, storage=storage)
    solid, frac = create_substrate_with_fraction(150, 150, 107, "ridge", 1.0, 22.5)
    s.set_solid(solid, solid_fraction=frac)
    phi0, _, u0 = create_fe_droplet_with_impact(150, 150, 107, center=(75, 75, 80),
                                                radius=22.5, xi=4.0, u_impact=(0, 0, -0.05))
    s.init_fields(phi0, u0)
    return s, (solid, frac, phi0, u0)


def cpu_baseline(inputs, seconds=15.0, threads=None):
    """The reference's CPU path, restated: oracle/torch_step.py (the step in
    torch on the CPU, pinned against the numpy oracle) on the same inputs,
    with torch.set_num_threads(os.cpu_count()), timed over whole steps of the
    full 150x150x107 case for about `seconds` of CPU work."""
    import torch
    from oracle.ac_oracle import OracleSolver, Params
    from oracle.torch_step import TorchStep
    solid, frac, phi0, u0 = inputs
    threads = threads or os.cpu_count() or 1
    p = Params(nx=150, ny=150, nz=107, rho_l=1.0, rho_g=1.0 / 828.0, sigma=0.01424, xi=4.0,
               tau_l=0.53, tau_g=0.53, theta_eq=162.0)
    o = OracleSolver(p, dtype=np.float32, geo_amplification=1.5, geometric_wetting=True)
    o.set_solid(solid, frac)
    o.init_fields(phi0, u0)
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        t = TorchStep(o)
        t0 = time.perf_counter()
        t.step()                                    # first step: warm-up, and the pace
        first = time.perf_counter() - t0
        nsteps = max(2, min(40, int(seconds / max(first, 1e-3))))
        t0 = time.perf_counter()
        for _ in range(nsteps):
            t.step()
        dt = time.perf_counter() - t0
    finally:
        torch.set_num_threads(prev)
    n = 150 * 150 * 107
    return {"value": round(n * nsteps / dt / 1e6, 4), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": f"{nsteps} steps (after 1 untimed) of the 150x150x107 ridge case: "
                      f"oracle/torch_step.py, torch {torch.__version__} on the CPU, float32, "
                      f"{threads} threads, {dt:.1f} s on {cpu_model()}"}


def hbm_stream(bytes_=1 << 30, reps=5):
    """Streaming bandwidth of the device in TB/s (lbm_hbm_stream: read, write
    and copy kernels over 1 GiB buffers, best of three grid sizes): the
    practical ceiling the step's measured traffic is compared with."""
    import ctypes
    from lbm import _native as nat
    out = (ctypes.c_double * 6)()
    nat.check(nat.lib.lbm_hbm_stream(bytes_, reps, nat.stream_ptr(), out), "lbm_hbm_stream")
    return dict(zip(("read", "read_nt", "write", "write_nt", "copy", "copy_nt"),
                    (round(v, 3) for v in out)))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def reference_cpu_timing():
    """The reference's own AllenCahnSolver.step() timed on the CPU in the
    build container (tests/golden/make_cfg2_reference.py writes the fixture):
    reported beside the port's figure, labelled as measured elsewhere."""
    path = os.path.join(ROOT, "tests", "golden", "reference_cpu_timing.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    return {"value": round(rec["MLUPS"], 4), "unit": "MLUPS", "cores": rec["torch_threads"],
            "kind": "reference",
            "sample": f"{rec['steps_timed']} steps after {rec['warmup']} of the same workload, "
                      f"the reference's own step (torch {rec['torch']}, CPU, float32, "
                      f"{rec['torch_threads']} threads) on {rec['cpu_model'].split(':', 1)[-1].strip()}; "
                      "measured in the build container, not on this host"}


def hbm_traffic(storage):
    """Per-launch HBM bytes of pass B measured with rocprofv3 PMC counters
    (FETCH_SIZE and WRITE_SIZE passes of this same command, each scaled by the
    reading of a device copy of known size), as committed in
    profiles/hbm_traffic.json; None when no measurement exists for `storage`."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
        t = rec["k_collide"][storage]
        return int(t["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def precision_report(steps=2000, thetas=(90.0, 120.0, 162.0), every=20):
    """Storage forms against an fp64 run of the same kernels on the 150x150x107
    ridge impact: relative L2 of phi and u over fluid nodes at steps
    500/1000/1500/2000, the max spread factor D_max/D0 (D = max(Dx, Dy) of the
    phi > 0.5 non-solid extents, sampled every `every` steps, D0 = 45) and the
    contact-line radius about the impact axis (75, 75)."""
    import torch
    cps = [c for c in (500, 1000, 1500, 2000) if c <= steps]
    forms = [("fp64", None, "float64"), ("fp32", None, "float32"),
             ("bf16", "bf16", "float32"), ("fp16", "fp16", "float32")]
    out = {}
    worst = 0.0
    for th in thetas:
        ref = None
        rows = {}
        for name, storage, dt in forms:
            s, inputs = cfg2(storage, dt, theta=th)
            fluid = ~inputs[0]
            snaps, dmax, rmax = {}, 0.0, 0.0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            k = 0
            while k < steps:
                n = min(every, steps - k)
                s.step(n)
                k += n
                m = s.measure()
                dmax = max(dmax, max(m["Dx"], m["Dy"]))
                r = s.contact_radius(75.0, 75.0)
                rmax = max(rmax, r)
                if k in cps:
                    snaps[k] = (s.phi.double().cpu().numpy()[fluid],
                                s.u.double().cpu().numpy()[:, fluid], r)
            wall = time.perf_counter() - t0
            rec = {"spread_factor_max": dmax / 45.0, "contact_radius_max": rmax,
                   "contact_radius": {str(c): snaps[c][2] for c in cps},
                   "nan": bool(m["nan"]), "wall_s": round(wall, 2)}
            if name == "fp64":
                ref = snaps, rec
            else:
                rs, rr = ref
                rec["phi_l2"] = {str(c): float(np.linalg.norm(snaps[c][0] - rs[c][0]) /
                                               np.linalg.norm(rs[c][0])) for c in cps}
                rec["u_l2"] = {str(c): float(np.linalg.norm(snaps[c][1] - rs[c][1]) /
                                             max(np.linalg.norm(rs[c][1]), 1e-300)) for c in cps}
                e_sf = abs(rec["spread_factor_max"] / rr["spread_factor_max"] - 1.0)
                e_rc = abs(rec["contact_radius_max"] - rr["contact_radius_max"]) / max(
                    rr["contact_radius_max"], 1e-300)
                rec["spread_factor_rel_err"] = e_sf
                rec["contact_radius_rel_err"] = e_rc
                worst = max(worst, e_sf, e_rc)
            rows[name] = rec
            print(f"[precision] theta {th:g} {name}: spread {rec['spread_factor_max']:.4f} "
                  f"r_cl {rec['contact_radius_max']:.3f} ({wall:.1f} s)", file=sys.stderr, flush=True)
            del s
            torch.cuda.empty_cache()
        out[f"{th:g}"] = rows
    return {"metric": "precision forms vs an fp64 run of the same kernels (150x150x107 ridge "
                      "impact, 2000 steps): worst relative error of max spread factor and "
                      "contact-line radius",
            "value": round(100 * worst, 4), "unit": "%", "higher_is_better": False,
            "gate_pct": 1.0, "steps": steps, "n_gpus": 1, "thetas": out}


def sweep_bench(steps=1500, every=50, storage="bf16", seed=2024, n=48):
    """The 48-case uneven impact sweep: batched (all cases device-resident,
    one launch per pass for all of them, graph-captured, autotuned launch
    shape, `storage` distributions) against the same cases run one after
    another through the plain fp32 single-case path (host generators,
    set_solid, init_fields, step loop), both with on-device diagnostics every
    `every` steps."""
    import torch
    from lbm.batch import BatchEngine, host_inputs, sweep_cases
    from lbm.fe_ac_solver import AllenCahnSolver
    cases = sweep_cases(seed, n, steps)
    lu = sum(c.nodes for c in cases) * steps

    # batched
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng = BatchEngine(storage)
    # inputs built on the device (the library's generators; the sequential
    # runs below take the numpy generators, which may differ in the last fp32
    # bit of a few tanh values)
    ids = [eng.add(c, inputs="device") for c in cases]
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    eng.step(1)
    eng.autotune(3)
    tune = eng.launch()
    done = eng.status(ids[0])[0]
    g = every if every % 2 == 0 else every + 1
    reps = (steps - done) // g
    eng.capture_graph(g)
    # diagnostics of every case written on the device in stream order after
    # each graph replay; the running maximum stays on the device (no host
    # round trip between replays)
    diag = torch.empty((len(ids), 10), dtype=torch.float64, device="cuda")
    best_b = torch.zeros(len(ids), dtype=torch.float64, device="cuda")
    t_step0 = time.perf_counter()
    for _ in range(reps):
        eng.run_graph(1)
        for j, cid in enumerate(ids):
            eng.measure_async(cid, diag[j])
        best_b = torch.maximum(best_b, diag[:, :2].amax(dim=1))
    rest = steps - eng.status(ids[0])[0]
    if rest > 0:
        eng.step(rest)
    for j, cid in enumerate(ids):
        eng.measure_async(cid, diag[j])
    best_b = torch.maximum(best_b, diag[:, :2].amax(dim=1))
    best_b = best_b.cpu().numpy()
    t_b = time.perf_counter() - t0
    t_b_step = time.perf_counter() - t_step0
    assert all(eng.status(cid) == (steps, steps) for cid in ids)
    spread_b = {cid: float(best_b[j]) for j, cid in enumerate(ids)}
    del eng
    torch.cuda.empty_cache()
    print(f"[sweep] batched {storage}: {t_b:.2f} s (generation {t_gen:.2f} s)",
          file=sys.stderr, flush=True)

    # sequential plain fp32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tts, spread_s, t_steps = [], [], 0.0
    for c in cases:
        nx, ny, nz = c.grid
        cfg = c.config()
        cfg.device = "cuda:0"
        sol = AllenCahnSolver(cfg, dtype=torch.float32, geo_amplification=c.geo_amplification,
                              geometric_wetting=c.geometric_wetting,
                              boundary_relax=c.boundary_relax)
        solid, frac, phi0, u0 = host_inputs(c)
        sol.set_solid(solid, solid_fraction=frac)
        sol.init_fields(phi0, u0)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        diag1 = torch.empty(10, dtype=torch.float64, device="cuda")
        best = torch.zeros((), dtype=torch.float64, device="cuda")
        k = 0
        while k < steps:
            m_ = min(every, steps - k)
            sol.step(m_)
            k += m_
            sol.measure_async(diag1)
            best = torch.maximum(best, diag1[:2].amax())
        best = float(best)
        t_steps += time.perf_counter() - ts
        spread_s.append(best)
        tts.append(time.perf_counter() - t0)
        del sol
    t_s = time.perf_counter() - t0
    torch.cuda.empty_cache()
    print(f"[sweep] sequential fp32: {t_s:.2f} s (stepping {t_steps:.2f} s)",
          file=sys.stderr, flush=True)
    sf_err = max(abs(spread_b[cid] / spread_s[i] - 1.0) for i, cid in enumerate(ids))
    return {
        "metric": "sweep solves/hour, 48 uneven D3Q19 impact cases x 1500 steps (batched, "
                  f"{storage} storage, graph-captured) vs the same cases run sequentially in "
                  "the plain fp32 form",
        "value": round(n / t_b * 3600.0, 1), "unit": "solves/h", "higher_is_better": True,
        "n_gpus": 1, "steps": steps,
        "speedup_vs_sequential": round(t_s / t_b, 3),
        "batched": {"wall_s": round(t_b, 3), "generation_s": round(t_gen, 3),
                    "stepping_s": round(t_b_step, 3), "MLUPS": round(lu / t_b / 1e6, 1),
                    "storage": storage, "graph_steps": g,
                    "launch": dict(zip(("block_phi", "block_collide", "cap_phi", "cap_collide",
                                        "split"), tune))},
        "sequential_fp32": {"wall_s": round(t_s, 3), "stepping_s": round(t_steps, 3),
                            "MLUPS": round(lu / t_s / 1e6, 1),
                            "solves_per_hour": round(n / t_s * 3600.0, 1),
                            "p50_time_to_solution_s": round(float(np.percentile(tts, 50)), 3),
                            "p99_time_to_solution_s": round(float(np.percentile(tts, 99)), 3)},
        "stepping_speedup": round(t_steps / t_b_step, 3),
        "spread_factor_max_rel_diff_batched_vs_sequential": round(sf_err, 5),
        "cases": {"seed": seed, "n": n, "lattice_updates": lu},
    }


def serve_bench(duration=20.0, rate=4.0, seed=5, storage="bf16", slots=32, chunk=50,
                cpu_sample_steps=20):
    """Continuous-batching serving of a seeded Poisson stream of impact solves
    (mean `rate` requests/s for `duration` s; grids 64^3 .. 256x256x160,
    500-3000 steps; `storage` distributions, fp32 compute): admission by HBM
    footprint and execution slots, lattices device-resident.  Reports p50/p99
    time-to-solution (arrival to result) and aggregate MLUPS; the step
    restated in torch on the host CPU (oracle/torch_step.py, all cores) is
    timed on a 64^3 sample in the same run."""
    import torch
    from lbm.serve import SweepServer, poisson_requests
    reqs = poisson_requests(seed, rate, duration)
    free, total = torch.cuda.mem_get_info()
    budget = int(0.9 * free)
    srv = SweepServer(storage, budget_bytes=budget, max_active=slots, chunk=chunk,
                      inputs="device")
    peak = {"active": 0, "bytes": 0}

    def on_event(kind, req):
        peak["active"] = max(peak["active"], len(srv.active))
        peak["bytes"] = max(peak["bytes"], srv.engine.bytes_used())
    done = srv.run(reqs, on_event=on_event)
    tts = np.array([r.t_done - r.t_arrival for r in done])
    wait = np.array([r.t_admit - r.t_arrival for r in done])
    lu = sum(r.lattice_updates for r in done)
    print(f"[serve] {len(done)} requests in {srv.wall:.1f} s, {lu / srv.wall / 1e6:.0f} MLUPS",
          file=sys.stderr, flush=True)
    # CPU: the step restated in torch on the host CPU, all cores, on a 64^3
    # flat case
    from oracle.ac_oracle import OracleSolver, Params
    from oracle.torch_step import TorchStep
    n = 64
    p = Params(nx=n, ny=n, nz=n, rho_l=1.0, rho_g=1.0 / 828.0, sigma=0.01424, xi=4.0,
               tau_l=0.53, tau_g=0.53, theta_eq=126.0)
    o = OracleSolver(p, dtype=np.float32, geo_amplification=1.5, geometric_wetting=True)
    solid = np.zeros((n, n, n), bool)
    solid[:, :, 0] = True
    o.set_solid(solid)
    X, Y, Z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((X - 32.0) ** 2 + (Y - 32.0) ** 2 + (Z - 30.0) ** 2)
    phi0 = np.clip(0.5 + 0.5 * np.tanh(2.0 * (9.6 - r) / 4.0), 0, 1).astype(np.float32)
    o.init_fields(phi0, np.zeros((3, n, n, n), np.float32))
    threads = os.cpu_count() or 1
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        ts = TorchStep(o)
        ts.step()                                   # untimed warm-up step
        t0 = time.perf_counter()
        for _ in range(cpu_sample_steps):
            ts.step()
        cpu_dt = time.perf_counter() - t0
    finally:
        torch.set_num_threads(prev)
    cpu_mlups = n ** 3 * cpu_sample_steps / cpu_dt / 1e6
    med_lu = float(np.median([r.lattice_updates for r in done]))
    return {
        "metric": "serving: p99 time-to-solution of a Poisson stream (4 req/s) of D3Q19 impact "
                  "solves, continuous batching, bf16 storage / fp32 compute",
        "value": round(float(np.percentile(tts, 99)), 3), "unit": "s", "higher_is_better": False,
        "n_gpus": 1,
        "p50_time_to_solution_s": round(float(np.percentile(tts, 50)), 3),
        "p99_time_to_solution_s": round(float(np.percentile(tts, 99)), 3),
        "p50_queue_wait_s": round(float(np.percentile(wait, 50)), 3),
        "p99_queue_wait_s": round(float(np.percentile(wait, 99)), 3),
        "aggregate_MLUPS": round(lu / srv.wall / 1e6, 1),
        "requests": len(done), "arrival_window_s": duration, "rate_per_s": rate,
        "offered_MLUPS": round(lu / duration / 1e6, 1), "wall_s": round(srv.wall, 2),
        "slots": slots, "chunk_steps": chunk, "hbm_budget_GB": round(budget / 1e9, 1),
        "peak_active": peak["active"], "peak_resident_GB": round(peak["bytes"] / 1e9, 2),
        "storage": storage, "any_nan": any(r.result["nan"] for r in done),
        "cpu_baseline": {"value": round(cpu_mlups, 3), "unit": "MLUPS", "cores": threads,
                         "kind": "port",
                         "sample": f"{cpu_sample_steps} steps (after 1 untimed) of a 64^3 flat "
                                   f"impact case, oracle/torch_step.py on the CPU (float32, "
                                   f"{threads} threads), {cpu_dt:.1f} s",
                         "median_request_projected_s": round(med_lu / (cpu_mlups * 1e6), 1)},
    }


SUSTAIN_STEPS = 500


def graph_chunk(steps, most=50):
    """Largest even g <= most dividing steps (steps - 1 when steps is odd);
    0 when fewer than two steps are timed."""
    n = steps - (steps & 1)
    for g in range(min(most, n) & ~1, 1, -2):
        if n % g == 0:
            return g
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--storage", default=None, choices=[None, "fp32", "bf16", "fp16"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of replays of a captured step graph")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape (256-thread blocks, no occupancy cap)")
    ap.add_argument("--launch", default=None,
                    help="fixed launch shape instead of the autotuner: block A, block B, "
                         "cap A, cap B, form of pass B (e.g. 256,256,0,4,2)")
    ap.add_argument("--fused-tile", default=None,
                    help="ty,tx,zlen of the fused step (form 3 of --launch; 0 = automatic)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU work of the torch-CPU baseline sample (whole steps)")
    ap.add_argument("--sweep-storage", default="fp16", choices=["fp32", "bf16", "fp16"],
                    help="distribution storage of the batched sweep (serving stays bf16)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the 48-case batched sweep in the default line")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the Poisson serving run in the default line")
    ap.add_argument("--mode", default="plain", choices=["plain", "precision", "sweep", "serve"],
                    help="plain: the roofline/MLUPS line (default); precision: storage forms "
                         "against fp64 on the same case; sweep: 48 uneven cases batched vs "
                         "sequential; serve: continuous batching of a Poisson request stream")
    ap.add_argument("--rate", type=float, default=4.0,
                    help="serve mode: mean request arrivals per second of the Poisson stream")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run a device copy of this many MiB (a known-size kernel "
                         "for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solver is a single-device code "
              "(multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.mode == "precision":
        print(json.dumps(precision_report()), flush=True)
        return
    if args.mode == "sweep":
        print(json.dumps(sweep_bench(storage=args.sweep_storage)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(serve_bench(rate=args.rate)), flush=True)
        return
    print(f"[bench] building case on {torch.cuda.get_device_name(0)}", file=sys.stderr,
          flush=True)
    if args.copy_probe_mb:
        n_el = args.copy_probe_mb * (1 << 20) // 4
        src = torch.rand(n_el, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    stream_tbs = hbm_stream()
    s, inputs = cfg2(args.storage)
    n = 150 * 150 * 107
    sixteen = args.storage in ("bf16", "fp16")
    b_step = B_STEP_16 if sixteen else B_STEP
    b_coll = B_COLLIDE_16 if sixteen else B_COLLIDE
    if args.launch:
        ba, bb, ca, cb, form = (int(v) for v in args.launch.split(","))
        s.set_split(form)
        s.set_occupancy(ca, cb)
        s.set_block_sizes(ba, bb)
        if args.fused_tile:
            s.set_fused_tile(*(int(v) for v in args.fused_tile.split(",")))
    elif not args.no_autotune:
        # 27 launch shapes x 11 real steps (297 steps): its timings decide
        # the shape, and they run before the warm-up steps
        s.autotune(10)
    s.step(max(args.warmup, 1))
    # the timed steps run as replays of a captured graph of g steps (even,
    # dividing the step count; an odd count runs its last step eagerly),
    # replayed untimed for at least SUSTAIN_STEPS steps first
    g = 0 if args.eager else graph_chunk(args.steps)
    sustain = 0
    if g:
        s.capture_graph(g)
        reps = max(1, -(-SUSTAIN_STEPS // g))
        s.run_graph(reps)
        sustain = reps * g
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    if g:
        s.run_graph(args.steps // g)
        if args.steps % g:
            s.step(args.steps % g)
    else:
        s.step(args.steps)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = start.elapsed_time(stop) / args.steps
    mlups = n / (ms_step * 1e-3) / 1e6
    print(f"[bench] {args.steps} steps: {ms_step * 1e3:.1f} us/step, {mlups:.0f} MLUPS "
          f"(wall {wall:.2f} s)", file=sys.stderr, flush=True)
    # per-kernel durations with events on the launch stream
    prof_steps = min(100, max(10, args.steps // 5))
    ta, tb, tstep = s.profile_steps(prof_steps)
    achieved = b_coll * n / (tb * 1e-3) / 1e9
    step_gbs = b_step * n / (ms_step * 1e-3) / 1e9
    mass0, mass = s.mass()
    meas = s.measure()
    rec = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage in (None, "fp32") else f"f32-compute/{args.storage}-storage",
        "data": "synthetic (deterministic droplet-impact initial condition, no RNG)",
        "config": {"workload": "ridge impact, D3Q19 150x150x107, README construction, "
                               "fp32, 2,407,500 nodes incl. solid",
                   "grid": [150, 150, 107], "substrate": "ridge R*=1", "theta_eq": 162.0,
                   "geo_amplification": 1.5, "boundary_relax": 0.3,
                   "storage": args.storage or "fp32"},
        "hbm_roofline_pct_step": round(100 * step_gbs / HBM_PEAK_GBS, 1),
        "bytes_per_lu_step": b_step,
        "roofline": {"bound": "hbm", "kernel": "k_collide (pass B)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": hbm_traffic(args.storage or "fp32"),
                     "bytes_per_lu": b_coll, "avg_ms": round(tb, 5),
                     "launch": dict(zip(("block_phi", "block_collide", "cap_phi",
                                         "cap_collide", "split"), s.launch()))},
        "hbm_stream_TBps": stream_tbs,
        "timing": (f"replays of a captured {g}-step graph" if g else "eager launches"),
        "sustain_steps": sustain,
        "kernels_ms": {"k_phi": round(ta, 5), "k_collide": round(tb, 5),
                       "step_events": round(tstep, 5)},
        "mass_drift": abs(mass - mass0) / mass0,
        "diag": {k: meas[k] for k in ("Dx", "Dy", "Dz", "k", "z_com", "nan")},
    }
    del s
    torch.cuda.empty_cache()
    if not args.no_sweep:
        print("[bench] sweep ...", file=sys.stderr, flush=True)
        sw = sweep_bench(storage=args.sweep_storage)
        rec["sweep"] = {k: sw[k] for k in ("value", "unit", "speedup_vs_sequential",
                                           "stepping_speedup")}
        rec["sweep"].update(
            workload="48 uneven D3Q19 impact cases (tests/golden/cfg4_cases.json) x 1500 "
                     f"steps, batched {args.sweep_storage} storage vs sequential plain fp32",
            storage=args.sweep_storage, launch=sw["batched"]["launch"],
            batched_MLUPS=sw["batched"]["MLUPS"], batched_wall_s=sw["batched"]["wall_s"],
            sequential_wall_s=sw["sequential_fp32"]["wall_s"],
            sequential_solves_per_hour=sw["sequential_fp32"]["solves_per_hour"],
            spread_max_rel_diff=sw["spread_factor_max_rel_diff_batched_vs_sequential"])
    if not args.no_serve:
        print("[bench] serve ...", file=sys.stderr, flush=True)
        sv = serve_bench(rate=args.rate)
        rec["serve"] = {k: sv[k] for k in ("p50_time_to_solution_s", "p99_time_to_solution_s",
                                           "aggregate_MLUPS", "offered_MLUPS", "requests",
                                           "rate_per_s", "p99_queue_wait_s", "peak_active",
                                           "storage", "any_nan", "cpu_baseline")}
        rec["serve"]["workload"] = ("seeded Poisson stream, 4 requests/s for 20 s, grids "
                                    "64^3 .. 256x256x160, 500-3000 steps, continuous batching")
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        rec["cpu_baseline"] = cpu_baseline(inputs, args.cpu_seconds)
        ref = reference_cpu_timing()
        if ref is not None:
            rec["cpu_baseline_reference"] = ref
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
