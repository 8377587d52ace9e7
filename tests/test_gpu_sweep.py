"""The 48-case uneven impact sweep at its stated sizes (grids 96^3,
128x128x96 and 150x150x107, 1500 steps each; tests/golden/cfg4_cases.json),
the batched replacement of the reference's case-by-case sweep
(scripts/run_thesis_sweep.py:57-284).

* all 48 cases resident in one bf16 batch, advanced by shared graph-captured
  launches; a subset covering every grid size and substrate, run alone, gives
  the same fields and the same spread history bit for bit;
* fp32 batch cases equal AllenCahnSolver runs of the same host-generated
  inputs bit for bit;
* the batched bf16 max spread factor (max over the 50-step checkpoints of
  max(Dx, Dy), scripts/run_thesis_sweep.py:188-195) lies within 1% of an
  fp64 run of the same case.
"""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EVERY = 50
# one case of each grid size and substrate (and the three R*)
SUBSET = ("c00", "c02", "c05", "c07", "c09", "c12")


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def cases():
    from lbm.batch import CaseSpec
    with open(os.path.join(GOLD, "cfg4_cases.json")) as f:
        rows = json.load(f)["cases"]
    return [CaseSpec(grid=tuple(r["grid"]), substrate=r["substrate"], R_star=r["R_star"],
                     R_drop=r["R_drop"], theta_eq=r["theta_eq"], u_impact=r["u_impact"],
                     steps=r["steps"], tag=r["tag"]) for r in rows]


def run_batch(specs, storage, graph=True):
    """Advance `specs` together to their 1500 steps, measuring every case
    every 50 steps; returns the engine, the ids and each case's max(Dx, Dy)
    history."""
    from lbm.batch import BatchEngine
    eng = BatchEngine(storage)
    ids = [eng.add(s) for s in specs]
    steps = specs[0].steps
    hist = {cid: [] for cid in ids}

    def measure():
        for cid in ids:
            m = eng.measure(cid)
            hist[cid].append(max(m["Dx"], m["Dy"]))
    eng.step(EVERY)
    measure()
    if graph:
        eng.capture_graph(EVERY)
    for _ in range(steps // EVERY - 1):
        if graph:
            eng.run_graph(1)
        else:
            eng.step(EVERY)
        measure()
    for cid, s in zip(ids, specs):
        assert eng.status(cid) == (s.steps, s.steps)
    return eng, ids, hist


def solver_hist(spec, dtype, storage=None):
    from lbm.batch import host_inputs
    from lbm.fe_ac_solver import AllenCahnSolver
    cfg = spec.config()
    cfg.device = "cuda"
    sol = AllenCahnSolver(cfg, dtype=dtype, geo_amplification=spec.geo_amplification,
                          geometric_wetting=spec.geometric_wetting,
                          boundary_relax=spec.boundary_relax, storage=storage)
    solid, frac, phi0, u0 = host_inputs(spec)
    sol.set_solid(solid, frac)
    sol.init_fields(phi0, u0)
    hist = []
    for _ in range(spec.steps // EVERY):
        sol.step(EVERY)
        m = sol.measure()
        hist.append(max(m["Dx"], m["Dy"]))
    return sol, hist


def test_sweep_48_cases_batched_equal_alone():
    need_gpu()
    print("[gpu] sweep: 48 cases, bf16 batch, 1500 steps", flush=True)
    allc = cases()
    assert len(allc) == 48
    assert {c.grid for c in allc} == {(96, 96, 96), (128, 128, 96), (150, 150, 107)}
    eng, ids, hist = run_batch(allc, "bf16")
    by_tag = {c.tag: (cid, c) for cid, c in zip(ids, allc)}
    sub = [by_tag[t] for t in SUBSET]
    assert {c.grid for _, c in sub} == {c.grid for c in allc}
    assert {c.substrate for _, c in sub} == {"flat", "ridge", "convex", "concave"}
    for cid, c in sub:
        print(f"[gpu] sweep: alone {c.tag} {c.grid} {c.substrate}", flush=True)
        a_eng, (a,), a_hist = run_batch([c], "bf16", graph=False)
        assert a_hist[a] == hist[cid]
        for name in ("phi", "u"):
            np.testing.assert_array_equal(a_eng.field(a, name).cpu().numpy(),
                                          eng.field(cid, name).cpu().numpy())
        del a_eng
    # every case ran to a finite, non-vanished state
    for cid in ids:
        m = eng.measure(cid)
        assert m["nan"] == 0 and m["C_max"] > 0.5


def test_sweep_fp32_batch_equals_solver():
    """One case of each grid size, batched in fp32, against AllenCahnSolver
    runs of the same host-generated inputs: bit for bit."""
    need_gpu()
    print("[gpu] sweep: fp32 batch vs solver", flush=True)
    by_tag = {c.tag: c for c in cases()}
    sub = [by_tag[t] for t in ("c00", "c05", "c07")]
    assert len({c.grid for c in sub}) == 3
    eng, ids, hist = run_batch(sub, "fp32")
    for cid, c in zip(ids, sub):
        sol, h = solver_hist(c, torch.float32)
        assert h == hist[cid]
        for name in ("phi", "u"):
            np.testing.assert_array_equal(getattr(sol, name).cpu().numpy(),
                                          eng.field(cid, name).cpu().numpy())
        del sol


def test_sweep_spread_within_1pct_of_fp64():
    """Max spread factor of the batched bf16 cases against fp64 runs of the
    same cases (the 1% precision gate)."""
    need_gpu()
    print("[gpu] sweep: bf16 spread vs fp64", flush=True)
    by_tag = {c.tag: c for c in cases()}
    sub = [by_tag[t] for t in SUBSET]
    eng, ids, hist = run_batch(sub, "bf16")
    worst = 0.0
    for cid, c in zip(ids, sub):
        _, h64 = solver_hist(c, torch.float64)
        d16, d64 = max(hist[cid]), max(h64)
        err = abs(d16 / d64 - 1.0)
        worst = max(worst, err)
        print(f"[gpu] sweep: {c.tag} {c.grid} {c.substrate} theta {c.theta_eq:g}: "
              f"D_max bf16 {d16:g} fp64 {d64:g} ({100 * err:.2f}%)", flush=True)
    assert worst <= 0.01
