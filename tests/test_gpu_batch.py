"""Batch engine: several cases of uneven size advanced by shared launches.

Each result must be the same alone and in the mix (bit for bit: every case keeps
its own reduction tree and buffers), cases stop at their own step targets, a
captured batch graph replays exactly, admission respects the HBM budget, and a
batch case matches the same case built through AllenCahnSolver from the numpy
generators.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def specs():
    from lbm.batch import CaseSpec
    return [
        CaseSpec(grid=(24, 24, 20), substrate="ridge", R_star=1.0, R_drop=5.0,
                 theta_eq=162.0, u_impact=-0.05, steps=30, tag="a"),
        CaseSpec(grid=(32, 28, 24), substrate="concave", R_star=1.0, R_drop=6.0,
                 theta_eq=126.0, u_impact=-0.03, steps=20, tag="b"),
        CaseSpec(grid=(20, 20, 18), substrate="flat", R_star=1.0, R_drop=4.0,
                 theta_eq=90.0, u_impact=-0.06, steps=30, tag="c"),
        CaseSpec(grid=(28, 24, 24), substrate="convex", R_star=0.5, R_drop=5.0,
                 theta_eq=144.0, u_impact=-0.04, steps=26, tag="d"),
    ]


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_same_alone_and_in_the_mix(storage):
    need_gpu()
    from lbm.batch import BatchEngine
    print(f"[gpu] batch alone vs mix {storage}", flush=True)
    sp = specs()
    mix = BatchEngine(storage)
    ids = [mix.add(s) for s in sp]
    mix.step(40)           # beyond every target: cases stop at their own
    for cid, s in zip(ids, sp):
        assert mix.status(cid) == (s.steps, s.steps)
    for k, (cid, s) in enumerate(zip(ids, sp)):
        alone = BatchEngine(storage)
        alone.set_split(k % 2)               # one launch or the chunk lists: same result
        a = alone.add(s)
        alone.step(s.steps)
        for name in ("phi", "u", "f"):
            np.testing.assert_array_equal(alone.field(a, name).cpu().numpy(),
                                          mix.field(cid, name).cpu().numpy())
        assert alone.measure(a) == mix.measure(cid)


def test_batch_graph_replay_matches_eager():
    need_gpu()
    from lbm.batch import BatchEngine
    sp = specs()
    for s in sp:
        s.steps = 21
    e, g = BatchEngine("fp32"), BatchEngine("fp32")
    ie = [e.add(s) for s in sp]
    ig = [g.add(s) for s in sp]
    e.step(1)
    g.step(1)
    g.capture_graph(4)
    for _ in range(5):
        e.step(4)
        g.run_graph(1)
        torch.cuda.synchronize()
    for a, b in zip(ie, ig):
        assert e.status(a) == g.status(b) == (21, 21)
        np.testing.assert_array_equal(e.field(a, "phi").cpu().numpy(),
                                      g.field(b, "phi").cpu().numpy())


def test_admission_by_budget():
    need_gpu()
    from lbm import _native as nat
    from lbm.batch import BatchEngine, footprint
    sp = specs()
    budget = footprint(sp[0], "bf16") + footprint(sp[2], "bf16") + 1000
    b = BatchEngine("bf16", budget_bytes=budget)
    b.add(sp[0])
    b.add(sp[2])
    with pytest.raises(nat.LbmError):
        b.add(sp[1])
    assert b.bytes_used() <= budget


def test_batch_case_matches_solver_api():
    """The device-built case equals the AllenCahnSolver case built from the
    numpy generators (device tanh may differ from numpy's in the last fp32 bit
    of a few initial values, hence a tolerance)."""
    need_gpu()
    from geometry.substrate import create_substrate_with_fraction
    from lbm.batch import BatchEngine, CaseSpec
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_droplet import create_fe_droplet_with_impact
    s = CaseSpec(grid=(24, 24, 20), substrate="ridge", R_star=1.0, R_drop=5.0,
                 theta_eq=162.0, u_impact=-0.05, steps=10)
    b = BatchEngine("fp32")
    cid = b.add(s)
    b.step(10)
    cfg = s.config()
    cfg.device = "cuda"
    sol = AllenCahnSolver(cfg, dtype=torch.float32, geo_amplification=1.5,
                          geometric_wetting=True)
    solid, frac = create_substrate_with_fraction(24, 24, 20, "ridge", 1.0, 5.0)
    sol.set_solid(solid, frac)
    # placement rule of scripts/run_thesis_sweep.py:146-157 on the solid that
    # is handed to set_solid (the volume-penalization solid)
    top = int(np.nonzero(solid[12, 12])[0].max())
    cz = min(top + 2.0 + 5.0, 20 - 5.0 - 2)
    phi0, _, u0 = create_fe_droplet_with_impact(24, 24, 20, center=(12.0, 12.0, cz), radius=5.0,
                                                xi=4.0, u_impact=(0, 0, -0.05))
    sol.init_fields(phi0, u0)
    sol.step(10)
    a = sol.phi.cpu().numpy()
    c = b.field(cid, "phi").cpu().numpy()
    assert np.linalg.norm(a - c) / np.linalg.norm(a) < 1e-5
