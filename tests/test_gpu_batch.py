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


@pytest.mark.parametrize("storage", ["fp32", "bf16", "fp16"])
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
        alone.set_split(k % 3)               # every form of pass B: same result
        alone.set_occupancy(0, (0, 4, 5)[k % 3])
        a = alone.add(s)
        alone.step(s.steps)
        for name in ("phi", "u", "f"):
            np.testing.assert_array_equal(alone.field(a, name).cpu().numpy(),
                                          mix.field(cid, name).cpu().numpy())
        assert alone.measure(a) == mix.measure(cid)


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_single_launch_form_and_autotune(storage):
    """The single-launch pass-B form of the batch (and whatever shape the
    autotuner picks) gives the default form's results bit for bit, cases in
    their first (post-stream form) step included."""
    need_gpu()
    from lbm.batch import BatchEngine
    sp = specs()
    ref = BatchEngine(storage)
    ir = [ref.add(s) for s in sp]
    ref.step(40)
    mx = BatchEngine(storage)
    im = [mx.add(s) for s in sp]
    mx.set_split(2)
    mx.step(3)
    mx.autotune(1)                       # 17 shapes x 2 steps
    launch = mx.launch()
    assert launch[4] in (0, 1, 2) and launch[1] in (128, 256, 512)
    mx.step(40)
    for a, b in zip(ir, im):
        assert mx.status(b) == ref.status(a)
        for name in ("phi", "u", "f", "g"):
            np.testing.assert_array_equal(ref.field(a, name).cpu().numpy(),
                                          mx.field(b, name).cpu().numpy())


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


def _solver_case(s, steps, storage=None, dtype=torch.float32):
    """The same case through AllenCahnSolver, built from the numpy generators
    and placed by the sweep driver's rule."""
    from lbm.batch import host_inputs
    from lbm.fe_ac_solver import AllenCahnSolver
    cfg = s.config()
    cfg.device = "cuda"
    sol = AllenCahnSolver(cfg, dtype=dtype, geo_amplification=s.geo_amplification,
                          geometric_wetting=s.geometric_wetting,
                          boundary_relax=s.boundary_relax, storage=storage)
    solid, frac, phi0, u0 = host_inputs(s)
    sol.set_solid(solid, frac)
    sol.init_fields(phi0, u0)
    sol.step(steps)
    return sol


@pytest.mark.parametrize("storage", ["fp32", "bf16", "fp16"])
def test_batch_case_matches_solver_api(storage):
    """A batch case built from the host generators equals the AllenCahnSolver
    case built from the same generators, bit for bit (same inputs, same
    kernels, its own reduction tree)."""
    need_gpu()
    from lbm.batch import BatchEngine, CaseSpec
    print(f"[gpu] batch case vs solver api {storage}", flush=True)
    s = CaseSpec(grid=(24, 24, 20), substrate="ridge", R_star=1.0, R_drop=5.0,
                 theta_eq=162.0, u_impact=-0.05, steps=10)
    b = BatchEngine(storage)
    cid = b.add(s)
    b.step(10)
    sol = _solver_case(s, 10, storage=None if storage == "fp32" else storage)
    for name in ("phi", "u"):
        np.testing.assert_array_equal(getattr(sol, name).cpu().numpy(),
                                      b.field(cid, name).cpu().numpy())


def test_device_built_case_close_to_host_built():
    """inputs='device' builds substrate and droplet on the device: the same
    case to the last fp32 bit of a few tanh values (the device tanh is not
    numpy's)."""
    need_gpu()
    from lbm.batch import BatchEngine, CaseSpec
    s = CaseSpec(grid=(24, 24, 20), substrate="ridge", R_star=1.0, R_drop=5.0,
                 theta_eq=162.0, u_impact=-0.05, steps=10)
    b = BatchEngine("fp32")
    ch = b.add(s)
    cd = b.add(s, inputs="device")
    b.step(10)
    a = b.field(ch, "phi").cpu().numpy()
    c = b.field(cd, "phi").cpu().numpy()
    assert np.linalg.norm(a - c) / np.linalg.norm(a) < 1e-5


def test_stale_batch_graph_is_refused():
    """A captured batch graph holds the case set's block maps and grid sizes:
    after the case set or the launch shape changes (and one eager step
    re-uploads the tables), replaying it must be refused, not run."""
    need_gpu()
    from lbm import _native as nat
    from lbm.batch import BatchEngine
    sp = specs()
    for s in sp:
        s.steps = 40
    b = BatchEngine("fp32")
    ids = [b.add(s) for s in sp[:2]]
    b.step(1)
    b.capture_graph(4)
    b.run_graph(1)                       # valid replay
    b.add(sp[2])
    b.step(1)
    with pytest.raises(nat.LbmError):
        b.run_graph(1)
    b.capture_graph(2)
    b.run_graph(1)
    b.set_split(False)
    b.step(1)
    with pytest.raises(nat.LbmError):
        b.run_graph(1)
    b.capture_graph(2)
    nat.check(nat.lib.lbm_batch_set_block_sizes(b._h, 128, 512), "set_block_sizes")
    b.step(1)
    with pytest.raises(nat.LbmError):
        b.run_graph(1)
    torch.cuda.synchronize()
    assert all(b.status(c)[0] > 0 for c in ids)


def test_batched_solver_handle_refuses_changes():
    """The solver handle of a batch case serves reads only: calls that would
    change the case behind the batch's back are refused."""
    need_gpu()
    import ctypes
    from lbm import _native as nat
    from lbm.batch import BatchEngine
    b = BatchEngine("fp32")
    cid = b.add(specs()[0])
    h = b._solver(cid)
    st = nat.stream_ptr()
    assert nat.lib.lbm_step(h, 1, st) == -1
    assert nat.lib.lbm_graph_capture(h, 2) == -1
    assert nat.lib.lbm_set_block_sizes(h, 256, 256) == -1
    assert nat.lib.lbm_destroy(h) == -1
    n = int(nat.lib.lbm_state_bytes(h))
    img = torch.empty(n, dtype=torch.uint8, device="cuda")
    nat.check(nat.lib.lbm_save_state(h, nat.ptr(img), st), "save")      # reads are fine
    assert nat.lib.lbm_load_state(h, nat.ptr(img), st) == -1
    out = (ctypes.c_double * 10)()
    nat.check(nat.lib.lbm_measure(h, out, st), "measure")
    b.step(2)
    assert b.status(cid)[0] == 2


def test_batch_refuses_image_with_flipped_buffer_set():
    """A stand-alone solver's image taken right after an injection has its
    buffer set out of step with the step parity; the batch, whose kernels
    take the set from the parity, refuses it."""
    need_gpu()
    from lbm import _native as nat
    from lbm.batch import BatchEngine, CaseSpec
    s = CaseSpec(grid=(24, 24, 20), substrate="ridge", R_star=1.0, R_drop=5.0,
                 theta_eq=162.0, u_impact=-0.05, steps=30)
    sol = _solver_case(s, 2)
    sol.inject_droplet((12.0, 12.0, 8.0), 3.0, 4.0)
    sol.step(1)                          # the set stays flipped after further steps
    img = sol.state_dict(device="cpu").pin_memory()
    b = BatchEngine("fp32")
    cid = b.add(s)
    with pytest.raises(nat.LbmError):
        b.load_case(cid, img)
    plain = _solver_case(s, 3)           # no injection: set follows the parity
    b.load_case(cid, plain.state_dict(device="cpu"))
    assert b.status(cid)[0] == 3
    b.step(4)
    plain.step(4)
    np.testing.assert_array_equal(plain.phi.cpu().numpy(), b.field(cid, "phi").cpu().numpy())
