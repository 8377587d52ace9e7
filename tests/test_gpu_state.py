"""Solver state snapshots: a run resumed from a snapshot (in a fresh solver,
from device or host memory) continues bit for bit like the uninterrupted run;
images for another grid are refused."""
import numpy as np
import pytest
import torch

from _cases import config_kwargs, load, solver_kwargs

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def make(name, storage=None, init=True):
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    z, m = load(name)
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float32, storage=storage, **solver_kwargs(m))
    s.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    if init:
        s.init_fields(z["phi_init"], z["u_init"])
    return s


@pytest.mark.parametrize("storage,where", [(None, "cuda"), (None, "cpu"), ("bf16", "cuda")])
def test_resume_is_bit_identical(storage, where):
    need_gpu()
    print(f"[gpu] state resume {storage} {where}", flush=True)
    name = "ridge162_readme_f32"
    a = make(name, storage)
    a.step(7)
    snap = a.state_dict(device=where)
    before = {k: getattr(a, k).cpu().numpy().copy() for k in ("phi", "u", "P", "f", "g")}
    a.step(9)
    b = make(name, storage, init=False)
    b.load_state_dict(snap)
    assert b.step_count == 7 and b.phi_mass_init == a.phi_mass_init
    for k, v in before.items():
        np.testing.assert_array_equal(getattr(b, k).cpu().numpy(), v)
    b.step(9)
    for k in ("phi", "u", "P", "f", "g", "mu_phi"):
        np.testing.assert_array_equal(getattr(a, k).cpu().numpy(), getattr(b, k).cpu().numpy())
    # the same solver rewound to the snapshot
    a.load_state_dict(snap)
    a.step(9)
    np.testing.assert_array_equal(a.phi.cpu().numpy(), b.phi.cpu().numpy())


def test_foreign_image_refused():
    need_gpu()
    from lbm import _native as nat
    a = make("ridge162_readme_f32")
    a.step(2)
    snap = a.state_dict()
    other = make("flat162_2d_f32", init=False)
    if tuple(other.shape) == tuple(a.shape):
        pytest.skip("same grid")
    with pytest.raises(nat.LbmError):
        other.load_state_dict(snap)
    junk = torch.zeros_like(snap)
    with pytest.raises(nat.LbmError):
        a.load_state_dict(junk)


def test_stale_solver_graph_is_refused():
    """A captured step graph reads the buffer set, chunk lists and fields of
    the moment of capture: after set_solid, init_fields or an odd number of
    eager steps the replay is refused instead of run on stale pointers; an
    even number of eager steps brings the state back to the captured parity."""
    need_gpu()
    from lbm import _native as nat
    z, m = load("ridge162_readme_f32")
    a = make("ridge162_readme_f32")
    a.step(1)
    a.capture_graph(2)
    a.run_graph(1)
    a.step(1)
    with pytest.raises(nat.LbmError):
        a.run_graph(1)
    a.step(1)
    a.run_graph(1)                      # parity and buffer set as captured
    a.set_solid(z["solid"], z["fraction"])
    a.init_fields(z["phi_init"], z["u_init"])
    a.step(1)
    with pytest.raises(nat.LbmError):
        a.run_graph(1)
    a.capture_graph(2)
    a.init_fields(z["phi_init"], z["u_init"])
    a.step(1)
    with pytest.raises(nat.LbmError):
        a.run_graph(1)
    torch.cuda.synchronize()


def test_replay_after_even_eager_steps_matches_eager():
    """Replays interleaved with an even number of eager steps give the eager
    trajectory bit for bit."""
    need_gpu()
    a = make("ridge162_readme_f32")
    b = make("ridge162_readme_f32")
    a.step(1)
    b.step(1)
    b.capture_graph(4)
    for _ in range(3):
        b.run_graph(1)
        b.step(2)
    a.step(18)
    np.testing.assert_array_equal(a.phi.cpu().numpy(), b.phi.cpu().numpy())
    np.testing.assert_array_equal(a.u.cpu().numpy(), b.u.cpu().numpy())
