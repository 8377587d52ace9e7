"""Mid-run droplet injection (AllenCahnSolver.inject_droplet,
lbm/fe_ac_solver.py:645-732) through the C ABI, against the oracle's
restatement on the same state.

fp64: every observable right after the injection within 1e-12 (relative to
its max) and the new mass target within 1e-13; the trajectory after it within
1e-10, as the step parity tests require.  fp32 and bf16 storage: 1e-6 right
after and for two steps (the fp32 step tolerance of the parity tests).
Cases: an injection after a few steps (state in the post-collision form), one
right after init_fields (post-stream form) and two in a row (the second sees
the first's state and the last step's P and mu); drops overlap existing
liquid and the solid substrate."""
import numpy as np
import pytest
import torch

from _cases import config_kwargs, load, rel, solver_kwargs
from oracle.ac_oracle import OracleSolver, Params

pytestmark = pytest.mark.gpu

DROPS = {
    3: [((6.0, 12.0, 12.0), 3.0, (0.0, 0.01, -0.05)),
        ((14.0, 6.0, 6.0), 3.0, (0.02, 0.0, -0.03))],
    2: [((24.0, 9.0), 4.0, (0.01, -0.04)),
        ((6.0, 14.0), 3.0, (0.0, -0.02))],
}


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def pair(name, storage=None):
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    z, m = load(name)
    f64 = m["dtype"] == "float64"
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float64 if f64 else torch.float32, storage=storage,
                        **solver_kwargs(m))
    frac = z["fraction"] if "fraction" in z.files else None
    s.set_solid(z["solid"], frac)
    s.init_fields(z["phi_init"], z["u_init"])
    o = OracleSolver(Params(**config_kwargs(m)), dtype=np.float64 if f64 else np.float32,
                     storage=storage, **solver_kwargs(m))
    o.set_solid(z["solid"], frac)
    o.init_fields(z["phi_init"], z["u_init"])
    return s, o, z, config_kwargs(m)["xi"]


def compare(s, o, fluid, tol, fields=("phi", "rho", "u", "P", "mu_phi", "f", "g")):
    omap = {"mu_phi": "mu"}
    for k in fields:
        got = getattr(s, k).cpu().numpy()
        want = getattr(o, omap.get(k, k))
        if want is None:
            continue
        mask = fluid if got.ndim == fluid.ndim else None
        e = rel(got, want, mask)
        assert e < tol, f"{k}: {e:.3e}"


@pytest.mark.parametrize("name,k0,n_inject", [
    ("ridge162_readme_f64", 3, 1),
    ("ridge162_readme_f64", 0, 2),
    ("ridge108_2d_f64", 4, 2),
])
def test_inject_fp64(name, k0, n_inject):
    need_gpu()
    print(f"[gpu] inject {name} k0={k0} n={n_inject}", flush=True)
    s, o, z, xi = pair(name)
    fluid = ~z["solid"]
    s.step(k0)
    for _ in range(k0):
        o.step()
    for center, radius, uimp in DROPS[s.ndim][:n_inject]:
        n_gpu = s.inject_droplet(center, radius, xi, uimp)
        n_ora = o.inject_droplet(center, radius, xi, uimp)
        assert n_gpu == n_ora > 0
        compare(s, o, fluid, 1e-12)
        assert abs(s.phi_mass_init - o.M0) <= 1e-13 * o.M0
    assert s.step_count == k0
    for n in range(6):
        s.step(1)
        o.step()
        compare(s, o, fluid, 1e-10, ("phi", "u", "P", "f", "g"))
    assert abs(s.mass()[0] - o.M0) <= 1e-13 * o.M0


@pytest.mark.parametrize("storage", [None, "bf16"])
def test_inject_fp32_compute(storage):
    need_gpu()
    print(f"[gpu] inject ridge120_f32 storage={storage}", flush=True)
    s, o, z, xi = pair("ridge120_f32", storage)
    fluid = ~z["solid"]
    s.step(2)
    o.step()
    o.step()
    center, radius, uimp = DROPS[3][0]
    assert s.inject_droplet(center, radius, xi, uimp) == o.inject_droplet(center, radius, xi,
                                                                          uimp)
    compare(s, o, fluid, 1e-6)
    for _ in range(2):
        s.step(1)
        o.step()
        compare(s, o, fluid, 1e-6, ("phi", "u", "f"))


def test_inject_state_round_trip_and_graph():
    """A snapshot taken right after an injection resumes identically (it
    carries the kept P and mu); a captured graph is dropped by the injection
    and can be captured again after a step."""
    need_gpu()
    from lbm import _native as nat
    print("[gpu] inject snapshot/graph", flush=True)
    s, _, z, xi = pair("ridge120_f32")
    s.step(3)
    s.capture_graph(2)
    s.inject_droplet(*DROPS[3][0][:2], xi, DROPS[3][0][2])
    P0, mu0 = s.P.cpu().numpy().copy(), s.mu_phi.cpu().numpy().copy()
    snap = s.state_dict()
    with pytest.raises(RuntimeError):
        s.run_graph(1)
    t, _, _, _ = pair("ridge120_f32")
    t.load_state_dict(snap)
    np.testing.assert_array_equal(t.P.cpu().numpy(), P0)
    np.testing.assert_array_equal(t.mu_phi.cpu().numpy(), mu0)
    s.step(3)
    t.step(3)
    for k in ("phi", "u", "f", "g", "P"):
        np.testing.assert_array_equal(getattr(s, k).cpu().numpy(), getattr(t, k).cpu().numpy())
    s.capture_graph(2)
    s.run_graph(1)
    t.step(2)
    np.testing.assert_array_equal(s.phi.cpu().numpy(), t.phi.cpu().numpy())
    assert nat.lib.lbm_step_count(s._h) == 8
