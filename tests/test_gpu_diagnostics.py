"""Device diagnostics (measure, contact_radius) against the oracle's
brute-force-pinned restatements applied to the same phi."""
import numpy as np
import pytest
import torch

from _cases import config_kwargs, load, solver_kwargs
from oracle.ac_oracle import base_radius, contact_radius, spread

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("name", ["ridge162_readme_f32", "concave126_f32", "flat90_f32",
                                  "flat162_2d_f32"])
def test_measure_and_contact_radius(name):
    need_gpu()
    print(f"[gpu] diagnostics {name}", flush=True)
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    z, m = load(name)
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float32, **solver_kwargs(m))
    s.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    s.init_fields(z["phi_init"], z["u_init"])
    for nsteps in (0, 7, 13):
        s.step(nsteps)
        phi = s.phi.cpu().numpy()
        solid = z["solid"]
        ref = spread(phi, solid)
        got = s.measure()
        for k in ("Dx", "Dy", "Dz", "C_max"):
            assert got[k] == ref[k], (k, got[k], ref[k])
        assert abs(got["volume"] - ref["volume"]) <= 1e-12 * max(1.0, ref["volume"])
        assert abs(got["z_com"] - ref["z_com"]) <= 1e-12 * max(1.0, ref["z_com"])
        if phi.ndim == 3:
            cx, cy = phi.shape[0] / 2.0, phi.shape[1] / 2.0
            assert s.contact_radius(cx, cy) == contact_radius(phi, solid, cx, cy)
            assert s.base_radius(cx, cy) == base_radius(phi, solid, cx, cy)
        # the non-blocking form writes the same record to device memory
        dev = s.measure_async()
        keys = ("Dx", "Dy", "Dz", "k", "volume", "C_max", "z_com", "n_liquid", "phi_sum", "nan")
        assert dev.is_cuda
        for k, v in zip(keys, dev.cpu().tolist()):
            assert v == got[k], (k, v, got[k])


def test_reference_measure_helpers_on_the_solver():
    """The reference-named host helpers read the HIP solver's observables:
    measure_spread_factor(solver.phi, solver.solid) gives the device
    measure()'s extents, and Monitor.record() fills its history."""
    need_gpu()
    print("[gpu] reference measure helpers", flush=True)
    from io_utils import Monitor
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_measure import measure_contact_angle, measure_spread_factor
    z, m = load("ridge162_readme_f32")
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float32, **solver_kwargs(m))
    s.set_solid(z["solid"], z["fraction"])
    s.init_fields(z["phi_init"], z["u_init"])
    mon = Monitor(cfg)
    for _ in range(3):
        s.step(5)
        mon.record(s)
        host = measure_spread_factor(s.phi, s.solid)
        dev = s.measure()
        for k in ("Dx", "Dy", "Dz", "k", "C_max"):
            assert host[k] == dev[k], k
        theta = measure_contact_angle(s.phi, s.solid)
        assert 0.0 <= theta <= 180.0
    assert mon.history["step"] == [5, 10, 15]
    assert mon.history["mass"][-1] == float(s.rho.cpu().numpy().sum())


def test_hbm_stream_probe():
    """The streaming-bandwidth probe returns plausible MI355X figures: every
    kind between 1 and 8 TB/s (the HBM3E peak), reads faster than writes."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import ctypes
    from lbm import _native as nat
    out = (ctypes.c_double * 6)()
    nat.check(nat.lib.lbm_hbm_stream(1 << 30, 3, nat.stream_ptr(), out), "lbm_hbm_stream")
    v = list(out)
    print(f"[gpu] stream TB/s read {v[0]:.2f}/{v[1]:.2f} write {v[2]:.2f}/{v[3]:.2f} "
          f"copy {v[4]:.2f}/{v[5]:.2f}", flush=True)
    assert all(1.0 < x < 8.0 for x in v), v
    assert max(v[0], v[1]) > max(v[2], v[3])
