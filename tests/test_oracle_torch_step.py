"""The torch-CPU restatement used as the benchmark's CPU baseline
(oracle/torch_step.py) against the numpy oracle's step on the same state."""
import numpy as np
import pytest
import torch

from _cases import config_kwargs, load, rel, solver_kwargs
from oracle.ac_oracle import OracleSolver, Params
from oracle.torch_step import TorchStep


@pytest.mark.parametrize("name,dtype,tol", [("ridge162_readme_f32", np.float32, 2e-6),
                                            ("ridge162_readme_f64", np.float64, 1e-12),
                                            ("ridge120_f32", np.float32, 2e-6),
                                            ("flat90_2d_f32", np.float32, 2e-6)])
def test_torch_step_follows_oracle(name, dtype, tol):
    z, m = load(name)
    kw = solver_kwargs(m)
    o = OracleSolver(Params(**config_kwargs(m)), dtype=dtype, **kw)
    o.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    o.init_fields(z["phi_init"], z["u_init"])
    t = TorchStep(o)
    torch.set_num_threads(2)
    fluid = ~z["solid"]
    for _ in range(6):
        o.step()
        t.step()
        assert rel(t.phi.numpy(), o.phi, fluid) < tol
        assert rel(t.u.numpy(), o.u, fluid, 1e-3) < 50 * tol
    np.testing.assert_allclose(t.f.numpy(), o.f, rtol=0, atol=50 * tol)
