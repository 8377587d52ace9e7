"""16-bit distribution storage (bf16 / fp16 deviations from the lattice-weight
equilibrium, fp32 collision, fp64 mass reductions).

The oracle models the same storage (oracle.ac_oracle.round16), so the HIP path
must follow it to 1e-6 for two steps; the error of each form against the fp64
run of the same kernels is bounded and printed."""
import numpy as np
import pytest
import torch

from _cases import config_kwargs, load, rel, solver_kwargs
from oracle.ac_oracle import OracleSolver, Params, round_bf16

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def gpu(name, storage):
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    z, m = load(name)
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    dt = torch.float64 if storage == "fp64" else torch.float32
    s = AllenCahnSolver(cfg, dtype=dt, storage=None if storage in ("fp32", "fp64") else storage,
                        **solver_kwargs(m, z, "cuda"))
    s.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    s.init_fields(z["phi_init"], z["u_init"])
    return s, z


def oracle(name, storage):
    z, m = load(name)
    o = OracleSolver(Params(**config_kwargs(m)), dtype=np.float32, storage=storage,
                     **solver_kwargs(m, z))
    o.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    o.init_fields(z["phi_init"], z["u_init"])
    return o


def test_round_bf16_matches_torch():
    x = np.random.default_rng(0).normal(size=10000).astype(np.float32) * 1e-3
    t = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(round_bf16(x), t)


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
@pytest.mark.parametrize("name", ["ridge162_readme_f32", "flat162_2d_f32", "ridge120_f32",
                                  "ridge162_ch_f32", "ridge144_direct_f32", "concave126_f32",
                                  "convex144_nofrac_f32", "ridge120_alphafield_f32"])
def test_16bit_follows_oracle_model(name, storage):
    need_gpu()
    print(f"[gpu] {storage} {name}", flush=True)
    s, z = gpu(name, storage)
    o = oracle(name, storage)
    fluid = ~z["solid"]
    np.testing.assert_array_equal(s.f.cpu().numpy(), o.f)
    for step in range(1, 3):
        s.step()
        o.step()
        assert rel(s.phi.cpu().numpy(), o.phi, fluid) < 1e-6
        assert rel(s.u.cpu().numpy(), o.u, fluid, 1e-3) < 1e-5
        assert rel(s.f.cpu().numpy(), o.f) < 1e-6


@pytest.mark.parametrize("storage,bound", [("bf16", 2e-2), ("fp16", 2e-3), ("fp32", 1e-4)])
def test_error_vs_fp64_same_kernels(storage, bound):
    need_gpu()
    name = "ridge162_readme_f32"
    s, z = gpu(name, storage)
    r, _ = gpu(name, "fp64")
    s.step(60)
    r.step(60)
    fluid = ~z["solid"]
    e_phi = rel(s.phi.cpu().numpy(), r.phi.cpu().numpy(), fluid)
    e_u = rel(s.u.cpu().numpy(), r.u.cpu().numpy(), fluid)
    print(f"[gpu] {storage} vs fp64 after 60 steps: phi {e_phi:.2e}, u {e_u:.2e}", flush=True)
    assert e_phi < bound and e_u < 20 * bound


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
@pytest.mark.parametrize("name", ["ridge162_readme_f32", "flat162_2d_f32", "concave126_f32"])
def test_16bit_pass_b_forms_agree(name, storage):
    """The three forms of pass B (one launch over all nodes, two chunk-list
    launches, the single mixed launch) give bit-identical 16-bit states; the
    autotuner picks the mixed one for 16-bit storage on wall-heavy cases."""
    need_gpu()
    print(f"[gpu] {storage} {name} pass-B forms", flush=True)
    runs = []
    for split in (0, 1, 2):
        s, _ = gpu(name, storage)
        s.set_split(split)
        assert s.launch()[4] == split
        s.step(9)
        runs.append({k: getattr(s, k).cpu().numpy() for k in ("phi", "u", "f", "g")})
    for other in runs[1:]:
        for k in ("phi", "u", "f", "g"):
            np.testing.assert_array_equal(other[k], runs[0][k])


# ---------------------------------------------------------------------------
# The precision gate at full size: the 150x150x107 ridge impact (README
# construction) over 2000 steps at theta in {90, 120, 162} degrees (no,
# quadratic and linear wetting), every storage form against an fp64 run of the
# same kernels: max spread factor and max base radius within 1%.
# ---------------------------------------------------------------------------
def _cfg3_run(theta, form, steps=2000, every=20):
    from geometry.substrate import create_substrate_with_fraction
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_droplet import create_fe_droplet_with_impact
    cfg = FEConfig(nx=150, ny=150, nz=107, rho_l=1.0, rho_g=1.0 / 828.0, sigma=0.01424,
                   xi=4.0, tau_l=0.53, tau_g=0.53, theta_eq=theta, device="cuda")
    dt = torch.float64 if form == "fp64" else torch.float32
    storage = form if form in ("bf16", "fp16") else None
    s = AllenCahnSolver(cfg, dtype=dt, geo_amplification=1.5, geometric_wetting=True,
                        storage=storage)
    solid, frac = create_substrate_with_fraction(150, 150, 107, "ridge", 1.0, 22.5)
    s.set_solid(solid, solid_fraction=frac)
    phi0, _, u0 = create_fe_droplet_with_impact(150, 150, 107, center=(75, 75, 80),
                                                radius=22.5, xi=4.0, u_impact=(0, 0, -0.05))
    s.init_fields(phi0, u0)
    dmax = rmax = 0.0
    for _ in range(steps // every):
        s.step(every)
        m = s.measure()
        assert m["nan"] == 0
        dmax = max(dmax, m["Dx"], m["Dy"])
        rmax = max(rmax, s.base_radius(75.0, 75.0))
    fluid = ~solid
    out = (dmax / 45.0, rmax, s.phi.double().cpu().numpy()[fluid],
           s.u.double().cpu().numpy()[:, fluid])
    del s
    torch.cuda.empty_cache()
    return out


@pytest.mark.parametrize("theta", [90.0, 120.0, 162.0])
def test_cfg3_full_size_gate(theta):
    need_gpu()
    print(f"[gpu] cfg3 150x150x107 theta {theta:g}: fp64 reference", flush=True)
    sf64, rb64, phi64, u64 = _cfg3_run(theta, "fp64")
    worst = {}
    for form in ("fp32", "bf16", "fp16"):
        sf, rb, phi, u = _cfg3_run(theta, form)
        e_sf = abs(sf / sf64 - 1.0)
        e_rb = abs(rb / rb64 - 1.0) if rb64 > 0 else abs(rb)
        e_phi = float(np.linalg.norm(phi - phi64) / np.linalg.norm(phi64))
        e_u = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
        print(f"[gpu] cfg3 theta {theta:g} {form}: spread {sf:.4f} (fp64 {sf64:.4f}, "
              f"{100 * e_sf:.2f}%), base radius {rb:.3f} (fp64 {rb64:.3f}, {100 * e_rb:.2f}%), "
              f"phi L2 {e_phi:.2e}, u L2 {e_u:.2e}", flush=True)
        worst[form] = (e_sf, e_rb)
    assert rb64 > 5.0          # the base radius measures a real footprint
    for form, (e_sf, e_rb) in worst.items():
        assert e_sf <= 0.01, (form, e_sf)
        assert e_rb <= 0.01, (form, e_rb)
