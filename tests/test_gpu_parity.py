"""Parity of the HIP solver with the oracle and the reference's golden vectors.

Every case of tests/golden is rebuilt through the package's public API
(AllenCahnSolver -> C ABI -> HIP kernels) and compared with:
  * the reference's own outputs (golden fixtures): geometry bit-exact, the
    operator outputs bit-exact, the step trajectory (fp64: <= 1e-10 relative);
  * the numpy oracle run on the same inputs, in the same precision.
fp32 tolerance: the reference's own fp32 trajectory departs from its fp64
trajectory by up to a few 1e-5 within a few steps (rounding-noise switch in
the gradient correction), so after step 2 the fp32 check is an envelope:
|gpu32 - oracle32| <= max(5e-5, 4 |oracle64 - oracle32|); steps 1-2 must agree
to 1e-6.  fp64 runs of the same kernels must agree with the fp64 oracle to
1e-10 over 60 steps.
"""
import numpy as np
import pytest
import torch

from _cases import case_names, config_kwargs, inject_event, load, rel, solver_kwargs
from oracle.ac_oracle import OracleSolver, Params

pytestmark = pytest.mark.gpu

CASES = case_names()


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def gpu_solver(name, dtype):
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    z, m = load(name)
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=dtype, **solver_kwargs(m, z, "cuda"))
    s.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    s.init_fields(z["phi_init"], z["u_init"])
    return s, z, m


def oracle(name, dtype):
    z, m = load(name)
    o = OracleSolver(Params(**config_kwargs(m)), dtype=dtype, **solver_kwargs(m, z))
    o.set_solid(z["solid"], z["fraction"] if "fraction" in z.files else None)
    o.init_fields(z["phi_init"], z["u_init"])
    return o


def np_(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", CASES)
def test_geometry_bit_exact(name):
    need_gpu()
    print(f"[gpu] geometry {name}", flush=True)
    z, m = load(name)
    dt = torch.float32 if m["dtype"] == "float32" else torch.float64
    s, z, m = gpu_solver(name, dt)
    np.testing.assert_array_equal(np_(s.solid), z["solid"])
    if "is_boundary" in z.files:
        w = s.wetting
        assert w is not None and w.mode == str(z["wet_mode"])
        np.testing.assert_array_equal(np_(w.is_boundary), z["is_boundary"])
        np.testing.assert_array_equal(np_(w.near_wall_mask), z["near_wall_mask"])
        np.testing.assert_array_equal(np_(w.wall_normal).view(np.uint32),
                                      z["wall_normal"].view(np.uint32))
        np.testing.assert_array_equal(np_(w.wall_weight).view(np.uint32),
                                      z["wall_weight"].view(np.uint32))
    else:
        assert s.wetting is None
    if "partial_mask" in z.files:
        np.testing.assert_array_equal(np_(s._partial_mask), z["partial_mask"])
        np.testing.assert_array_equal(np_(s._partial_eps), z["partial_eps"])


@pytest.mark.parametrize("name", CASES)
def test_init_and_operator(name):
    need_gpu()
    print(f"[gpu] init/operator {name}", flush=True)
    z, m = load(name)
    f32 = m["dtype"] == "float32"
    dt = torch.float32 if f32 else torch.float64
    s, z, m = gpu_solver(name, dt)
    tol = 1e-6 if f32 else 1e-14
    assert abs(s.phi_mass_init - float(z["M0"])) / float(z["M0"]) < tol
    np.testing.assert_array_equal(np_(s.phi), z["s0_phi"])
    o = oracle(name, np.float32 if f32 else np.float64)
    # f_eq / g_eq / init mu: identical arithmetic to the oracle
    np.testing.assert_array_equal(np_(s.f), o.f)
    np.testing.assert_array_equal(np_(s.g), o.g)
    np.testing.assert_array_equal(np_(s.mu_phi), o.mu)
    assert rel(np_(s.f), z["s0_f"]) < 10 * tol
    # the private operator on the reference's step-1 phi: bit-exact
    gr, lp = s._lattice_gradient_and_laplacian(z["op_phi"])
    np.testing.assert_array_equal(np_(gr), z["op_grad"])
    np.testing.assert_array_equal(np_(lp), z["op_lap"])


@pytest.mark.parametrize("name", CASES)
def test_trajectory(name):
    need_gpu()
    print(f"[gpu] trajectory {name}", flush=True)
    z, m = load(name)
    f32 = m["dtype"] == "float32"
    s, z, m = gpu_solver(name, torch.float32 if f32 else torch.float64)
    o = oracle(name, np.float32 if f32 else np.float64)
    t64 = oracle(name, np.float64) if f32 else None
    fluid = ~z["solid"]
    ev = inject_event(m)
    for step in range(1, max(m["save_steps"]) + 1):
        s.step()
        o.step()
        if t64 is not None:
            t64.step()
        if ev is not None and step == ev[0]:
            s.inject_droplet(*ev[1])
            o.inject_droplet(*ev[1])
            if t64 is not None:
                t64.inject_droplet(*ev[1])
            assert rel(np_(s.phi), z["inj_phi"]) < (1e-6 if f32 else 1e-12)
        if step not in m["save_steps"]:
            continue
        for key, a, ob, fl, mk in (("phi", "phi", "phi", 0.0, fluid),
                                   ("u", "u", "u", 1e-3, fluid),
                                   ("P", "P", "P", 1e-5, None),
                                   ("rho", "rho", "rho", 0.0, None),
                                   ("mu", "mu_phi", "mu", 1e-4, None)):
            g = np_(getattr(s, a))
            e = rel(g, getattr(o, ob), mk, fl)
            if f32:
                if step <= 2:
                    assert e < 1e-6, (step, key, e)
                elif step <= 20:
                    env = rel(getattr(t64, ob), getattr(o, ob), mk, fl)
                    assert e < max(5e-5, 4 * env), (step, key, e, env)
            else:
                assert e < 1e-10, (step, key, e)
                gk = f"s{step}_{'mu' if key == 'mu' else key}"
                assert rel(g, z[gk], mk, fl) < 1e-9, (step, key, "vs reference")
        if f"s{step}_f" in z.files:
            tol = 1e-6 if f32 else 1e-10
            assert rel(np_(s.f), o.f) < tol
            assert rel(np_(s.g), o.g, floor=1e-5) < tol


@pytest.mark.parametrize("name", ["ridge162_readme_f32", "flat162_2d_f32"])
def test_fp32_vs_fp64_same_kernels(name):
    """Error of the fp32 form against an fp64 run of the same kernels stays at
    the level of the reference's own fp32-vs-fp64 distance."""
    need_gpu()
    print(f"[gpu] fp32 vs fp64 {name}", flush=True)
    s32, z, m = gpu_solver(name, torch.float32)
    s64, _, _ = gpu_solver(name, torch.float64)
    s32.step(20)
    s64.step(20)
    fluid = ~z["solid"]
    e = rel(np_(s32.phi), np_(s64.phi), fluid)
    assert e < 1e-5, e


def test_mass_and_determinism():
    need_gpu()
    print("[gpu] mass/determinism", flush=True)
    # production construction (boundary_relax = 0): the mass factor keeps
    # sum(phi) at M0 up to the clamp at 1 (scripts/validate_solver.py:29 gate)
    a, z, m = gpu_solver("ridge162_prod_f32", torch.float32)
    b, _, _ = gpu_solver("ridge162_prod_f32", torch.float32)
    for _ in range(30):
        a.step()
        b.step()
        M0 = a.phi_mass_init
        mass = float(np_(a.phi).astype(np.float64).sum())
        assert abs(mass - M0) / M0 < 1e-4
    for key in ("phi", "u", "f", "g"):
        np.testing.assert_array_equal(np_(getattr(a, key)), np_(getattr(b, key)))


def test_graph_replay_matches_eager():
    """Captured steps replayed several times, with eager work and syncs between
    replays, equal eager stepping bit for bit."""
    need_gpu()
    print("[gpu] graph replay", flush=True)
    e, z, m = gpu_solver("ridge120_f32", torch.float32)
    g, _, _ = gpu_solver("ridge120_f32", torch.float32)
    e.step(1)
    g.step(1)
    g.capture_graph(4)
    for rep in range(3):
        e.step(4)
        g.run_graph(1)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()   # eager work between
        for key in ("phi", "u", "f", "P"):
            np.testing.assert_array_equal(np_(getattr(e, key)), np_(getattr(g, key)))
    assert e.step_count == g.step_count == 13


def test_division_by_constants_is_ieee_exact():
    """The kernels divide by cs2, 2 cs2 and rho_eff cs2 as one fp64 product
    with the divisor's fp64 reciprocal, and by xi with an fp64 Markstein
    correction; over all 2^32 fp32 inputs both equal IEEE division (the
    product form only for divisors that are not even integers, which is
    checked for the general form alone)."""
    need_gpu()
    import ctypes
    from lbm import _native as nat
    # cs2, 2 cs2, xi values, and rho_eff * cs2 over its range [0.5/3, 1/3]
    for d in (np.float32(1.0 / 3.0), np.float32(2.0 / 3.0), np.float32(4.0), np.float32(5.0),
              np.float32(6.0), np.float32(3.0), np.float32(1.0),
              np.float32(0.5) * np.float32(1.0 / 3.0), np.float32(0.7312) * np.float32(1.0 / 3.0),
              np.float32(0.9999) * np.float32(1.0 / 3.0)):
        bad = ctypes.c_ulonglong(0)
        nat.check(nat.lib.lbm_check_division(ctypes.c_float(float(d)), ctypes.byref(bad),
                                             nat.stream_ptr()), "lbm_check_division")
        print(f"[gpu] division by {float(d)!r}: {bad.value} mismatches", flush=True)
        assert bad.value == 0


def test_block_sizes_do_not_change_results():
    """Every launch shape of the autotuner (block sizes, occupancy caps, the
    three forms of pass B) gives bit-identical states (the mass reduction tree
    depends on node indices only)."""
    need_gpu()
    print("[gpu] block-size invariance", flush=True)
    ref, z, m = gpu_solver("ridge162_readme_f32", torch.float32)
    assert ref.launch()[4] == 2          # fp32 default: the single mixed pass-B launch
    ref.step(12)
    for bsa, bsb, ca, cb, sp in ((128, 512, 0, 0, 0), (512, 128, 0, 0, 1), (256, 256, 6, 4, 0),
                                 (128, 256, 6, 5, 1), (256, 128, 0, 4, 2), (512, 256, 6, 0, 2),
                                 (128, 512, 0, 5, 2), (256, 512, 0, 0, 3), (128, 256, 6, 4, 3)):
        s, _, _ = gpu_solver("ridge162_readme_f32", torch.float32)
        s.set_block_sizes(bsa, bsb)
        s.set_occupancy(ca, cb)
        s.set_split(sp)
        assert s.launch() == (bsa, bsb, ca, cb, sp)
        assert (s.fused_tile()[0] > 0) == (sp == 3)
        s.step(12)
        for key in ("phi", "u", "f", "g"):
            np.testing.assert_array_equal(np_(getattr(s, key)), np_(getattr(ref, key)))


def test_autotune_picks_valid_sizes():
    need_gpu()
    s, _, _ = gpu_solver("ridge120_f32", torch.float32)
    a, b = s.autotune(3)
    assert a in (128, 256, 512) and b in (128, 256, 512)
    la, lb, ca, cb, sp = s.launch()
    assert (la, lb) == (a, b) and ca in (0, 6) and cb in (0, 4, 5) and sp in (0, 1, 2)
    assert s.step_count == 27 * 4


@pytest.mark.parametrize("split", [1, 2, 3])
@pytest.mark.parametrize("grid", [(19, 21, 17), (37, 29, 1)])
def test_ragged_grid_matches_oracle(grid, split):
    """Node counts that are multiples of neither the 64-node wave nor the
    256-node distribution block (6,783 and 1,073 nodes): the padded tail of
    the last block must stay out of every pull, sum and stencil.  Ridge with
    fraction, impact at theta = 162 with amplification 1.5, README
    construction; fp64 kernels against the fp64 oracle over 25 steps, with
    pass B as two chunk-list launches, as the single mixed launch and (3D) as
    the fused step."""
    need_gpu()
    from geometry.substrate import create_substrate_with_fraction
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_droplet import create_fe_droplet_with_impact
    nx, ny, nz = grid
    print(f"[gpu] ragged grid {grid} split {split}", flush=True)
    two_d = nz == 1
    centre = (nx / 2.0, 11.0) if two_d else (nx / 2.0, ny / 2.0, 10.0)
    u_imp = (0.0, -0.04) if two_d else (0.0, 0.0, -0.04)
    solid, frac = create_substrate_with_fraction(nx, ny, nz, substrate_type="ridge",
                                                 R_star=1.0, R_d=5.0)
    C0, _, u0 = create_fe_droplet_with_impact(nx, ny, nz, center=centre, radius=4.5, xi=4.0,
                                              rho_l=1.0, rho_g=1.0 / 828.0, u_impact=u_imp)
    m = {"grid": [nx, ny, nz], "theta": 162.0, "tau_h": None,
         "kw": {"geo_amplification": 1.5, "geometric_wetting": True}}
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float64, **solver_kwargs(m, None, "cuda"))
    s.set_solid(solid, frac)
    s.init_fields(C0, u0)
    s.set_split(split)
    o = OracleSolver(Params(**config_kwargs(m)), dtype=np.float64, **solver_kwargs(m))
    o.set_solid(solid, frac)
    o.init_fields(C0, u0)
    assert s.phi_mass_init == pytest.approx(o.M0, rel=1e-15)
    for step in range(1, 26):
        s.step()
        o.step()
        if step in (1, 5, 25):
            for key in ("phi", "u", "P"):
                e = rel(np_(getattr(s, key)), getattr(o, key), floor=1e-5 if key == "u" else 0.0)
                assert e < 1e-10, (step, key, e)
            assert rel(np_(s.f), o.f) < 1e-10
            assert rel(np_(s.g), o.g, floor=1e-5) < 1e-10


def _ragged_fp32(grid=(19, 21, 17), storage=None):
    from geometry.substrate import create_substrate_with_fraction
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_droplet import create_fe_droplet_with_impact
    nx, ny, nz = grid
    solid, frac = create_substrate_with_fraction(nx, ny, nz, substrate_type="ridge",
                                                 R_star=1.0, R_d=5.0)
    C0, _, u0 = create_fe_droplet_with_impact(nx, ny, nz, center=(nx / 2.0, ny / 2.0, 10.0),
                                              radius=4.5, xi=4.0, rho_l=1.0, rho_g=1.0 / 828.0,
                                              u_impact=(0.0, 0.0, -0.04))
    m = {"grid": [nx, ny, nz], "theta": 162.0, "tau_h": None,
         "kw": {"geo_amplification": 1.5, "geometric_wetting": True}}
    cfg = FEConfig(**config_kwargs(m), device="cuda")
    s = AllenCahnSolver(cfg, dtype=torch.float32, storage=storage, **solver_kwargs(m, None, "cuda"))
    s.set_solid(solid, frac)
    s.init_fields(C0, u0)
    return s


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
def test_fused_form_elsewhere_runs_form_2(storage):
    """Form 3 on a solver the fused step does not cover (16-bit storage) runs
    the single mixed pass-B launch: no fused plan, identical states."""
    need_gpu()
    a = _ragged_fp32(storage=storage)
    a.set_split(2)
    b = _ragged_fp32(storage=storage)
    b.set_split(3)
    assert b.fused_tile() == (0, 0, 0, 0, 0, 0)
    a.step(6)
    b.step(6)
    for key in ("phi", "u", "f", "g"):
        np.testing.assert_array_equal(np_(getattr(b, key)), np_(getattr(a, key)))


@pytest.mark.parametrize("bs,tile", [(256, (0, 0, 0)), (256, (1, 1, 1)), (256, (21, 1, 2)),
                                     (128, (5, 3, 4)), (256, (7, 9, 17)), (256, (16, 16, 0)),
                                     (512, (21, 19, 3))])
def test_fused_tiles_bit_identical(bs, tile):
    """The fused step (pass A and pass B in one launch, phi stencil from an
    LDS ring around each tile) gives the same fp32 states as the two-pass step
    for every tile shape: single columns, whole rows, tiles larger than the
    grid's remainder (ragged edge tiles), one plane per segment and one
    segment for the whole height, on a 19 x 21 x 17 grid whose periodic wraps
    cross every tile edge.  Also replayed from a captured graph."""
    need_gpu()
    print(f"[gpu] fused tile bs {bs} {tile}", flush=True)
    ref = _ragged_fp32()
    ref.set_split(2)
    s = _ragged_fp32()
    s.set_split(3)
    s.set_block_sizes(256, bs)
    s.set_fused_tile(*tile)
    plan = s.fused_tile()
    assert plan[0] > 0 and plan[0] * plan[1] <= bs
    if tile[0]:
        assert plan[:2] == (min(tile[0], 21), min(tile[1], 19))
    ref.step(1)
    s.step(1)            # first step after init_fields: the chunk-list launches
    ref.step(6)
    s.step(6)
    for key in ("phi", "u", "f", "g", "P"):
        np.testing.assert_array_equal(np_(getattr(s, key)), np_(getattr(ref, key)))
    s.capture_graph(2)
    for _ in range(2):
        ref.step(2)
        s.run_graph(1)
        torch.cuda.synchronize()
        for key in ("phi", "u", "f", "g"):
            np.testing.assert_array_equal(np_(getattr(s, key)), np_(getattr(ref, key)))
    m_ref, m_s = ref.mass()[1], s.mass()[1]
    assert abs(m_s - m_ref) <= 1e-12 * abs(m_ref)


def test_fused_tile_rejects_oversized_tiles():
    need_gpu()
    from lbm._native import LbmError
    s = _ragged_fp32()
    s.set_split(3)
    s.set_block_sizes(256, 256)
    with pytest.raises(LbmError):
        s.set_fused_tile(21, 19, 0)      # 399 columns > 256 threads
    plan = s.fused_tile()
    assert plan[0] > 0                   # the previous (automatic) plan stays
    s.set_fused_tile(16, 16, 0)          # fits 256 threads ...
    with pytest.raises(LbmError):
        s.set_block_sizes(256, 128)      # ... but not 128: refused, nothing changes
    assert s.launch()[1] == 256 and s.fused_tile()[:2] == (16, 16)
    with pytest.raises(LbmError):
        s.set_occupancy(6, 9)
    assert s.launch()[2:4] == (0, 0)
    s.step(3)


@pytest.mark.parametrize("rho_l", [1.0, 8.0])
def test_division_forms_match_oracle_bitwise(rho_l):
    """rho_eff cs2 reaches 2 or more for liquid densities above 5.7, where the
    kernels divide the Guo forcing terms by it with the general (Markstein)
    form instead of the one-product form; both must reproduce the oracle's
    IEEE fp32 divisions bit for bit (24x24x20 ridge impact, fp32, 3 steps)."""
    need_gpu()
    from geometry.substrate import create_substrate_with_fraction
    from lbm.fe_ac_solver import AllenCahnSolver
    from lbm.fe_config import FEConfig
    from lbm.fe_droplet import create_fe_droplet_with_impact
    print(f"[gpu] division forms, rho_l = {rho_l}", flush=True)
    kw = dict(nx=24, ny=24, nz=20, rho_l=rho_l, rho_g=rho_l / 828.0, sigma=0.01424, xi=4.0,
              tau_l=0.53, tau_g=0.53, theta_eq=162.0)
    solid, frac = create_substrate_with_fraction(24, 24, 20, "ridge", 1.0, 6.0)
    phi0, _, u0 = create_fe_droplet_with_impact(24, 24, 20, center=(12, 12, 13), radius=5.0,
                                                xi=4.0, u_impact=(0, 0, -0.05))
    s = AllenCahnSolver(FEConfig(**kw, device="cuda"), dtype=torch.float32,
                        geo_amplification=1.5, geometric_wetting=True)
    s.set_solid(solid, solid_fraction=frac)
    s.init_fields(phi0, u0)
    o = OracleSolver(Params(**kw), dtype=np.float32, geo_amplification=1.5,
                     geometric_wetting=True)
    o.set_solid(solid, frac)
    o.init_fields(phi0, u0)
    for step in range(1, 4):
        s.step()
        o.step()
        for key in ("phi", "u", "P", "f", "g"):
            np.testing.assert_array_equal(np_(getattr(s, key)), getattr(o, key),
                                          err_msg=f"step {step} {key}")
