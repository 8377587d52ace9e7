"""Continuous batching: requests admitted while others run, slots and memory
reused as cases finish, and every result the same alone and in the mix."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def small_stream():
    from lbm.batch import CaseSpec
    from lbm.serve import Request
    specs = [((32, 32, 28), "ridge", 1.0, 162.0, -0.05, 40),
             ((28, 28, 24), "flat", 1.0, 90.0, -0.03, 60),
             ((36, 32, 30), "concave", 1.0, 126.0, -0.06, 30),
             ((24, 24, 24), "convex", 0.5, 144.0, -0.04, 50),
             ((30, 30, 26), "ridge", 0.5, 108.0, -0.07, 36),
             ((26, 26, 22), "flat", 1.0, 150.0, -0.02, 44)]
    out = []
    for i, (g, sub, rs, th, u, n) in enumerate(specs):
        spec = CaseSpec(grid=g, substrate=sub, R_star=rs, R_drop=0.15 * min(g[0], g[2]),
                        theta_eq=th, u_impact=u, steps=n, tag=f"t{i}")
        # all arrive at once: three fill the slots, the rest wait for a slot
        out.append(Request(spec=spec, t_arrival=0.0))
    return out


@pytest.mark.parametrize("storage", ["bf16", "fp32"])
def test_served_results_equal_alone(storage):
    need_gpu()
    print(f"[gpu] serve alone vs mix {storage}", flush=True)
    from lbm.batch import BatchEngine
    from lbm.serve import SweepServer
    events = []
    srv = SweepServer(storage, max_active=3, chunk=8)
    done = srv.run(small_stream(), keep_fields=True,
                   on_event=lambda k, r: events.append((k, r.spec.tag, len(srv.active))))
    assert len(done) == 6
    # continuous batching: more than one case advanced together, and a waiting
    # request admitted after an earlier one finished (its slot reused)
    assert max(n for k, _, n in events if k == "admit") >= 2
    first_done = next(i for i, e in enumerate(events) if e[0] == "done")
    assert any(e[0] == "admit" for e in events[first_done:])
    for req in done:
        alone = BatchEngine(storage)
        cid = alone.add(req.spec)
        alone.step(req.spec.steps)
        phi = alone.field(cid, "phi").cpu()
        np.testing.assert_array_equal(phi.numpy(), req.result["phi"].numpy())
        m = alone.measure(cid)
        for k in ("Dx", "Dy", "Dz", "volume"):
            assert m[k] == req.result[k]
        assert alone.contact_radius(cid) == req.result["contact_radius"]


def test_admission_respects_hbm_budget():
    need_gpu()
    from lbm.batch import footprint
    from lbm.serve import SweepServer
    reqs = small_stream()
    fp = sorted(footprint(r.spec, "bf16") for r in reqs)
    budget = fp[-1] + fp[-2] + 1024          # room for two of the largest
    peak = []
    srv = SweepServer("bf16", budget_bytes=budget, max_active=8, chunk=8)
    done = srv.run(reqs, on_event=lambda k, r: peak.append(srv.engine.bytes_used()))
    assert len(done) == 6
    assert max(peak) <= budget


def test_preempted_server_resumes_identically(tmp_path):
    """A server stopped mid-run (preemption) and restarted with the same result
    store and checkpoint directory skips finished labels, resumes resident
    cases from their checkpoints, and ends with the results of an
    uninterrupted run, bit for bit."""
    need_gpu()
    print("[gpu] serve preempt/resume", flush=True)
    from lbm.serve import ResultStore, SweepServer
    ref = {r.spec.tag: r for r in SweepServer("bf16", max_active=3, chunk=8).run(
        small_stream(), keep_fields=True)}
    store_path = str(tmp_path / "results.json")
    ckpt = str(tmp_path / "ckpt")
    first = SweepServer("bf16", max_active=3, chunk=8, store=ResultStore(store_path),
                        checkpoint_dir=ckpt, checkpoint_every=2)
    first.run(small_stream(), max_chunks=7)
    n_first = len(first.done)
    assert 0 < n_first < 6                        # stopped with work in flight
    events = []
    second = SweepServer("bf16", max_active=3, chunk=8, store=ResultStore(store_path),
                         checkpoint_dir=ckpt, checkpoint_every=2)
    second.run(small_stream(), keep_fields=True,
               on_event=lambda k, r: events.append((k, r.spec.tag)))
    assert len(second.skipped) == n_first
    assert any(k == "resume" for k, _ in events)  # resumed from a checkpoint
    final = ResultStore(store_path)
    assert sorted(final.labels) == sorted(ref)
    for rec in final.results:
        want = ref[rec["label"]].result
        for k in ("Dx", "Dy", "Dz", "z_com", "volume", "contact_radius"):
            assert rec[k] == want[k], (rec["label"], k)
    for req in second.done:
        np.testing.assert_array_equal(req.result["phi"].numpy(),
                                      ref[req.spec.tag].result["phi"].numpy())
    import os
    assert not [f for f in os.listdir(ckpt) if f.endswith(".npy")]


def test_failed_case_is_isolated():
    """A case whose state turns to NaN (forced from the host: a block of its
    distributions overwritten through a state image at admission) is stopped
    after its first chunk and recorded unstable; every other request still
    gets the result of the undisturbed run, bit for bit."""
    need_gpu()
    print("[gpu] serve failure isolation", flush=True)
    from lbm.serve import SweepServer
    ref = {r.spec.tag: r for r in SweepServer("bf16", max_active=3, chunk=8).run(
        small_stream(), keep_fields=True)}
    srv = SweepServer("bf16", max_active=3, chunk=8)
    events = []

    def on_event(kind, req):
        events.append((kind, req.spec.tag))
        if kind == "admit" and req.spec.tag == "t1":
            img = srv.engine.save_case(req.case_id)
            img[256:256 + 4096] = 0xFF           # NaN in the first distribution block
            srv.engine.load_case(req.case_id, img)
    done = {r.spec.tag: r for r in srv.run(small_stream(), keep_fields=True, on_event=on_event)}
    assert ("failed", "t1") in events
    bad = done["t1"]
    assert bad.result["stable"] is False and bad.result["nan"] == 1.0
    assert bad.result["steps_done"] < bad.spec.steps
    for tag, req in done.items():
        if tag == "t1":
            continue
        assert req.result["stable"] is True
        np.testing.assert_array_equal(req.result["phi"].numpy(), ref[tag].result["phi"].numpy())
        for k in ("Dx", "Dy", "Dz", "volume", "contact_radius"):
            assert req.result[k] == ref[tag].result[k]


def test_poisson_stream_at_real_sizes():
    """A seeded Poisson stream (4 requests/s) of impact solves at the stated
    grid sizes (nx = ny from 64 to 192, nz from 64 to 160; steps cut to
    500-1500 so the test stays short), bf16 storage: requests are admitted by
    HBM footprint against a budget that holds only a few of the large ones,
    arrive and are admitted while others run, and every result equals the same
    request run alone, bit for bit.  The reference serves a sweep one process
    per case from a queue file (scripts/night_queue.py:28-50)."""
    need_gpu()
    print("[gpu] serve: Poisson stream at real sizes", flush=True)
    from lbm.batch import BatchEngine, footprint
    from lbm.serve import SweepServer, poisson_requests
    reqs = poisson_requests(seed=11, rate=4.0, duration=3.0, steps=(500, 1500),
                            grid_xy=(64, 192), grid_z=(64, 160))
    assert len(reqs) >= 8
    sizes = [r.spec.nodes for r in reqs]
    assert max(sizes) >= 3_000_000 and min(sizes) <= 1_500_000
    fps = sorted(footprint(r.spec, "bf16") for r in reqs)
    budget = fps[-1] + fps[-2]                     # about two large cases
    slots = 8
    log = []
    srv = SweepServer("bf16", budget_bytes=budget, max_active=slots, chunk=50)
    held_by_memory = []
    fits = srv._fits

    def fits_logged(req):
        ok = fits(req)
        if not ok and len(srv.active) < slots:
            held_by_memory.append(req.spec.tag)
        return ok
    srv._fits = fits_logged

    def on_event(kind, req):
        log.append((kind, req.spec.tag, len(srv.active), srv.engine.bytes_used()))
    done = srv.run(reqs, keep_fields=True, on_event=on_event)
    assert len(done) == len(reqs)
    assert max(b for *_, b in log) <= budget
    admits = [e for e in log if e[0] == "admit"]
    # admitted while other cases were on the device
    assert any(n >= 2 for _, _, n, _ in admits)
    # memory, not slots, held some request back while slots were free
    assert held_by_memory, "no request ever waited for device memory"
    tts = np.array([r.t_done - r.t_arrival for r in done])
    print(f"[gpu] serve: {len(done)} requests, p50 {np.percentile(tts, 50):.2f} s, "
          f"p99 {np.percentile(tts, 99):.2f} s, {srv.lattice_updates / srv.wall / 1e6:.0f} MLUPS",
          flush=True)
    for req in done:
        alone = BatchEngine("bf16")
        cid = alone.add(req.spec)
        alone.step(req.spec.steps)
        np.testing.assert_array_equal(alone.field(cid, "phi").cpu().numpy(),
                                      req.result["phi"].numpy())
        m = alone.measure(cid)
        for k in ("Dx", "Dy", "Dz", "volume"):
            assert m[k] == req.result[k], (req.spec.tag, k)
        del alone
        torch.cuda.empty_cache()
