"""Continuous-batching request loop for impact sweeps on one GPU.

The reference runs sweep cases one process at a time from a queue file
(scripts/night_queue.py:28-50, scripts/run_pending_sweep.py:207-279).  Here a
single loop owns one BatchEngine: requests that arrive while others are on
the GPU are admitted as soon as an execution slot and their HBM footprint
are free, join the resident batch at the next chunk boundary, advance inside
the same multi-case launches as everything else, and leave (measured, then
evicted) when they reach their own step target; their memory and slot go to
the next waiting request.  Lattices stay device-resident from admission to
completion.

Every case keeps its own buffers, reduction tree and step counter, so a
request's result does not depend on what else shares the launches.

Crash safety (the reference resumes a sweep by skipping labels already in its
results file, scripts/run_pending_sweep.py:207-240, 271-279): a ResultStore
keeps that label-keyed JSON list, rewritten atomically after every completed
request, and optional checkpoints hold the full device state of every
resident case; a restarted server skips finished labels and resumes
unfinished ones from their last checkpoint, with results identical to an
uninterrupted run.

Failure isolation (the reference's drivers stop a case whose phi holds a NaN
or has vanished, max(phi) < 0.01, at every 50-step measurement,
scripts/run_thesis_sweep.py:214-216): after every chunk each resident case's
diagnostics are reduced on the device and read back in one transfer; a failed
case is measured, recorded with stable = False and evicted, and the cases
sharing its launches are unaffected.
"""
from __future__ import annotations

import json
import os
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

from .batch import BatchEngine, CaseSpec, footprint, host_inputs

SUBSTRATE_TYPES = ("flat", "ridge", "convex", "concave")


@dataclass
class Request:
    spec: CaseSpec
    t_arrival: float                 # seconds after the start of the run
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    case_id: Optional[int] = None
    result: Dict = field(default_factory=dict)

    @property
    def lattice_updates(self):
        return self.spec.nodes * self.spec.steps


def poisson_requests(seed=5, rate=4.0, duration=20.0, steps=(500, 3000),
                     grid_xy=(64, 256), grid_z=(64, 160)):
    """A seeded Poisson stream of impact solves: exponential inter-arrival
    times of mean 1/rate over `duration` seconds; per request nx = ny ~
    U{grid_xy}, nz ~ U{grid_z}, substrate ~ U(4 types), R* ~ U[0.5, 2],
    theta ~ U[90, 162], |u_impact| ~ U[0.02, 0.08], steps ~ U{steps}; droplet
    radius 0.15 min(nx, nz) (0.15 nx, the sweep's rule, would not fit the
    lower domains), placed by the sweep driver's rule.  R* is capped where the
    substrate top (R* R for ridge and cavity rim, 2 R* R for the convex
    sphere) plus the 2-node gap and the drop would not fit under the lid."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out, t, i = [], 0.0, 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        nxy = int(rng.integers(grid_xy[0], grid_xy[1] + 1))
        nz = int(rng.integers(grid_z[0], grid_z[1] + 1))
        sub = SUBSTRATE_TYPES[int(rng.integers(0, 4))]
        rs = float(rng.uniform(0.5, 2.0))
        th = float(rng.uniform(90.0, 162.0))
        u = -float(rng.uniform(0.02, 0.08))
        n = int(rng.integers(steps[0], steps[1] + 1))
        r = 0.15 * min(nxy, nz)
        room = nz - 2.0 * r - 4.0            # height left for the substrate
        if sub in ("ridge", "concave"):
            rs = min(rs, room / r)
        elif sub == "convex":
            rs = min(rs, room / (2.0 * r))
        spec = CaseSpec(grid=(nxy, nxy, nz), substrate=sub, R_star=rs, R_drop=r,
                        theta_eq=th, u_impact=u, steps=n, tag=f"q{i:03d}")
        out.append(Request(spec=spec, t_arrival=t))
        i += 1
    return out


class ResultStore:
    """Label-keyed results file: a JSON list of dicts with a "label" key.
    Labels present are done; add() appends a result and rewrites the file
    through a temporary file and an atomic rename, so a crash leaves either
    the old or the new list."""

    def __init__(self, path):
        self.path = path
        self.results = []
        if os.path.exists(path):
            with open(path) as f:
                self.results = json.load(f)
        self.labels = {r["label"] for r in self.results}

    def done(self, label):
        return label in self.labels

    def add(self, result):
        if result["label"] in self.labels:
            raise ValueError(f"duplicate label {result['label']!r}")
        self.results.append(result)
        self.labels.add(result["label"])
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.results, f, indent=2, default=str)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, self.path)


class SweepServer:
    """Admission by HBM footprint and execution slots; continuous batching.

    budget_bytes: device bytes the resident cases may take (0: no limit);
    max_active: execution slots (cases advanced together); chunk: steps per
    batch launch sequence between admission/eviction points (shortened so
    that no resident case overshoots its target)."""

    def __init__(self, storage="bf16", budget_bytes=0, max_active=32, chunk=50,
                 device=None, store: Optional[ResultStore] = None,
                 checkpoint_dir: Optional[str] = None, checkpoint_every=0,
                 health_check=True, inputs="host"):
        self.engine = BatchEngine(storage, budget_bytes=budget_bytes, device=device)
        self.storage = storage
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.chunk = int(chunk)
        self.active: Dict[int, Request] = {}
        self.done: List[Request] = []
        self.skipped: List[Request] = []
        self.store = store
        self.ckpt_dir = checkpoint_dir
        self.ckpt_every = int(checkpoint_every)
        self.health_check = bool(health_check)
        if inputs not in ("host", "device"):
            raise ValueError("inputs must be 'host' or 'device'")
        self.inputs = inputs
        self._prepared: Dict[str, tuple] = {}
        if checkpoint_dir:
            os.makedirs(checkpoint_dir, exist_ok=True)

    # -- checkpoints of resident cases ----------------------------------------
    def _ckpt_path(self, req):
        return os.path.join(self.ckpt_dir, f"{req.spec.tag}.state.npy")

    def _checkpoint(self):
        import numpy as np
        for cid, req in self.active.items():
            path = self._ckpt_path(req)
            tmp = path + ".tmp.npy"
            np.save(tmp, self.engine.save_case(cid).numpy())
            os.replace(tmp, path)

    def _resume(self, req):
        if not self.ckpt_dir or not os.path.exists(self._ckpt_path(req)):
            return False
        import numpy as np
        import torch
        state = torch.from_numpy(np.load(self._ckpt_path(req)))
        self.engine.load_case(req.case_id, state)
        return True

    def _unstable(self):
        """Resident requests whose phi holds a NaN or has vanished: every
        case's device diagnostics, one read-back."""
        import torch
        ids = list(self.active)
        if not ids:
            return []
        rec = torch.empty((len(ids), 10), dtype=torch.float64, device=self.engine.device)
        for j, cid in enumerate(ids):
            self.engine.measure_async(cid, rec[j])
        rows = rec.cpu().numpy()            # columns 5: max(phi), 9: NaN flag
        return [self.active[cid] for j, cid in enumerate(ids)
                if rows[j, 9] != 0 or not rows[j, 5] >= 0.01]

    def _inputs_of(self, req):
        """Inputs for engine.add: prepared host arrays, a fresh host build,
        or "device" (the library builds them from the spec)."""
        if self.inputs == "device":
            return "device"
        arrays = self._prepared.pop(req.spec.tag, None) if req.spec.tag else None
        return arrays if arrays is not None else host_inputs(req.spec)

    def _prepare(self, waiting, pending, t0):
        """Host inputs of the next request to be admitted (the head of the
        waiting queue, else the next arrival due within a second)."""
        if self.inputs != "host":
            return
        nxt = waiting[0] if waiting else None
        if nxt is None and pending and pending[0].t_arrival <= time.perf_counter() - t0 + 1.0:
            nxt = pending[0]
        if nxt is not None and nxt.spec.tag and nxt.spec.tag not in self._prepared:
            self._prepared[nxt.spec.tag] = host_inputs(nxt.spec)

    def _fits(self, req):
        if len(self.active) >= self.max_active:
            return False
        if self.budget <= 0:
            return True
        return self.engine.bytes_used() + footprint(req.spec, self.storage) <= self.budget

    def _finish(self, req, now, stable=True):
        cid = req.case_id
        m = self.engine.measure(cid)
        req.result = {"Dx": m["Dx"], "Dy": m["Dy"], "Dz": m["Dz"], "z_com": m["z_com"],
                      "volume": m["volume"], "nan": m["nan"],
                      "contact_radius": self.engine.contact_radius(cid),
                      "stable": stable, "steps_done": self.engine.status(cid)[0]}
        if self.keep_fields:
            req.result["phi"] = self.engine.field(cid, "phi").cpu()
        req.t_done = now
        self.engine.remove(cid)
        del self.active[cid]
        self.done.append(req)
        if self.store is not None:
            rec = {k: v for k, v in req.result.items() if k != "phi"}
            rec["label"] = req.spec.tag
            rec["steps"] = req.spec.steps
            self.store.add(rec)
        if self.ckpt_dir and os.path.exists(self._ckpt_path(req)):
            os.remove(self._ckpt_path(req))

    def run(self, requests, keep_fields=False, on_event=None, max_chunks=None):
        """Serve `requests` (sorted by arrival) in real time; returns them
        with admission / completion times and results.  Requests whose label
        the store already holds are skipped; max_chunks stops the loop early
        (a preemption), leaving resident cases to their last checkpoint."""
        import torch
        self.keep_fields = keep_fields
        if self.store is not None or self.ckpt_dir:
            tags = [r.spec.tag for r in requests]
            if "" in tags or len(set(tags)) != len(tags):
                raise ValueError("a result store or checkpoints need unique, non-empty tags")
        todo = []
        for r in requests:
            if self.store is not None and self.store.done(r.spec.tag):
                self.skipped.append(r)
            else:
                todo.append(r)
        pending = deque(sorted(todo, key=lambda r: r.t_arrival))
        chunks = 0
        waiting: deque = deque()
        t0 = time.perf_counter()
        self.lattice_updates = 0
        while pending or waiting or self.active:
            now = time.perf_counter() - t0
            while pending and pending[0].t_arrival <= now:
                waiting.append(pending.popleft())
            # FIFO admission: the head of the queue waits for room
            while waiting and self._fits(waiting[0]):
                req = waiting.popleft()
                req.case_id = self.engine.add(req.spec, self._inputs_of(req))
                req.t_admit = time.perf_counter() - t0
                self.active[req.case_id] = req
                resumed = self._resume(req)
                if on_event:
                    on_event("resume" if resumed else "admit", req)
            if not self.active:
                if pending:
                    time.sleep(max(0.0, min(0.05, pending[0].t_arrival - now)))
                continue
            left = min(r.spec.steps - self.engine.status(cid)[0]
                       for cid, r in self.active.items())
            n = max(1, min(self.chunk, left))
            if max_chunks is not None and chunks >= max_chunks:
                break
            self.engine.step(n)
            chunks += 1
            self.lattice_updates += n * sum(r.spec.nodes for r in self.active.values())
            # while the chunk runs, build the host inputs of the head of the
            # queue (the next admission) in its shadow
            self._prepare(waiting, pending, t0)
            torch.cuda.current_stream(self.engine.device).synchronize()
            now = time.perf_counter() - t0
            if self.health_check:
                for req in self._unstable():
                    self._finish(req, now, stable=False)
                    if on_event:
                        on_event("failed", req)
            for cid, req in list(self.active.items()):
                k, target = self.engine.status(cid)
                if k >= target:
                    self._finish(req, now)
                    if on_event:
                        on_event("done", req)
            if self.ckpt_dir and self.ckpt_every > 0 and chunks % self.ckpt_every == 0:
                self._checkpoint()
        self.wall = time.perf_counter() - t0
        return self.done
