"""Serving loop: continuous batching of assemble-and-solve requests on one GPU.

A request names a registered mesh, or carries a mesh of its own (its topology, built on
the device when the request is placed and dropped once it is answered), and its own
fields (diffusivity and source, or velocity flux) and solver kind.  The server keeps *slots* resident in HBM: per mesh
and solver kind, a matrix on the mesh's pattern, its IC(0) or ILU(0) preconditioner
(with the batched-solve plan cached on it), an assembler whose field buffers are reused,
and right-hand-side / solution vectors.  Serving a request is: upload its fields,
assemble on the device, refactor, and solve as a member of a batch.

Scheduling:
- Admission.  A waiting request is admitted when a free slot of its mesh and kind exists,
  or a new one fits the HBM budget; idle slots of other meshes are evicted
  (least recently used first) to make room.  A request that cannot be placed keeps
  waiting, and the requests behind it are still considered.
- Lanes.  Up to ``lanes`` batches are in flight at once, each a single kernel launch on a
  stream of its own, so batches overlap on the device.  Whenever a lane is free and
  requests are waiting, the admissible ones form a new batch: requests that arrive
  while others run join at the next free lane instead of waiting for everything in
  flight to drain.
- Completion.  Every system of a batch has its own completion flag, so a small request
  is answered as soon as its own solve ends, not when the batch's largest member does.
  Its slot returns to the pool when the whole batch is done.

Each result is independent of what it was batched with (one workgroup solves one system,
with fixed reduction order), so a request gives the same bits alone and in the mix.

The server drives ``3 * lanes + 2`` streams (per lane: preparation and a batch stream per
solver kind; plus allocation and result copies); with fewer hardware queues than streams
(GPU_MAX_HW_QUEUES, read when the HIP runtime starts) some of them share a queue and wait
on each other.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

import fvg


# ---------------------------------------------------------------------------------
# requests and statistics (pure host logic)
# ---------------------------------------------------------------------------------

@dataclass
class Request:
    """One solve: mesh key, solver kind ('cg' or 'bicgstab') and the equation's fields."""
    rid: int
    mesh: str
    kind: str
    prog: list
    face_fields: list
    cell_fields: list
    bc_i: np.ndarray
    bc_e: np.ndarray
    arrival: float = 0.0          # seconds after the start of the run
    admitted: float = -1.0
    launched: float = -1.0
    done: float = -1.0
    x: np.ndarray | None = None
    info: object = None
    error: object = None
    batch_size: int = 0
    raw: dict | None = None       # a mesh of the request's own (topology), when not registered

    @property
    def latency(self):
        return self.done - self.arrival


def poisson_arrivals(n, qps, seed=0):
    """Arrival times (seconds) of n requests from a Poisson process of rate qps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / qps, n))


def latency_summary(requests, wall):
    """p50 / p99 / mean / max latency in milliseconds and throughput in solves/s."""
    lat = np.array([r.latency for r in requests if r.done >= 0.0]) * 1e3
    if lat.size == 0:
        return dict(count=0)
    return dict(count=int(lat.size), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()), throughput=float(lat.size / wall))


def form_group(pending, can_place, max_batch, allow_large=True, large_rows=100_000, rows_of=None):
    """Pick the requests for the next batch, in arrival order.

    The first placeable request fixes the solver kind; later requests of that kind join
    while ``can_place(req)`` says a slot is available for them, up to max_batch.  With
    allow_large False, requests on meshes of at least large_rows cells are left waiting.
    Returns the chosen requests (removed from nothing: the caller removes them).
    """
    group, kind = [], None
    for r in pending:
        if len(group) >= max_batch:
            break
        if kind is not None and r.kind != kind:
            continue
        if not allow_large and rows_of is not None and rows_of(r) >= large_rows:
            continue
        if not can_place(r):
            continue
        group.append(r)
        kind = r.kind
    return group


# ---------------------------------------------------------------------------------
# device side
# ---------------------------------------------------------------------------------

def raw_cells(raw):
    """Cells of a mesh given as topology arrays."""
    return int(raw["cell_ptr"].shape[0] - 1)


def raw_nnz_bound(raw):
    """An upper bound of the CSR entries of a mesh given as topology arrays: the diagonal
    plus one entry per cell-face incidence."""
    return raw_cells(raw) + int(raw["cell_ptr"][-1])


class _MeshEntry:
    def __init__(self, ctx, key, raw, transient=False):
        self.key = key
        self.raw = raw
        self.mesh = fvg.Mesh(ctx, raw)
        self.n = self.mesh.n_cells
        self.nnz = self.mesh.nnz
        self.slot_bytes = None    # measured when the first slot is made
        self.transient = transient   # a request's own mesh: dropped with its last slot


class _Slot:
    def __init__(self, server, entry, kind, req):
        ctx = server.ctx
        before = ctx.bytes_in_use()
        self.entry, self.kind = entry, kind
        self.A = fvg.DistributedMatrix.from_mesh(ctx, entry.mesh)
        self.b = fvg.DistributedVector(ctx, entry.n)
        self.x = fvg.DistributedVector(ctx, entry.n)
        self.asm = fvg.Assembler(entry.mesh, req.prog, req.face_fields, req.cell_fields, req.bc_i, req.bc_e)
        self.asm.run(self.A, self.b, sync=True)
        # slots are only ever solved in batches, whose kernels do not read the blocked sweep
        # layout of device-wide solves: skip its analysis (most of the preconditioner's set-up
        # time, e.g. 116 of 139 ms on 511,104 tets) and its device arrays
        prev = ctx.get_tuning(fvg.TUNE_SWEEP_BLOCKS)
        ctx.set_tuning(fvg.TUNE_SWEEP_BLOCKS, 1)
        try:
            self.M = (fvg.IncompleteCholesky if kind == "cg" else fvg.IncompleteLowerUpper).from_matrix(self.A, 0)
        finally:
            ctx.set_tuning(fvg.TUNE_SWEEP_BLOCKS, prev)
        # make the batched-solve plan now, so the slot's whole footprint is counted here
        probe = fvg.Batch(ctx, [self.A], [self.M], kind, server.options)
        probe.close()
        self.bytes = ctx.bytes_in_use() - before
        entry.slot_bytes = self.bytes
        self.busy = False
        self.fresh = True          # fields of the creating request are already assembled
        self.last_used = time.perf_counter()

    def transfers(self, req):
        """(vectors, arrays) that carry req's fields into this slot's assembler."""
        if self.fresh:
            return [], []
        a = self.asm
        vecs = a.fvecs + a.cvecs + ([a.bi, a.be] if a.bi is not None else [])
        arrs = list(req.face_fields) + list(req.cell_fields) + ([req.bc_i, req.bc_e] if a.bi is not None else [])
        return vecs, arrs

    def prepare(self, ctx):
        """Assemble on ctx's stream without waiting (fields uploaded); returns True when
        the preconditioner needs refactoring."""
        fresh = self.fresh
        if not fresh:
            self.asm.run(self.A, self.b, sync=False, ctx=ctx)
        self.fresh = False
        self.x.fill(0.0, ctx=ctx)
        return not fresh

    def close(self):
        for o in (self.M, self.A, self.b, self.x):
            o.close()
        for v in self.asm.fvecs + self.asm.cvecs + [v for v in (self.asm.bi, self.asm.be) if v is not None]:
            v.close()


@dataclass
class _Lane:
    ctx: object = None                            # this lane's stream for request preparation
    batch: object = None                          # the batch in flight, if any
    objs: dict = field(default_factory=dict)      # solver kind -> this lane's reusable fvg.Batch
    reqs: list = field(default_factory=list)
    slots: list = field(default_factory=list)
    reported: np.ndarray | None = None
    large: bool = False


class Server:
    """Continuous-batching solve server on one device context.

    budget_bytes: HBM the server may hold (matrices, factors, plans, vectors);
    lanes: batches in flight at once; max_batch: systems per batch;
    options: fvg.LinearSolverOptions of every solve (mixed precision by default).
    """

    def __init__(self, budget_bytes=200 * 10**9, lanes=4, max_batch=128, options=None, large_rows=100_000,
                 device=0, max_inflight=192, threads=1024, cluster_rows=8192):
        self.ctx = fvg.Context(device, int(budget_bytes))
        self.io = fvg.Context(device)            # a second stream for result downloads
        self.budget = int(budget_bytes)
        self.options = options or fvg.LinearSolverOptions(1e-8, 0.0, 5000, precision=fvg.PRECISION_MIXED)
        # each lane prepares its requests (uploads, assembly, refactorisation) on a stream of
        # its own, so one lane's preparation never waits behind another's
        self.lanes = [_Lane(ctx=fvg.Context(device)) for _ in range(lanes)]
        self.max_batch = max_batch
        self.large_rows = large_rows
        # systems solving at once: each holds a compute unit, so keep some free for the
        # assembly and factorisation of the next batch
        self.max_inflight = max_inflight
        self.threads = threads
        # mixed precision: a system runs on ceil(n / cluster_rows) workgroups; each lane's
        # launches keep at most its share of the device's resident workgroups, so the
        # lanes' batches can all be resident at once
        self.cluster_rows = cluster_rows
        self.meshes = {}
        self.slots = []
        self.pending = []
        self.finished = []
        self.peak_bytes = 0
        self.evictions = 0
        self.timeline = []          # one record per batch launch (diagnostics)
        # host seconds spent per activity, and counts (diagnostics)
        self.stats = dict(slot_create_s=0.0, slots_created=0, prepare_s=0.0, batch_setup_s=0.0, batches=0,
                          complete_s=0.0)
        self.t0 = time.perf_counter()

    # -- registration / submission -------------------------------------------------
    def register_mesh(self, key, raw, transient=False):
        self.meshes[key] = _MeshEntry(self.ctx, key, raw, transient)
        return self.meshes[key]

    def _rows(self, req):
        e = self.meshes.get(req.mesh)
        return e.n if e is not None else raw_cells(req.raw)

    def _need(self, req):
        """HBM bytes a new slot for req is expected to take (measured once a slot of its
        mesh exists; else about 1200 bytes per CSR entry, plus the mesh of its own)."""
        e = self.meshes.get(req.mesh)
        if e is not None:
            return e.slot_bytes if e.slot_bytes is not None else 1200 * e.nnz
        return 1600 * raw_nnz_bound(req.raw)

    def _entry(self, req):
        """The mesh entry of req, built on the device from the request's own topology the
        first time (the build counts in the request's latency)."""
        e = self.meshes.get(req.mesh)
        if e is None:
            if req.raw is None:
                raise KeyError(f"mesh {req.mesh} is not registered and the request carries none")
            t = time.perf_counter()
            e = self.register_mesh(req.mesh, req.raw, transient=True)
            self.stats["mesh_build_s"] = self.stats.get("mesh_build_s", 0.0) + time.perf_counter() - t
            self.stats["meshes_built"] = self.stats.get("meshes_built", 0) + 1
        return e

    def _drop_transient(self, slots):
        """Close the idle slots of requests' own meshes, and those meshes once unused."""
        for s in slots:
            if s.entry.transient and not s.busy and s in self.slots:
                s.close()
                self.slots.remove(s)
        for key in [k for k, e in self.meshes.items() if e.transient]:
            if not any(s.entry.key == key for s in self.slots):
                self.meshes.pop(key).mesh.close()

    def submit(self, req):
        self.pending.append(req)

    def now(self):
        return time.perf_counter() - self.t0

    # -- slots and budget ------------------------------------------------------------
    def _free_slot(self, req):
        for s in self.slots:
            if not s.busy and s.entry.key == req.mesh and s.kind == req.kind:
                return s
        return None

    def _evict_for(self, need):
        """Destroy idle slots, least recently used first, until `need` more bytes fit."""
        idle = sorted((s for s in self.slots if not s.busy), key=lambda s: s.last_used)
        for s in idle:
            if self.ctx.bytes_in_use() + need <= self.budget:
                return True
            s.close()
            self.slots.remove(s)
            self.evictions += 1
        return self.ctx.bytes_in_use() + need <= self.budget

    def reserve(self, template, count):
        """Provision `count` idle resident slots for the template request's mesh and kind
        (within the budget), so requests arriving later find them ready."""
        have = sum(1 for s in self.slots if s.entry.key == template.mesh and s.kind == template.kind)
        for _ in range(max(0, count - have)):
            entry = self.meshes[template.mesh]
            need = self._need(template)
            if self.ctx.bytes_in_use() + need > self.budget:
                break
            t = time.perf_counter()
            s = _Slot(self, entry, template.kind, template)
            self.stats["slot_create_s"] += time.perf_counter() - t
            self.stats["slots_created"] += 1
            s.fresh = False          # its fields belong to no waiting request
            self.slots.append(s)
        self.peak_bytes = max(self.peak_bytes, self.ctx.bytes_in_use())

    def drop_idle_slots(self):
        """Destroy every slot not in a batch in flight."""
        for s in [s for s in self.slots if not s.busy]:
            s.close()
            self.slots.remove(s)

    def _place(self, req):
        """A slot for req (existing, new, or after eviction), or None to keep it waiting."""
        s = self._free_slot(req)
        if s is not None:
            return s
        if not self._evict_for(self._need(req)):
            return None
        try:
            entry = self._entry(req)
            t = time.perf_counter()
            s = _Slot(self, entry, req.kind, req)
            self.stats["slot_create_s"] += time.perf_counter() - t
            self.stats["slots_created"] += 1
        except fvg.BudgetError:
            self.ctx.sync()
            return None
        self.slots.append(s)
        self.peak_bytes = max(self.peak_bytes, self.ctx.bytes_in_use())
        return s

    # -- the loop --------------------------------------------------------------------
    def _complete(self, lane):
        t = time.perf_counter()
        try:
            return self._complete_inner(lane)
        finally:
            self.stats["complete_s"] += time.perf_counter() - t

    def _complete_inner(self, lane):
        flags, infos = lane.batch.poll()
        for k in np.nonzero(flags & ~lane.reported)[0]:
            k = int(k)
            r = lane.reqs[k]
            r.info = infos[k]
            if r.info.status != 0:   # pivot / symmetry failure of its factorisation: not solved
                r.error = fvg.FvgError(r.info.status, "factorisation failed its checks")
            x = np.empty(lane.slots[k].entry.n)
            fvg.check(fvg.lib().fvg_vec_download(self.io.h, lane.slots[k].x.h, fvg._ptr(x)))
            r.x = x
            r.done = self.now()
            lane.reported[k] = True
            self.finished.append(r)
        if lane.reported.all():
            lane.batch.finish(history=False)
            for s in lane.slots:
                s.busy = False
                s.last_used = time.perf_counter()
            self._drop_transient(lane.slots)
            lane.batch, lane.reqs, lane.slots, lane.reported = None, [], [], None
            return True
        return bool(flags.any())

    def _launch(self, lane, group, slots, large):
        t = time.perf_counter()
        vecs, arrs = [], []
        for r, s in zip(group, slots):
            r.admitted = self.now()
            v, a = s.transfers(r)
            vecs += v
            arrs += a
        fvg.upload_many(lane.ctx, vecs, arrs)
        stale = [s for s in slots if s.prepare(lane.ctx)]
        fvg.refactor_many(lane.ctx, [s.M for s in stale], [s.A for s in stale])
        t1 = time.perf_counter()
        kind = group[0].kind
        b = lane.objs.get(kind)
        t2 = time.perf_counter()
        if b is None:
            b = fvg.Batch(self.ctx, [s.A for s in slots], [s.M for s in slots], kind, self.options)
            b.threads = self.threads
            if b.rows_per_group > 0:
                b.set_cluster(self.cluster_rows, max(1, b.workers // len(self.lanes)))
            b.reserve(self.max_batch)
            lane.objs[kind] = b
        else:
            b.rebind([s.A for s in slots], [s.M for s in slots])
        self.stats["rebind_s"] = self.stats.get("rebind_s", 0.0) + time.perf_counter() - t2
        self.peak_bytes = max(self.peak_bytes, self.ctx.bytes_in_use())
        b.launch([s.b for s in slots], [s.x for s in slots], ctx=lane.ctx)
        self.stats["prepare_s"] += t1 - t
        self.stats["batch_setup_s"] += time.perf_counter() - t1
        self.stats["batches"] += 1
        now = self.now()
        for r in group:
            r.batch_size = len(group)
            r.launched = now
            self.pending.remove(r)
        self.timeline.append(dict(t=now, n=len(group), kind=kind, lane=self.lanes.index(lane),
                                  rows=max(s.entry.n for s in slots)))
        lane.batch, lane.reqs, lane.slots = b, list(group), slots
        lane.reported = np.zeros(len(group), bool)
        lane.large = large

    def step(self):
        """Collect finished systems, then admit waiting requests into free lanes.
        Returns True if anything happened."""
        progressed = False
        for lane in self.lanes:
            if lane.batch is not None:
                progressed |= self._complete(lane)
        large_busy = sum(1 for l in self.lanes if l.batch is not None and l.large)
        inflight = sum(len(l.reqs) for l in self.lanes if l.batch is not None)
        for lane in self.lanes:
            if lane.batch is not None or not self.pending:
                continue
            room = min(self.max_batch, self.max_inflight - inflight)
            if room <= 0:
                break
            allow_large = large_busy < max(1, len(self.lanes) // 2)
            reserved = set()

            def can_place(r):
                s = self._free_slot(r)
                if s is not None and id(s) not in reserved:
                    reserved.add(id(s))
                    return True
                if s is None:
                    idle = sum(x.bytes for x in self.slots if not x.busy and id(x) not in reserved)
                    return self.ctx.bytes_in_use() + self._need(r) <= self.budget + idle
                return False

            group = form_group(self.pending, can_place, room, allow_large, self.large_rows, self._rows)
            if not group:
                if all(l.batch is None for l in self.lanes) and not any(s.busy for s in self.slots):
                    r = self.pending[0]
                    raise RuntimeError(f"request {r.rid} on mesh {r.mesh} does not fit the HBM budget")
                break
            # can_place only estimated; place for real, marking each slot busy at once
            placed, slots = [], []
            for r in group:
                s = self._place(r)
                if s is None:
                    break
                s.busy = True
                placed.append(r)
                slots.append(s)
            if not placed:
                if all(l.batch is None for l in self.lanes) and not any(s.busy for s in self.slots):
                    raise RuntimeError(f"request {group[0].rid} on mesh {group[0].mesh} does not fit the HBM budget")
                break
            large = any(self._rows(r) >= self.large_rows for r in placed)
            self._launch(lane, placed, slots, large)
            inflight += len(placed)
            if large:
                large_busy += 1
            progressed = True
        return progressed

    def run(self, requests, poll_sleep=2e-5):
        """Serve requests whose .arrival times (seconds from now) are set; returns when
        all are answered.  Latency counts from each request's scheduled arrival."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        i, n = 0, len(reqs)
        start_finished = len(self.finished)
        while len(self.finished) - start_finished < n:
            now = self.now()
            while i < n and reqs[i].arrival <= now:
                self.submit(reqs[i])
                i += 1
            if not self.step():
                time.sleep(poll_sleep)
        return self.now()

    def close(self):
        for lane in self.lanes:
            if lane.batch is not None:
                lane.batch.finish(history=False)
                lane.batch = None
            for b in lane.objs.values():
                b.close()
            lane.objs = {}
            lane.ctx.close()
        for s in self.slots:
            s.close()
        self.slots = []
        for e in self.meshes.values():
            e.mesh.close()
        self.io.close()
        self.ctx.close()
