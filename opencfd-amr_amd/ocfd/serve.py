"""Request loop serving independent cases on one GPU.

Each request is a whole case: a flow case (a level, its three state
MultiFabs and a device time control block) run to its stop time, or a batch
of 0D reactors integrated to t_end.  State stays resident on the device from
admission to completion.  Admission is by device-memory footprint against a
budget and by free slots (one HIP stream per slot); requests that arrive while
others are on the GPU are admitted as soon as a slot and memory free up, and
their work overlaps with the running cases' on the device (continuous
batching).  A flow case advances in chunks of device-driven time steps
replayed from a hipGraph captured at admission, so the host only enqueues
graph launches and polls a pinned copy of the control block; a finished case
turns the rest of its chunk into no-ops on the device.

The reference runs one case per process (Source/main.cpp); this loop is the
serving form of its per-level advance (Source/Advance.cpp:81-177) and time
loop (Source/EBR.cpp:189-306).
"""
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import (CASE_CHEM_CHANNEL, CTL_BYTES, NCONS, NGROW, NSPEC, PLANAR, Level, Reactors, StepCtl, default_parm,
               ignition_batch, new_ctl, REACTOR_NCOMP)

MiB = 1 << 20


@dataclass
class FlowRequest:
    """A flow case: `make()` returns (boxes, geom, init kind) as in ocfd.cases,
    or (boxes, geom, kind, eb spec) for an embedded body, or for a reacting
    multispecies case (kind CASE_CHEM_CHANNEL) (boxes, geom, kind, eb spec,
    (state, partial densities)) -- the uniform initial data.  A reacting case
    steps with the Strang-split multispecies advance (scheme 0 Euler or
    1 RK2 for its flow part) and the CHEM build's dt = cfl * 1e-8."""
    name: str
    make: object
    flags: int = 0
    cfl: float = 0.5
    stop_time: float = -1.0
    max_steps: int = -1
    scheme: int = 2
    seed: int = 0
    parm_kw: dict = field(default_factory=dict)
    arrival: float = 0.0          # seconds after the loop starts

    def made(self):
        """The case set-up, made once (it may run device work)."""
        if getattr(self, "_made", None) is None:
            self._made = self.make()
        return self._made

    def is_chem(self):
        return self.made()[2] == CASE_CHEM_CHANNEL

    def footprint(self, boxes=None):
        """Device bytes while resident.  Flow cases: 3 state MultiFabs (4 ghosts,
        5 comps, FP64), the level's FP32+FP64 primitive scratch (grow 4, 6
        comps), the partial divergence scratch (5 comps valid) and the small
        tables.  Reacting cases: 2 state and 2 species MultiFabs (9 comps), the
        same level scratch, the multispecies scratch (19 primitive/transport
        and 3 x 14 face-flux comps per grown cell) and, bounded by every box,
        the embedded-boundary fields and scratch (flags, 4 geometry, 7
        primitive and 3 x 5 face-flux comps)."""
        chem = boxes is None and self.is_chem()
        if boxes is None:
            boxes = self.made()[0]
        ng = nv = 0
        for lo, hi in boxes:
            n = [hi[d] - lo[d] + 1 for d in range(3)]
            nv += n[0] * n[1] * n[2]
            ng += (n[0] + 2 * NGROW) * (n[1] + 2 * NGROW) * (n[2] + 2 * NGROW)
        tables = 128 * len(boxes) * 16 + MiB
        if chem:
            return (2 * 8 * NCONS * ng + 2 * 8 * NSPEC * ng + 12 * 6 * ng + 8 * NCONS * nv + 8 * (19 + 3 * 14) * ng +
                    (4 + 8 * (4 + 7 + 3 * NCONS)) * ng + tables)
        return 3 * 8 * NCONS * ng + 12 * 6 * ng + 8 * NCONS * nv + tables


@dataclass
class ReactorRequest:
    """A batch of n constant-volume reactors integrated to t_end."""
    name: str
    n: int
    seed: int = 7
    t_end: float = 1e-3
    rtol: float = 1e-4
    atol: float = 1e-9
    arrival: float = 0.0
    method: str = "rosenbrock"     # or "adaptive" (step doubling on the linearised-implicit step)

    def footprint(self):
        return self.n * (REACTOR_NCOMP * 8 + 3 * 8 + 2 * 4) + MiB


class _Groups:
    """Launch groups of one case in flight on its stream, each closed by a
    copy of the case's progress to its own pinned buffer and an event.  With
    depth 2 a second group is queued so the stream has work while the host
    looks at the first; a group launched after the end does nothing on the
    device."""

    def __init__(self, torch, depth):
        self.depth = depth
        self.events = [torch.cuda.Event() for _ in range(depth)]
        self.queue = deque()   # buffer indices in launch order
        self.n = 0

    def can_launch(self):
        return len(self.queue) < self.depth

    def next(self):
        i = self.n % self.depth
        self.n += 1
        return i

    def record(self, i, stream):
        self.events[i].record(stream)
        self.queue.append(i)

    def completed(self):
        """Buffer indices of the groups that completed since the last call, in order."""
        out = []
        while self.queue and self.events[self.queue[0]].query():
            out.append(self.queue.popleft())
        return out

    def drained(self):
        return not self.queue or all(self.events[i].query() for i in self.queue)


class _FlowJob:
    # two graphs of `chunk` steps queued (served mix, same box: p50 58 -> 57 ms,
    # mean 163 -> 160 ms)
    DEPTH = 2

    def __init__(self, req, stream, chunk, torch):
        self.req = req
        self.stream = stream
        made = req.made()
        boxes, geom, kind = made[:3]
        eb = made[3] if len(made) > 3 else None
        chem = kind == CASE_CHEM_CHANNEL
        self.level = Level(boxes, geom, eb=eb)
        if req.flags & PLANAR:
            self.level.set_planar(True)
        self.parm = default_parm(**req.parm_kw)
        with torch.cuda.stream(stream):
            self.A = self.level.new_mf()
            self.B = self.level.new_mf()
            self.C = None if chem else self.level.new_mf()
            self.As = self.level.new_spec_mf() if chem else None
            self.Bs = self.level.new_spec_mf() if chem else None
            self.ctl = new_ctl()
            # one pinned copy of the control block per launch group in flight
            self.host = [torch.zeros(CTL_BYTES, dtype=torch.uint8).pin_memory() for _ in range(self.DEPTH)]
        stream.synchronize()   # the zero fills above ran on this stream
        if chem:
            state, spec = made[4]
            self.level.set_inflow_species(spec)
            self.level.init_species_uniform(state, spec, self.A, self.As)
            self.level.ctl_start_chem(self.A, self.ctl, self.parm, req.cfl, 0.0, req.stop_time, req.max_steps,
                                      stream)
            self.graph = self.level.capture_steps_chem(self.A, self.As, self.B, self.Bs, self.ctl, self.parm, chunk,
                                                       req.scheme, 0, req.flags, True, stream)
        else:
            self.level.init_state(kind, self.A, self.parm, req.seed)
            self.level.ctl_start(self.A, self.ctl, self.parm, req.cfl, 0.0, req.stop_time, req.max_steps, stream)
            self.graph = self.level.capture_steps(self.A, self.B, self.C, self.ctl, self.parm, chunk, req.scheme,
                                                  req.flags, stream)
        self.groups = _Groups(torch, self.DEPTH)
        self.cells = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in boxes)

    def launch(self, torch):
        i = self.groups.next()
        self.graph.launch(self.stream)
        with torch.cuda.stream(self.stream):
            self.host[i].copy_(self.ctl, non_blocking=True)
        self.groups.record(i, self.stream)

    def finished(self):
        """True once a completed launch group found the run done; a run that
        stopped on a non-finite state raises."""
        done = False
        for i in self.groups.completed():
            c = StepCtl.from_buffer_copy(self.host[i].numpy().tobytes())
            self.steps, self.time = c.steps, c.time
            if c.status:
                raise RuntimeError(f"served case {self.req.name} stopped at step {c.steps}: status {c.status} "
                                   "(non-finite state)")
            done = done or bool(c.done)
        return done

    def result(self):
        return self.A if self.As is None else (self.A, self.As)

    def release(self):
        """Keep what the device may still be reading until the loop ends:
        a level's destructor frees with hipFree, which waits for the device."""
        g = (self.graph, self.level)
        self.A = self.B = self.C = self.As = self.Bs = self.ctl = None
        self.level = self.graph = None
        return g


class _ReactorJob:
    """A reactor batch integrated in short resumable launches (a few substeps
    per reactor each), so the batch shares its hardware queue with the flow
    cases instead of holding it for its whole integration.  One group of
    chunks at a time: a second queued group shortens the batches (p99 509 ->
    440 ms) but makes the small flow cases wait behind them (p50 58 -> 74 ms)."""

    DEPTH = 1

    def __init__(self, req, stream, torch, budget=16, ahead=8):
        self.req = req
        self.stream = stream
        self.budget = budget
        self.ahead = ahead     # launches queued per poll (launches after the end do nothing)
        T0, p0, phi = ignition_batch(req.n, req.seed)
        with torch.cuda.stream(stream):
            self.r = Reactors(T0, p0, phi, stream)
            self.r.begin_chunks(1e-8, stream, method=req.method)
            self.host = [torch.zeros(1, dtype=torch.int32).pin_memory() for _ in range(self.DEPTH)]
        self.groups = _Groups(torch, self.DEPTH)
        self.cells = 0

    def launch(self, torch):
        i = self.groups.next()
        for _ in range(self.ahead):
            self.r.chunk(self.req.t_end, self.budget, rtol=self.req.rtol, atol=self.req.atol, stream=self.stream)
        with torch.cuda.stream(self.stream):
            self.host[i].copy_(self.r.remaining, non_blocking=True)
        self.groups.record(i, self.stream)

    def finished(self):
        self.steps = None
        return any(int(self.host[i][0]) == 0 for i in self.groups.completed())

    def result(self):
        return self.r.state

    def release(self):
        r, self.r = self.r, None
        return (r,)


class Server:
    """Continuous-batching request loop over one GPU.

    hbm_budget: bytes of device memory the resident cases may hold together.
    slots: streams, i.e. the number of cases resident at once.
    chunk: device-driven time steps per graph launch of a flow case.
    keep_results: copy every finished case's state to the host (for checks).

    HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
    default) and runs the kernels of one queue in order, so with more slots
    than queues a small case waits behind another slot's kernels.  A serving
    process should set GPU_MAX_HW_QUEUES to `slots` before the HIP runtime
    starts (before the first GPU call).
    """

    def __init__(self, hbm_budget=64 << 30, slots=8, chunk=8, keep_results=False):
        import torch
        self.torch = torch
        self.budget = hbm_budget
        self.slots = slots
        self.chunk = chunk
        self.keep_results = keep_results
        self.streams = [torch.cuda.Stream() for _ in range(slots)]   # default priority
        self.records = []
        self.results = {}
        self.peak_used = 0     # largest admitted footprint sum
        self.peak_active = 0

    def run(self, requests, real_time=True):
        """Serve `requests` (their `arrival` offsets in seconds; with
        real_time False they are all present at the start).  Returns per-case
        records (name, arrival, admit, done, latency, steps, cells) and the
        loop's wall time."""
        torch = self.torch
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        pending = []           # arrived, not admitted (FIFO with first-fit)
        active = {}            # slot -> (job, record)
        free_slots = list(range(self.slots))
        used = 0
        graveyard = []
        nxt = 0
        t0 = time.perf_counter()
        while nxt < len(order) or pending or active:
            now = time.perf_counter() - t0
            while nxt < len(order) and (not real_time or requests[order[nxt]].arrival <= now):
                pending.append(order[nxt])
                nxt += 1
            # admission: first pending request that fits in memory and a slot
            admitted = True
            while admitted and pending and free_slots:
                admitted = False
                for pi, ri in enumerate(pending):
                    req = requests[ri]
                    fp = req.footprint()
                    if fp > self.budget:
                        raise RuntimeError(f"request {req.name} needs {fp} B, above the budget {self.budget} B")
                    if used + fp <= self.budget:
                        slot = free_slots.pop(0)
                        s = self.streams[slot]
                        t_adm = time.perf_counter() - t0
                        job = (_FlowJob(req, s, self.chunk, torch) if isinstance(req, FlowRequest)
                               else _ReactorJob(req, s, torch))
                        used += fp
                        self.peak_used = max(self.peak_used, used)
                        self.peak_active = max(self.peak_active, len(active) + 1)
                        rec = dict(name=req.name, arrival=req.arrival if real_time else 0.0, admit=t_adm,
                                   footprint=fp, cells=job.cells)
                        active[slot] = (job, rec)
                        pending.pop(pi)
                        admitted = True
                        break
            # keep every case's stream fed, collect finished cases, and free
            # a finished case's slot and memory once its queued groups drained
            for slot, (job, rec) in list(active.items()):
                if "done" not in rec:
                    while job.groups.can_launch():
                        job.launch(torch)
                    if not job.finished():
                        while job.groups.can_launch():
                            job.launch(torch)
                        continue
                    now = time.perf_counter() - t0
                    rec["done"] = now
                    rec["latency"] = now - rec["arrival"]
                    rec["steps"] = job.steps
                    if self.keep_results:
                        with torch.cuda.stream(job.stream):
                            res = job.result()
                            self.results[rec["name"]] = (tuple(x.cpu() for x in res) if isinstance(res, tuple)
                                                         else res.cpu())
                    self.records.append(rec)
                if job.groups.drained():
                    graveyard.append(job.release())
                    used -= rec["footprint"]
                    del active[slot]
                    free_slots.append(slot)
            if real_time and not active and not pending and nxt < len(order):
                wait = requests[order[nxt]].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        del graveyard
        return self.records, wall


def latency_stats(records):
    lat = np.array([r["latency"] for r in records])
    return dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3), n=len(records))


def poisson_mix(n, rate, seed=99, kinds=None):
    """n requests with exponential inter-arrival times at `rate` per second
    over the served case mix: 2D Sod/DMR slabs (64^2 .. 1024x256, FP32 flux),
    3D periodic KH boxes (64^3 .. 128^3, viscous, FP32 flux), the 2D reacting
    H2/air channel past a cylinder (512x128, 9 species, viscous, RK2, cfl 20
    as in Exec/9_CHEMEB/inputs, 10 steps) and 0D ignition batches
    (1k .. 16k reactors)."""
    from . import FLUX_FP32, PLANAR, VISC, cases
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    kinds = kinds or ("sod", "dmr", "kh", "chemeb", "reactors")
    for q in range(n):
        t += rng.exponential(1.0 / rate)
        kind = kinds[int(rng.integers(len(kinds)))]
        if kind == "sod":
            nx = int(rng.choice([64, 128, 256, 512]))
            ny = max(8, nx // 8)
            out.append(FlowRequest(f"sod{q}", (lambda nx=nx, ny=ny: cases.sod(nx, ny, 2, max_grid=(128, 128, 2))),
                                   flags=FLUX_FP32 | PLANAR, cfl=0.5, stop_time=0.2, arrival=t))
        elif kind == "dmr":
            ny = int(rng.choice([64, 128, 256]))
            nx = 4 * ny
            out.append(FlowRequest(f"dmr{q}", (lambda nx=nx, ny=ny: cases.dmr(nx, ny, 2, sizes=(32, 64, 128),
                                                                             seed=1234 + q)),
                                   flags=FLUX_FP32 | PLANAR, cfl=0.5, stop_time=0.01, arrival=t))
        elif kind == "kh":
            n3 = int(rng.choice([64, 96, 128]))
            out.append(FlowRequest(f"kh{q}", (lambda n3=n3: cases.kh3d(n3)), flags=FLUX_FP32 | VISC, cfl=0.5,
                                   max_steps=10, seed=42, arrival=t))
        elif kind == "chemeb":
            out.append(FlowRequest(f"chemeb{q}", (lambda: cases.chem_channel(512, 128, 2, max_grid=64)), flags=VISC,
                                   cfl=20.0, scheme=1, max_steps=10, arrival=t))
        else:
            nr = int(rng.choice([1024, 4096, 16384]))
            out.append(ReactorRequest(f"ign{q}", nr, seed=int(rng.integers(1 << 30)), t_end=1e-4, arrival=t))
    return out
