"""Continuous batching of independent MD segments on one GPU.

A request is an independent water-box MD segment (its own atoms, box, seed
and step count).  The server keeps every admitted segment's state resident in
HBM in its own library context (no global state is shared between contexts)
with its own HIP stream.  Each segment advances by replaying its captured
step graph (one neighbour-list period of whole MD steps: force path, middle
Langevin leap-frog, SETTLE) on its stream, with a few periods queued ahead, so
the kernels of all admitted segments overlap on the device.  Whenever a
segment's last period completes it leaves, and arrived requests are admitted
while slots and the HBM budget allow (admission by an upper estimate of the
footprint; the measured one is reported).  Window shapes of the pair kernel,
and how its blocks hand the window forces on, are tuned per grid class on
admission.  A segment's arithmetic does not
depend on what runs beside it (integer fixed-point sums, counter-based
noise), so its result is the same alone and in the mix.

Request streams follow the serving workload: atom counts log-uniform in
[n_min, n_max], waters at 1.0 g/cm^3 (29.915 A^3 per water), cutoff 9 A,
PME grid spacing about 1 A, 2 fs steps, Poisson arrivals.
"""
import time

import numpy as np

from . import autotune, systems
from .lib import Nonbonded, StepBatch

WATER_VOLUME = 29.915      # A^3 per TIP3P water at 1.0 g/cm^3
MAX_ATOMS_IN_CELL = 256    # the largest per-cell capacity the windowed list builder keeps
BATCH_SHAPE = (2, 2, 2)    # window shape of windowed members of a step batch
NS_PER_STEP = 2.0e-6       # 2 fs


class Request:
    def __init__(self, rid, n_atoms, steps, arrival, seed):
        self.rid, self.n_atoms, self.steps = rid, int(n_atoms), int(steps)
        self.arrival, self.seed = float(arrival), int(seed)
        self.sysd = None
        self.admitted = self.finished = None
        self.result = None

    def system(self, cutoff=9.0):
        """The segment's water box (generated once, before it arrives)."""
        if self.sysd is None:
            nw = max(2, int(round(self.n_atoms / 3.0)))
            box = (nw * WATER_VOLUME) ** (1.0 / 3.0)
            self.sysd = systems.tip3p_box(nw, box, seed=self.seed, cutoff=cutoff)
        return self.sysd


def make_requests(count=256, n_min=1500, n_max=50000, rate=40.0, steps=1000, seed=20260217):
    """Seeded request stream: log-uniform atom counts, Poisson arrivals at
    `rate` per second."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count))
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    return [Request(k, n[k], steps, t[k], seed + 1 + k) for k in range(count)]


class Segment:
    """One admitted request: its library context, stream, thermostat state
    and the captured graph of `chunk` whole MD steps."""

    def __init__(self, req, chunk=10, tune=True, overlap=False, temperature=300.0, plain=False):
        import torch
        self.torch = torch
        self.req = req
        self.chunk = chunk
        s = req.system()
        free0 = torch.cuda.mem_get_info()[0]
        self.stream = torch.cuda.Stream()
        # the last cell of each axis takes the remainder of box / cell width
        # and can be nearly twice as wide, so a corner cell can hold ~8x the
        # mean occupancy (~17 atoms here): the per-cell capacity key is raised
        # from its default of 150
        nb = Nonbonded(s, cutoff=s["cutoff"], refresh_interval=chunk, stream=self.stream,
                       max_atom_in_grid_numbers=MAX_ATOMS_IN_CELL)
        nb.set_pme_overlap(overlap)
        if not plain:
            nb.set_export(False)   # nothing reads the reference-layout list here
        self.nb = nb
        # window shape and hand-off mode of the window forces, tuned per grid class
        tuned = autotune.tune_window(nb, interval=chunk) if tune and not plain \
            else (None, None, None)
        self.shape, self.direct_j = tuned[0], tuned[2]
        with torch.cuda.stream(self.stream):
            nb.md_initial(s, temperature=temperature, tau=0.1, seed=req.seed, dt=0.002)
            # the first period runs eagerly (it builds the list from scratch)
            for k in range(chunk):
                nb.md_step(step=k, forced_nl=(k == 0))
        self.done = chunk          # steps queued so far
        self.inflight = []         # events of queued periods, oldest first
        self.graph = None
        if req.steps > chunk:
            nb.sync()
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph, stream=self.stream):
                for k in range(chunk):
                    nb.md_step(step=chunk + k)
        torch.cuda.synchronize()
        self.footprint = max(0, free0 - torch.cuda.mem_get_info()[0])

    def advance(self):
        """Queues the next period on the segment's stream."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            self.graph.replay()
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.inflight.append(ev)
        self.done += self.chunk

    def retire(self):
        """Drops the events of completed periods; returns how many are queued."""
        while self.inflight and self.inflight[0].query():
            self.inflight.pop(0)
        return len(self.inflight)

    @property
    def queued_all(self):
        return self.done >= self.req.steps

    def collect(self):
        nb = self.nb
        nb.sync()
        self.req.result = dict(crd=nb.crd.cpu().numpy(), vel=nb.vel.cpu().numpy(),
                               status=nb.status(), shape=self.shape, direct_j=self.direct_j,
                               footprint=self.footprint)
        nb.close()
        self.nb = self.graph = None


def pretune(requests, chunk=10):
    """Tunes the pair kernel's window shape once per grid class of the
    request stream (before serving starts), so admissions reuse the choice."""
    import torch
    seen = {}
    for r in sorted(requests, key=lambda r: r.n_atoms):
        s = r.system()
        nb = Nonbonded(s, cutoff=s["cutoff"], refresh_interval=chunk,
                       max_atom_in_grid_numbers=MAX_ATOMS_IN_CELL)
        nb.set_export(False)
        key = autotune.grid_key(nb)
        if key not in seen:
            seen[key] = autotune.tune_window(nb, interval=chunk)
        nb.close()
    torch.cuda.synchronize()
    return seen


def estimate_footprint(n_atoms, max_nb=1200):
    """Upper estimate of a segment's device memory before it is built (the
    padded public list and its 16-bit copy dominate), used for admission."""
    return int(n_atoms * (max_nb * 6 + 4096) + 64 * 2 ** 20)


class Member:
    """An admitted request of the batched launch: its library context, built
    and run for its first neighbour-list period (which builds the list) on
    its own stream; `ready` fires when that period is done.  Its later
    periods run inside the server's step batch."""

    def __init__(self, req, chunk=10, tune=True, temperature=300.0):
        import torch
        self.req = req
        s = req.system()
        self.stream = torch.cuda.Stream()
        nb = Nonbonded(s, cutoff=s["cutoff"], refresh_interval=chunk, stream=self.stream,
                       max_atom_in_grid_numbers=MAX_ATOMS_IN_CELL)
        nb.set_pme_overlap(False)  # one stream per member: the batch runs the lanes
        nb.set_export(False)
        self.nb = nb
        # In a step batch the pair dispatch holds the workgroups of every
        # member, so fewer, larger windows pay less per-workgroup set-up and
        # window hand-off: 2 x 2 x 2 where the grid takes it (measured on two
        # 8-member mixes: 4.47 -> 4.08 and 3.00 -> 2.95 ms per period against
        # the shapes tuned alone; every other uniform shape was slower), with
        # the library's default hand-off (atomic adds below 65,536 atoms, the
        # tuner's choice for every serving grid class); the tuner's shape
        # otherwise.  One shape change per admission: each reallocates the
        # window buffers.
        if tune and nb.set_window_shape(*BATCH_SHAPE):
            self.shape, self.direct_j = BATCH_SHAPE, -1
        else:
            tuned = autotune.tune_window(nb, interval=chunk) if tune else (None, None, None)
            self.shape, self.direct_j = tuned[0], tuned[2]
        with torch.cuda.stream(self.stream):
            nb.md_initial(s, temperature=temperature, tau=0.1, seed=req.seed, dt=0.002)
            for k in range(chunk):
                nb.md_step(step=k, forced_nl=(k == 0))
            self.ready = torch.cuda.Event()
            self.ready.record(self.stream)
        self.done = chunk       # steps queued so far
        self.last = self.ready  # fires when the last queued step is done
        self.joined = False

    def collect(self):
        nb = self.nb
        self.req.result = dict(crd=nb.crd.cpu().numpy(), vel=nb.vel.cpu().numpy(),
                               status=nb.status(), shape=self.shape, direct_j=self.direct_j,
                               footprint=None)


class Server:
    """Continuous batching: up to `slots` segments in flight, admitted in
    arrival order while the HBM budget allows.

    launch="batched" (default): the admitted segments advance together, one
    neighbour-list period (chunk steps) at a time, as one step batch
    (lib.StepBatch): each kernel of a step is one dispatch over all segments,
    the reciprocal pipeline on a second stream beside the direct-space work,
    and the period is one captured HIP graph, re-planned whenever a segment
    joins or leaves (while a segment is admitted, the current graph keeps
    the other members' periods queued).  launch="streams": each segment
    replays its own captured period graph on its own stream.  Either way up
    to `depth` periods are queued ahead, so the device does not wait for the
    host's bookkeeping, and a segment's result is the same as when it runs
    alone.  plain=True (streams launch) runs each segment with the library's
    defaults, as a drop-in user would: default window shape, no tuner, the
    reference-layout list written at every rebuild."""

    def __init__(self, slots=8, hbm_budget=64 * 2 ** 30, chunk=10, tune=True, overlap=False,
                 depth=2, launch="batched", free_every=8 * 2 ** 30, plain=False):
        if launch not in ("batched", "streams"):
            raise ValueError("launch is 'batched' or 'streams'")
        if plain and launch != "streams":
            raise ValueError("plain runs use the streams launch")
        self.slots, self.hbm_budget, self.chunk = slots, hbm_budget, chunk
        self.tune, self.overlap, self.depth = tune, overlap, depth
        self.launch, self.free_every, self.plain = launch, free_every, plain

    def run(self, requests, realtime=True, log=None):
        """Serves the requests (their .steps must be a multiple of chunk);
        returns a summary dict.  With realtime, a request is admitted no
        earlier than its arrival time on the wall clock."""
        import torch
        for r in requests:
            if r.steps <= 0 or r.steps % self.chunk:
                raise ValueError("steps must be a positive multiple of the chunk")
            r.system()
        if self.launch == "batched":
            return self._run_batched(requests, realtime, log)
        pending = sorted(requests, key=lambda r: r.arrival)
        active = []
        used = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        periods = 0
        while pending or active:
            now = clock()
            while (pending and pending[0].arrival <= now and len(active) < self.slots and
                   used + estimate_footprint(pending[0].n_atoms) <= self.hbm_budget):
                r = pending.pop(0)
                r.admitted = time.perf_counter() - t0
                seg = Segment(r, self.chunk, self.tune, self.overlap, plain=self.plain)
                used += estimate_footprint(r.n_atoms)
                active.append(seg)
            if not active:
                if pending:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival - clock())))
                continue
            busy = False
            for seg in list(active):
                queued = seg.retire()
                if seg.queued_all:
                    if queued == 0:
                        seg.req.finished = time.perf_counter() - t0
                        seg.collect()
                        used -= estimate_footprint(seg.req.n_atoms)
                        active.remove(seg)
                        if log:
                            log(seg.req)
                    continue
                while queued < self.depth and not seg.queued_all:
                    seg.advance()
                    queued += 1
                    periods += 1
                busy = True
            if busy:
                time.sleep(0)
        torch.cuda.synchronize()
        return self._summary(requests, t0, realtime, periods)

    def _summary(self, requests, t0, realtime, periods, **extra):
        makespan = time.perf_counter() - t0
        lat = np.array([r.finished - r.arrival if realtime else r.finished for r in requests])
        sim_ns = sum(r.steps for r in requests) * NS_PER_STEP
        out = dict(requests=len(requests), makespan_s=makespan, periods=periods,
                   aggregate_ns_per_day=sim_ns / makespan * 86400.0,
                   p50_latency_s=float(np.percentile(lat, 50)),
                   p99_latency_s=float(np.percentile(lat, 99)),
                   atoms_total=int(sum(r.n_atoms for r in requests)),
                   slots=self.slots, launch=self.launch)
        out.update(extra)
        return out

    def _run_batched(self, requests, realtime, log):
        import torch
        chunk = self.chunk
        pending = sorted(requests, key=lambda r: r.arrival)
        active = []      # members in the step batch
        leaving = []     # members with every step queued, not yet complete
        retired = []     # finished members and old batches / graphs, freed in groups
        retired_bytes = 0
        used = 0
        stream = torch.cuda.Stream()
        graph = None
        members = []     # the members the current graph advances
        dirty = False    # membership changed: re-plan
        stale = False    # a member of the current graph left: no more replays
        inflight = []    # events of queued periods, oldest first
        periods = plans = 0
        launches = merged = 0
        host = dict(admit=0.0, collect=0.0, record=0.0, plan=0.0, capture=0.0, free=0.0)

        def tick(key, t):
            host[key] += time.perf_counter() - t
            return time.perf_counter()

        def free_retired():
            # Freeing device memory waits for the device, so finished
            # contexts, old plans and their graphs are freed in groups (after
            # the queued periods, which may still use them), not one by one.
            nonlocal retired_bytes
            t = time.perf_counter()
            stream.synchronize()
            for x in retired:
                if isinstance(x, torch.cuda.CUDAGraph):
                    x.reset()
            for x in retired:
                if not isinstance(x, torch.cuda.CUDAGraph):
                    x.close()
            retired.clear()
            retired_bytes = 0
            tick("free", t)

        def top_up():
            # Queues periods of the current graph.  A member whose every step
            # is queued leaves, and the graph is not replayed after that; a
            # member that joined is not in it and waits for the re-plan.
            nonlocal periods, dirty, stale
            busy = False
            while graph is not None and not stale and len(inflight) < self.depth:
                with torch.cuda.stream(stream):
                    graph[1].replay()
                    ev = torch.cuda.Event()
                    ev.record(stream)
                inflight.append(ev)
                periods += 1
                busy = True
                for m in members:
                    m.done += chunk
                    m.last = ev
                    if m.done >= m.req.steps:
                        active.remove(m)
                        leaving.append(m)
                        dirty = stale = True
            return busy

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        while pending or active or leaving:
            now = clock()
            while (pending and pending[0].arrival <= now and
                   len(active) + len(leaving) < self.slots and
                   used + retired_bytes + estimate_footprint(pending[0].n_atoms) <=
                   self.hbm_budget):
                r = pending.pop(0)
                r.admitted = time.perf_counter() - t0
                t = time.perf_counter()
                m = Member(r, chunk, self.tune)
                tick("admit", t)
                used += estimate_footprint(r.n_atoms)
                if m.done >= r.steps:
                    leaving.append(m)    # its first period was all of it
                else:
                    active.append(m)
                    dirty = True
                # An admission takes milliseconds of host time: keep the
                # current members' periods queued meanwhile.
                while inflight and inflight[0].query():
                    inflight.pop(0)
                top_up()
            if (pending and pending[0].arrival <= now and retired and
                    len(active) + len(leaving) < self.slots):
                free_retired()   # memory, not slots, holds the next request back
            for m in list(leaving):
                if m.last.query():
                    m.req.finished = time.perf_counter() - t0
                    t = time.perf_counter()
                    m.collect()
                    tick("collect", t)
                    leaving.remove(m)
                    used -= estimate_footprint(m.req.n_atoms)
                    retired.append(m.nb)
                    retired_bytes += estimate_footprint(m.req.n_atoms)
                    if log:
                        log(m.req)
            if retired_bytes > self.free_every:
                free_retired()
            while inflight and inflight[0].query():
                inflight.pop(0)
            if dirty and active:
                # re-plan: each member records its next period
                t = time.perf_counter()
                batch = StepBatch()
                for m in active:
                    batch.record(m.nb, m.done, chunk)
                t = tick("record", t)
                batch.plan()
                t = tick("plan", t)
                info = batch.info()
                launches, merged = info["launches"], info["merged"]
                for m in active:
                    if not m.joined:
                        stream.wait_event(m.ready)
                        m.joined = True
                g = torch.cuda.CUDAGraph()
                with torch.cuda.stream(stream):
                    g.capture_begin()
                    batch.launch(stream)
                    g.capture_end()
                tick("capture", t)
                if graph is not None:
                    retired.extend(graph)
                graph = (batch, g)
                members = list(active)
                plans += 1
                dirty = stale = False
            busy = top_up()
            if not busy:
                if not active and not leaving and pending:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival - clock())))
                else:
                    time.sleep(0)
        torch.cuda.synchronize()
        out = self._summary(requests, t0, realtime, periods, plans=plans,
                            last_plan={"launches": launches, "member_launches": merged},
                            host_s={k: round(v, 3) for k, v in host.items()})
        if graph is not None:
            retired.extend(graph)
        free_retired()
        return out


def main(argv=None):
    """Serves a seeded request stream and prints the summary as one JSON line:
    python -m sponge_amd.serving --requests 16 --steps 200 --slots 8"""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--rate", type=float, default=40.0)
    ap.add_argument("--overlap", action="store_true",
                    help="streams launch: reciprocal pipeline of each segment on its own stream")
    ap.add_argument("--launch", choices=("batched", "streams"), default="batched")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--seed", type=int, default=20260217)
    a = ap.parse_args(argv)
    reqs = make_requests(count=a.requests, steps=a.steps, rate=a.rate, seed=a.seed)
    if not a.no_tune:
        pretune(reqs)
    out = Server(slots=a.slots, tune=not a.no_tune, overlap=a.overlap, launch=a.launch).run(reqs)
    out["status_flags"] = int(sum(r.result["status"] for r in reqs))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
