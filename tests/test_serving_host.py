"""Host-side logic of the serving layer (no GPU): the seeded request stream
and the segment systems it describes."""
import numpy as np
import pytest
from scipy import stats

from sponge_amd import serving


def test_request_stream_distributions():
    """Atom counts log-uniform in [1500, 50000] (KS test on the logs), Poisson
    arrivals at 40 per second (mean gap and exponential KS test), steps as
    asked, seeded (the same stream twice)."""
    reqs = serving.make_requests(count=4000, rate=40.0, steps=1000, seed=7)
    n = np.array([r.n_atoms for r in reqs], np.float64)
    assert n.min() >= 1500 and n.max() <= 50000
    u = (np.log(n) - np.log(1500)) / (np.log(50000) - np.log(1500))
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    t = np.array([r.arrival for r in reqs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 40.0) < 5 * (1 / 40.0) / np.sqrt(len(gaps))
    assert stats.kstest(gaps * 40.0, "expon").pvalue > 1e-3
    assert all(r.steps == 1000 for r in reqs)
    again = serving.make_requests(count=4000, rate=40.0, steps=1000, seed=7)
    assert [r.n_atoms for r in again] == [r.n_atoms for r in reqs]


def test_segment_system_matches_request():
    """A request's water box holds round(N/3) waters at 29.915 A^3 each
    (1.0 g/cm^3) with the 9 A cutoff of the serving workload."""
    r = serving.Request(0, 4000, 10, 0.0, 3)
    s = r.system()
    nw = int(round(4000 / 3.0))
    assert s["n"] == 3 * nw and s["n_waters"] == nw
    assert abs(float(s["box"][0]) ** 3 - nw * serving.WATER_VOLUME) < 1e-3 * nw * 29.915
    assert s["cutoff"] == 9.0
    assert r.system() is s
    assert serving.estimate_footprint(50000) > serving.estimate_footprint(1500) > 0


def test_grid_key_matches_cell_rule():
    """GRIDS::Initial (neighbor_list.cpp:58-60): floor(L / ((rc + skin) / 2)) cells per axis,
    in float; a 27.35 A box with rc 9, skin 2 has 4 cells (27.35 / 5.5 = 4.97)."""
    from sponge_amd import autotune

    class Ctx:
        cutoff = 9.0
        box = np.array([27.35, 33.0, 55.0], np.float32)
        n = 2000

    assert autotune.grid_key(Ctx()) == (4, 6, 10)
    assert autotune.grid_key(Ctx(), occupancy=True) == (4, 6, 10, 8)


class _Fake:
    """A stand-in device for the batched server's host loop: graph replays
    apply their members' steps at once, and an event reports completion
    after `lag` polls, so periods stay queued for a while as on a device."""

    def __init__(self, lag):
        self.lag = lag
        self.capturing = None
        self.replays = 0
        self.replays_while_admitting = 0
        self.unjoined = set()   # members built and not yet recorded into a plan
        self.live = 0           # members built and not yet collected
        self.slots = 0
        self.errors = []

    def install(self, monkeypatch):
        import contextlib
        import torch
        fake = self

        class Event:
            def __init__(self):
                self.left = 0

            def record(self, stream=None):
                self.left = fake.lag

            def query(self):
                self.left -= 1
                return self.left < 0

        class Stream:
            def __init__(self):
                self.waited = set()

            def wait_event(self, ev):
                self.waited.add(ev)

            def synchronize(self):
                pass

        class Graph:
            def __init__(self):
                self.batch = None

            def capture_begin(self):
                fake.capturing = self

            def capture_end(self):
                fake.capturing = None

            def replay(self):
                fake.replays += 1
                if fake.unjoined:
                    fake.replays_while_admitting += 1
                for nb in self.batch.rec:
                    if nb.closed or nb.collected:
                        fake.errors.append("replay of a retired member %d" % nb.req.rid)
                    if nb.member.ready not in self.batch.stream.waited:
                        fake.errors.append("member %d replayed before joining" % nb.req.rid)
                    nb.steps_done += nb.chunk
                    if nb.steps_done > nb.req.steps:
                        fake.errors.append("member %d stepped past its end" % nb.req.rid)

            def reset(self):
                self.batch = None

        class Nb:
            def __init__(self, req, chunk, member):
                self.req, self.chunk, self.member = req, chunk, member
                self.steps_done = chunk          # the eager first period
                self.closed = self.collected = False

            def close(self):
                self.closed = True

        class Member:
            def __init__(self, req, chunk=10, tune=True):
                self.req = req
                self.nb = Nb(req, chunk, self)
                self.ready = Event()
                self.ready.record()
                self.done = chunk
                self.last = self.ready
                self.joined = False
                fake.unjoined.add(self.nb)
                fake.live += 1
                if fake.live > fake.slots:
                    fake.errors.append("%d members in flight" % fake.live)

            def collect(self):
                if self.nb.collected or self.nb.steps_done != self.req.steps:
                    fake.errors.append("member %d collected twice or early" % self.req.rid)
                self.nb.collected = True
                fake.live -= 1
                self.req.result = dict(steps=self.nb.steps_done, status=0)

        class Batch:
            def __init__(self):
                self.rec = []
                self.stream = None

            def record(self, nb, first_step, steps):
                if nb.steps_done != first_step:
                    fake.errors.append("member %d recorded at step %d, has run %d" % (
                        nb.req.rid, first_step, nb.steps_done))
                fake.unjoined.discard(nb)
                self.rec.append(nb)

            def plan(self):
                pass

            def info(self):
                return dict(launches=1, merged=len(self.rec))

            def launch(self, stream):
                self.stream = stream
                fake.capturing.batch = self

            def close(self):
                pass

        monkeypatch.setattr(torch.cuda, "Event", Event)
        monkeypatch.setattr(torch.cuda, "Stream", Stream)
        monkeypatch.setattr(torch.cuda, "CUDAGraph", Graph)
        monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
        monkeypatch.setattr(torch.cuda, "stream", lambda s: contextlib.nullcontext())
        monkeypatch.setattr(serving, "Member", Member)
        monkeypatch.setattr(serving, "StepBatch", Batch)


def test_batched_loop_steps_every_member_exactly(monkeypatch):
    """The batched server's host loop (serving.py, Server._run_batched) on a
    stand-in device: every request runs exactly its steps (one-period
    requests included) and is collected once, after its last period; no
    member is replayed before its first period is done or after it left;
    each re-plan records a member at the step it has reached; the current
    graph keeps being replayed while a new member is admitted; no more than
    `slots` members are in flight."""
    import torch  # noqa: F401  (the stand-ins replace attributes of torch.cuda)
    fake = _Fake(lag=3)
    fake.install(monkeypatch)
    fake.slots = 4
    rng = np.random.default_rng(5)
    steps = [10, 20, 30, 50, 120, 10, 200, 40, 60, 90, 10, 150, 70, 30, 20, 250]
    reqs = [serving.Request(k, 300, s, 0.001 * k, 100 + k) for k, s in enumerate(steps)]
    rng.shuffle(reqs)
    reqs.sort(key=lambda r: r.arrival)
    for r in reqs:
        r.sysd = {"n": 300}   # the stand-in members never read the system
    out = serving.Server(slots=4, chunk=10, depth=2, launch="batched").run(reqs, realtime=False)
    assert fake.errors == []
    assert all(r.result is not None and r.result["steps"] == r.steps for r in reqs)
    assert out["requests"] == len(reqs) and out["plans"] >= 2
    assert fake.replays_while_admitting > 0
    assert out["periods"] == fake.replays and fake.live == 0


def test_server_rejects_step_counts_off_the_period():
    """Step counts are whole neighbour-list periods (at least one)."""
    for bad in (0, 15):
        r = serving.Request(0, 300, bad, 0.0, 1)
        r.sysd = {"n": 300}
        with pytest.raises(ValueError):
            serving.Server(chunk=10).run([r], realtime=False)


def test_streams_loop_steps_every_segment_exactly(monkeypatch):
    """The per-segment streams loop (serving.py, Server.run with
    launch="streams") with stand-in segments: each segment queues exactly the
    periods after its eager first one (none for a one-period request), never
    more than `depth` at a time, and is collected once, after the last one
    has completed; no more than `slots` segments are in flight."""
    import torch
    errors, live = [], [0]
    depth, slots, lag = 2, 3, 2

    class Segment:
        def __init__(self, req, chunk=10, tune=True, overlap=False, plain=False):
            self.req, self.chunk = req, chunk
            self.done = chunk
            self.inflight = []      # polls left per queued period
            self.collected = False
            live[0] += 1
            if live[0] > slots:
                errors.append("%d segments in flight" % live[0])

        def advance(self):
            if self.done >= self.req.steps:
                errors.append("segment %d advanced past its end" % self.req.rid)
            self.inflight.append(lag)
            self.done += self.chunk
            if len(self.inflight) > depth:
                errors.append("segment %d queued %d periods" % (self.req.rid, len(self.inflight)))

        def retire(self):
            self.inflight = [k - 1 for k in self.inflight]
            while self.inflight and self.inflight[0] < 0:
                self.inflight.pop(0)
            return len(self.inflight)

        @property
        def queued_all(self):
            return self.done >= self.req.steps

        def collect(self):
            if self.collected or self.inflight or self.done != self.req.steps:
                errors.append("segment %d collected twice or early" % self.req.rid)
            self.collected = True
            live[0] -= 1
            self.req.result = dict(steps=self.done, status=0)

    monkeypatch.setattr(serving, "Segment", Segment)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    steps = [10, 40, 20, 90, 10, 30, 60, 10, 120, 50]
    reqs = [serving.Request(k, 300, s, 0.001 * k, 50 + k) for k, s in enumerate(steps)]
    for r in reqs:
        r.sysd = {"n": 300}
    out = serving.Server(slots=slots, chunk=10, depth=depth, launch="streams").run(
        reqs, realtime=False)
    assert errors == [] and live[0] == 0
    assert all(r.result["steps"] == r.steps for r in reqs)
    assert out["periods"] == sum(s // 10 - 1 for s in steps)
