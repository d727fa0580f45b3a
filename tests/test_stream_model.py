"""CPU model of the two-slot job stream of the wavefront renderer (no GPU).

The renderer (pt_amd/csrc/trace.hip render_views, k_refill / k_refill_commit, k_slot_reset,
k_slot_free) streams jobs of paths through one path pool: each job is loaded into one of two
slots, the refill stage issues the older slot's remaining paths first, the shade stage ends
paths and counts them per slot, and the host reads the counters one iteration late (a pinned
snapshot copied at the end of every iteration) to decide when a job has ended, accumulates
it and frees its slot. This model restates that protocol step by step - device counters,
refill plan with its unsigned arithmetic, lagged snapshots, slot reuse - and checks over
random schedules that:

  * no path is issued past its job's end, and every path of every job is issued and ends
    exactly once;
  * a job is accumulated only after all its paths have ended, none of them still in the pool;
  * jobs are accumulated in the order they were loaded;
  * the stream drains (every job finishes);
  * every shadow ray is traced exactly once, and a job is accumulated only after the shadow
    rays of all its paths are traced: pipelined iterations (the production path with graphs)
    trace an iteration's shadow rays in the next iteration, before its shade stage; other
    iterations trace them in their own, after tracing any a pipelined one left (flush).

A variant with the defect found when the two-slot design first ran on the GPU (a freed slot
kept its start counter and the refill's unsigned `paths - started` wrapped) must violate the
first property, and variants that skip the flush or trace the queue a non-pipelined iteration
already traced must violate the last, which shows the checks see those classes of error.
"""
import random

import pytest

U32 = 0xFFFFFFFF


class Violation(AssertionError):
    pass


class Device:
    """The device side: job tables, counters, the pool of in-flight paths."""

    def __init__(self, capacity, rng, clear_on_free=True, clamp=True):
        self.capacity = capacity
        self.rng = rng
        self.clear_on_free = clear_on_free
        self.clamp = clamp
        self.job = [{"paths": 0, "seq": 0, "id": None}, {"paths": 0, "seq": 0, "id": None}]
        self.next = [0, 0]    # CTR(5 + s): next path to start
        self.ended = [0, 0]   # CTR(7 + s): paths ended
        self.pool = []        # (slot, job id, path) of the paths in the queue
        self.snapshot = [None, None]  # per queue parity: (count, ended0, ended1, next0, next1)
        self.q = 0
        self.issued = {}      # job id -> set of issued paths
        self.done = {}        # job id -> set of ended paths
        self.shadow = []      # the shadow queue: (job id, path, ray number), until it is cleared
        self.queued = {}      # job id -> shadow rays ever queued for its paths
        self.traced = set()   # shadow rays traced
        self.rays = 0

    def trace_shadow(self):
        for ray in self.shadow:
            if ray in self.traced:
                raise Violation("shadow ray %r traced twice" % (ray,))
            self.traced.add(ray)

    # k_slot_reset: a job was loaded into slot s
    def slot_reset(self, s, paths, seq, job_id):
        self.job[s] = {"paths": paths, "seq": seq, "id": job_id}
        self.next[s] = 0
        self.ended[s] = 0

    # k_slot_free: the job of slot s was accumulated
    def slot_free(self, s):
        self.job[s]["paths"] = 0
        if self.clear_on_free:
            self.next[s] = 0
            self.ended[s] = 0

    def iteration(self, pipelined=False, consume=False, flush=False):
        # the previous iteration's shadow rays: with this iteration's extension rays (consume)
        # or before it (flush); a queue not handed over is cleared by k_prepare
        if consume or flush:
            self.trace_shadow()
        self.shadow = []  # k_clear_shadow after tracing them, else k_prepare
        # shade: every queued path continues or ends (ended paths counted per slot), most
        # with a shadow ray towards a light
        cont = []
        for slot, jid, p in self.pool:
            if self.rng.random() < 0.7:
                self.shadow.append((jid, p, self.rays))
                self.queued.setdefault(jid, set()).add((jid, p, self.rays))
                self.rays += 1
            if self.rng.random() < 0.45:
                cont.append((slot, jid, p))
            else:
                self.ended[slot] += 1
                self.done.setdefault(jid, set())
                if p in self.done[jid]:
                    raise Violation("path %d of job %s ended twice" % (p, jid))
                self.done[jid].add(p)
        # refill_plan: the older slot issues first; unsigned arithmetic as on the device
        first = 1 if ((self.job[1]["seq"] - self.job[0]["seq"]) & U32) >= 0x80000000 else 0
        room = self.capacity - len(cont)
        n = [0, 0]
        for s in (first, 1 - first):
            paths, nxt = self.job[s]["paths"], self.next[s]
            if self.clamp:
                n[s] = min(room, paths - nxt) if paths > nxt else 0
            else:
                n[s] = min(room, (paths - nxt) & U32)
            room -= n[s]
        for s in (first, 1 - first):
            jid, paths = self.job[s]["id"], self.job[s]["paths"]
            for k in range(n[s]):
                p = self.next[s] + k
                if p >= paths:
                    raise Violation("path %d issued past the end of job %s (%d paths)" % (p, jid, paths))
                iss = self.issued.setdefault(jid, set())
                if p in iss:
                    raise Violation("path %d of job %s issued twice" % (p, jid))
                iss.add(p)
                cont.append((s, jid, p))
        if not pipelined:
            self.trace_shadow()
        # k_refill_commit
        for s in (0, 1):
            self.next[s] += n[s]
        self.pool = cont
        self.snapshot[self.q] = (len(cont), self.ended[0], self.ended[1], self.next[0], self.next[1])
        self.q = 1 - self.q


def run_stream(rng, calls, capacity, clear_on_free=True, clamp=True, max_iters=100000, modes=None, flush=True,
               consume_always=False):
    """Host side of render_views over a list of render calls (each a list of job sizes, traced
    pipelined or not as `modes` says, default not); returns the job ids in load order and in
    the order they were accumulated."""
    dev = Device(capacity, rng, clear_on_free, clamp)
    slots = [dict(busy=False, paths=0, seq=0, first=0, issued=0, ended=0, id=None) for _ in range(2)]
    st = {"launched": 0, "seq": 0, "q": 0, "pending": False}
    finished, loaded = [], []
    next_id = [0]

    def idle():
        return not slots[0]["busy"] and not slots[1]["busy"]

    def load(s, paths):
        st["seq"] += 1
        jid = next_id[0]
        next_id[0] += 1
        dev.slot_reset(s, paths, st["seq"], jid)
        slots[s].update(busy=True, paths=paths, seq=st["seq"], first=st["launched"], issued=0, ended=0, id=jid)
        loaded.append(jid)

    def finish(s):
        jid = slots[s]["id"]
        paths = slots[s]["paths"]
        if len(dev.done.get(jid, ())) != paths:
            raise Violation("job %s accumulated with %d of %d paths ended" % (jid, len(dev.done.get(jid, ())), paths))
        if any(j == jid for _, j, _ in dev.pool):
            raise Violation("job %s accumulated while paths of it are in the pool" % jid)
        if not dev.queued.get(jid, set()) <= dev.traced:
            raise Violation("job %s accumulated before its shadow rays were traced" % jid)
        finished.append(jid)
        dev.slot_free(s)
        slots[s]["busy"] = False

    def render(jobs, drain, pipelined=False):
        jobs = list(jobs)
        iters = 0
        while True:
            for s in (0, 1):
                if jobs and not slots[s]["busy"]:
                    load(s, jobs.pop(0))
            if idle():
                return
            if not drain and not jobs:
                if all(not sl["busy"] or sl["issued"] == sl["paths"] for sl in slots):
                    return
            consume = pipelined and (st["pending"] or consume_always)
            dev.iteration(pipelined, consume, flush and not pipelined and st["pending"])
            st["pending"] = pipelined
            st["q"] = 1 - st["q"]
            st["launched"] += 1
            iters += 1
            if iters > max_iters:
                raise Violation("the stream does not drain")
            if st["launched"] < 2:
                continue
            counted = st["launched"] - 2
            snap = dev.snapshot[st["q"]]
            for s in (0, 1):
                if not slots[s]["busy"] or counted < slots[s]["first"]:
                    continue
                slots[s]["ended"] = snap[1 + s]
                slots[s]["issued"] = snap[3 + s]
            for _ in range(2):
                busy = [s for s in (0, 1) if slots[s]["busy"]]
                if not busy:
                    break
                oldest = min(busy, key=lambda s: slots[s]["seq"])
                if slots[oldest]["ended"] != slots[oldest]["paths"]:
                    break
                finish(oldest)

    modes = modes or [False] * len(calls)
    for jobs, pipelined in zip(calls, modes):
        render(jobs, drain=False, pipelined=pipelined)
        if rng.random() < 0.3:  # an option change settles the stream first
            render([], drain=True, pipelined=pipelined)
    render([], drain=True, pipelined=modes[-1] if modes else False)  # pt_synchronize / readback
    for jid in loaded:
        if len(dev.issued.get(jid, ())) != len(dev.done.get(jid, ())):
            raise Violation("job %s: issued and ended paths differ" % jid)
    return loaded, finished


def _schedule(rng):
    calls = []
    for _ in range(rng.randint(1, 6)):
        calls.append([rng.randint(1, 400) for _ in range(rng.randint(1, 4))])
    return calls, rng.choice([8, 16, 37, 64, 256])


def test_stream_protocol_holds_over_random_schedules():
    rng = random.Random(12345)
    for _ in range(600):
        calls, cap = _schedule(rng)
        loaded, finished = run_stream(rng, calls, cap)
        assert finished == loaded  # every job, in load order


def test_model_detects_the_unclamped_refill_of_a_reused_slot():
    rng = random.Random(7)
    seen = False
    for _ in range(300):
        calls, cap = _schedule(rng)
        try:
            run_stream(rng, calls, cap, clear_on_free=False, clamp=False)
        except Violation as e:
            assert "past the end" in str(e) or "issued twice" in str(e)
            seen = True
            break
    assert seen


@pytest.mark.parametrize("clear,clamp", [(True, False), (False, True)])
def test_either_fix_alone_keeps_the_protocol(clear, clamp):
    # clearing the counters of a freed slot, or clamping the subtraction, each removes the wrap
    rng = random.Random(99)
    for _ in range(200):
        calls, cap = _schedule(rng)
        loaded, finished = run_stream(rng, calls, cap, clear_on_free=clear, clamp=clamp)
        assert finished == loaded


def test_pipelined_shadow_rays_over_random_mode_changes():
    rng = random.Random(2024)
    for _ in range(600):
        calls, cap = _schedule(rng)
        modes = [rng.random() < 0.6 for _ in calls]
        loaded, finished = run_stream(rng, calls, cap, modes=modes)
        assert finished == loaded


@pytest.mark.parametrize("defect", ["no flush", "consume a traced queue"])
def test_model_detects_mishandled_shadow_queues(defect):
    rng = random.Random(5)
    seen = False
    for _ in range(300):
        calls, cap = _schedule(rng)
        modes = [rng.random() < 0.5 for _ in calls]
        try:
            run_stream(rng, calls, cap, modes=modes, flush=defect != "no flush",
                       consume_always=defect == "consume a traced queue")
        except Violation as e:
            assert "shadow" in str(e)
            seen = True
            break
    assert seen
