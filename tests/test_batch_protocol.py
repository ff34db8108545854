"""A CPU model of the persistent batched kernel's task protocol, run under
random wave interleavings.

Each workgroup of the batched kernel (taco_amd/csrc/batched.hip,
k_batch_spmv / k_batch_spmm) has 4 waves.  Per task, wave 0 takes a slot
from an atomic queue counter and writes it to LDS; a workgroup barrier;
every wave reads the slot; a second barrier; a negative slot ends the queue.
The shipped form (for_each_task) walks one queue per XCD, its own first.  An
earlier form added, per task, a third barrier and a 64-bit atomic add of the
task's duration read from the s_memrealtime clock; its one GPU run did not
return.  The model runs both protocols at the configuration of that run
(48 instances of 100..4,000 rows, 256 rows per task, one workgroup per task
up to 8 per CU) and checks that every interleaving tried ends with every
wave finished and every task run exactly once, so the queue and barrier
protocol is not what stopped that run.  A barrier here waits for every wave
of the workgroup that has not ended, as s_barrier does."""
import numpy as np

from taco_amd import batch as BT

WAVES = 4


def tasks_of(m, rows_per_task):
    return [(i, r) for i, mi in enumerate(m) for r in range(0, int(mi), rows_per_task)]


class Workgroup:
    """Wave programs as generators; yield 'barrier' to wait at a barrier."""

    def __init__(self, wg, program):
        self.slot = None
        self.waves = [program(self, wg, w) for w in range(WAVES)]
        self.state = ["run"] * WAVES  # run / barrier / done
        self.arrived = set()

    def step(self, w):
        try:
            ev = next(self.waves[w])
        except StopIteration:
            self.state[w] = "done"
            self.release_if_complete()
            return
        if ev == "barrier":
            self.state[w] = "barrier"
            self.arrived.add(w)
            self.release_if_complete()

    def release_if_complete(self):
        live = [w for w in range(WAVES) if self.state[w] != "done"]
        if live and all(self.state[w] == "barrier" for w in live):
            for w in live:
                self.state[w] = "run"
            self.arrived.clear()

    def runnable(self):
        return [w for w in range(WAVES) if self.state[w] == "run"]


def simulate(nwg, program, rng, max_steps=10_000_000):
    wgs = [Workgroup(g, program) for g in range(nwg)]
    for _ in range(max_steps):
        choices = [(g, w) for g, wg in enumerate(wgs) for w in wg.runnable()]
        if not choices:
            break
        g, w = choices[rng.integers(len(choices))]
        wgs[g].step(w)
    stuck = [g for g, wg in enumerate(wgs) if any(s != "done" for s in wg.state)]
    return stuck


def timed_protocol(ntasks, executed, counter):
    """The removed timed kernel: queue counter, barrier, read, barrier, body,
    then a third barrier and wave 0's time accumulation."""
    def program(wg, g, w):
        while True:
            if w == 0:
                t = counter[0]
                counter[0] += 1
                wg.slot = t
            yield "barrier"
            t = wg.slot
            yield "barrier"
            if t >= ntasks:
                return
            if w == 0:
                executed[t] += 1
            yield "body"
            yield "barrier"  # TIMING: all waves done with the task
            if w == 0:
                yield "time-atomic"
    return program


def shipped_protocol(bounds, executed, counters, nwg):
    """for_each_task: per queue (home queue first) take slots until the queue
    is exhausted (slot -1), two barriers per slot."""
    nq = len(bounds) - 1

    def program(wg, g, w):
        home = g % nq if nq > 1 else 0
        for k in range(nq):
            qi = (home + k) % nq
            lo, hi = bounds[qi], bounds[qi + 1]
            while True:
                if w == 0:
                    c = counters[qi]
                    counters[qi] += 1
                    wg.slot = lo + c if c < hi - lo else -1
                yield "barrier"
                slot = wg.slot
                yield "barrier"
                if slot < 0:
                    break
                if w == 0:
                    executed[slot] += 1
                yield "body"
    return program


def autotune_config():
    m, _ = BT.instance_plan(48, 7, 100, 4000)
    return tasks_of(m, 256)


def test_timed_protocol_terminates_under_random_interleavings():
    tasks = autotune_config()
    nwg = min(len(tasks), 256 * 8)
    for seed in range(5):
        executed = np.zeros(len(tasks), np.int64)
        counter = [0]
        stuck = simulate(nwg, timed_protocol(len(tasks), executed, counter),
                         np.random.default_rng(seed))
        assert stuck == [], seed
        assert np.all(executed == 1), seed
        assert counter[0] == len(tasks) + nwg


def test_shipped_protocol_terminates_under_random_interleavings():
    tasks = autotune_config()
    for nq, nwg in ((8, min(len(tasks), 2048)), (8, 3), (1, 17)):
        bounds = np.linspace(0, len(tasks), nq + 1).astype(int).tolist()
        for seed in range(3):
            executed = np.zeros(len(tasks), np.int64)
            counters = [0] * nq
            stuck = simulate(nwg, shipped_protocol(bounds, executed, counters, nwg),
                             np.random.default_rng(seed))
            assert stuck == [], (nq, nwg, seed)
            assert np.all(executed == 1), (nq, nwg, seed)


def test_model_reports_waves_that_never_end():
    """The checks above are not vacuous: wave 0 waits at a barrier that the
    other waves, spinning, never reach, and the model reports both
    workgroups stuck."""
    def program(wg, g, w):
        if w == 0:
            yield "barrier"
            return
        while True:
            yield "spin"

    stuck = simulate(2, program, np.random.default_rng(0), max_steps=10_000)
    assert stuck == [0, 1]
