"""Admission policy of the serving loop, on the CPU: arrival order, slot cap, HBM budget,
refusal of requests that can never fit."""
import collections

from cellato_amd.serve import Request, admit, serving_workload


def req(i, nbytes):
    return Request(i, "game-of-life", 64, 64, 10, float(i), i, nbytes)


def test_budget_and_order():
    w = collections.deque([req(0, 40), req(1, 50), req(2, 5)])
    admitted, refused, used = admit(w, 0, 0, budget=100, max_live=10)
    assert [r.id for r in admitted] == [0, 1, 2] and used == 95 and refused == 0
    w = collections.deque([req(3, 60), req(4, 5)])
    admitted, refused, used = admit(w, 3, 95, budget=100, max_live=10)
    assert admitted == [] and used == 95 and [r.id for r in w] == [3, 4]   # head of line waits


def test_slots_and_refusal():
    w = collections.deque([req(0, 500), req(1, 10), req(2, 10), req(3, 10)])
    admitted, refused, used = admit(w, 1, 0, budget=100, max_live=3)
    assert refused == 1 and [r.id for r in admitted] == [1, 2] and used == 20 and len(w) == 1


def test_workload_shape():
    reqs = serving_workload(n=200, rate=4.0, seed=1)
    assert len({r.rule for r in reqs}) == 10
    assert all(128 <= r.width <= 8192 and r.width % 64 == 0 and 64 <= r.steps <= 2048 for r in reqs)
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr) and abs(len(reqs) / arr[-1] - 4.0) < 1.0
    assert max(r.bytes for r in reqs) <= 2 * 5 * 8192 * 8192 // 8


class FakeEngine:
    """Ticks that stay in flight for `lag` polls of done() after being recorded, as GPU work
    does while the host runs ahead; times are tick counts."""

    def __init__(self, lag=3):
        self.lag, self.ticks, self.start_called = lag, 0, False

    def start(self):
        self.start_called = True

    def tick(self, arr, n):
        self.ticks += 1
        return [arr[i].steps % 2 for i in range(n)]   # one pass per generation: odd counts flip

    def record(self):
        return {"tick": self.ticks, "polls": 0, "done": False}

    def done(self, ev):
        ev["polls"] += 1
        ev["done"] = ev["done"] or ev["polls"] > self.lag
        return ev["done"]

    def wait(self, ev):
        ev["done"] = True

    def seconds(self, ev):
        return float(ev["tick"])


class FakeGrid:
    def __init__(self, r):
        from cellato_amd.rules import rule as get_rule
        self.rule, self.cur, self.bytes = get_rule(r.rule), 0, r.bytes

    def fill_random(self, seed):
        pass

    def ptrs(self, which):
        return [0x1000 * (which + 1)] * self.rule.planes


def test_budget_holds_until_the_finishing_tick_retires():
    """The bytes of grids actually alive (allocated and not yet released) never exceed the
    budget, although ticks stay in flight after their jobs finish on the host."""
    from cellato_amd.serve import Server
    reqs = [Request(i, "game-of-life" if i % 2 else "wire", 64, 64, 4 + 4 * (i % 3), 0.0, i, 100)
            for i in range(40)]
    alive, peak = [0], [0]

    def allocate(r):
        alive[0] += r.bytes
        peak[0] = max(peak[0], alive[0])
        assert alive[0] <= 350, alive[0]
        return FakeGrid(r)

    def release(g):
        alive[0] -= g.bytes

    eng = FakeEngine(lag=3)
    srv = Server(hbm_budget=350, max_live=64, default_T=4, tick_passes=1, allocate=allocate, release=release,
                 engine=eng)
    rep = srv.run(reqs, clock=lambda: 1.0, sleep=lambda s: None)
    assert eng.start_called and len(rep.completions) == 40
    assert alive[0] == 0 and peak[0] == 300 and rep.peak_bytes <= 350
    # completion order follows the ticks: a job finishes in the tick of its last pass
    assert all(rep.completions[r.id] >= r.steps // 4 for r in reqs)
