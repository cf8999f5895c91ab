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
