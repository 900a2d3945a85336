"""Host logic of the serving loop (admission, continuous batching, completion
accounting) on a stand-in ensemble with a simulated clock."""
import numpy as np
import pytest

from physicell_amd import scenarios as S
from physicell_amd.serving import Job, Server, estimate_bytes, poisson_jobs


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, d):
        self.t += max(0.0, d)


class FakeEnsemble:
    """Counts the periods each member ran; one batched period costs 1 ms."""

    def __init__(self, clock):
        self.clock = clock
        self.live = {}
        self.next = 0
        self.max_live = 0
        self.log = []

    def add(self, spec, cells):
        mid = self.next
        self.next += 1
        self.live[mid] = 0
        self.max_live = max(self.max_live, len(self.live))
        self.log.append(("add", mid, self.clock()))
        return mid

    def remove(self, mid):
        del self.live[mid]
        self.log.append(("remove", mid, self.clock()))

    def run_periods(self, k, use_graph=True):
        for m in self.live:
            self.live[m] += k
        self.clock.t += 1e-3 * k

    def wait(self):
        pass

    def density(self, mid):
        return np.array([self.live[mid]], np.float64)

    def field(self, mid, f):
        return np.array([float(self.live[mid])])


def tiny_jobs(n, gap, periods):
    sc, c = S.config0()
    return [Job(i, sc, c, p, i * gap) for i, p in zip(range(n), periods)]


def test_every_job_runs_exactly_its_steps():
    clk = Clock()
    jobs = tiny_jobs(6, 0.002, [5, 1, 8, 3, 3, 10])
    fake = FakeEnsemble(clk)
    srv = Server(lambda: fake, budget_bytes=1e12, max_members=64, clock=clk, sleep=clk.sleep,
                 bytes_of=lambda j: 1)
    res, st = srv.run(jobs)
    assert sorted(res) == list(range(6))
    for j in jobs:
        assert res[j.id].density[0] == j.periods
        assert res[j.id].admitted >= j.arrival
        assert res[j.id].finished > res[j.id].admitted
    assert st.agent_steps == sum(j.cells.n * j.periods for j in jobs)
    assert st.admitted_while_busy >= 1            # arrivals joined running jobs
    assert not fake.live


def test_slots_and_budget_limit_admission():
    clk = Clock()
    jobs = tiny_jobs(8, 0.0, [4] * 8)
    fake = FakeEnsemble(clk)
    srv = Server(lambda: fake, budget_bytes=1e12, max_members=3, clock=clk, sleep=clk.sleep,
                 bytes_of=lambda j: 1)
    res, st = srv.run(jobs)
    assert fake.max_live == 3 and st.max_live == 3
    clk2 = Clock()
    fake2 = FakeEnsemble(clk2)
    srv2 = Server(lambda: fake2, budget_bytes=250, max_members=64, clock=clk2, sleep=clk2.sleep,
                  bytes_of=lambda j: 100)
    res2, st2 = srv2.run(jobs)
    assert fake2.max_live == 2 and st2.max_bytes == 200
    # arrival order is kept: admissions happen in job order
    adds = [m for op, m, _ in fake2.log if op == "add"]
    assert adds == sorted(adds)
    assert all(res2[j.id].density[0] == 4 for j in jobs)


def test_oversized_job_is_refused():
    clk = Clock()
    srv = Server(lambda: FakeEnsemble(clk), budget_bytes=10, clock=clk, sleep=clk.sleep, bytes_of=lambda j: 11)
    with pytest.raises(ValueError, match="more than the whole budget"):
        srv.run(tiny_jobs(1, 0.0, [1]))


def test_idle_server_waits_for_arrivals():
    clk = Clock()
    fake = FakeEnsemble(clk)
    jobs = tiny_jobs(2, 5.0, [2, 2])
    srv = Server(lambda: fake, clock=clk, sleep=clk.sleep, bytes_of=lambda j: 1)
    res, st = srv.run(jobs)
    assert res[1].admitted >= 5.0
    assert st.periods_run == 4


def test_stream_draws_and_estimate():
    jobs = poisson_jobs(3, 4.0, seed=1000, member_kw={"edges": (18, 24), "cells": (100, 300)})
    again = poisson_jobs(3, 4.0, seed=1000, member_kw={"edges": (18, 24), "cells": (100, 300)})
    assert [j.periods for j in jobs] == [j.periods for j in again]
    assert all(500 <= j.periods <= 5000 for j in jobs)
    assert all(b.arrival > a.arrival for a, b in zip(jobs, jobs[1:]))
    j = jobs[0]
    assert estimate_bytes(j.spec, j.cells.n, True) < estimate_bytes(j.spec, j.cells.n, False)
