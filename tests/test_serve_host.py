"""Host logic of the serving loop (ccfd_amd/serve.py) with a stand-in batch
and a simulated clock: FIFO admission within the HBM budget, continuous
batching (admission while others run), retirement on convergence or the
step cap, latency accounting. Runs without a GPU."""
import math

import numpy as np

from ccfd_amd import serve as S


class FakeMesh:
    def __init__(self, n):
        self.n_elems = n


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeBatch:
    """Capacity in elements; the residual of an instance falls as
    exp(-steps / tau) with tau set by its size; a step costs simulated time
    in proportion to the active elements."""

    def __init__(self, cap, clock, stall=()):
        self.cap, self.clock, self.stall = cap, clock, set(stall)
        self.used, self.inst, self.next = 0, {}, 0
        self.max_used, self.max_active, self.admit_order = 0, 0, []

    def fits(self, mesh):
        return self.used + mesh.n_elems <= self.cap

    def add(self, mesh, prim, cfl=0.9):
        assert self.fits(mesh)
        k = self.next
        self.next += 1
        self.inst[k] = {"n": mesh.n_elems, "steps": 0, "rid": prim}
        self.used += mesh.n_elems
        self.max_used = max(self.max_used, self.used)
        self.max_active = max(self.max_active, len(self.inst))
        self.admit_order.append(prim)
        return k

    def remove(self, k):
        self.used -= self.inst.pop(k)["n"]

    def step(self, n):
        cells = sum(i["n"] for i in self.inst.values())
        self.clock.t += n * cells * 1e-7
        for i in self.inst.values():
            i["steps"] += n

    def poll(self):
        out = {}
        for k, i in self.inst.items():
            tau = 20.0 + i["n"] / 50.0
            drop = 1.0 if i["rid"] in self.stall else math.exp(-i["steps"] / tau)
            out[k] = {"steps": i["steps"], "res": np.array([drop, 0, 0, 0]),
                      "res0": np.array([1.0, 0, 0, 0]), "nonfinite": False}
        return out

    def prim(self, k):
        return None


def _requests(n, rate, seed=1):
    rng = np.random.default_rng(seed)
    arr = S.poisson_arrivals(n, rate, seed)
    sizes = np.exp(rng.uniform(np.log(500), np.log(5000), n)).astype(int)
    reqs = []
    for j in range(n):
        size = int(sizes[j])
        reqs.append(S.Request(j, arr[j],
                              lambda size=size, j=j: (FakeMesh(size), j, 0.9),
                              {"size": size}))
    return reqs


def test_all_requests_complete_within_budget_in_fifo_order():
    clock = Clock()
    fb = FakeBatch(cap=12000, clock=clock)
    srv = S.Server(fb, target_drop=1e-6, max_steps=100000, steps_per_poll=25,
                   prep_workers=2, clock=clock, sleep=clock.sleep)
    reqs = _requests(60, rate=40.0)
    res = srv.run(reqs)
    assert [r["id"] for r in res] == list(range(60))
    assert fb.max_used <= fb.cap
    assert fb.admit_order == sorted(fb.admit_order)       # FIFO admission
    assert fb.max_active > 1                              # requests overlap
    for r in res:
        assert r["arrival"] <= r["admitted"] <= r["done"]
        assert r["converged"] and r["drop"] <= 1e-6
        assert math.isclose(r["latency"], r["done"] - r["arrival"])
    s = S.summary(res)
    assert s["solves"] == 60 and s["converged"] == 60
    assert s["latency_p50_s"] <= s["latency_p99_s"] <= s["latency_max_s"]
    assert s["solves_per_s"] > 0


def test_step_cap_retires_stalled_solves():
    clock = Clock()
    fb = FakeBatch(cap=20000, clock=clock, stall={3, 7})
    srv = S.Server(fb, target_drop=1e-6, max_steps=5000, steps_per_poll=50,
                   prep_workers=1, clock=clock, sleep=clock.sleep)
    res = srv.run(_requests(10, rate=100.0, seed=3))
    assert len(res) == 10
    stalled = [r for r in res if r["id"] in (3, 7)]
    assert all(not r["converged"] and r["steps"] >= 5000 for r in stalled)
    assert all(r["converged"] for r in res if r["id"] not in (3, 7))
    assert fb.used == 0 and not fb.inst


def test_request_larger_than_the_arena_waits_for_room():
    clock = Clock()
    fb = FakeBatch(cap=6000, clock=clock)
    srv = S.Server(fb, steps_per_poll=20, prep_workers=1, clock=clock,
                   sleep=clock.sleep)
    reqs = [S.Request(0, 0.0, lambda: (FakeMesh(4000), 0, 0.9)),
            S.Request(1, 0.001, lambda: (FakeMesh(4000), 1, 0.9)),
            S.Request(2, 0.002, lambda: (FakeMesh(1000), 2, 0.9))]
    res = srv.run(reqs)
    # the second cannot join the first (4000 + 4000 > 6000); FIFO keeps the
    # third behind it
    assert res[1]["admitted"] >= res[0]["done"]
    assert res[2]["admitted"] >= res[1]["admitted"]
    assert fb.max_used <= 6000


def test_poisson_arrivals_have_the_requested_rate():
    a = S.poisson_arrivals(20000, 50.0, seed=5)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 50.0) < 2.0


def test_one_at_a_time_mode_never_overlaps():
    clock = Clock()
    fb = FakeBatch(cap=10 ** 6, clock=clock)
    srv = S.Server(fb, steps_per_poll=25, prep_workers=2, max_active=1,
                   clock=clock, sleep=clock.sleep)
    res = srv.run(_requests(12, rate=1000.0, seed=7))
    assert fb.max_active == 1
    order = sorted(res, key=lambda r: r["admitted"])
    for a, b in zip(order, order[1:]):
        assert b["admitted"] >= a["done"]
