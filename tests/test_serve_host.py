"""Host logic of the serving loop (the C ABI of include/ccfd_serve.h,
ccfd_amd/csrc/serve.cpp, driven by ccfd_amd/serve.py) with stand-in arenas
reached through callbacks and a simulated clock: FIFO admission within the
HBM budget, continuous batching (admission while others run), retirement
on convergence or the step cap, precision promotion, latency accounting,
argument checks. Runs without a GPU."""
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
    assert s["requests"] == 60 and s["solves"] == 60
    assert s["failed_step_cap"] == s["failed_nonfinite"] == 0
    assert s["latency_p50_s"] <= s["latency_p99_s"] <= s["latency_max_s"]
    assert s["solves_per_s"] == s["retirements_per_s"] > 0


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
    # capped requests are failures: they do not count as solves
    s = S.summary(res)
    assert s["requests"] == 10 and s["solves"] == 8
    assert s["failed_step_cap"] == 2 and s["failed_nonfinite"] == 0
    assert math.isclose(s["solves_per_s"], 0.8 * s["retirements_per_s"])


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


class FloorBatch(FakeBatch):
    """main arena whose residual floors (like FP32 flux arithmetic)"""

    def __init__(self, cap, clock, floor):
        super().__init__(cap, clock)
        self.floor = floor

    def poll(self):
        out = super().poll()
        for k, i in self.inst.items():
            out[k]["res"][0] = max(out[k]["res"][0], self.floor)
        return out

    def cons(self, k):
        i = self.inst[k]
        tau = 20.0 + i["n"] / 50.0
        i["mesh"].level = max(math.exp(-i["steps"] / tau), self.floor)
        return np.ones((2, 4))

    def add(self, mesh, prim, cfl=0.9):
        k = super().add(mesh, prim, cfl)
        self.inst[k]["mesh"] = mesh
        return k


class PolishBatch(FakeBatch):
    """FP64 arena: continues from the level the main arena handed over"""

    def add(self, mesh, prim, cfl=0.9):
        k = super().add(mesh, prim, cfl)
        self.inst[k]["level"] = mesh.level
        return k

    def poll(self):
        out = {}
        for k, i in self.inst.items():
            r = i["level"] * math.exp(-i["steps"] / 30.0)
            out[k] = {"steps": i["steps"], "res": np.array([r, 0, 0, 0]),
                      "res0": np.array([i["level"], 0, 0, 0]),
                      "nonfinite": False}
        return out


def test_precision_promotion_reaches_the_target_below_the_floor():
    clock = Clock()
    main = FloorBatch(cap=20000, clock=clock, floor=3e-5)
    pol = PolishBatch(cap=8000, clock=clock)
    srv = S.Server(main, target_drop=1e-6, max_steps=100000,
                   steps_per_poll=25, prep_workers=2, polish=pol,
                   switch_drop=1e-4, stall_polls=4, clock=clock,
                   sleep=clock.sleep)
    res = srv.run(_requests(20, rate=200.0, seed=11))
    assert len(res) == 20
    for r in res:
        assert r["converged"] and r["drop"] <= 1e-6
        assert 0 < r["polish_steps"] < r["steps"]
    assert main.used == 0 and pol.used == 0
    assert pol.max_used <= pol.cap


def test_without_polish_the_floor_caps_the_solves():
    clock = Clock()
    main = FloorBatch(cap=20000, clock=clock, floor=3e-5)
    srv = S.Server(main, target_drop=1e-6, max_steps=3000,
                   steps_per_poll=25, prep_workers=2, clock=clock,
                   sleep=clock.sleep)
    res = srv.run(_requests(6, rate=200.0, seed=12))
    assert all(not r["converged"] and r["steps"] >= 3000 for r in res)


def test_c_abi_checks_its_arguments():
    """ccfd_serve_create refuses incomplete backends and bad settings; an
    empty server's round does nothing"""
    import ctypes as C
    L = S._lib()
    clock = Clock()
    fb = FakeBatch(cap=100, clock=clock)
    a = S._PyArena(fb, {})
    cfg = S.ServeCfg(1e-6, 100, 10, 0, 1e-4, 4, 1.4, 0, S._CLOCK(lambda _: 0.0),
                     None)
    h = C.c_void_p()
    empty = S.Backend()
    assert L.ccfd_serve_create(C.byref(h), C.byref(empty), None,
                               C.byref(cfg)) != 0
    bad = S.ServeCfg(1e-6, 100, 0, 0, 1e-4, 4, 1.4, 0, cfg.clock, None)
    assert L.ccfd_serve_create(C.byref(h), C.byref(a.be), None,
                               C.byref(bad)) != 0
    assert L.ccfd_serve_create(C.byref(h), C.byref(a.be), None,
                               C.byref(cfg)) == 0
    n = [C.c_int(-1) for _ in range(3)]
    assert L.ccfd_serve_round(h, *[C.byref(x) for x in n]) == 0
    assert [x.value for x in n] == [0, 0, 0]
    assert L.ccfd_serve_submit(h, 1, None, 10, None, 0.9) != 0
    out = np.zeros(4)
    assert L.ccfd_serve_take_state(h, 1, out.ctypes.data_as(
        C.POINTER(C.c_double))) != 0
    L.ccfd_serve_destroy(h)


def test_request_larger_than_an_empty_arena_is_an_error():
    clock = Clock()
    fb = FakeBatch(cap=1000, clock=clock)
    srv = S.Server(fb, steps_per_poll=20, prep_workers=1, clock=clock,
                   sleep=clock.sleep)
    import pytest
    with pytest.raises(ValueError, match="does not fit the empty arena"):
        srv.run([S.Request(0, 0.0, lambda: (FakeMesh(4000), 0, 0.9))])
