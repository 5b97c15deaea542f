"""Batched explicit solves (include/ccfd_batch.h) on the GPU.

Every instance of a batch runs the single-mesh operator over its own
blocks with its own time step and residual, so its results must be
bit-identical to a single-mesh context, alone and in any mix of other
instances, eager or replayed from a captured graph."""
import numpy as np
import pytest

from ccfd_amd import batch as BT
from ccfd_amd import gpu as G
from ccfd_amd import mesh as M
from oracle import bindings as B
from tests import cases

pytestmark = pytest.mark.gpu

MB = 1 << 20


def _scheme(prec=G.FP64, flux=5, lim=1, nrk=3, torder=2, order=2):
    return G.params(flux=flux, order=order, limiter=lim, venk_k=5.0, nrk=nrk,
                    torder=torder, precision=prec)


def _cases(ns=(14, 22, 31), seed0=4):
    out = []
    for k, n in enumerate(ns):
        mesh = cases.delaunay_mesh(n, seed=seed0 + k, periodic=False,
                                   bc_code=401)
        out.append((mesh, BT.riemann4_state(mesh.bary, seed=100 + k)))
    return out


@pytest.mark.parametrize("prec", [G.FP64, G.FP32_FLUX], ids=["fp64", "fp32"])
@pytest.mark.parametrize("variant", [G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE],
                         ids=["split", "fused", "fused_face"])
def test_instances_equal_single_contexts(prec, variant):
    sch = _scheme(prec)
    cs = _cases()
    b = BT.Batch(sch, max_sides=3, hbm_bytes=64 * MB)
    b.set_variant(variant)
    assert b.variant() == variant
    ids = [b.add(m, p) for m, p in cs]
    b.step(7)          # three graph replays of two steps + one eager step
    b.sync()
    for k, (m, p) in zip(ids, cs):
        s = G.Solver(m, sch)
        s.set_variant(variant)
        s.set_prim(p)
        t, r = s.advance(7, 0.0, 1e300, 1e300)
        inf = b.info(k)
        print("instance %d: %d cells t %.6e res %s" % (k, m.n_elems, inf["t"],
                                                       inf["res"]))
        assert inf["steps"] == 7
        assert inf["t"] == t
        assert np.array_equal(b.cons(k), s.cons())
        assert np.array_equal(b.prim(k), s.prim())
        assert np.array_equal(inf["res"], r)
        s.close()
    b.close()


@pytest.mark.parametrize("variant", [G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE],
                         ids=["split", "fused", "fused_face"])
def test_four_sided_arena_mixes_triangle_and_mixed_meshes(variant):
    """a maxSides=4 arena holds a mixed triangle/quadrilateral mesh next to
    an all-triangle one (padded to four slots); both stay bit-identical to
    their single-mesh contexts, which lay the triangle mesh out with three
    slots"""
    sch = _scheme(G.FP64)
    mixed, _ = cases.mixed_mesh()
    tm, tp = _cases((16,))[0]
    mp = cases.random_state(mixed.n_elems, seed=23, mach=0.5)
    b = BT.Batch(sch, max_sides=4, hbm_bytes=64 * MB)
    b.set_variant(variant)
    assert b.variant() == variant
    ids = [b.add(mixed, mp), b.add(tm, tp)]
    b.step(5)
    b.sync()
    for k, (m, p) in zip(ids, [(mixed, mp), (tm, tp)]):
        s = G.Solver(m, sch)
        s.set_variant(variant)
        s.set_prim(p)
        t, r = s.advance(5, 0.0, 1e300, 1e300)
        inf = b.info(k)
        print("instance %d: %d cells t %.6e" % (k, m.n_elems, inf["t"]),
              flush=True)
        assert inf["t"] == t
        assert np.array_equal(b.cons(k), s.cons())
        assert np.array_equal(inf["res"], r)
        s.close()
    b.close()


@pytest.mark.parametrize("order", [1, 2])
def test_first_and_second_order_against_oracle(order):
    """FP64 batch instances against the oracle's explicit steps."""
    sch = _scheme(G.FP64, order=order)
    cs = _cases((12, 19))
    b = BT.Batch(sch, max_sides=3, hbm_bytes=32 * MB)
    ids = [b.add(m, p) for m, p in cs]
    b.step(3)
    for k, (m, p) in zip(ids, cs):
        orc = B.Oracle(m, B.make_cfg(flux=5, order=order, limiter=1, nrk=3,
                                     torder=2))
        orc.set_prim(p)
        t = 0.0
        for _ in range(3):
            dt, _, _ = orc.calc_dt(t, 1e300, 1e300)
            orc.explicit_step(t, dt)
            t += dt
        assert b.info(k)["t"] == t
        assert np.array_equal(b.cons(k), orc.cons())
    b.close()


def test_same_alone_and_in_a_changing_mix():
    """Continuous batching: an instance admitted while others run, with
    removals and admissions around it, ends bit-identical to the same solve
    alone."""
    sch = _scheme(G.FP32_FLUX)
    cs = _cases((20, 12, 28, 17))
    tm, tp = cs[0]
    alone = BT.Batch(sch, hbm_bytes=64 * MB)
    ka = alone.add(tm, tp)
    alone.step(10)
    mix = BT.Batch(sch, hbm_bytes=64 * MB)
    others = [mix.add(m, p) for m, p in cs[1:3]]
    mix.step(3)
    kb = mix.add(tm, tp)
    mix.step(4)
    mix.remove(others[0])
    mix.add(*cs[3])
    mix.step(6)
    assert mix.info(kb)["steps"] == 10
    assert np.array_equal(mix.cons(kb), alone.cons(ka))
    assert np.array_equal(mix.info(kb)["res"], alone.info(ka)["res"])
    alone.close()
    mix.close()


def test_graph_replays_equal_eager_steps():
    sch = _scheme(G.FP64)
    cs = _cases()
    g = BT.Batch(sch, hbm_bytes=64 * MB)
    e = BT.Batch(sch, hbm_bytes=64 * MB)
    e.set_graph(False)
    ids = [(g.add(m, p), e.add(m, p)) for m, p in cs]
    for rep in range(4):
        g.step(4)
        e.step(4)
        g.sync()
        e.sync()
        for kg, ke in ids:
            assert np.array_equal(g.cons(kg), e.cons(ke)), rep
    g.close()
    e.close()


def test_admission_is_bounded_by_the_budget():
    sch = _scheme(G.FP64)
    mesh, prim = _cases((30,))[0]     # ~1.7k cells
    b = BT.Batch(sch, hbm_bytes=6 * MB, max_instances=64)
    need = b.bytes_for(mesh)
    assert need > 0
    ids = []
    while b.fits(mesh):
        ids.append(b.add(mesh, prim))
        assert len(ids) < 64
    assert len(ids) >= 2
    with pytest.raises(BT.NoRoom):
        b.add(mesh, prim)
    used, cap, n = b.usage()
    assert n == len(ids) and used <= cap
    b.remove(ids[1])
    assert b.fits(mesh)
    k = b.add(mesh, prim)
    b.step(2)
    assert b.info(k)["steps"] == 2
    assert b.info(ids[0])["steps"] == 2
    b.close()


def test_autotune_keeps_the_state():
    sch = _scheme(G.FP32_FLUX)
    cs = _cases()
    b = BT.Batch(sch, hbm_bytes=64 * MB)
    ids = [b.add(m, p) for m, p in cs]
    b.step(2)
    before = [b.cons(k) for k in ids]
    info0 = [b.info(k) for k in ids]
    v, ms = b.autotune(4)
    print("autotune: variant %d, ms per step %s" % (v, ms))
    assert v in (G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE)
    assert b.variant() == v
    assert (ms[1:] > 0).all()
    for k, c, i0 in zip(ids, before, info0):
        assert np.array_equal(b.cons(k), c)
        assert b.info(k)["steps"] == i0["steps"]
        assert b.info(k)["t"] == i0["t"]
    b.close()


def test_local_time_stepping_reaches_a_steady_state():
    """Channel with a bump at Mach 0.5, AUSMD + VK, local time steps at CFL
    2: the residual drops by 1e-6, and the oracle's operator on the final
    state agrees that it is steady."""
    raw = M.channel_bump(48, 16)
    bcs = [M.BC(101), M.BC(501, rho=1.0, mach=0.5, alpha=0.0, pressure=1.0),
           M.BC(801, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    prim = np.tile([1.0, 0.5 * c, 0.0, 1.0], (mesh.n_elems, 1))
    sch = G.params(flux=9, order=2, limiter=2, venk_k=5.0, nrk=4, torder=2)
    b = BT.Batch(sch, max_sides=4, local_dt=True, hbm_bytes=32 * MB)
    k = b.add(mesh, prim, cfl=2.0)
    drop = 1.0
    for it in range(40):
        b.step(500)
        inf = b.info(k)
        drop = inf["res"][0] / inf["res0"][0]
        if drop < 1e-6:
            break
    print("steps %d, density residual drop %.2e" % (inf["steps"], drop))
    assert drop < 1e-6
    orc = B.Oracle(mesh, B.make_cfg(flux=9, order=2, limiter=2, venk_k=5.0))
    orc.set_prim(b.prim(k))
    ut = orc.fvtd(0.0)
    r = orc.residual(ut)
    orc.set_prim(prim)
    r0 = orc.residual(orc.fvtd(0.0))
    print("oracle residual on the final state: %.2e of the initial" %
          (r[0] / r0[0]))
    assert r[0] / r0[0] < 1e-5
    b.close()


def test_wide_halo_instance_switches_to_the_split_operator():
    """An instance whose block halo exceeds the LDS tile is admitted; the
    batch runs the split operator while it is present, with unchanged
    results."""
    sch = _scheme(G.FP64)
    cs = _cases((16, 24))
    b = BT.Batch(sch, hbm_bytes=32 * MB, max_halo=16)
    ref = BT.Batch(sch, hbm_bytes=32 * MB)
    ids = [b.add(m, p) for m, p in cs]
    rids = [ref.add(m, p) for m, p in cs]
    assert b.variant() == G.KV_SPLIT
    assert ref.variant() == G.KV_FUSED_FACE
    b.step(4)
    ref.step(4)
    for k, r in zip(ids, rids):
        assert np.array_equal(b.cons(k), ref.cons(r))
    b.close()
    ref.close()


@pytest.mark.parametrize("sides,prec", [(3, G.FP32_FLUX), (4, G.FP32_FLUX),
                                        (3, G.FP64)])
def test_arena_capacity_fits_32_bit_indices(sides, prec):
    """the kernels index the [4][S][E] side-state arrays with 32-bit ints:
    however large the budget, an arena holds at most E elements with
    4 S E < 2^31 (a 64 GiB FP32 triangle arena would otherwise hold 214M
    elements, whose split-operator indices wrap)"""
    b = BT.Batch(G.params(flux=5, order=2, limiter=1, precision=prec),
                 max_sides=sides, hbm_bytes=64 << 30, max_instances=8)
    _, cap, _ = b.usage()
    b.close()
    print("capacity %d elements" % cap)
    assert cap > 0 and 4 * sides * cap < 2 ** 31


def test_split_steps_after_a_freed_context_match_a_fresh_arena():
    """a context that ran and was freed does not change what an arena
    computes afterwards, on every operator implementation the autotuner
    times (the split operator's side states included)"""
    mesh = cases.delaunay_mesh(60, seed=3, periodic=False)
    prim = BT.riemann4_state(mesh.bary, 11)
    prm = G.params(flux=5, order=2, limiter=1, nrk=3, torder=2, cfl=0.9,
                   precision=G.FP32_FLUX)
    big = cases.delaunay_mesh(300, seed=4, periodic=True)
    s = G.Solver(big, G.params(flux=5, order=2, limiter=2, venk_k=5.0))
    s.set_prim(M.smooth_state(big.bary, seed=1))
    s.advance(3, 0.0, 1e300, 1e300)
    s.close()
    out = []
    for variant in (G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE):
        b = BT.Batch(prm, max_sides=3, hbm_bytes=64 << 30, max_instances=8)
        b.set_variant(variant)
        k = b.add(mesh, prim)
        b.step(20)
        out.append(b.prim(k))
        b.close()
    assert np.isfinite(out[0]).all()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
