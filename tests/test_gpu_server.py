"""GPU tests of the continuous-batching server: resident results are
bit-exact vs the oracle, identical alone and in the mix, cache hits and
coalescing, LRU eviction under an HBM budget, and the streaming simulation."""
import sys

import numpy as np
import pytest

import voxel_amd as V
from test_gpu_parity import C2, assert_mesh_equal, oterrain

pytestmark = pytest.mark.gpu

HMAP = {e: V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=11, octaves=4, frequency=1 / 32,
                     height_scale=float(e), height_max=e - 1) for e in (16, 32, 64)}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def mesh_of(server, coord, edge, tid):
    got = server.lookup(coord, edge, tid)
    assert got is not None, f"{coord}/{edge}/terrain {tid} not resident"
    return got


def test_server_results_bit_exact(ctx, oracle_mod):
    say("server bit-exact")
    tid = ctx.terrain(C2)
    rng = np.random.default_rng(3)
    reqs = [((int(x), 0, int(z)), e, tid, V.NB_NONE)
            for (x, z), e in zip(rng.integers(-40, 40, size=(60, 2)), rng.choice([16, 32, 64], 60))]
    with V.Server(ctx, max_batch=16, max_inflight=3,
                  formats=V.FMT_VERTICES | V.FMT_INSTANCES) as s:
        tickets = s.submit(reqs)
        s.drain()
        comps = {c.ticket: c for c in s.poll()}
        assert sorted(comps) == sorted(tickets)
        seen = set()
        for (c, e, _, _), t in zip(reqs, tickets):
            comp = comps[t]
            assert comp.status == V.OK
            if (c, e) in seen:
                continue
            seen.add((c, e))
            chunk, mesh = mesh_of(s, c, e, tid)
            ids = oracle_mod.chunk_fill(oterrain(oracle_mod, C2), *c, e)
            assert (chunk.ids() == ids).all()
            assert comp.faces == mesh.faces
            assert_mesh_equal(oracle_mod, mesh, ids, e, c, (None,) * 4)


def test_same_alone_and_in_the_mix(ctx, oracle_mod):
    """A request's resident mesh is byte-identical when served alone and when
    admitted together with 300 other requests of mixed sizes."""
    say("server alone vs mix")
    tids = {e: ctx.terrain(HMAP[e]) for e in (16, 32, 64)}
    target = ((5, 0, -3), 32)
    with V.Server(ctx) as alone:
        alone.submit([(target[0], 32, tids[32], V.NB_NONE)])
        alone.drain()
        _, m_alone = mesh_of(alone, *target, tids[32])
        v_alone, i_alone = m_alone.vertices(), m_alone.indices()
    rng = np.random.default_rng(7)
    mix = [((int(x), 0, int(z)), int(e), tids[int(e)], V.NB_NONE)
           for (x, z), e in zip(rng.integers(-30, 30, size=(300, 2)), rng.choice([16, 32, 64], 300))]
    mix.insert(150, (target[0], 32, tids[32], V.NB_NONE))
    with V.Server(ctx, max_batch=64, max_inflight=4) as s:
        for k in range(0, len(mix), 37):  # arrivals interleaved with device work
            s.submit(mix[k:k + 37])
            s.pump(50)
        s.drain()
        _, m_mix = mesh_of(s, *target, tids[32])
        assert (m_mix.vertices().view(np.uint32) == v_alone.view(np.uint32)).all()
        assert (m_mix.indices() == i_alone).all()
        st = s.stats()
        assert st.batches > 1 and st.failed == 0


def test_resident_neighbour_snapshot(ctx, oracle_mod):
    """NB_RESIDENT meshes use the resident neighbours at admission; the
    snapshot is reported and reproduces the mesh alone."""
    say("server neighbour snapshot")
    tid = ctx.terrain(C2)
    t = oterrain(oracle_mod, C2)
    with V.Server(ctx, max_batch=8) as s:
        s.submit([((10, 0, 10), 32, tid, V.NB_NONE), ((11, 0, 10), 32, tid, V.NB_NONE)])
        s.drain()
        s.poll()
        s.submit([((10, 0, 11), 32, tid, V.NB_RESIDENT)])
        s.drain()
        (comp,) = s.poll()
        assert list(comp.nb_present) == [0, 0, 1, 0]  # only -z (10,0,10) was resident
        _, mesh = mesh_of(s, (10, 0, 11), 32, tid)
        ids = oracle_mod.chunk_fill(t, 10, 0, 11, 32)
        adj = (None, None, oracle_mod.chunk_fill(t, 10, 0, 10, 32), None)
        assert_mesh_equal(oracle_mod, mesh, ids, 32, (10, 0, 11), adj)


def test_two_terrains_same_coordinate(ctx, oracle_mod):
    """One coordinate and edge under two terrains is two chunks: neither is
    served from the other's cache entry or flight, and a resident-neighbour
    snapshot only takes neighbours of the request's own terrain."""
    say("server two terrains at one coordinate")
    hm = HMAP[32]
    ta, tb = ctx.terrain(C2), ctx.terrain(hm)
    oa, ob = oterrain(oracle_mod, C2), oterrain(oracle_mod, hm)
    with V.Server(ctx, max_batch=8) as s:
        # same flight: the two requests must not coalesce
        s.submit([((3, 0, 4), 32, ta, V.NB_NONE), ((3, 0, 4), 32, tb, V.NB_NONE)])
        s.drain()
        assert s.stats().coalesced == 0 and s.stats().generated == 2
        # later requests: each is a cache hit on its own terrain's entry
        s.submit([((3, 0, 4), 32, tb, V.NB_NONE), ((3, 0, 4), 32, ta, V.NB_NONE)])
        s.pump(0)
        s.drain()
        # (4,0,4) is resident under terrain a only: a terrain-b request next
        # to it must not take it as its -z neighbour
        s.submit([((4, 0, 4), 32, ta, V.NB_NONE)])
        s.drain()
        s.poll()
        s.submit([((3, 0, 5), 32, tb, V.NB_RESIDENT), ((4, 0, 5), 32, tb, V.NB_RESIDENT)])
        s.drain()
        comps = {((c.coord.x, c.coord.y, c.coord.z), c.edge): c for c in s.poll()}
        assert list(comps[((3, 0, 5), 32)].nb_present) == [0, 0, 1, 0]  # b's (3,0,4)
        assert list(comps[((4, 0, 5), 32)].nb_present) == [0, 0, 0, 0]  # a's (4,0,4) is not b's
        for tid, ot in ((ta, oa), (tb, ob)):
            chunk, mesh = mesh_of(s, (3, 0, 4), 32, tid)
            ids = oracle_mod.chunk_fill(ot, 3, 0, 4, 32)
            assert (chunk.ids() == ids).all(), f"terrain {tid}"
            assert_mesh_equal(oracle_mod, mesh, ids, 32, (3, 0, 4), (None,) * 4)
        _, mesh = mesh_of(s, (3, 0, 5), 32, tb)
        adj = (None, None, oracle_mod.chunk_fill(ob, 3, 0, 4, 32), None)
        assert_mesh_equal(oracle_mod, mesh, oracle_mod.chunk_fill(ob, 3, 0, 5, 32), 32,
                          (3, 0, 5), adj)
        _, mesh = mesh_of(s, (4, 0, 5), 32, tb)
        assert_mesh_equal(oracle_mod, mesh, oracle_mod.chunk_fill(ob, 4, 0, 5, 32), 32,
                          (4, 0, 5), (None,) * 4)
        assert s.lookup((4, 0, 4), 32, tb) is None
        st = s.stats()
        assert st.cache_hits == 2 and st.failed == 0


def test_cache_hits_and_coalescing(ctx):
    say("server hits/coalescing")
    tid = ctx.terrain(C2)
    with V.Server(ctx) as s:
        t1 = s.submit([((0, 0, 0), 32, tid, V.NB_NONE)] * 3)  # duplicates in one submission
        s.drain()
        c1 = s.poll()
        assert len(c1) == 3 and len({c.faces for c in c1}) == 1
        st = s.stats()
        assert st.generated == 1 and st.coalesced == 2
        s.submit([((0, 0, 0), 32, tid, V.NB_NONE)])
        s.pump(0)
        (c2,) = s.poll()
        assert c2.cache_hit == 1 and c2.faces == c1[0].faces
        assert len(set(t1)) == 3


def test_lru_eviction_under_budget(oracle_mod):
    say("server budget")
    budget = 96 << 20
    with V.Context(0, hbm_budget=budget) as c:
        tid = c.terrain(C2)
        with V.Server(c, max_batch=32, max_inflight=2) as s:
            reqs = [((x, 0, z), 32, tid, V.NB_NONE) for x in range(20) for z in range(20)]
            for k in range(0, len(reqs), 40):
                s.submit(reqs[k:k + 40])
                s.pump(100)
            s.drain()
            comps = s.poll()
            assert len(comps) == 400 and all(cm.status == V.OK for cm in comps)
            st = s.stats()
            assert st.evictions > 0
            assert st.peak_resident_bytes <= budget
            # the most recent chunk is still resident and correct
            chunk, mesh = s.lookup((19, 0, 19), 32, tid)
            ids = oracle_mod.chunk_fill(oterrain(oracle_mod, C2), 19, 0, 19, 32)
            assert (chunk.ids() == ids).all()
            assert_mesh_equal(oracle_mod, mesh, ids, 32, (19, 0, 19), (None,) * 4)


def test_streaming_simulation_smoke(ctx):
    say("server simulation")
    tids = [ctx.terrain(HMAP[e]) for e in (16, 32, 64)]
    r = V.simulate(ctx, seed=3, rate=2000, duration=1.0,
                   classes=((16, 1.0, tids[0]), (32, 1.0, tids[1]), (64, 1.0, tids[2])))
    assert r.completed == r.requests > 1000
    assert r.generated > 0 and r.p99_ms > 0
    assert r.p50_ms <= r.p99_ms <= r.max_ms


def test_streaming_under_binding_budget(oracle_mod):
    """Continuous batching at 20,000 req/s for 0.4 s (a camera walk, mixed
    16/32/64 sizes) under a 128 MiB budget that the working set (about
    0.6 GB) overflows, so LRU eviction runs throughout.  Every request
    completes OK; a completion's chunk is still resident when the client
    polls it after the pump that posted it (admissions in that pump do not
    evict it); the meshes of a sample of misses match the oracle; one
    request's mesh equals its mesh when served alone; the budget is never
    exceeded."""
    import time
    say("server under a binding budget")
    budget = 128 << 20
    rate, duration = 20000.0, 0.4
    rng = np.random.default_rng(3)
    with V.Context(0, hbm_budget=budget) as c:
        tids = {e: c.terrain(HMAP[e]) for e in (16, 32, 64)}
        target = ((3, 0, -2), 32)
        with V.Server(c, max_batch=64, max_inflight=4) as alone:
            alone.submit([(target[0], 32, tids[32], V.NB_NONE)])
            alone.drain()
            _, m = mesh_of(alone, *target, tids[32])
            v_alone, i_alone = m.vertices(), m.indices()
            m.release()
        checked, target_seen = 0, False
        with V.Server(c, max_batch=64, max_inflight=4) as s:
            t0, sent, cx, cz, heading = time.perf_counter(), 0, 0.0, 0.0, 0.0
            meta, target_ticket = {}, None
            while True:
                el = time.perf_counter() - t0
                due = min(int(el * rate), int(duration * rate)) - sent
                if due > 0:
                    heading += rng.normal(0, 0.05)
                    cx, cz = cx + 0.5 * np.cos(heading), cz + 0.5 * np.sin(heading)
                    reqs = []
                    for _ in range(due):
                        r, a = 16 * 32 * np.sqrt(rng.random()), 2 * np.pi * rng.random()
                        e = int(rng.choice([16, 32, 64]))
                        coord = (int(np.floor((cx + r * np.cos(a)) / e)), 0,
                                 int(np.floor((cz + r * np.sin(a)) / e)))
                        reqs.append((coord, e, tids[e], V.NB_NONE))
                    if target_ticket is None and sent > rate * duration / 2:
                        reqs.append((target[0], 32, tids[32], V.NB_NONE))
                    tickets = s.submit(reqs)
                    for tk, rq in zip(tickets, reqs):
                        meta[tk] = rq
                    if target_ticket is None and sent > rate * duration / 2:
                        target_ticket = tickets[-1]
                    sent += due
                s.pump(0)
                for comp in s.poll():
                    assert comp.status == V.OK, comp.status
                    coord, e, tid, _ = meta[comp.ticket]
                    if comp.ticket == target_ticket:
                        ch, m = mesh_of(s, coord, e, tid)  # looked up before any further pump
                        assert (m.vertices().view(np.uint32) == v_alone.view(np.uint32)).all()
                        assert (m.indices() == i_alone).all()
                        ch.release()
                        m.release()
                        target_seen = True
                    elif not comp.cache_hit and checked < 8 and comp.ticket % 97 == 0:
                        ch, m = mesh_of(s, coord, e, tid)
                        ids = oracle_mod.chunk_fill(oterrain(oracle_mod, HMAP[e]), *coord, e)
                        assert (ch.ids() == ids).all()
                        assert_mesh_equal(oracle_mod, m, ids, e, coord, (None,) * 4)
                        ch.release()
                        m.release()
                        checked += 1
                if sent >= int(duration * rate) and s.stats().completed == s.stats().submitted:
                    break
            st = s.stats()
            assert st.evictions > 0, "the budget did not bind"
            assert st.failed == 0 and st.completed == st.submitted == sent + 1
            assert st.peak_resident_bytes <= budget
        assert target_seen and checked >= 3
