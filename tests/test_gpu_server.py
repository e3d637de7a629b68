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


def mesh_of(server, coord, edge):
    got = server.lookup(coord, edge)
    assert got is not None, f"{coord}/{edge} not resident"
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
            chunk, mesh = mesh_of(s, c, e)
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
        _, m_alone = mesh_of(alone, *target)
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
        _, m_mix = mesh_of(s, *target)
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
        _, mesh = mesh_of(s, (10, 0, 11), 32)
        ids = oracle_mod.chunk_fill(t, 10, 0, 11, 32)
        adj = (None, None, oracle_mod.chunk_fill(t, 10, 0, 10, 32), None)
        assert_mesh_equal(oracle_mod, mesh, ids, 32, (10, 0, 11), adj)


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
            chunk, mesh = s.lookup((19, 0, 19), 32)
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
