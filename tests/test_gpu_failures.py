"""Failure handling forced from the host: the HBM budget refuses a batch or a
server request (a status comes back, nothing crashes, the context keeps
working), a server is destroyed with requests still queued and in flight,
and chunk/mesh handles outlive the cache entries and batches they came
from.  Every kernel in these tests runs to completion normally; the
failures are allocation refusals decided on the host."""
import gc

import numpy as np
import pytest

import voxel_amd as V
from test_gpu_parity import C2, assert_mesh_equal, batch_vs_oracle, oterrain, say

pytestmark = pytest.mark.gpu


def context_bytes():
    """Device bytes a fresh context holds (tables, terrain slots)."""
    with V.Context(0) as p:
        return p.memory()[0]


def test_batch_over_budget_is_refused_and_context_recovers(oracle_mod):
    """A 4096-chunk batch under a 16 MiB budget: vx_batch_create returns
    VX_ERR_BUDGET, frees what it had taken, and a batch that fits then runs
    bit-exact on the same context."""
    say("batch over budget")
    budget = 16 << 20
    with V.Context(0, hbm_budget=budget) as c:
        used0, _ = c.memory()
        coords = [(x, 0, z) for x in range(64) for z in range(64)]
        with pytest.raises(V.VoxelError) as e:
            V.Batch(c, coords, 32, c.terrain(C2), V.grid_neighbours(64, 64),
                    formats=V.FMT_VERTICES)
        assert e.value.status == V.ERR_BUDGET
        used1, b1 = c.memory()
        assert used1 == used0 and b1 == budget
        b = batch_vs_oracle(c, oracle_mod, [(0, 0, 0), (0, 0, 1), (1, 0, 0), (1, 0, 1)], [32] * 4,
                            [C2] * 4, V.grid_neighbours(2, 2))
        b.close()
        assert c.memory()[0] <= budget


def test_server_request_beyond_budget_fails_with_status(oracle_mod):
    """A 64^3 request whose mesh (24,132 faces, 2.5 MB) cannot fit in the
    budget even with the cache empty completes with VX_ERR_BUDGET (counted
    in stats.failed); 16^3 requests on the same server then complete and
    match the oracle."""
    say("server request beyond budget")
    budget = context_bytes() + (1 << 20)  # the 64^3 store (~0.3 MiB) fits, its mesh does not
    with V.Context(0, hbm_budget=budget) as c:
        tid = c.terrain(C2)
        with V.Server(c, max_batch=8, max_inflight=2) as s:
            (t_big,) = s.submit([((0, 0, 0), 64, tid, V.NB_NONE)])
            s.drain()
            comps = s.poll()
            assert [(cm.ticket, cm.status) for cm in comps] == [(t_big, V.ERR_BUDGET)]
            assert s.stats().failed == 1
            small = [((x, 0, 0), 16, tid, V.NB_NONE) for x in range(3)]
            tickets = s.submit(small)
            s.drain()
            comps = {cm.ticket: cm for cm in s.poll()}
            assert all(comps[t].status == V.OK for t in tickets)
            assert s.stats().peak_resident_bytes <= budget
            chunk, mesh = s.lookup((2, 0, 0), 16, tid)
            ids = oracle_mod.chunk_fill(oterrain(oracle_mod, C2), 2, 0, 0, 16)
            assert (chunk.ids() == ids).all()
            assert_mesh_equal(oracle_mod, mesh, ids, 16, (2, 0, 0), (None,) * 4)


def test_server_destroyed_with_work_in_flight(ctx, oracle_mod):
    """Destroying a server with queued and in-flight requests waits for its
    streams, releases every flight and cache entry, and leaves the context
    usable with no memory held."""
    say("server destroyed in flight")
    tid = ctx.terrain(C2)
    gc.collect()  # handles left by earlier tests release their blocks now
    used0, _ = ctx.memory()
    s = V.Server(ctx, max_batch=16, max_inflight=3)
    s.submit([((x, 0, z), 32, tid, V.NB_NONE) for x in range(12) for z in range(12)])
    s.pump(0)  # admits the first batches onto the device
    s.close()
    gc.collect()
    assert ctx.memory()[0] <= used0
    b = batch_vs_oracle(ctx, oracle_mod, [(3, 0, 4)], [32], [C2], np.full((1, 4), -1, np.int32))
    b.close()


def test_handles_outlive_eviction_and_batch(oracle_mod):
    """A chunk and mesh looked up from the server stay valid after the
    cache evicts their entry (refcounted views keep the device block), and a
    chunk taken from a batch stays valid after the batch is closed."""
    say("handles outlive eviction")
    budget = context_bytes() + (24 << 20)
    with V.Context(0, hbm_budget=budget) as c:
        tid = c.terrain(C2)
        t = oterrain(oracle_mod, C2)
        with V.Server(c, max_batch=16, max_inflight=2) as s:
            s.submit([((0, 0, 0), 32, tid, V.NB_NONE)])
            s.drain()
            s.poll()
            chunk, mesh = s.lookup((0, 0, 0), 32, tid)
            for k in range(8):  # enough 32^3 requests to cycle the cache several times
                s.submit([((x, 0, 10 + k), 32, tid, V.NB_NONE) for x in range(40)])
                s.drain()
                s.poll()
            st = s.stats()
            assert st.evictions > 0
            assert s.lookup((0, 0, 0), 32, tid) is None  # evicted from the cache
            ids = oracle_mod.chunk_fill(t, 0, 0, 0, 32)
            assert (chunk.ids() == ids).all()
            assert_mesh_equal(oracle_mod, mesh, ids, 32, (0, 0, 0), (None,) * 4)
            chunk.release()
            mesh.release()
        b = V.Batch(c, [(5, 0, 5), (5, 0, 6)], 32, tid, V.grid_neighbours(1, 2),
                    formats=V.FMT_VERTICES)
        b.run()
        b.wait()
        ch = b.chunk(1)
        b.close()
        assert (ch.ids() == oracle_mod.chunk_fill(t, 5, 0, 6, 32)).all()
        ch.release()


def test_arena_growth_refused_then_recovers(oracle_mod):
    """A batch created with room for 16 faces overflows its mesh arena; the
    grown arena is refused by the budget (a ballast buffer holds the rest),
    so vx_batch_wait returns VX_ERR_BUDGET and the batch holds no readable
    mesh.  A run with the ballast still held is refused the same way; once
    the ballast is freed the next run grows the arena, re-captures its graph
    and matches the oracle byte for byte."""
    say("arena growth refused by the budget")
    budget = 48 << 20
    coords = [(x, 0, z) for x in range(4) for z in range(4)]
    nb = V.grid_neighbours(4, 4)
    with V.Context(0, hbm_budget=budget) as c:
        tid = c.terrain(C2)
        b = V.Batch(c, coords, 32, tid, nb, formats=V.FMT_VERTICES, face_capacity=16)
        # 16 chunks of this terrain hold about 44,000 faces (4.6 MB of
        # vertices and indices): hold all but 1-2 MiB of the budget in
        # 1 MiB blocks (a size class of the heap, so nothing is rounded up)
        ballast = []
        while budget - c.memory()[0] > (2 << 20):
            ballast.append(V.DeviceBuffer(c, 1 << 20))
        for attempt in range(2):
            b.run()
            with pytest.raises(V.VoxelError) as e:
                b.wait()
            assert e.value.status == V.ERR_BUDGET, attempt
            with pytest.raises(V.VoxelError) as e:
                b.chunk_faces(0)
            assert e.value.status == V.ERR_NOT_READY
        for blk in ballast:
            blk.free()
        b.run()
        total = b.wait()
        assert total > 16
        t = oterrain(oracle_mod, C2)
        ids = [oracle_mod.chunk_fill(t, *cc, 32) for cc in coords]
        verts, idxs = b.mesh(V.FMT_VERTICES), b.mesh(V.FMT_INDICES)
        run = 0
        for i, cc in enumerate(coords):
            adj = [ids[j] if j >= 0 else None for j in nb[i]]
            faces, _, vert, idx = oracle_mod.mesh_buffers(ids[i], 32, tuple(v * 32 for v in cc), adj)
            first, cnt = b.chunk_faces(i)
            assert (first, cnt) == (run, len(faces))
            assert (verts[4 * first:4 * (first + cnt)].view(np.uint32) == vert.view(np.uint32)).all()
            assert (idxs[6 * first:6 * (first + cnt)] == idx).all()
            run += cnt
        assert run == total
        b.close()
