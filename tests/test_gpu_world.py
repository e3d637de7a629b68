"""vx_world (the library's World + ChunkGenerator, jobs on the GPU) against
the CPU restatement oracle/world.py, tick for tick on a scripted walk:
resident chunks and meshes, rebuild queue, chunks in process, counters, and
every resident mesh's vertices bit for bit."""
import numpy as np
import pytest

import voxel_amd as V

C5 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=5, octaves=4, frequency=1 / 64, lacunarity=2.0,
               gain=0.5, y_scale=1 / 128)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def camera(pos, look):
    import camera as T
    pos = np.asarray(pos, np.float64)
    return T.colmajor(T.perspective(np.radians(70), 16 / 9, 0.1, 1024) @
                      T.look_at(pos, pos + np.asarray(look, np.float64)))


@pytest.mark.gpu
@pytest.mark.parametrize("delay", [0, 1, 3])
def test_world_matches_restatement_tick_for_tick(ctx, oracle_mod, delay):
    import world as W
    print("[case] world delay", delay, flush=True)
    t = oracle_mod.terrain(kind=3, seed=5, octaves=4, frequency=1 / 64, lacunarity=2.0, gain=0.5,
                           height_scale=0, height_max=0, y_scale=1 / 128)
    start = (16.0, 40.0, 16.0)
    gw = V.World(ctx, ctx.terrain(C5), start, render_distance=5, result_delay=delay)
    mw = W.WorldModel(t, start, render_distance=5, result_delay=delay)
    pos = np.array(start)
    # walk +x/+z; stand (the backlog of loads drains, one arrival per tick,
    # then rebuilds run: the reference rebuilds only while nothing loads);
    # turn; jump far (unloads); stand.  130 ticks.
    look = (7.0, -8.0, 3.5)
    for k in range(130):
        if k < 25:
            step = (7.0, 0.0, 3.0)
        elif k < 90:
            step = (0.0, 0.0, 0.0)
        elif k < 100:
            step = (-4.0, 0.0, 6.0)
        elif k == 100:
            step = (400.0, 0.0, 0.0)
        else:
            step = (0.0, 0.0, 0.0)
        if step[0] or step[2]:
            look = (step[0], -8.0, step[2] + 0.5)
        pos = pos + np.array(step)
        f = V.Frustum(camera(pos, look))
        gw.tick(pos.astype(np.float32), f)
        mw.tick(pos.astype(np.float32), f.planes)
        st, ms = gw.stats(), mw.state()
        assert gw.list(V.WORLD_CHUNKS) == ms["chunks"], k
        assert gw.list(V.WORLD_MESHES) == ms["meshes"], k
        assert gw.list(V.WORLD_REBUILD_QUEUE) == ms["rebuild_queue"], k
        assert gw.list(V.WORLD_IN_PROCESS) == ms["in_process"], k
        assert st["player_chunk"] == ms["player_chunk"], k
        for name in ("jobs_loaded", "jobs_rebuilt", "received_chunks", "received_meshes", "unloaded"):
            assert st[name] == ms[name], (k, name)
        assert st["tick"] == k + 1
        check_all = k % 15 == 14 or k == 129
        for c in ms["meshes"]:
            m = gw.mesh(c)
            assert m.faces == len(mw.meshes[c]), (k, c)
            if check_all:
                _, vert, _ = oracle_mod.expand(mw.meshes[c])
                assert (bits(m.vertices()) == bits(vert)).all(), (k, c)
            m.release()
    s = gw.stats()
    assert s["jobs_loaded"] > 80 and s["jobs_rebuilt"] > 20 and s["unloaded"] > 50
    gw.close()
