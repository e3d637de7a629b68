"""The CPU restatement of World + ChunkGenerator (oracle/world.py) against
what the reference's code fixes on scripted runs (src/chunk_builder.rs,
src/world.rs): ring order and bounds, one load per tick, the primary ring
on entering a chunk, arrival-order rebuilds, sorted de-duplicated rebuild
queue, the unload box, the result delay.  CPU only."""
import numpy as np
import pytest

ALL = np.zeros(24, np.float32)  # every plane 0: dist = 0 is never < 0, all chunks pass


def small_terrain(O):
    # air everywhere: chunk fill and meshing stay trivial, scheduling is what is tested
    return O.terrain(kind=0, seed=0, octaves=1, frequency=1 / 32, lacunarity=2.0, gain=0.5,
                     height_scale=0, height_max=0, y_scale=0)


def ring_order(px, pz, rd):
    """enqueue_chunks_in_frustum's visiting order (chunk_builder.rs:214-253)."""
    out = []
    for r in range(2, rd):
        for z in range(-r, r + 1):
            out += [(px + r, 0, pz + z), (px - r, 0, pz + z)]
        for x in range(-r + 1, r):
            out += [(px + x, 0, pz + r), (px + x, 0, pz - r)]
    return out


def test_stationary_player_loads_rings_one_per_tick(oracle_mod):
    import world as W
    m = W.WorldModel(small_terrain(oracle_mod), (16.0, 40.0, 16.0), render_distance=5,
                     result_delay=1)
    order = ring_order(0, 0, 5)
    assert len(order) == 8 * (2 + 3 + 4)
    m.tick((16.0, 40.0, 16.0), ALL)
    # no primary loads: the player has not entered a new chunk (player.rs:24-27)
    assert m.jobs_loaded == 1 and m.in_process == [order[0]] and m.load_queue == []
    for k in range(1, len(order) + 1):
        m.tick((16.0, 40.0, 16.0), ALL)
    assert sorted(m.chunks) == sorted(order)
    assert m.jobs_loaded == len(order)
    # arrival order: the k-th chunk arrives at tick k (delay 1), one per tick
    assert m.received_chunks == len(order)


def test_entering_a_chunk_loads_the_primary_ring_in_order(oracle_mod):
    import world as W
    m = W.WorldModel(small_terrain(oracle_mod), (16.0, 40.0, 16.0), render_distance=2,
                     result_delay=3)
    m.tick((48.0, 40.0, 16.0), ALL)  # into chunk x = 1
    assert m.player_chunk == (1, 1, 0)
    expect = [(1, 0, 0), (2, 0, -1), (0, 0, -1), (2, 0, 0), (0, 0, 0), (2, 0, 1), (0, 0, 1),
              (1, 0, 1), (1, 0, -1)]
    assert m.in_process == expect
    # results arrive one per tick from tick 3 on (delay 3): none yet
    m.tick((48.0, 40.0, 16.0), ALL)
    m.tick((48.0, 40.0, 16.0), ALL)
    assert m.chunks == {}
    m.tick((48.0, 40.0, 16.0), ALL)
    assert list(m.chunks) == [expect[0]]


def test_arrivals_queue_sorted_deduplicated_rebuilds(oracle_mod):
    import world as W
    m = W.WorldModel(small_terrain(oracle_mod), (16.0, 40.0, 16.0), render_distance=2,
                     result_delay=1)
    m.tick((48.0, 40.0, 16.0), ALL)  # 9 primary jobs
    seen = []
    for _ in range(9):
        m.tick((48.0, 40.0, 16.0), ALL)
        q = m.rebuild_queue
        assert q == sorted(set(q))          # clean_up_queues: sort + dedup
        seen.append(list(q))
    # the 2nd arrival (2,0,-1) has no resident neighbour yet except none: (1,0,0) is not adjacent
    assert seen[1] == []
    # the 4th arrival (2,0,0) touches (1,0,0) and (2,0,-1): both queued
    assert seen[3] == [(1, 0, 0), (2, 0, -1)]
    # nothing rebuilds while loads are in process (chunk_builder.rs:95)
    assert m.jobs_rebuilt == 0
    for _ in range(40):
        m.tick((48.0, 40.0, 16.0), ALL)
    assert m.rebuild_queue == [] and m.jobs_rebuilt > 0


def test_unload_box_and_stale_rebuilds(oracle_mod):
    import world as W
    m = W.WorldModel(small_terrain(oracle_mod), (16.0, 40.0, 16.0), render_distance=2,
                     result_delay=0)
    m.tick((48.0, 40.0, 16.0), ALL)
    for _ in range(12):
        m.tick((48.0, 40.0, 16.0), ALL)
    assert len(m.chunks) == 9
    # jump 10 chunks in x: every chunk leaves the +-2 box (chunk_builder.rs:334-352)
    m.tick((48.0 + 320.0, 40.0, 16.0), ALL)
    assert all(abs(p[0] - 11) <= 2 for p in m.chunks)
    assert all(abs(p[0] - 11) <= 2 for p in m.meshes)  # their meshes went with them
    assert m.unloaded == 9


def test_delay_zero_receives_in_the_same_tick(oracle_mod):
    import world as W
    m = W.WorldModel(small_terrain(oracle_mod), (16.0, 40.0, 16.0), render_distance=3,
                     result_delay=0)
    m.tick((16.0, 40.0, 16.0), ALL)
    assert m.received_chunks == 1 and m.in_process == []


def test_chunk_coordinate_conversion():
    """to_chunk_coord: floor then Rust's saturating `as i32`, then >> 5
    (floor division by 32 for in-range values, clamped at the i32 range)."""
    import world as W
    assert W.to_chunk_coord((0.0, 31.9, 32.0)) == (0, 0, 1)
    assert W.to_chunk_coord((-0.5, -32.0, -32.5)) == (-1, -1, -2)
    assert W.to_chunk_coord((3e9, -3e9, float("nan"))) == ((2 ** 31 - 1) >> 5, -(2 ** 26), 0)
