"""Parity at the configured full sizes: every chunk's face count against the
oracle's multithreaded driver, byte-exact meshes for a seeded sample, the
batch's prefix structure, and full-size cull masks."""
import os

import numpy as np
import pytest

import voxel_amd as V
from test_gpu_parity import C5, bits, oterrain, say

pytestmark = pytest.mark.gpu

C2 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0,
               gain=0.5, y_scale=1 / 128)
THREADS = min(16, os.cpu_count() or 1)


def check_sample(O, b, t, coords, edges, nb, idx, verts, idxs):
    """Byte-exact vertices and indices of chunks `idx` (with their in-batch
    neighbours) against the oracle."""
    for i in idx:
        e = int(edges[i])
        ids = O.chunk_fill(t, *map(int, coords[i]), e)
        adj = [O.chunk_fill(t, *map(int, coords[j]), e) if j >= 0 else None for j in nb[i]]
        faces, _, vert, ix = O.mesh_buffers(ids, e, tuple(int(v) * e for v in coords[i]), adj)
        first, cnt = b.chunk_faces(int(i))
        assert cnt == len(faces), i
        assert (bits(verts[4 * first:4 * (first + cnt)]) == bits(vert)).all(), i
        assert (idxs[6 * first:6 * (first + cnt)] == ix).all(), i


@pytest.mark.parametrize("mode", [V.BATCH_STAGED, V.BATCH_FUSED, V.BATCH_PIPELINED, V.BATCH_SPLIT,
                                  "autotuned"])
def test_config2_full_grid(ctx, oracle_mod, mode):
    """4096 chunks of 32^3 (64 x 64 grid, seed 7): every chunk's face count
    equals the oracle's, bases are the exclusive prefix in request order, 16
    seeded chunks match byte for byte with oracle meshes built one by one,
    and the whole vertex and index buffers (11.2 M faces, 1.17 GB) equal the
    oracle driver's, laid out at the same bases.  "autotuned": the path the
    benchmark times, vx_batch_autotune and then VX_BATCH_AUTO."""
    O = oracle_mod
    say("config 2 full", mode)
    coords = np.array([(x, 0, z) for x in range(64) for z in range(64)], np.int32)
    nb = V.grid_neighbours(64, 64)
    b = V.Batch(ctx, coords, 32, ctx.terrain(C2), nb, formats=V.FMT_VERTICES)
    if mode == "autotuned":
        b.run()
        b.wait()
        b.autotune(2)
    else:
        b.set_mode(mode)
    if mode == V.BATCH_PIPELINED:
        b.set_grid(4 * 256)  # persistent: four workgroups per CU
    b.run()
    total = b.wait()
    t = oterrain(O, C2)
    otot, ofaces = O.batch_run(t, coords, nb, 32, THREADS)
    got = np.array([b.chunk_faces(i)[1] for i in range(len(coords))])
    firsts = np.array([b.chunk_faces(i)[0] for i in range(len(coords))])
    assert total == otot == int(got.sum())
    assert (got == ofaces).all()
    assert (firsts == np.concatenate([[0], np.cumsum(got)[:-1]])).all()
    rng = np.random.default_rng(22)
    verts, idxs = b.mesh(V.FMT_VERTICES), b.mesh(V.FMT_INDICES)
    check_sample(O, b, t, coords, [32] * len(coords), nb, rng.choice(len(coords), 16, replace=False),
                 verts, idxs)
    _, _, overt, oidx = O.batch_mesh(t, coords, nb, 32, THREADS, firsts, total)
    assert np.array_equal(bits(verts), bits(overt))
    assert np.array_equal(idxs, oidx)


@pytest.mark.parametrize("mode", [V.BATCH_STAGED, V.BATCH_FUSED, "autotuned"])
def test_config3_full_ragged(ctx, oracle_mod, mode):
    """2048 chunks, edges uniform over {16, 32, 64}, classes 30% air / 10%
    solid / 60% 4-octave heightmap surface (one seeded stream, seed 11), no
    neighbours: every chunk's faces byte for byte, prefix order.  The
    15.0 M faces exceed the default arena (4 N^2 per chunk, 14.5 M), so the
    run also exercises the grow-and-re-emit path after either form.
    "autotuned": vx_batch_autotune on the fresh batch (its candidates
    outgrow the arena too), then VX_BATCH_AUTO."""
    O = oracle_mod
    say("config 3 full", mode)
    rng = np.random.default_rng(11)
    n = 2048
    edges = rng.choice([16, 32, 64], size=n)
    cls = rng.choice(3, size=n, p=[0.3, 0.1, 0.6])
    terrains = {}
    tids = []
    for e, k in zip(edges, cls):
        key = (int(k), int(e))
        if key not in terrains:
            if k == 0:
                tv = V.Terrain(kind=V.TERRAIN_AIR)
            elif k == 1:
                tv = V.Terrain(kind=V.TERRAIN_SOLID)
            else:
                tv = V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=11, octaves=4, frequency=1 / 32,
                               height_scale=float(e), height_max=int(e) - 1)
            terrains[key] = (tv, ctx.terrain(tv))
        tids.append(terrains[key][1])
    coords = np.array([(i % 64, 0, i // 64) for i in range(n)], np.int32)
    nb = np.full((n, 4), -1, np.int32)
    b = V.Batch(ctx, coords, edges, tids, nb, formats=V.FMT_VERTICES)
    if mode == "autotuned":
        b.autotune(2)
    else:
        b.set_mode(mode)
    b.run()
    total = b.wait()
    assert total == 15027984
    got = np.array([b.chunk_faces(i)[1] for i in range(n)])
    firsts = np.array([b.chunk_faces(i)[0] for i in range(n)])
    assert total == int(got.sum())
    assert (firsts == np.concatenate([[0], np.cumsum(got)[:-1]])).all()
    # per-class closed forms: air 0 faces, solid without neighbours 6 N^2
    for i in range(n):
        if cls[i] == 0:
            assert got[i] == 0
        elif cls[i] == 1:
            assert got[i] == 6 * edges[i] ** 2
    # every surface and solid chunk byte for byte against the oracle (air
    # chunks have no faces, checked above): the whole 15.0 M-face mesh
    verts, idxs = b.mesh(V.FMT_VERTICES), b.mesh(V.FMT_INDICES)
    vbits = bits(verts)
    for i in np.nonzero(cls != 0)[0]:
        t = oterrain(O, terrains[(int(cls[i]), int(edges[i]))][0])
        e = int(edges[i])
        ids = O.chunk_fill(t, *map(int, coords[i]), e)
        faces, _, vert, ix = O.mesh_buffers(ids, e, tuple(int(v) * e for v in coords[i]))
        assert got[i] == len(faces), i
        f0 = int(firsts[i])
        assert np.array_equal(vbits[20 * f0:20 * (f0 + got[i])], bits(vert)), i
        assert np.array_equal(idxs[6 * f0:6 * (f0 + got[i])], ix), i


def test_config5_full_resident_grid_and_cull(ctx, oracle_mod):
    """65,536 resident chunks (256 x 256, seed 5): face counts of all chunks
    against the oracle; the whole mesh (183 M faces: 14.7 GB of vertices,
    past every 32-bit byte offset, and 4.4 GB of indices) byte for byte
    against the oracle driver's meshes at the same bases, compared in slices;
    and the cull mask of every AABB for three cameras."""
    import camera as T
    O = oracle_mod
    say("config 5 full")
    side = 256
    coords = np.array([(x, 0, z) for x in range(side) for z in range(side)], np.int32)
    nb = V.grid_neighbours(side, side)
    b = V.Batch(ctx, coords, 32, ctx.terrain(C5), nb, formats=V.FMT_VERTICES,
                face_capacity=200_000_000)
    b.run()
    total = b.wait()
    otot, ofaces = O.batch_run(oterrain(O, C5), coords, nb, 32, THREADS)
    got = np.array([b.chunk_faces(i)[1] for i in range(len(coords))])
    assert total == otot and (got == ofaces).all()
    firsts = np.array([b.chunk_faces(i)[0] for i in range(len(coords))])
    assert (firsts == np.concatenate([[0], np.cumsum(got)[:-1]])).all()
    _, _, overt, oidx = O.batch_mesh(oterrain(O, C5), coords, nb, 32, THREADS, firsts, total)
    step = 1 << 28
    gv, ov = bits(b.mesh(V.FMT_VERTICES)), bits(overt)
    assert len(gv) == len(ov) == 20 * total
    for k in range(0, len(gv), step):
        assert np.array_equal(gv[k:k + step], ov[k:k + step]), k // 20
    del gv, ov, overt
    gi = b.mesh(V.FMT_INDICES)
    assert len(gi) == len(oidx) == 6 * total
    for k in range(0, len(gi), step):
        assert np.array_equal(gi[k:k + step], oidx[k:k + step]), k // 6
    del gi, oidx
    cb = V.DeviceBuffer.of(ctx, coords)
    mb = V.DeviceBuffer(ctx, len(coords))
    rng = np.random.default_rng(5)
    for _ in range(3):
        eye = np.array([rng.uniform(0, 8192), 60.0, rng.uniform(0, 8192)])
        d = np.array([np.cos(rng.uniform(0, 6.3)), -0.2, np.sin(rng.uniform(0, 6.3))])
        f = V.Frustum(T.colmajor(T.perspective(np.radians(70), 16 / 9, 0.1, 1024) @
                                 T.look_at(eye, eye + d)))
        ctx.check(ctx.L.vx_frustum_mask_device(ctx.h, V.C.byref(f.c), cb.ptr, len(coords), 32,
                                               mb.ptr), "vx_frustum_mask_device")
        got_m = mb.download(np.uint8, len(coords))
        ref_m = O.frustum_contains(f.planes, coords, 32)
        assert (got_m == ref_m).all() and got_m.sum() > 0
    # The frame: 512 new chunks (columns x = 256, 257) meshed beside the
    # resident grid -- their -x neighbours are the resident column x = 255,
    # +-z in the same batch -- with the cull of all 65,536 resident AABBs in
    # the same run, in every form; every new chunk byte for byte.
    say("config 5 frame at full size")
    t5 = oterrain(O, C5)
    ext = [b.chunk((side - 1) * side + z) for z in range(side)]
    nc = [(side + dx, 0, z) for dx in range(2) for z in range(side)]
    fnb = []
    for dx in range(2):
        for z in range(side):
            i = dx * side + z
            fnb.append([(-2 - z) if dx == 0 else (i - side), (i + side) if dx == 0 else -1,
                        i - 1 if z > 0 else -1, i + 1 if z < side - 1 else -1])
    ids = {c: O.chunk_fill(t5, *c, 32, fast=True) for c in nc}
    ids.update({(side - 1, 0, z): O.chunk_fill(t5, side - 1, 0, z, 32, fast=True)
                for z in range(side)})
    eye = np.array([8000.0, 60.0, 4000.0])
    f = V.Frustum(T.colmajor(T.perspective(np.radians(70), 16 / 9, 0.1, 1024) @
                             T.look_at(eye, eye + np.array([-1.0, -0.2, 0.3]))))
    ref_m = O.frustum_contains(f.planes, coords, 32)
    for mode in (V.BATCH_STAGED, V.BATCH_FUSED, V.BATCH_PIPELINED):
        B = V.Batch(ctx, nc, 32, ctx.terrain(C5), fnb, formats=V.FMT_VERTICES, ext=ext)
        B.set_mode(mode)
        B.set_cull(f, cb, len(coords), 32, mb)
        B.run()
        ftot = B.wait()
        assert (mb.download(np.uint8, len(coords)) == ref_m).all(), mode
        verts, idxs = B.mesh(V.FMT_VERTICES), B.mesh(V.FMT_INDICES)
        run = 0
        for i, c in enumerate(nc):
            adj = [ids.get((c[0] - 1, 0, c[2])), ids.get((c[0] + 1, 0, c[2])),
                   ids.get((c[0], 0, c[2] - 1)), ids.get((c[0], 0, c[2] + 1))]
            _, _, vert, ix = O.mesh_buffers(ids[c], 32, (c[0] * 32, 0, c[2] * 32), adj)
            first, cnt = B.chunk_faces(i)
            assert (first, cnt) == (run, len(ix) // 6), (mode, i)
            assert (bits(verts[4 * first:4 * (first + cnt)]) == bits(vert)).all(), (mode, i)
            assert (idxs[6 * first:6 * (first + cnt)] == ix).all(), (mode, i)
            run += cnt
        assert run == ftot
        B.close()
    for c in ext:
        c.release()
    cb.free()
    mb.free()
