"""GPU parity: every output of the HIP path, called through the C ABI, is
compared with the CPU oracle on the same seeded inputs -- bit for bit
(block ids, occupancy, noise values, vertex/index buffers, InstanceRaw
records, cull masks)."""
import itertools
import sys

import numpy as np
import pytest

import voxel_amd as V

pytestmark = pytest.mark.gpu

C1 = V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=42, octaves=1, frequency=1 / 32, height_scale=16.0,
               height_max=15)
C2 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0,
               gain=0.5, y_scale=1 / 128)
C5 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=5, octaves=4, frequency=1 / 64, y_scale=1 / 128)
REF_NOISE = V.Terrain(kind=V.TERRAIN_DENSITY, seed=1, octaves=3, frequency=0.09, lacunarity=0.5,
                      gain=2.5, y_scale=0.0)  # the reference's own fbm parameters, threshold 0
# the reference's own rule and axis order (noise at (wx, wz, wy) < 0 -> grass,
# src/perlin_noise.rs:6-20, src/chunk.rs:38-42) with its parameters
REF_RULE = V.Terrain(kind=V.TERRAIN_REFERENCE, seed=1, octaves=3, frequency=0.09, lacunarity=0.5,
                     gain=2.5)
REF_RULE5 = V.Terrain(kind=V.TERRAIN_REFERENCE, seed=9, octaves=5, frequency=1 / 40,
                      lacunarity=2.0, gain=0.5)  # > 4 octaves: the octave-outer loop
# octave counts beyond the configs' 1, 3 and 4: 2 runs the octave-inner row
# loop, 5 and 8 (the maximum) the octave-outer loop; non-power-of-two gain
# and lacunarity
OCT2 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=23, octaves=2, frequency=1 / 40, lacunarity=1.9,
                 gain=0.6, y_scale=1 / 96)
OCT5 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=31, octaves=5, frequency=1 / 100, lacunarity=2.0,
                 gain=0.5, y_scale=1 / 128)
OCT8 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=47, octaves=8, frequency=1 / 300, lacunarity=2.1,
                 gain=0.45, y_scale=1 / 90)
# octave frequencies 0.3 .. 2.4 near the ends of the i32 world range: x and
# z lattice coordinates reach 5e9, past the range of an int conversion (their
# fractions are 0 there, so the noise comes from y and the cells' hashes)
BIGF = V.Terrain(kind=V.TERRAIN_DENSITY, seed=3, octaves=4, frequency=0.3, lacunarity=2.0,
                 gain=0.5, y_scale=1 / 128)
HMAPBIG = V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=9, octaves=4, frequency=1.0, lacunarity=2.0,
                    gain=0.5, height_scale=32.0, height_max=31)
HMAP8 = V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=5, octaves=8, frequency=1 / 50, lacunarity=1.8,
                  gain=0.55, height_scale=32.0, height_max=31)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def oterrain(O, t: V.Terrain):
    return O.terrain(kind=t.kind, seed=t.seed, octaves=t.octaves, frequency=t.frequency,
                     lacunarity=t.lacunarity, gain=t.gain, height_scale=t.height_scale,
                     height_max=t.height_max, y_scale=t.y_scale)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_mesh_equal(O, mesh, ids, edge, coord, adj_ids):
    origin = tuple(int(c) * edge for c in coord)
    faces, inst, vert, idx = O.mesh_buffers(ids, edge, origin, adj_ids)
    assert mesh.faces == len(faces)
    if mesh.formats & V.FMT_INSTANCES:
        got = mesh.instances()
        assert (bits(got["m"]) == bits(inst["m"])).all()
        assert (got["tex"] == inst["tex"]).all()
    if mesh.formats & V.FMT_VERTICES:
        gv = mesh.vertices()
        assert (bits(gv["pos"]) == bits(vert["pos"])).all()
        assert (bits(gv["uv"]) == bits(vert["uv"])).all()
        assert (mesh.indices() == idx).all()


# ------------------------------------------------------------------ noise
@pytest.mark.parametrize("terrain,coord,edge", [
    (C2, (5, 0, 9), 32), (C2, (-3, -1, 7), 32), (C2, (0, 0, 0), 16), (C2, (2, 1, -2), 64),
    (C5, (100, 0, -120), 32), (REF_NOISE, (3, 0, -4), 32), (C1, (3, 0, 1), 16),
    (V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=11, octaves=4, frequency=1 / 32,
               height_scale=64.0, height_max=63), (-2, 0, 6), 64),
    (OCT2, (1, 0, -1), 32), (OCT2, (4, -1, 2), 16), (OCT5, (-6, 0, 3), 32), (OCT5, (0, 0, 1), 64),
    (OCT8, (2, 0, 2), 32), (OCT8, (-1, 1, -9), 16), (HMAP8, (7, 0, -3), 32),
    (BIGF, (2 ** 26 - 2, 0, 1 - 2 ** 26), 32), (BIGF, (1 - 2 ** 26, 3, 2 ** 26 - 2), 32),
    (HMAPBIG, (2 ** 26 - 2, 0, 2 ** 26 - 2), 32),
    (REF_RULE, (0, 0, 0), 32), (REF_RULE, (-2, 1, 3), 16), (REF_RULE, (1, 0, -1), 64),
    (REF_RULE5, (3, -1, 2), 32),
])
def test_noise_bit_exact(ctx, oracle_mod, terrain, coord, edge):
    say("noise", terrain.kind, terrain.seed, coord, edge)
    got = ctx.noise(terrain, coord, edge)
    ref = oracle_mod.noise_block(oterrain(oracle_mod, terrain), *coord, edge)
    assert (bits(got) == bits(ref)).all()


# ------------------------------------------------------------------ fill
@pytest.mark.parametrize("terrain,coord,edge", [
    (C1, (0, 0, 0), 16), (C1, (3, 0, 2), 16), (C1, (1, 1, 1), 16),
    (C2, (5, 0, 9), 32), (C2, (63, 0, 63), 32), (C2, (-7, -1, 3), 32), (C2, (4, 0, 4), 16),
    (C2, (1, 0, 1), 64), (REF_NOISE, (0, 0, 0), 32),
    (V.Terrain(kind=V.TERRAIN_AIR), (0, 0, 0), 32), (V.Terrain(kind=V.TERRAIN_SOLID), (2, 0, 1), 64),
    (OCT2, (3, 0, 3), 32), (OCT5, (1, 0, -2), 32), (OCT8, (0, 0, 0), 64), (HMAP8, (-2, 0, 4), 16),
    (REF_RULE, (0, 0, 0), 32), (REF_RULE, (3, -1, 2), 64), (REF_RULE, (-5, 0, 4), 16),
    (REF_RULE5, (0, 0, 1), 32),
])
def test_fill_ids_and_occupancy(ctx, oracle_mod, terrain, coord, edge):
    say("fill", terrain.kind, terrain.seed, coord, edge)
    ch = V.Chunk.new(ctx, terrain, coord, edge)
    ref = oracle_mod.chunk_fill(oterrain(oracle_mod, terrain), *coord, edge)
    assert (ch.ids() == ref).all()
    assert (ch.occupancy() == oracle_mod.occupancy(ref, edge)).all()


# ------------------------------------------------------------------ mesh
def neighbour_sets(rng, edge):
    yield (None, None, None, None)
    yield tuple((rng.random(edge ** 3) < 0.5).astype(np.uint8) for _ in range(4))
    yield (np.ones(edge ** 3, np.uint8), None, np.zeros(edge ** 3, np.uint8), None)


@pytest.mark.parametrize("edge", [16, 32, 64])
def test_mesh_from_ids_random(ctx, oracle_mod, edge):
    rng = np.random.default_rng(edge)
    for density in (0.3, 0.7):
        ids = (rng.random(edge ** 3) < density).astype(np.uint8)
        for adj in neighbour_sets(rng, edge):
            say("mesh random", edge, density, [a is None for a in adj])
            coord = (int(rng.integers(-50, 50)), int(rng.integers(-3, 3)), int(rng.integers(-50, 50)))
            ch = V.Chunk.from_ids(ctx, coord, edge, ids)
            nbs = [None if a is None else V.Chunk.from_ids(ctx, (0, 0, 0), edge, a) for a in adj]
            mesh = ch.create_mesh(nbs)
            assert_mesh_equal(oracle_mod, mesh, ids, edge, coord, adj)


def test_mesh_mixed_block_types(ctx, oracle_mod):
    """Grass and dirt in one chunk: texture layers come from each voxel."""
    rng = np.random.default_rng(5)
    edge = 32
    ids = rng.integers(0, 3, size=edge ** 3).astype(np.uint8)
    ch = V.Chunk.from_ids(ctx, (1, 0, -1), edge, ids)
    assert_mesh_equal(oracle_mod, ch.create_mesh(), ids, edge, (1, 0, -1), (None,) * 4)
    dirt = np.where(ids > 0, 2, 0).astype(np.uint8)  # single non-grass type
    ch2 = V.Chunk.from_ids(ctx, (0, 0, 0), edge, dirt)
    assert_mesh_equal(oracle_mod, ch2.create_mesh(), dirt, edge, (0, 0, 0), (None,) * 4)


@pytest.mark.parametrize("edge", [16, 32, 64])
def test_mesh_closed_forms(ctx, oracle_mod, edge):
    say("closed forms", edge)
    solid = np.ones(edge ** 3, np.uint8)
    air = np.zeros(edge ** 3, np.uint8)
    s = V.Chunk.from_ids(ctx, (0, 0, 0), edge, solid)
    a = V.Chunk.from_ids(ctx, (0, 0, 0), edge, air)
    assert s.create_mesh().faces == 6 * edge * edge
    assert s.create_mesh([s, s, s, s]).faces == 2 * edge * edge
    assert a.create_mesh().faces == 0
    y, z, x = np.meshgrid(np.arange(edge), np.arange(edge), np.arange(edge), indexing="ij")
    checker = (((x + y + z) % 2) == 0).astype(np.uint8).reshape(-1)
    c = V.Chunk.from_ids(ctx, (2, -1, 3), edge, checker)
    m = c.create_mesh([c, c, c, c])
    assert m.faces == 3 * edge ** 3  # the worst case: more than one LDS tile of faces
    assert_mesh_equal(oracle_mod, m, checker, edge, (2, -1, 3), (checker,) * 4)


def test_single_voxels_at_every_corner(ctx, oracle_mod):
    edge = 16
    for x in (0, edge - 1):
        for y in (0, edge - 1):
            for z in (0, edge - 1):
                ids = np.zeros(edge ** 3, np.uint8)
                ids[x + edge * z + edge * edge * y] = 1
                ch = V.Chunk.from_ids(ctx, (-1, 0, 1), edge, ids)
                assert_mesh_equal(oracle_mod, ch.create_mesh(), ids, edge, (-1, 0, 1), (None,) * 4)


def test_faces_near_world_zero_bit_exact(ctx, oracle_mod):
    """Single voxels at world coordinates -1, 0 and +1 on every axis, each
    alone in a 16^3 chunk: InstanceRaw and vertex records bit-exact against
    the oracle and, for the non-zero matrix entries and every vertex
    position, against the independent textbook evaluation of quad.rs:14-54
    and vertex.vert:24 (the ~4.4e-8 rotation terms set the low bits here)."""
    from test_oracle import expected_vertices
    edge = 16
    for p in itertools.product((-1, 0, 1), repeat=3):
        coord = tuple(v // edge for v in p)
        local = tuple(v - c * edge for v, c in zip(p, coord))
        ids = np.zeros(edge ** 3, np.uint8)
        ids[local[0] + edge * local[2] + edge * edge * local[1]] = 1
        say("face bits near zero", p)
        mesh = V.Chunk.from_ids(ctx, coord, edge, ids).create_mesh()
        assert_mesh_equal(oracle_mod, mesh, ids, edge, coord, (None,) * 4)
        inst, vert = mesh.instances(), mesh.vertices()
        assert mesh.faces == 6
        for d in range(6):  # one voxel: faces in direction order L R B F D U
            R, t, ev = expected_vertices(d, p)
            E = np.zeros((4, 4), np.float32)
            E[:3, :3] = R.T
            E[3, :3] = t
            E[3, 3] = 1
            m = inst[d]["m"]
            nz = E != 0
            assert (bits(m[nz]) == bits(E[nz])).all() and (m[~nz] == 0).all(), (p, d)
            assert (bits(vert[4 * d:4 * d + 4]["pos"]) == bits(ev)).all(), (p, d)


def test_generated_chunk_mesh_with_generated_neighbours(ctx, oracle_mod):
    t = oterrain(oracle_mod, C2)
    c = (6, 0, 9)
    say("generated mesh", c)
    ch = V.Chunk.new(ctx, C2, c, 32)
    coords = [(c[0] - 1, 0, c[2]), (c[0] + 1, 0, c[2]), (c[0], 0, c[2] - 1), (c[0], 0, c[2] + 1)]
    nbs = [V.Chunk.new(ctx, C2, q, 32) for q in coords]
    ids = oracle_mod.chunk_fill(t, *c, 32)
    adj = [oracle_mod.chunk_fill(t, *q, 32) for q in coords]
    assert_mesh_equal(oracle_mod, ch.create_mesh(nbs), ids, 32, c, adj)
    # neighbour snapshot with some neighbours absent (chunk_builder.rs:134-169)
    assert_mesh_equal(oracle_mod, ch.create_mesh([nbs[0], None, None, nbs[3]]), ids, 32, c,
                      [adj[0], None, None, adj[3]])


# ------------------------------------------------------------------ batches
def apply_mode(b, mode):
    """mode: a VX_BATCH_* value or (mode, grid[, ratio]), grid = persistent
    workgroups, ratio = one pipelined mesher workgroup in `ratio`."""
    m, g, r = (tuple(mode) + (0,))[:3] if isinstance(mode, tuple) else (mode, 0, 0)
    if m in (V.BATCH_PIPELINED, V.BATCH_SPLIT) and int(np.max(b.edges)) > 32:
        with pytest.raises(V.VoxelError) as e:
            b.set_mode(m)
        assert e.value.status == V.ERR_INVALID
        pytest.skip("the pipelined and split forms take edges up to 32 (refused as specified)")
    b.set_mode(m)
    b.set_grid(g)
    b.set_roles(r)


def batch_vs_oracle(ctx, O, coords, edges, terrains, nb, formats=V.FMT_VERTICES | V.FMT_INSTANCES,
                    face_capacity=0, per_chunk=False, mode=V.BATCH_AUTO):
    tids = [ctx.terrain(t) for t in terrains]
    b = V.Batch(ctx, coords, edges, tids, nb, formats=formats, face_capacity=face_capacity)
    apply_mode(b, mode)
    return check_batch(b, O, coords, edges, terrains, nb, formats, per_chunk)


def check_batch(b, O, coords, edges, terrains, nb, formats, per_chunk=False):
    """Runs batch `b` and checks its ids and meshes against the oracle."""
    if per_chunk:
        b.run_per_chunk()
    else:
        b.run()
    total = b.wait()
    ids = [O.chunk_fill(oterrain(O, t), *map(int, c), int(e)) for c, e, t in zip(coords, edges, terrains)]
    verts = b.mesh(V.FMT_VERTICES) if formats & V.FMT_VERTICES else None
    idxs = b.mesh(V.FMT_INDICES) if formats & V.FMT_VERTICES else None
    insts = b.mesh(V.FMT_INSTANCES) if formats & V.FMT_INSTANCES else None
    run = 0
    for i in range(len(coords)):
        assert (b.ids(i) == ids[i]).all()
        adj = [ids[j] if j >= 0 else None for j in nb[i]]
        e = int(edges[i])
        faces, inst, vert, idx = O.mesh_buffers(ids[i], e, tuple(int(v) * e for v in coords[i]), adj)
        first, cnt = b.chunk_faces(i)
        assert first == run and cnt == len(faces)
        if verts is not None:
            assert (bits(verts[4 * first:4 * (first + cnt)]) == bits(vert)).all()
            assert (idxs[6 * first:6 * (first + cnt)] == idx).all()
        if insts is not None:
            assert (bits(insts[first:first + cnt]) == bits(inst)).all()
        run += cnt
    assert total == run
    return b


# fused with one workgroup per chunk; fused with 3 persistent workgroups that
# take every chunk ticket in turn; staged; pipelined with one workgroup per
# chunk, with 3 workgroups (two fill, one meshes from the start), with 4
# workgroups of which two mesh from the start, and with a single workgroup
# (a filler: it fills every chunk, then meshes them all); the staged kernels
# split over two streams
MODES = [pytest.param((V.BATCH_FUSED, 0), id="fused"),
         pytest.param((V.BATCH_FUSED, 3), id="fused_persistent3"),
         pytest.param((V.BATCH_STAGED, 0), id="staged"),
         pytest.param((V.BATCH_PIPELINED, 0), id="pipelined"),
         pytest.param((V.BATCH_PIPELINED, 3), id="pipelined_persistent3"),
         pytest.param((V.BATCH_PIPELINED, 4, 2), id="pipelined_persistent4_roles2"),
         pytest.param((V.BATCH_PIPELINED, 1), id="pipelined_single"),
         pytest.param((V.BATCH_SPLIT, 0), id="split")]


@pytest.mark.parametrize("mode", MODES)
def test_batch_config1_grid(ctx, oracle_mod, mode):
    """Config 1: 4x4 grid of 16^3 heightmap chunks, all in-grid neighbours."""
    say("batch C1", mode)
    coords = [(x, 0, z) for x in range(4) for z in range(4)]
    batch_vs_oracle(ctx, oracle_mod, coords, [16] * 16, [C1] * 16, V.grid_neighbours(4, 4),
                    mode=mode)


@pytest.mark.parametrize("mode,per_chunk", [(V.BATCH_FUSED, False), (V.BATCH_STAGED, False),
                                            (V.BATCH_STAGED, True), ((V.BATCH_FUSED, 1), False),
                                            (V.BATCH_PIPELINED, False),
                                            ((V.BATCH_PIPELINED, 7), False),
                                            ((V.BATCH_PIPELINED, 9, 4), False),
                                            ((V.BATCH_PIPELINED, 16, 8), False)])
def test_batch_config2_subgrid(ctx, oracle_mod, mode, per_chunk):
    """Config 2 terrain on an 8x8 sub-grid of 32^3 chunks."""
    say("batch C2 8x8", mode, per_chunk)
    coords = [(x + 20, 0, z + 30) for x in range(8) for z in range(8)]
    batch_vs_oracle(ctx, oracle_mod, coords, [32] * 64, [C2] * 64, V.grid_neighbours(8, 8),
                    per_chunk=per_chunk, mode=mode)


def test_batch_autotune_keeps_results(ctx, oracle_mod):
    """The autotuner times the staged, fused and pipelined forms (the last
    with one mesher workgroup in three) and keeps the fastest; the tuned
    batch's ids and meshes are the oracle's, byte for byte.  Role splits
    outside 2..8 are refused."""
    say("batch autotune")
    coords = [(x + 20, 0, z + 30) for x in range(8) for z in range(8)]
    nb = V.grid_neighbours(8, 8)
    fmt = V.FMT_VERTICES | V.FMT_INSTANCES
    b = V.Batch(ctx, coords, [32] * 64, [ctx.terrain(C2)] * 64, nb, formats=fmt)
    r = b.autotune(reps=2)
    kinds = {(c["mode"], c["ratio"]) for c in r["candidates"]}
    assert {("staged", 0), ("fused", 0), ("pipelined", 3)} <= kinds, r
    # 64 chunks: the fill and the emit split chunks, so staged tile shapes
    # other than the rule are among the candidates
    shapes = {tuple(c["splits"]) for c in r["candidates"] if c["mode"] == "staged"}
    assert (0, 0) in shapes and any(f for f, _ in shapes) and any(e for _, e in shapes), r
    assert r["mode"] in ("staged", "fused", "pipelined")
    chosen = [c for c in r["candidates"]
              if (c["mode"], c["grid"], c["ratio"], c["splits"]) ==
              (r["mode"], r["grid"], r["ratio"], r["splits"])]
    assert chosen and chosen[0]["ms"] == min(c["ms"] for c in r["candidates"]), r
    for bad in (1, 9, -1):
        with pytest.raises(V.VoxelError) as e:
            b.set_roles(bad)
        assert e.value.status == V.ERR_INVALID
    check_batch(b, oracle_mod, coords, [32] * 64, [C2] * 64, nb, fmt)
    b.close()


@pytest.mark.parametrize("splits", [(1, 1), (2, 4), (4, 2), (8, 3), (0, 1), (1, 0)])
@pytest.mark.parametrize("edge", [32, 64])
@pytest.mark.parametrize("terrain", [C2, C1], ids=["density", "heightmap"])
def test_batch_staged_tile_shapes(ctx, oracle_mod, splits, edge, terrain):
    """The staged form at set tile shapes (vx_batch_set_splits): fill
    workgroups per chunk splitting the density rows (a heightmap chunk is
    filled whole by the first of its workgroups), emit workgroups per item
    splitting its tiles, on a 3x3 grid with in-grid neighbours; ids and
    meshes byte for byte (at edge 32 a fill split of 8 uses 4 of them).
    Values outside 1, 2, 4, 8 and 1..4 are refused."""
    say("batch tile shapes", splits, edge, terrain.kind)
    coords = [(x - 1, 0, z + 2) for x in range(3) for z in range(3)]
    nb = V.grid_neighbours(3, 3)
    fmt = V.FMT_VERTICES | V.FMT_INSTANCES
    b = V.Batch(ctx, coords, edge, ctx.terrain(terrain), nb, formats=fmt)
    b.set_mode(V.BATCH_STAGED)
    for bad in ((3, 0), (16, 0), (-1, 0), (0, 5), (0, -1)):
        with pytest.raises(V.VoxelError) as e:
            b.set_splits(*bad)
        assert e.value.status == V.ERR_INVALID
    b.set_splits(*splits)
    check_batch(b, oracle_mod, coords, [edge] * 9, [terrain] * 9, nb, fmt)
    b.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("edge", [16, 64])
def test_batch_density_grid_other_edges(ctx, oracle_mod, mode, edge):
    """3-D density with in-grid neighbours at edges 16 and 64 (halo lines of
    every width), including negative coordinates."""
    say("batch density grid", mode, edge)
    coords = [(x - 2, -1 if edge == 16 else 0, z - 1) for x in range(3) for z in range(3)]
    batch_vs_oracle(ctx, oracle_mod, coords, [edge] * 9, [C5] * 9, V.grid_neighbours(3, 3),
                    mode=mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("terrain", [pytest.param(OCT2, id="oct2"), pytest.param(OCT5, id="oct5"),
                                     pytest.param(OCT8, id="oct8"), pytest.param(HMAP8, id="hmap8")])
def test_batch_octave_counts(ctx, oracle_mod, mode, terrain):
    """3x3 grids of 32^3 chunks with 2, 5 and 8 octaves (density) and 8
    octaves (heightmap): fill, halo recompute and meshes byte for byte."""
    say("batch octaves", terrain.octaves, terrain.kind, mode)
    coords = [(x + 1, 0, z - 2) for x in range(3) for z in range(3)]
    batch_vs_oracle(ctx, oracle_mod, coords, [32] * 9, [terrain] * 9, V.grid_neighbours(3, 3),
                    mode=mode)


@pytest.mark.parametrize("n", [40, 200, 600])
def test_batch_rows_split_over_workgroups(ctx, oracle_mod, n):
    """Staged batches of fewer than 1024 chunks fill each chunk with 8 (at
    most 128 chunks, 64^3 chunks present: edge 64 in 8 parts, smaller edges
    in 4), 4 (n < 512) or 2 workgroups: density terrains of 1 to 4 octaves
    split the chunk's rows among them, other terrains are filled whole by the
    first.  Mixed edges 16/32/64 and terrains of 1, 2, 3, 4, 5 octaves and a
    heightmap; no neighbours."""
    say("batch split fill", n)
    rng = np.random.default_rng(n)
    one = V.Terrain(kind=V.TERRAIN_DENSITY, seed=13, octaves=1, frequency=1 / 24, y_scale=1 / 64)
    pool = [one, OCT2, REF_NOISE, C2, OCT5, C1, REF_RULE]
    edges = [(16, 16, 16, 16, 16, 16, 32, 32, 64)[i % 9] for i in range(n)]
    terrains = [pool[int(k)] for k in rng.integers(0, len(pool), n)]
    cells = rng.choice(80 * 4 * 80, n, replace=False)
    coords = [(int(c // 320) - 40, int(c // 80 % 4) - 2, int(c % 80) - 40) for c in cells]
    b = batch_vs_oracle(ctx, oracle_mod, coords, edges, terrains, np.full((n, 4), -1, np.int32),
                        formats=V.FMT_VERTICES, mode=V.BATCH_STAGED)
    b.close()


@pytest.mark.parametrize("shape", [pytest.param((4, 4, 32), id="4x4x32"),
                                   pytest.param((32, 32, 16), id="32x32x16"),
                                   pytest.param((3, 3, 64), id="3x3x64")])
def test_batch_reference_rule(ctx, oracle_mod, shape):
    """Grids in the reference's own terrain rule (noise at (wx, wz, wy) < 0,
    its seed 1 / f 0.09 / gain 2.5 / lacunarity 0.5 / 3 octaves) with in-grid
    neighbours: 16 chunks (rows split over 4 fill workgroups), 1024 chunks
    (one fill workgroup each) and 64^3 chunks (8 parts); ids and meshes byte
    for byte.  The fused form is refused for this terrain."""
    gx, gz, edge = shape
    say("batch reference rule", shape)
    coords = [(x - gx // 2, 0, z - 1) for x in range(gx) for z in range(gz)]
    nb = V.grid_neighbours(gx, gz)
    b = batch_vs_oracle(ctx, oracle_mod, coords, [edge] * len(coords), [REF_RULE] * len(coords),
                        nb, formats=V.FMT_VERTICES | V.FMT_INSTANCES)
    assert b.mode == V.BATCH_STAGED
    with pytest.raises(V.VoxelError) as e:
        b.set_mode(V.BATCH_FUSED)
    assert e.value.status == V.ERR_INVALID
    b.close()


def test_coordinates_outside_the_i32_world_are_refused(ctx, oracle_mod):
    """A chunk whose voxels or one-voxel halo would leave the i32 world range
    is refused (VX_ERR_INVALID) by every entry point that builds chunks; the
    last accepted chunk on each side fills bit-exact."""
    tid = ctx.terrain(BIGF)
    for coord in ((2 ** 26 - 1, 0, 0), (0, -(2 ** 26), 0), (0, 0, 2 ** 30)):
        with pytest.raises(V.VoxelError) as e:
            V.Chunk.new(ctx, BIGF, coord, 32)
        assert e.value.status == V.ERR_INVALID
        with pytest.raises(V.VoxelError) as e:
            V.Batch(ctx, [coord], 32, tid, np.full((1, 4), -1, np.int32))
        assert e.value.status == V.ERR_INVALID
    with V.Server(ctx) as s:
        with pytest.raises(V.VoxelError) as e:
            s.submit([((2 ** 26 - 1, 0, 0), 32, tid, V.NB_NONE)])
        assert e.value.status == V.ERR_INVALID
    for coord in ((2 ** 26 - 2, 0, 0), (1 - 2 ** 26, 0, 0)):
        ch = V.Chunk.new(ctx, BIGF, coord, 32)
        assert (ch.ids() == oracle_mod.chunk_fill(oterrain(oracle_mod, BIGF), *coord, 32)).all()


def test_batch_sizes_outside_the_range_are_refused(ctx):
    """A batch holds 1 to VX_BATCH_MAX_CHUNKS (2^26) requests: an empty batch
    and a count one above the maximum are refused with VX_ERR_INVALID (the
    count is checked before any request is read), and a one-chunk batch
    runs."""
    import ctypes as C
    tid = ctx.terrain(C1)
    with pytest.raises(V.VoxelError) as e:
        V.Batch(ctx, np.zeros((0, 3), np.int32), 16, tid)
    assert e.value.status == V.ERR_INVALID
    one = (V.BatchReq * 1)()
    one[0].edge = 16
    one[0].terrain_id = tid
    for k in range(4):
        one[0].nb[k] = -1
    out = C.c_void_p()
    st = ctx.L.vx_batch_create(ctx.h, C.addressof(one), (1 << 26) + 1, V.FMT_VERTICES, 0,
                               C.byref(out))
    assert st == V.ERR_INVALID and not out.value
    b = V.Batch(ctx, [(0, 0, 0)], 16, tid)
    b.run()
    assert b.wait() > 0
    b.close()


def test_terrain_validation(ctx):
    """Octave counts outside 1..8 and octave frequencies below 2^-20 are
    refused with VX_ERR_INVALID."""
    import dataclasses
    for bad in (dataclasses.replace(OCT2, octaves=0), dataclasses.replace(OCT2, octaves=9),
                dataclasses.replace(OCT8, frequency=1e-6, lacunarity=0.5)):
        with pytest.raises(V.VoxelError) as e:
            ctx.terrain(bad)
        assert e.value.status == V.ERR_INVALID
    assert ctx.terrain(dataclasses.replace(OCT2, octaves=8)) >= 0


def test_fused_is_refused_for_non_adjacent_neighbours(ctx, oracle_mod):
    """A neighbour index that is not the adjacent chunk cannot be recomputed:
    the batch runs staged and still matches the oracle."""
    coords = [(0, 0, 0), (5, 0, 5), (9, 0, 1)]
    nb = [[-1, 1, -1, 2], [0, -1, -1, -1], [-1, -1, 0, -1]]
    b = V.Batch(ctx, coords, 32, ctx.terrain(C2), nb)
    assert b.mode == V.BATCH_STAGED
    with pytest.raises(V.VoxelError):
        b.set_mode(V.BATCH_FUSED)
    batch_vs_oracle(ctx, oracle_mod, coords, [32] * 3, [C2] * 3, nb)


@pytest.mark.parametrize("mode", MODES)
def test_batch_ragged(ctx, oracle_mod, mode):
    """Config 3 shape: edges from {16,32,64}, 30% air / 10% solid / 60%
    heightmap surface (seed 11), no neighbours, longest-first schedule."""
    say("batch ragged")
    rng = np.random.default_rng(11)
    n = 48
    edges = rng.choice([16, 32, 64], size=n)
    cls = rng.choice(3, size=n, p=[0.3, 0.1, 0.6])
    terrains = []
    for e, k in zip(edges, cls):
        if k == 0:
            terrains.append(V.Terrain(kind=V.TERRAIN_AIR))
        elif k == 1:
            terrains.append(V.Terrain(kind=V.TERRAIN_SOLID))
        else:
            terrains.append(V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=11, octaves=4,
                                      frequency=1 / 32, height_scale=float(e), height_max=int(e) - 1))
    coords = [(i % 8, 0, i // 8) for i in range(n)]
    batch_vs_oracle(ctx, oracle_mod, coords, edges, terrains, np.full((n, 4), -1), mode=mode)


def test_batch_ragged_autotune_keeps_results(ctx, oracle_mod):
    """The autotuner on a small ragged batch holding 64^3 chunks (slab emit
    items, an 8-way fill split by rule): staged tile shapes and the fused
    form are timed, and the tuned batch's ids and meshes are the oracle's,
    byte for byte."""
    say("batch ragged autotune")
    rng = np.random.default_rng(11)
    n = 48
    edges = rng.choice([16, 32, 64], size=n)
    cls = rng.choice(3, size=n, p=[0.3, 0.1, 0.6])
    terrains = []
    for e, k in zip(edges, cls):
        if k == 0:
            terrains.append(V.Terrain(kind=V.TERRAIN_AIR))
        elif k == 1:
            terrains.append(V.Terrain(kind=V.TERRAIN_SOLID))
        else:
            terrains.append(C2)
    coords = [(i % 8, 0, i // 8) for i in range(n)]
    nb = np.full((n, 4), -1)
    fmt = V.FMT_VERTICES | V.FMT_INSTANCES
    b = V.Batch(ctx, coords, edges, [ctx.terrain(t) for t in terrains], nb, formats=fmt)
    r = b.autotune(reps=2)
    shapes = {(c["mode"], tuple(c["splits"])) for c in r["candidates"]}
    assert ("staged", (0, 0)) in shapes and ("fused", (0, 0)) in shapes, r
    assert any(m == "staged" and f for m, (f, _) in shapes), r
    assert any(m == "staged" and e for m, (_, e) in shapes), r
    check_batch(b, oracle_mod, coords, edges, terrains, nb, fmt)
    b.close()


def test_batch_split_with_counting_fills(ctx, oracle_mod):
    """The split form on 2,048 chunks of 16^3 with no neighbour inside the
    batch (the fills count; each half has 1,024 whole-chunk fills) and on
    a 64-chunk grid whose first half borders the second: ids and meshes
    byte for byte."""
    say("batch split")
    coords = [(x, 0, z) for x in range(32) for z in range(64)]
    batch_vs_oracle(ctx, oracle_mod, coords, [16] * len(coords), [C2] * len(coords),
                    np.full((len(coords), 4), -1), formats=V.FMT_VERTICES,
                    mode=(V.BATCH_SPLIT, 0))
    coords = [(x, 0, z) for x in range(8) for z in range(8)]
    batch_vs_oracle(ctx, oracle_mod, coords, [32] * 64, [C2] * 64, V.grid_neighbours(8, 8),
                    mode=(V.BATCH_SPLIT, 0))


def test_batch_fill_counts_against_resident_neighbours(ctx, oracle_mod):
    """1,024 chunks of 16^3 (one fill workgroup each) with no neighbour
    inside the batch: the staged fill counts the faces itself, so the batch
    has no count kernel.  The first column's -x neighbours are resident
    chunks of an earlier batch, every other side is absent.  Ids, counts,
    bases and every mesh equal the oracle's; a replay gives the same; the
    per-chunk path (count kernel) agrees."""
    O = oracle_mod
    say("batch fill counts")
    t = V.Terrain(kind=V.TERRAIN_DENSITY, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0,
                  gain=0.5, y_scale=1 / 128)
    tid = ctx.terrain(t)
    side = 32
    res = V.Batch(ctx, [(0, 0, z) for z in range(side)], 16, tid, np.full((side, 4), -1),
                  formats=V.FMT_VERTICES)
    res.run()
    res.wait()
    ext = [res.chunk(z) for z in range(side)]
    coords = [(x, 0, z) for x in range(1, side + 1) for z in range(side)]
    nb = np.full((len(coords), 4), -1, np.int32)
    for i, (x, _, z) in enumerate(coords):
        if x == 1:
            nb[i, 0] = -2 - z
    b = V.Batch(ctx, coords, 16, tid, nb, formats=V.FMT_VERTICES | V.FMT_INSTANCES, ext=ext)
    # halves of 512 chunks cannot have whole-chunk fills: the split form is refused
    with pytest.raises(V.VoxelError) as e:
        b.set_mode(V.BATCH_SPLIT)
    assert e.value.status == V.ERR_INVALID
    b.set_mode(V.BATCH_STAGED)
    ot = oterrain(O, t)
    ids = {c: O.chunk_fill(ot, *c, 16) for c in coords}
    ids.update({(0, 0, z): O.chunk_fill(ot, 0, 0, z, 16) for z in range(side)})
    for run in range(3):
        if run < 2:
            b.run()
        else:
            b.run_per_chunk()
        total = b.wait()
        verts, idxs, insts = b.mesh(V.FMT_VERTICES), b.mesh(V.FMT_INDICES), b.mesh(V.FMT_INSTANCES)
        first_expected = 0
        for i, c in enumerate(coords):
            if run == 0:
                assert (b.ids(i) == ids[c]).all(), i
            adj = [ids[(0, 0, c[2])] if c[0] == 1 else None, None, None, None]
            faces, inst, vert, ix = O.mesh_buffers(ids[c], 16, (c[0] * 16, 0, c[2] * 16), adj)
            first, cnt = b.chunk_faces(i)
            assert (first, cnt) == (first_expected, len(faces)), (run, i)
            if run == 0 or i % 97 == 0:
                assert (bits(verts[4 * first:4 * (first + cnt)]) == bits(vert)).all(), (run, i)
                assert (idxs[6 * first:6 * (first + cnt)] == ix).all(), (run, i)
                assert (bits(insts[first:first + cnt]) == bits(inst)).all(), (run, i)
            first_expected += cnt
        assert total == first_expected
    b.close()
    for c in ext:
        c.release()
    res.close()


@pytest.mark.parametrize("mode", MODES)
def test_batch_overflow_grows_arena(ctx, oracle_mod, mode):
    say("batch overflow", mode)
    coords = [(x, 0, z) for x in range(3) for z in range(3)]
    batch_vs_oracle(ctx, oracle_mod, coords, [32] * 9, [C2] * 9, V.grid_neighbours(3, 3),
                    face_capacity=100, mode=mode)


@pytest.mark.parametrize("mode", MODES)
def test_batch_replay_is_stable(ctx, oracle_mod, mode):
    """hipGraph replays give identical results, with eager work between."""
    coords = [(x, 0, z) for x in range(4) for z in range(4)]
    b = V.Batch(ctx, coords, 32, ctx.terrain(C2), V.grid_neighbours(4, 4))
    apply_mode(b, mode)
    b.run()
    t0 = b.wait()
    v0 = b.mesh(V.FMT_VERTICES).copy()
    for _ in range(3):
        ctx.noise(C2, (0, 0, 0), 16)  # eager work between replays
        b.run()
        assert b.wait() == t0
        assert (bits(b.mesh(V.FMT_VERTICES)) == bits(v0)).all()


def test_same_alone_and_in_the_mix(ctx, oracle_mod):
    """A request's mesh is byte-identical whether it runs alone (per-chunk
    API) or inside a batch of other requests, given the same neighbours."""
    coords = [(x, 0, z) for x in range(3) for z in range(3)]
    b = V.Batch(ctx, coords, 32, ctx.terrain(C2), V.grid_neighbours(3, 3),
                formats=V.FMT_VERTICES | V.FMT_INSTANCES)
    b.run()
    b.wait()
    verts, insts = b.mesh(V.FMT_VERTICES), b.mesh(V.FMT_INSTANCES)
    chunks = [V.Chunk.new(ctx, C2, c, 32) for c in coords]
    nb = V.grid_neighbours(3, 3)
    for i in range(len(coords)):
        m = chunks[i].create_mesh([chunks[j] if j >= 0 else None for j in nb[i]])
        first, cnt = b.chunk_faces(i)
        assert m.faces == cnt
        assert (bits(m.vertices()) == bits(verts[4 * first:4 * (first + cnt)])).all()
        assert (bits(m.instances()) == bits(insts[first:first + cnt])).all()


# ------------------------------------------------------------------ frustum
def test_frustum_mask_matches_oracle(ctx, oracle_mod):
    import camera as T
    rng = np.random.default_rng(12)
    for _ in range(5):
        eye = rng.uniform(-500, 500, 3)
        m = T.colmajor(T.perspective(np.radians(70), 1.6, 0.1, 1024) @
                       T.look_at(eye, eye + rng.normal(size=3)))
        f = V.Frustum(m)
        coords = rng.integers(-64, 64, size=(20000, 3)).astype(np.int32)
        got = f.mask(ctx, coords, 32)
        ref = oracle_mod.frustum_contains(f.planes, coords, 32)
        assert (got == ref).all()
        assert got.sum() > 0


# ------------------------------------------------------------------ frame
@pytest.mark.parametrize("mode", MODES)
def test_frame_new_chunks_beside_resident_grid_with_cull(ctx, oracle_mod, mode):
    """Config 5 shape at small scale: a resident 8x8 grid; a frame meshes 8
    new chunks in the column x = 8 (their -x neighbours are resident chunks,
    +-z neighbours are in the same batch) and culls the resident AABBs in
    the same run."""
    import camera as T
    say("frame", mode)
    t = oterrain(oracle_mod, C5)
    tid = ctx.terrain(C5)
    rc = [(x, 0, z) for x in range(8) for z in range(8)]
    R = V.Batch(ctx, rc, 32, tid, V.grid_neighbours(8, 8))
    R.run()
    R.wait()
    ext = [R.chunk(7 * 8 + z) for z in range(8)]  # resident column x = 7
    nc = [(8, 0, z) for z in range(8)]
    nb = [[-2 - z, -1, z - 1 if z > 0 else -1, z + 1 if z < 7 else -1] for z in range(8)]
    B = V.Batch(ctx, nc, 32, tid, nb, formats=V.FMT_VERTICES | V.FMT_INSTANCES, ext=ext)
    apply_mode(B, mode)
    coords = np.array(rc, np.int32)
    cbuf = V.DeviceBuffer.of(ctx, coords)
    mbuf = V.DeviceBuffer(ctx, len(rc))
    eye = np.array([-40.0, 30.0, 100.0])
    for frame in range(2):  # the camera moves: new planes, same captured graph
        m = T.colmajor(T.perspective(np.radians(70), 1.5, 0.1, 1024) @
                       T.look_at(eye, eye + np.array([1.0, -0.3, -0.5 - frame])))
        f = V.Frustum(m)
        B.set_cull(f, cbuf, len(rc), 32, mbuf)
        B.run()
        total = B.wait()
        mask = mbuf.download(np.uint8, len(rc))
        assert (mask == oracle_mod.frustum_contains(f.planes, coords, 32)).all()
        assert 0 < mask.sum() < len(rc)
    ids = {c: oracle_mod.chunk_fill(t, *c, 32) for c in nc + [(7, 0, z) for z in range(8)]}
    verts, insts = B.mesh(V.FMT_VERTICES), B.mesh(V.FMT_INSTANCES)
    run = 0
    for i, c in enumerate(nc):
        adj = [ids[(7, 0, c[2])], None, ids.get((8, 0, c[2] - 1)), ids.get((8, 0, c[2] + 1))]
        faces, inst, vert, idx = oracle_mod.mesh_buffers(ids[c], 32, (c[0] * 32, 0, c[2] * 32), adj)
        first, cnt = B.chunk_faces(i)
        assert first == run and cnt == len(faces)
        assert (bits(verts[4 * first:4 * (first + cnt)]) == bits(vert)).all()
        assert (bits(insts[first:first + cnt]) == bits(inst)).all()
        run += cnt
    assert total == run
    B.close()
    R.close()
    cbuf.free()
    mbuf.free()


# ------------------------------------------------------------------ draw lists
def oracle_draws(O, planes, coords, faces, first, edge, mode):
    """World::render's loop (world.rs:34-38): one draw per mesh whose chunk
    passes Frustum::contains, in mesh order, as indirect arguments."""
    mask = O.frustum_contains(planes, coords, edge).astype(bool)
    f = np.asarray(faces, np.uint64)[mask]
    s = np.asarray(first, np.uint64)[mask]
    out = np.zeros(int(mask.sum()), V.DRAW_DTYPE)
    if mode == V.DRAW_INSTANCED:
        out["index_count"], out["instance_count"], out["first_instance"] = 6, f, s
    else:
        out["index_count"], out["instance_count"] = 6 * f, 1
        out["first_index"], out["base_vertex"] = 6 * s, (4 * s).astype(np.int32)
    return out


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 1025, 70000])
@pytest.mark.parametrize("mode", [V.DRAW_INSTANCED, V.DRAW_EXPANDED])
def test_draw_list_matches_oracle(ctx, oracle_mod, n, mode):
    import camera as T
    say("draw list", n, mode)
    rng = np.random.default_rng(n + 7 * mode)
    eye = rng.uniform(-300, 300, 3)
    f = V.Frustum(T.colmajor(T.perspective(np.radians(70), 1.6, 0.1, 1024) @
                             T.look_at(eye, eye + rng.normal(size=3))))
    coords = rng.integers(-40, 40, size=(n, 3)).astype(np.int32)
    faces = rng.integers(0, 6000, size=n).astype(np.uint32)
    first = np.concatenate([[0], np.cumsum(faces, dtype=np.uint64)[:-1]]).astype(np.uint64) if n else \
        np.zeros(0, np.uint64)
    got = f.draw_list(ctx, coords, faces, first, 32, mode)
    ref = oracle_draws(oracle_mod, f.planes, coords, faces, first, 32, mode)
    assert got.tobytes() == ref.tobytes()
    if n >= 1024:
        assert 0 < len(got) < n


def test_draw_list_reports_offsets_beyond_32_bits(ctx):
    f = V.Frustum(np.eye(4, dtype=np.float32).reshape(16))  # clip-space cube
    coords = np.zeros((1, 3), np.int32)
    with pytest.raises(V.VoxelError):
        f.draw_list(ctx, coords, np.array([5], np.uint32), np.array([2**32 - 3], np.uint64), 1,
                    V.DRAW_INSTANCED)


@pytest.mark.parametrize("mode", [V.DRAW_INSTANCED, V.DRAW_EXPANDED])
def test_batch_draw_list(ctx, oracle_mod, mode):
    """A resident grid's culled draw list references exactly the visible
    chunks' slices of the batch arena."""
    import camera as T
    coords = [(x, 0, z) for x in range(8) for z in range(8)]
    b = V.Batch(ctx, coords, 32, ctx.terrain(C5), V.grid_neighbours(8, 8),
                formats=V.FMT_VERTICES | V.FMT_INSTANCES)
    b.run()
    b.wait()
    eye = np.array([-20.0, 40.0, 120.0])
    f = V.Frustum(T.colmajor(T.perspective(np.radians(70), 1.5, 0.1, 1024) @
                             T.look_at(eye, eye + np.array([1.0, -0.4, 0.2]))))
    fc = [b.chunk_faces(i) for i in range(len(coords))]
    got = b.draw_list(f, mode)
    ref = oracle_draws(oracle_mod, f.planes, np.array(coords, np.int32), [c for _, c in fc],
                       [s for s, _ in fc], 32, mode)
    assert got.tobytes() == ref.tobytes()
    assert 0 < len(got) < len(coords)
