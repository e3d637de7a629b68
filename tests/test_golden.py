"""Frozen golden vectors (tests/golden/golden.json, written by
tests/golden/make_golden.py from the oracle): the oracle must keep
reproducing them (CPU), and the HIP path must reproduce the same digests
through the C ABI (GPU)."""
import hashlib
import json
import os

import numpy as np
import pytest

import voxel_amd as V

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def oracle_terrain(O, name):
    return O.terrain(**GOLDEN["terrains"][name])


def lib_terrain(name):
    kw = dict(GOLDEN["terrains"][name])
    return V.Terrain(**kw)


@pytest.mark.parametrize("name", sorted(GOLDEN["noise"]))
def test_oracle_noise_matches_golden(oracle_mod, name):
    kw = GOLDEN["terrains"][name]
    p = oracle_mod.perm(kw["seed"])
    g = GOLDEN["noise"][name]
    for (x, y, z), b in zip(g["points"], g["bits"]):
        if kw["kind"] == 3:
            v = oracle_mod.fbm3(p, float(x), float(y), float(z), kw["octaves"], kw["frequency"],
                                kw["lacunarity"], kw["gain"])
        elif kw["kind"] == 4:  # the reference's axis order: noise at (x, z, y)
            v = oracle_mod.fbm3(p, float(x), float(z), float(y), kw["octaves"], kw["frequency"],
                                kw["lacunarity"], kw["gain"])
        else:
            v = oracle_mod.fbm2(p, float(x), float(z), kw["octaves"], kw["frequency"],
                                kw["lacunarity"], kw["gain"])
        assert int(np.float32(v).view(np.uint32)) == b


@pytest.mark.parametrize("name", sorted(GOLDEN["batches"]))
def test_oracle_batches_match_golden(oracle_mod, name):
    g = GOLDEN["batches"][name]
    t = oracle_terrain(oracle_mod, g["terrain"])
    ids = [oracle_mod.chunk_fill(t, *c, e) for c, e in zip(g["coords"], g["edges"])]
    inst, vert, idx, faces = [], [], [], []
    for i, (c, e) in enumerate(zip(g["coords"], g["edges"])):
        adj = [ids[j] if j >= 0 else None for j in g["neighbours"][i]]
        f, a, b, d = oracle_mod.mesh_buffers(ids[i], e, tuple(v * e for v in c), adj)
        inst.append(a), vert.append(b), idx.append(d), faces.append(len(f))
    assert faces == g["faces"]
    assert sha(*ids) == g["ids_sha256"]
    assert sha(*inst) == g["instances_sha256"]
    assert sha(*vert) == g["vertices_sha256"]
    assert sha(*idx) == g["indices_sha256"]


def frustum_setup():
    f = GOLDEN["frustum"]
    m = np.array(f["matrix_bits"], np.uint32).view(np.float32)
    coords = np.array([(x, y, z) for y in (-1, 0) for x in range(64) for z in range(64)], np.int32)
    return f, m, coords


def test_oracle_frustum_matches_golden(oracle_mod):
    f, m, coords = frustum_setup()
    mask = oracle_mod.frustum_contains(oracle_mod.frustum_planes(m), coords, f["edge"])
    assert int(mask.sum()) == f["visible"] and sha(mask) == f["mask_sha256"]


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("mode,grid", [(V.BATCH_FUSED, 0), (V.BATCH_FUSED, 2), (V.BATCH_STAGED, 0),
                                       (V.BATCH_PIPELINED, 0), (V.BATCH_PIPELINED, 2)])
@pytest.mark.parametrize("name", sorted(GOLDEN["batches"]))
def test_library_batches_match_golden(ctx, name, mode, grid):
    g = GOLDEN["batches"][name]
    tid = ctx.terrain(lib_terrain(g["terrain"]))
    b = V.Batch(ctx, g["coords"], g["edges"], tid, g["neighbours"],
                formats=V.FMT_VERTICES | V.FMT_INSTANCES)
    ref_rule = GOLDEN["terrains"][g["terrain"]]["kind"] == V.TERRAIN_REFERENCE
    if (mode == V.BATCH_FUSED and ref_rule) or \
            (mode == V.BATCH_PIPELINED and (ref_rule or max(g["edges"]) > 32)):
        with pytest.raises(V.VoxelError):  # this batch is not eligible for the form
            b.set_mode(mode)
        b.close()
        return
    b.set_mode(mode)
    b.set_grid(grid)
    b.run()
    assert b.wait() == sum(g["faces"])
    assert [b.chunk_faces(i)[1] for i in range(len(g["coords"]))] == g["faces"]
    assert sha(*[b.ids(i) for i in range(len(g["coords"]))]) == g["ids_sha256"]
    assert sha(b.mesh(V.FMT_INSTANCES)) == g["instances_sha256"]
    assert sha(b.mesh(V.FMT_VERTICES)) == g["vertices_sha256"]
    # indices are chunk-local: the batch buffer is the concatenation
    assert sha(b.mesh(V.FMT_INDICES)) == g["indices_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLDEN["noise"]))
def test_library_noise_matches_golden(ctx, name):
    kw = GOLDEN["terrains"][name]
    t = lib_terrain(name)
    g = GOLDEN["noise"][name]
    N = 32
    for (x, y, z), b in list(zip(g["points"], g["bits"]))[:16]:
        c = (x // N, y // N, z // N)
        field = ctx.noise(t, c, N)
        lx, ly, lz = x - c[0] * N, y - c[1] * N, z - c[2] * N
        v = field[lx + N * lz + N * N * ly] if kw["kind"] in (3, 4) else field[lx + N * lz]
        assert int(np.float32(v).view(np.uint32)) == b


@pytest.mark.gpu
def test_library_frustum_matches_golden(ctx):
    f, m, coords = frustum_setup()
    mask = V.Frustum(m).mask(ctx, coords, f["edge"])
    assert int(mask.sum()) == f["visible"] and sha(mask) == f["mask_sha256"]
