"""Pins the CPU oracle (oracle/voxel_oracle.c) to what the reference and the
mathematics fix: published known answers, closed forms, invariants, an
exhaustive brute force on tiny chunks and an independent NumPy
implementation.  CPU only."""
import itertools
import math

import numpy as np
import pytest

import np_noise
from camera import colmajor, look_at, perspective

# Ken Perlin's permutation from his reference ImprovedNoise implementation.
PERLIN_P = [
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69,
    142, 8, 99, 37, 240, 21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252, 219,
    203, 117, 35, 11, 32, 57, 177, 33, 88, 237, 149, 56, 87, 174, 20, 125, 136, 171, 168, 68, 175,
    74, 165, 71, 134, 139, 48, 27, 166, 77, 146, 158, 231, 83, 111, 229, 122, 60, 211, 133, 230,
    220, 105, 92, 41, 55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216, 80, 73, 209, 76,
    132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100, 109, 198, 173,
    186, 3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82, 85, 212, 207, 206,
    59, 227, 47, 16, 58, 17, 182, 189, 28, 42, 223, 183, 170, 213, 119, 248, 152, 2, 44, 154, 163,
    70, 221, 153, 101, 155, 167, 43, 172, 9, 129, 22, 39, 253, 19, 98, 108, 110, 79, 113, 224, 232,
    178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238, 210, 144, 12, 191, 179, 162,
    241, 81, 51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31, 181, 199, 106, 157, 184, 84, 204,
    176, 115, 121, 50, 45, 127, 4, 150, 254, 138, 236, 205, 93, 222, 114, 67, 29, 24, 72, 243, 141,
    128, 195, 78, 66, 215, 61, 156, 180]

NORMALS = {0: (-1, 0, 0), 1: (1, 0, 0), 2: (0, 0, -1), 3: (0, 0, 1), 4: (0, -1, 0), 5: (0, 1, 0)}


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


# ---------------------------------------------------------------- noise
def test_perlin_table_is_a_permutation():
    assert sorted(PERLIN_P) == list(range(256))


def test_perlin_known_answer(oracle_mod):
    """Perlin's reference implementation gives noise(3.14, 42, 7) =
    0.136919958784 with his permutation (double precision)."""
    p = oracle_mod.perm_from_table(PERLIN_P)
    assert oracle_mod.noise3d(p, 3.14, 42.0, 7.0) == pytest.approx(0.13691995878400012, abs=1e-15)
    # the float instantiation of the same source text agrees to float accuracy
    assert oracle_mod.noise3(p, 3.14, 42.0, 7.0) == pytest.approx(0.136919958784, abs=2e-6)


def test_splitmix64_known_answer(oracle_mod):
    """First outputs of splitmix64 for seed 1234567 (Vigna's reference)."""
    assert oracle_mod.splitmix64(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
        16408922859458223821]


@pytest.mark.parametrize("seed", [0, 1, 5, 7, 11, 42, 2**32 - 1])
def test_perm_matches_independent_shuffle(oracle_mod, seed):
    p = oracle_mod.perm(seed)
    assert sorted(p[:256].tolist()) == list(range(256))
    assert (p[:256] == p[256:]).all()
    assert (p.astype(np.int64) == np_noise.permutation(seed)).all()


def test_noise_vanishes_on_lattice(oracle_mod):
    p = oracle_mod.perm(7)
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(-300, 300, size=(200, 3)):
        assert oracle_mod.noise3(p, float(x), float(y), float(z)) == 0.0
        assert oracle_mod.noise2(p, float(x), float(z)) == 0.0


def test_noise2_linearises_to_gradient(oracle_mod):
    """Near a lattice point noise2 = g . (dx, dy) + O(d^3) with g the gradient
    selected by p[p[X]+Y] & 7 (catches a wrong hash index or gradient)."""
    p = oracle_mod.perm(42)
    grads = [(1, 1), (-1, 1), (1, -1), (-1, -1), (1, 0), (-1, 0), (0, 1), (0, -1)]
    d = 1.0 / 1024
    for X, Y in [(0, 0), (3, 5), (17, 200), (255, 1)]:
        g = grads[int(p[int(p[X]) + Y]) & 7]
        for dx, dy in [(d, 0), (0, d), (d, d)]:
            n = oracle_mod.noise2(p, X + dx, Y + dy)
            assert n == pytest.approx(g[0] * dx + g[1] * dy, abs=1e-7)


def test_noise3_linearises_to_gradient(oracle_mod):
    p = oracle_mod.perm(7)
    edges = [(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
             (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (0, -1, 1),
             (-1, 1, 0), (0, -1, -1)]
    d = 1.0 / 1024
    for X, Y, Z in [(0, 0, 0), (4, 9, 2), (100, 3, 77)]:
        h = int(p[int(p[int(p[X]) + Y]) + Z]) & 15
        g = edges[h]
        for v in [(d, 0, 0), (0, d, 0), (0, 0, d)]:
            n = oracle_mod.noise3(p, X + v[0], Y + v[1], Z + v[2])
            assert n == pytest.approx(g[0] * v[0] + g[1] * v[1] + g[2] * v[2], abs=1e-7)


def test_noise_bit_exact_vs_numpy(oracle_mod):
    rng = np.random.default_rng(3)
    p = oracle_mod.perm(7)
    pn = np_noise.permutation(7)
    pts = np.concatenate([rng.uniform(-500, 500, size=(300, 3)),
                          rng.integers(-4096, 4096, size=(100, 3)) / 64.0]).astype(np.float32)
    ours = np.array([oracle_mod.noise3(p, *map(float, q)) for q in pts], np.float32)
    ref = np_noise.noise3(pn, pts[:, 0], pts[:, 1], pts[:, 2])
    assert (bits(ours) == bits(ref)).all()
    ours2 = np.array([oracle_mod.noise2(p, float(q[0]), float(q[2])) for q in pts], np.float32)
    ref2 = np_noise.noise2(pn, pts[:, 0], pts[:, 2])
    assert (bits(ours2) == bits(ref2)).all()


@pytest.mark.parametrize("cfg", [
    dict(kind=3, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0, gain=0.5, coord=(5, 0, 9), edge=32),
    dict(kind=3, seed=5, octaves=4, frequency=1 / 64, lacunarity=2.0, gain=0.5, coord=(-3, 1, -7), edge=16),
    dict(kind=3, seed=1, octaves=3, frequency=0.09, lacunarity=0.5, gain=2.5, coord=(2, -1, 4), edge=16),
    dict(kind=2, seed=42, octaves=1, frequency=1 / 32, lacunarity=2.0, gain=0.5, coord=(3, 0, 1), edge=16),
    dict(kind=2, seed=11, octaves=4, frequency=1 / 32, lacunarity=2.0, gain=0.5, coord=(-2, 0, 6), edge=64),
])
def test_noise_block_bit_exact_vs_numpy(oracle_mod, cfg):
    t = oracle_mod.terrain(kind=cfg["kind"], seed=cfg["seed"], octaves=cfg["octaves"],
                           frequency=cfg["frequency"], lacunarity=cfg["lacunarity"], gain=cfg["gain"])
    N = cfg["edge"]
    ours = oracle_mod.noise_block(t, *cfg["coord"], N)
    pn = np_noise.permutation(cfg["seed"])
    o = np.array(cfg["coord"]) * N
    if cfg["kind"] == 3:
        y, z, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        ref = np_noise.fbm3(pn, (x + o[0]).astype(np.float32), (y + o[1]).astype(np.float32),
                            (z + o[2]).astype(np.float32), cfg["octaves"], cfg["frequency"],
                            cfg["lacunarity"], cfg["gain"]).reshape(-1)
    else:
        z, x = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
        ref = np_noise.fbm2(pn, (x + o[0]).astype(np.float32), (z + o[2]).astype(np.float32),
                            cfg["octaves"], cfg["frequency"], cfg["lacunarity"],
                            cfg["gain"]).reshape(-1)
    assert (bits(ours) == bits(ref)).all()


def test_avx2_rows_match_scalar(oracle_mod):
    p = oracle_mod.perm(7)
    rng = np.random.default_rng(5)
    for x0, y, z in rng.integers(-2000, 2000, size=(60, 3)):
        for f0, lac, gain, octs in [(1 / 64, 2.0, 0.5, 4), (0.09, 0.5, 2.5, 3)]:
            row = oracle_mod.fbm3_row8(p, int(x0), int(y), int(z), octs, f0, lac, gain)
            ref = [oracle_mod.fbm3(p, float(x0 + i), float(y), float(z), octs, f0, lac, gain)
                   for i in range(8)]
            assert (bits(row) == bits(np.array(ref, np.float32))).all()


# ---------------------------------------------------------------- fill
def test_fast_fill_matches_scalar(oracle_mod):
    t = oracle_mod.terrain()
    for c in [(0, 0, 0), (7, 0, -3), (-5, -1, 2)]:
        assert (oracle_mod.chunk_fill(t, *c, 32) == oracle_mod.chunk_fill(t, *c, 32, fast=True)).all()


def test_density_rule(oracle_mod):
    t = oracle_mod.terrain(seed=7)
    N, c = 16, (4, 0, 2)
    ids = oracle_mod.chunk_fill(t, *c, N)
    d = oracle_mod.noise_block(t, *c, N).reshape(N, N, N)
    y = (np.arange(N) + c[1] * N).astype(np.float32)[:, None, None]
    expect = (d - y * np.float32(1 / 128)) > 0
    assert (ids.reshape(N, N, N) == expect.astype(np.uint8)).all()


def test_heightmap_columns_are_prefixes(oracle_mod):
    t = oracle_mod.terrain(kind=2, seed=42, octaves=1, frequency=1 / 32, height_scale=16.0,
                           height_max=15)
    for c in [(0, 0, 0), (3, 0, 2), (-4, 0, 9)]:
        N = 16
        ids = oracle_mod.chunk_fill(t, *c, N).reshape(N, N, N)  # [y][z][x]
        n = oracle_mod.noise_block(t, *c, N).reshape(N, N)       # [z][x]
        h = np.clip(np.floor(np.float32(16) * (n + np.float32(1)) / np.float32(2)), 0, 15)
        for z in range(N):
            for x in range(N):
                col = ids[:, z, x]
                assert (col == (np.arange(N) <= h[z, x])).all()
    # a chunk above the surface band is all air, the configured clamp holds
    assert oracle_mod.chunk_fill(t, 0, 1, 0, 16).sum() == 0


REF_PARAMS = dict(seed=1, octaves=3, frequency=0.09, lacunarity=0.5, gain=2.5)


@pytest.mark.parametrize("params", [REF_PARAMS,
                                    dict(seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0,
                                         gain=0.5)])
def test_reference_rule_noise_is_density_noise_transposed(oracle_mod, params):
    """The reference rule samples the noise at (wx, wz, wy)
    (perlin_noise.rs:6-13): its noise block at chunk (cx, cy, cz) is the
    density rule's block at (cx, cz, cy) with y and z swapped, bit for bit,
    and a voxel is grass iff that value is < 0 (chunk.rs:38-42)."""
    O = oracle_mod
    tr = O.terrain(kind=O.TERRAIN_REFERENCE, **params)
    td = O.terrain(kind=O.TERRAIN_DENSITY, y_scale=0.0, **params)
    for (cx, cy, cz), N in [((0, 0, 0), 16), ((-1, 2, 3), 16), ((5, -1, -7), 32)]:
        nr = O.noise_block(tr, cx, cy, cz, N).reshape(N, N, N)           # [y][z][x]
        nd = O.noise_block(td, cx, cz, cy, N).reshape(N, N, N)           # [z][y][x] here
        assert (bits(nr) == bits(nd.transpose(1, 0, 2))).all()
        ids = O.chunk_fill(tr, cx, cy, cz, N).reshape(N, N, N)
        assert (ids == (nr < 0).astype(np.uint8)).all()
        assert (O.chunk_fill(tr, cx, cy, cz, N, fast=True).reshape(N, N, N) == ids).all()
        # and the complement of the density rule with threshold 0, except
        # where the noise is exactly 0 (air under both rules)
        dens = O.chunk_fill(td, cx, cz, cy, N).reshape(N, N, N).transpose(1, 0, 2)
        assert ((ids == 1) == ((dens == 0) & (nr != 0))).all()
    # the reference's parameters give a mixed chunk near the origin
    ids0 = O.chunk_fill(tr, 0, 0, 0, 32)
    assert 0 < ids0.sum() < ids0.size


def test_reference_rule_noise_vs_numpy(oracle_mod):
    """The reference rule's noise block against the independent NumPy
    implementation evaluated at (wx, wz, wy)."""
    O = oracle_mod
    t = O.terrain(kind=O.TERRAIN_REFERENCE, **REF_PARAMS)
    pn = np_noise.permutation(REF_PARAMS["seed"])
    for c, N in [((2, -1, 4), 16), ((0, 0, 0), 32)]:
        o = np.array(c) * N
        y, z, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        ref = np_noise.fbm3(pn, (x + o[0]).astype(np.float32), (z + o[2]).astype(np.float32),
                            (y + o[1]).astype(np.float32), REF_PARAMS["octaves"],
                            REF_PARAMS["frequency"], REF_PARAMS["lacunarity"],
                            REF_PARAMS["gain"]).reshape(-1)
        assert (bits(O.noise_block(t, *c, N)) == bits(ref)).all()


def test_air_and_solid_kinds(oracle_mod):
    assert oracle_mod.chunk_fill(oracle_mod.terrain(kind=0), 0, 0, 0, 16).sum() == 0
    assert (oracle_mod.chunk_fill(oracle_mod.terrain(kind=1), 3, 1, 2, 16) == 1).all()


def test_occupancy_bits(oracle_mod):
    rng = np.random.default_rng(9)
    for N in (16, 32, 64):
        ids = (rng.random(N ** 3) < 0.4).astype(np.uint8)
        occ = oracle_mod.occupancy(ids, N)
        W = (N + 31) // 32
        grid = ids.reshape(N, N, N)  # [y][z][x]
        for (y, z) in [(0, 0), (N - 1, N - 1), (3, 7)]:
            for x in range(N):
                bit = (int(occ[(y * N + z) * W + x // 32]) >> (x % 32)) & 1
                assert bit == grid[y, z, x]


# ---------------------------------------------------------------- meshing
def brute_faces(ids, N, origin, adj):
    """Independent formulation of the face rule: pad the chunk with its
    neighbours' facing cells (absent neighbour or vertical outside -> air),
    then emit, voxel by voxel in y, z, x order, each of L R B F D U whose
    padded neighbour cell is air."""
    g = ids.reshape(N, N, N)  # [y][z][x]
    pad = np.zeros((N + 2, N + 2, N + 2), np.uint8)  # [y][z][x] with 1-cell halo
    pad[1:-1, 1:-1, 1:-1] = g
    if adj[0] is not None:
        pad[1:-1, 1:-1, 0] = adj[0].reshape(N, N, N)[:, :, N - 1]
    if adj[1] is not None:
        pad[1:-1, 1:-1, N + 1] = adj[1].reshape(N, N, N)[:, :, 0]
    if adj[2] is not None:
        pad[1:-1, 0, 1:-1] = adj[2].reshape(N, N, N)[:, N - 1, :]
    if adj[3] is not None:
        pad[1:-1, N + 1, 1:-1] = adj[3].reshape(N, N, N)[:, 0, :]
    steps = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]  # (dy,dz,dx)
    tex = {1: [0, 0, 0, 0, 1, 2], 2: [1, 1, 1, 1, 1, 1]}
    out = []
    for y in range(N):
        for z in range(N):
            for x in range(N):
                b = g[y, z, x]
                if b == 0:
                    continue
                for d, (dy, dz, dx) in enumerate(steps):
                    if pad[y + 1 + dy, z + 1 + dz, x + 1 + dx] == 0:
                        out.append((x + origin[0], y + origin[1], z + origin[2], d, tex[b][d]))
    return out


def as_tuples(faces):
    return [(int(f["x"]), int(f["y"]), int(f["z"]), int(f["dir"]), int(f["tex"])) for f in faces]


def test_single_voxel_six_faces_in_order(oracle_mod):
    N = 8
    ids = np.zeros(N ** 3, np.uint8)
    ids[3 + N * 4 + N * N * 5] = 1
    f = oracle_mod.mesh_faces(ids, N, (80, -16, 8))
    assert as_tuples(f) == [(83, -11, 12, d, t) for d, t in enumerate([0, 0, 0, 0, 1, 2])]


@pytest.mark.parametrize("N", [2, 16, 32])
def test_closed_form_counts(oracle_mod, N):
    solid = np.ones(N ** 3, np.uint8)
    air = np.zeros(N ** 3, np.uint8)
    assert len(oracle_mod.mesh_faces(air, N, (0, 0, 0))) == 0
    assert len(oracle_mod.mesh_faces(solid, N, (0, 0, 0))) == 6 * N * N
    assert len(oracle_mod.mesh_faces(solid, N, (0, 0, 0), (solid,) * 4)) == 2 * N * N
    # absent neighbours and air neighbours are the same for the face rule
    assert len(oracle_mod.mesh_faces(solid, N, (0, 0, 0), (air,) * 4)) == 6 * N * N
    y, z, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    checker = (((x + y + z) % 2) == 0).astype(np.uint8).reshape(-1)
    assert len(oracle_mod.mesh_faces(checker, N, (0, 0, 0))) == 3 * N ** 3
    assert len(oracle_mod.mesh_faces(checker, N, (0, 0, 0), (checker,) * 4)) == 3 * N ** 3


def test_exhaustive_n2_brute_force(oracle_mod):
    """All 256 occupancies of a 2^3 chunk x all 16 neighbour-presence
    patterns (neighbour contents seeded), incl. dirt blocks."""
    N = 2
    rng = np.random.default_rng(2)
    for occ in range(256):
        ids = np.array([(occ >> i) & 1 for i in range(8)], np.uint8)
        ids = np.where(ids == 1, rng.integers(1, 3, size=8), 0).astype(np.uint8)
        for pres in range(16):
            adj = tuple(rng.integers(0, 3, size=8).astype(np.uint8) if (pres >> k) & 1 else None
                        for k in range(4))
            origin = tuple(int(v) for v in rng.integers(-40, 40, size=3))
            got = as_tuples(oracle_mod.mesh_faces(ids, N, origin, adj))
            assert got == brute_faces(ids, N, origin, adj)


def test_random_chunks_vs_brute_force(oracle_mod):
    rng = np.random.default_rng(4)
    for N in (4, 8, 16):
        for _ in range(4):
            ids = (rng.random(N ** 3) < rng.uniform(0.1, 0.9)).astype(np.uint8)
            adj = tuple((rng.random(N ** 3) < 0.5).astype(np.uint8) if rng.random() < 0.6 else None
                        for _ in range(4))
            got = as_tuples(oracle_mod.mesh_faces(ids, N, (N, 0, -N), adj))
            assert got == brute_faces(ids, N, (N, 0, -N), adj)


# ---------------------------------------------------------------- quads
def test_rotations_orthonormal_and_face_normals(oracle_mod):
    for d, n in NORMALS.items():
        R = oracle_mod.rotation(d).astype(np.float64)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-6
        assert abs(np.linalg.det(R) - 1) < 1e-6
        assert np.abs(R @ np.array([0, 0, 1.0]) - np.array(n)).max() < 1e-6


def test_instance_translation_and_vertices(oracle_mod):
    """Random faces: the translation column is the voxel, the texture index
    passes through, every vertex equals the independent evaluation bit for
    bit, sits on the face's side of the voxel, and the uv and index patterns
    are quad.rs:66-89's."""
    rng = np.random.default_rng(6)
    pts = rng.integers(-5000, 5000, size=(40, 3))
    faces = unit_faces(oracle_mod, pts)
    faces["tex"] = 7
    inst, vert, idx = oracle_mod.expand(faces)
    assert (inst["tex"] == 7).all()
    for i, f in enumerate(faces):
        m = inst[i]["m"]  # [col][row]
        assert tuple(m[3][:3]) == (float(f["x"]), float(f["y"]), float(f["z"]))
        assert m[3][3] == 1.0 and (m[:3, 3] == 0).all()
        p = (int(f["x"]), int(f["y"]), int(f["z"]))
        _, _, ev = expected_vertices(int(f["dir"]), p)
        corners = vert[4 * i:4 * i + 4]
        assert (bits(corners["pos"]) == bits(ev)).all()
        # the face lies on the voxel's side named by its direction
        n = np.array(NORMALS[int(f["dir"])], np.float64)
        rel = corners["pos"].astype(np.float64) - np.array(p, np.float64)
        assert (rel @ n == 0.5).all()
        assert [tuple(v["uv"]) for v in corners] == [(0, 0), (1, 0), (0, 1), (1, 1)]
        assert len({tuple(np.round(r, 3)) for r in rel}) == 4
    assert (idx.reshape(-1, 6) == (np.arange(len(faces))[:, None] * 4 + [0, 2, 1, 3, 1, 2])).all()


# Exact bits of Quad::new's rotations (quad.rs:14-46).  Rust's
# f32::to_radians is deg * (PI_f32 / 180_f32) in f32; sin and cos of that f32
# angle, correctly rounded to f32, are computed here from Python's f64
# math.sin / math.cos (each within 1 f64 ulp of the true value), after
# checking that no f64 result lies within a few f64 ulps of an f32 rounding
# midpoint (so the f32 rounding of the true value is the one taken).
F32_DIRS = {5: ("roll", -90.0), 4: ("roll", 90.0), 0: ("pitch", -90.0), 1: ("pitch", 90.0),
            3: ("none", 0.0), 2: ("pitch", 180.0)}  # UP DOWN LEFT RIGHT FRONT BACK


def f32_correctly_rounded(v: float) -> np.float32:
    r = np.float32(v)
    if float(r) != v:
        other = np.nextafter(r, np.float32(math.copysign(np.inf, v - float(r))))
        mid = (float(r) + float(other)) / 2
        assert abs(v - mid) > 8 * math.ulp(v), "too close to an f32 midpoint to decide"
    return r


def f32_sin_cos_deg(deg: float):
    k = np.float32(np.float32(math.pi) / np.float32(180.0))
    a = np.float32(np.float32(deg) * k)
    return f32_correctly_rounded(math.sin(float(a))), f32_correctly_rounded(math.cos(float(a)))


def textbook_rotation(d: int) -> np.ndarray:
    """from_euler_angles(roll, pitch, 0) of a pure roll (about x) or pure
    pitch (about y): Rx = [[1,0,0],[0,c,-s],[0,s,c]], Ry = [[c,0,s],[0,1,0],
    [-s,0,c]], with the correctly rounded f32 (s, c).  Every product of the
    general formula with the other angle's (sin, cos) = (0, 1) reduces to one
    of these entries exactly, whatever its grouping; only the signs of the
    zero entries depend on nalgebra's expression order."""
    axis, deg = F32_DIRS[d]
    s, c = f32_sin_cos_deg(deg)
    one, zero = np.float32(1), np.float32(0)
    if axis == "roll":
        return np.array([[one, zero, zero], [zero, c, -s], [zero, s, c]], np.float32)
    if axis == "pitch":
        return np.array([[c, zero, s], [zero, one, zero], [-s, zero, c]], np.float32)
    return np.eye(3, dtype=np.float32)


def test_rotation_constants_are_the_f32_sin_cos():
    """The closed-form values the rotation tables hold: sin/cos of the f32
    angles of +-90 and 180 degrees."""
    assert f32_sin_cos_deg(90.0) == (np.float32(1.0), np.float32(-4.371139e-08))
    assert f32_sin_cos_deg(-90.0) == (np.float32(-1.0), np.float32(-4.371139e-08))
    assert f32_sin_cos_deg(180.0) == (np.float32(-8.742278e-08), np.float32(-1.0))
    assert bits(f32_sin_cos_deg(90.0)[1]) == 0xB33BBD2E


def test_rotation_tables_bit_exact(oracle_mod):
    """Every non-zero entry of the oracle's six rotations equals the
    textbook matrix's bit for bit (no snapped cos(90) = 0, no transposed or
    sign-flipped term); every zero entry is a zero of either sign."""
    for d in range(6):
        R = oracle_mod.rotation(d)
        E = textbook_rotation(d)
        nz = E != 0
        assert (bits(R[nz]) == bits(E[nz])).all(), f"dir {d}"
        assert (R[~nz] == 0).all(), f"dir {d}"
        assert np.count_nonzero(R) == np.count_nonzero(E)


def expected_vertices(d: int, p) -> tuple:
    """InstanceRaw matrix (column-major [col][row]) and the four vertex
    positions of a face of direction d at voxel p, evaluated independently
    of the oracle: M = T * R with the textbook R, then vertex.vert:24's
    instance * (corner, 1) one f32 operation at a time, left to right."""
    R = textbook_rotation(d)
    t = [np.float32(v) for v in p]
    corners = [(-0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (-0.5, -0.5, 0.5), (0.5, -0.5, 0.5)]
    verts = []
    for cx, cy, cz in corners:
        pos = []
        for r in range(3):
            acc = R[r, 0] * np.float32(cx)
            acc = acc + R[r, 1] * np.float32(cy)
            acc = acc + R[r, 2] * np.float32(cz)
            acc = acc + t[r] * np.float32(1)
            pos.append(np.float32(acc))
        verts.append(pos)
    return R, t, np.array(verts, np.float32)


def unit_faces(oracle_mod, pts):
    faces = np.zeros(len(pts) * 6, oracle_mod.FACE_DTYPE)
    for i, p in enumerate(pts):
        for d in range(6):
            f = faces[6 * i + d]
            f["x"], f["y"], f["z"], f["dir"], f["tex"] = p[0], p[1], p[2], d, d
    return faces


def test_instances_and_vertices_bit_exact_near_zero(oracle_mod):
    """Faces at world coordinates -1, 0 and +1 on every axis, in all six
    directions: where the translation is 0 or +-1 the ~4.4e-8 rotation terms
    reach the vertices' low bits (0.5 - 2.2e-8 rounds to 0.49999997).  The
    oracle's InstanceRaw non-zero entries and vertex positions equal the
    independent evaluation bit for bit; only zero signs are left open."""
    pts = list(itertools.product((-1, 0, 1), repeat=3)) + [(1000, -77, 5), (-3, 2048, -9)]
    faces = unit_faces(oracle_mod, pts)
    inst, vert, _ = oracle_mod.expand(faces)
    low_bits_seen = False
    for i, f in enumerate(faces):
        d = int(f["dir"])
        R, t, ev = expected_vertices(d, (int(f["x"]), int(f["y"]), int(f["z"])))
        m = inst[i]["m"]  # [col][row]
        E = np.zeros((4, 4), np.float32)
        E[:3, :3] = R.T
        E[3, :3] = t
        E[3, 3] = 1
        nz = E != 0
        assert (bits(m[nz]) == bits(E[nz])).all(), (i, d)
        assert (m[~nz] == 0).all(), (i, d)
        got = vert[4 * i:4 * i + 4]["pos"]
        assert (bits(got) == bits(ev)).all(), (i, d, got, ev)
        low_bits_seen |= bool((np.abs(ev - np.round(ev * 2) / 2) > 0).any())
    assert low_bits_seen  # the cases do exercise the small rotation terms


# ---------------------------------------------------------------- frustum

def test_frustum_closed_forms(oracle_mod):
    vp = perspective(np.radians(70), 1.0, 0.1, 1024) @ look_at((0, 0, 0), (0, 0, -1))
    pl = oracle_mod.frustum_planes(colmajor(vp))
    assert oracle_mod.frustum_contains(pl, [(-1, -1, -3)])[0] == 1   # straight ahead
    assert oracle_mod.frustum_contains(pl, [(-1, -1, 2)])[0] == 0    # behind
    assert oracle_mod.frustum_contains(pl, [(-1, -1, -40)])[0] == 0  # beyond far
    # camera inside a chunk, narrow view: all 8 corners outside -> false
    vp2 = perspective(np.radians(20), 1.0, 0.1, 1024) @ look_at((16, 16, 16), (16, 16, -100))
    pl2 = oracle_mod.frustum_planes(colmajor(vp2))
    assert oracle_mod.frustum_contains(pl2, [(0, 0, 0)])[0] == 0


def test_frustum_vs_f64_brute_force(oracle_mod):
    rng = np.random.default_rng(8)
    checked = 0
    for _ in range(10):
        eye = rng.uniform(-200, 200, 3)
        tgt = eye + rng.normal(size=3)
        m = colmajor(perspective(np.radians(rng.uniform(40, 100)), rng.uniform(0.5, 2), 0.1,
                                 rng.uniform(100, 1024)) @ look_at(eye, tgt))
        pl = oracle_mod.frustum_planes(m)
        M = m.astype(np.float64).reshape(4, 4)  # [col][row]
        rows = [M[:, r] for r in range(4)]
        planes = np.array([rows[3] + rows[2], rows[3] - rows[2], rows[3] + rows[0],
                           rows[3] - rows[0], rows[3] - rows[1], rows[3] + rows[1]])
        coords = rng.integers(-12, 12, size=(300, 3))
        got = oracle_mod.frustum_contains(pl, coords)
        for c, g in zip(coords, got):
            corners = np.array([[c[0] * 32 + dx, c[1] * 32 + dy, c[2] * 32 + dz, 1.0]
                                for dx, dz, dy in [(0, 0, 0), (32, 0, 0), (0, 32, 0), (32, 32, 0),
                                                   (0, 0, 32), (32, 0, 32), (0, 32, 32), (32, 32, 32)]])
            dist = corners @ planes.T  # [corner][plane]
            scale = np.abs(corners) @ np.abs(planes).T + 1
            if (np.abs(dist) < 1e-4 * scale).any():
                continue  # a tie: float rounding decides
            checked += 1
            assert g == int((dist >= 0).all(axis=1).any())
    assert checked > 2000


# ---------------------------------------------------------------- driver
def test_batch_driver_counts(oracle_mod):
    import voxel_amd as V  # host helper only: builds the neighbour table
    t = oracle_mod.terrain()
    coords = np.array([(x, 0, z) for x in range(3) for z in range(3)], np.int32)
    nb = V.grid_neighbours(3, 3)
    total, per = oracle_mod.batch_run(t, coords, nb, 16, 4)
    ids = [oracle_mod.chunk_fill(t, *c, 16) for c in coords]
    for i, c in enumerate(coords):
        adj = tuple(ids[j] if j >= 0 else None for j in nb[i])
        assert per[i] == len(oracle_mod.mesh_faces(ids[i], 16, tuple(c * 16), adj))
    assert total == per.sum()


def test_batch_driver_meshes_match_per_chunk(oracle_mod):
    """The driver's kept meshes (the full-size GPU comparison's reference)
    equal the per-chunk restatement at the bases given, in a shuffled order
    of bases; bases that do not fit are refused."""
    import voxel_amd as V  # host helper only: builds the neighbour table
    O = oracle_mod
    t = O.terrain()
    coords = np.array([(x, 0, z) for x in range(3) for z in range(3)], np.int32)
    nb = V.grid_neighbours(3, 3)
    total, per = O.batch_run(t, coords, nb, 16, 4)
    # lay the chunks out in reverse request order to show bases are honoured
    order = np.arange(len(coords))[::-1]
    bases = np.zeros(len(coords), np.int64)
    run = 0
    for i in order:
        bases[i] = run
        run += per[i]
    tot2, per2, vert, idx = O.batch_mesh(t, coords, nb, 16, 4, bases, int(total))
    assert tot2 == total and (per2 == per).all()
    ids = [O.chunk_fill(t, *c, 16) for c in coords]
    for i, c in enumerate(coords):
        adj = tuple(ids[j] if j >= 0 else None for j in nb[i])
        _, _, v, ix = O.mesh_buffers(ids[i], 16, tuple(c * 16), adj)
        f0, n = int(bases[i]), int(per[i])
        assert (vert[4 * f0:4 * (f0 + n)].view(np.uint32) == v.view(np.uint32)).all()
        assert (idx[6 * f0:6 * (f0 + n)] == ix).all()
    with pytest.raises(ValueError):
        O.batch_mesh(t, coords, nb, 16, 4, bases, int(total) - 1)
