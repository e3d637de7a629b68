"""ctypes binding of the CPU oracle (oracle/liboracle_voxel.so).

TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline.  Only the
tests, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this module.  The product library never calls into it.

Every function here is a thin wrapper; the restated algorithm (and the
reference file:line each part follows) lives in ``voxel_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_voxel.so")

AIR, GRASS, DIRT = 0, 1, 2
LEFT, RIGHT, BACK, FRONT, DOWN, UP = range(6)
TERRAIN_AIR, TERRAIN_SOLID, TERRAIN_HEIGHTMAP, TERRAIN_DENSITY, TERRAIN_REFERENCE = range(5)


class Terrain(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("seed", C.c_uint32),
        ("octaves", C.c_int32),
        ("frequency", C.c_float),
        ("lacunarity", C.c_float),
        ("gain", C.c_float),
        ("height_scale", C.c_float),
        ("height_max", C.c_int32),
        ("y_scale", C.c_float),
    ]


FACE_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("dir", "<u4"), ("tex", "<u4")])
INSTANCE_DTYPE = np.dtype([("m", "<f4", (4, 4)), ("tex", "<u4")])  # 68 B, m[col][row]
VERTEX_DTYPE = np.dtype([("pos", "<f4", (3,)), ("uv", "<f4", (2,))])  # 20 B
assert INSTANCE_DTYPE.itemsize == 68 and VERTEX_DTYPE.itemsize == 20

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc; no GPU needed)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        f32, i32, u32, i64, u64 = C.c_float, C.c_int32, C.c_uint32, C.c_int64, C.c_uint64
        sig = {
            "vxo_splitmix64_next": (u64, [C.POINTER(u64)]),
            "vxo_perm": (None, [u64, P]),
            "vxo_perm_from_table": (None, [P, P]),
            "vxo_noise3": (f32, [P, f32, f32, f32]),
            "vxo_noise3d": (C.c_double, [P, C.c_double, C.c_double, C.c_double]),
            "vxo_noise2": (f32, [P, f32, f32]),
            "vxo_fbm3": (f32, [P, f32, f32, f32, C.c_int, f32, f32, f32]),
            "vxo_fbm2": (f32, [P, f32, f32, C.c_int, f32, f32, f32]),
            "vxo_fbm3_row8": (None, [P, i32, i32, i32, C.c_int, f32, f32, f32, P]),
            "vxo_have_avx2": (C.c_int, []),
            "vxo_chunk_fill": (None, [C.POINTER(Terrain), i32, i32, i32, i32, P]),
            "vxo_chunk_fill_fast": (None, [C.POINTER(Terrain), i32, i32, i32, i32, P]),
            "vxo_occupancy": (None, [P, i32, P]),
            "vxo_noise_block": (None, [C.POINTER(Terrain), i32, i32, i32, i32, P]),
            "vxo_mesh_faces": (i64, [P, i32, i32, i32, i32, P, P, i64]),
            "vxo_texture_layer": (u32, [C.c_uint8, u32]),
            "vxo_rotation": (None, [u32, P]),
            "vxo_instance_of": (None, [P, P]),
            "vxo_vertices_of": (None, [P, P]),
            "vxo_indices_of": (None, [i64, P]),
            "vxo_expand": (None, [P, i64, P, P, P]),
            "vxo_frustum_planes": (None, [P, P]),
            "vxo_frustum_contains": (C.c_int, [P, i32, i32, i32, i32]),
            "vxo_batch_run": (i64, [C.POINTER(Terrain), P, P, i32, i32, C.c_int, P]),
            "vxo_batch_mesh": (i64, [C.POINTER(Terrain), P, P, i32, i32, C.c_int, P, P, P, i64, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def terrain(kind=TERRAIN_DENSITY, seed=7, octaves=4, frequency=1.0 / 64, lacunarity=2.0,
            gain=0.5, height_scale=16.0, height_max=15, y_scale=1.0 / 128) -> Terrain:
    return Terrain(kind, seed, octaves, frequency, lacunarity, gain, height_scale, height_max,
                   y_scale)


def perm(seed: int) -> np.ndarray:
    out = np.zeros(512, np.uint8)
    lib().vxo_perm(seed, _p(out))
    return out


def perm_from_table(table) -> np.ndarray:
    t = np.ascontiguousarray(np.asarray(table, np.uint8))
    out = np.zeros(512, np.uint8)
    lib().vxo_perm_from_table(_p(t), _p(out))
    return out


def splitmix64(seed: int, n: int):
    s = C.c_uint64(seed)
    return [lib().vxo_splitmix64_next(C.byref(s)) for _ in range(n)]


def noise3(p, x, y, z) -> float:
    return lib().vxo_noise3(_p(p), x, y, z)


def noise3d(p, x, y, z) -> float:
    return lib().vxo_noise3d(_p(p), x, y, z)


def noise2(p, x, y) -> float:
    return lib().vxo_noise2(_p(p), x, y)


def fbm3(p, x, y, z, octaves, f0, lac, gain) -> float:
    return lib().vxo_fbm3(_p(p), x, y, z, octaves, f0, lac, gain)


def fbm2(p, x, y, octaves, f0, lac, gain) -> float:
    return lib().vxo_fbm2(_p(p), x, y, octaves, f0, lac, gain)


def fbm3_row8(p, x0, y, z, octaves, f0, lac, gain) -> np.ndarray:
    out = np.zeros(8, np.float32)
    lib().vxo_fbm3_row8(_p(p), x0, y, z, octaves, f0, lac, gain, _p(out))
    return out


def have_avx2() -> bool:
    return bool(lib().vxo_have_avx2())


def chunk_fill(t: Terrain, cx: int, cy: int, cz: int, edge: int, fast: bool = False) -> np.ndarray:
    ids = np.zeros(edge ** 3, np.uint8)
    fn = lib().vxo_chunk_fill_fast if fast else lib().vxo_chunk_fill
    fn(C.byref(t), cx, cy, cz, edge, _p(ids))
    return ids


def noise_block(t: Terrain, cx: int, cy: int, cz: int, edge: int) -> np.ndarray:
    n = edge ** 3 if t.kind in (TERRAIN_DENSITY, TERRAIN_REFERENCE) else edge ** 2
    out = np.zeros(n, np.float32)
    lib().vxo_noise_block(C.byref(t), cx, cy, cz, edge, _p(out))
    return out


def occupancy(ids: np.ndarray, edge: int) -> np.ndarray:
    w = (edge + 31) // 32
    occ = np.zeros(edge * edge * w, np.uint32)
    lib().vxo_occupancy(_p(np.ascontiguousarray(ids, np.uint8)), edge, _p(occ))
    return occ


def max_faces(edge: int) -> int:
    return 3 * edge * edge * (edge + 1)


def mesh_faces(ids: np.ndarray, edge: int, origin, adj=(None, None, None, None)) -> np.ndarray:
    ids = np.ascontiguousarray(ids, np.uint8)
    keep = [None if a is None else np.ascontiguousarray(a, np.uint8) for a in adj]
    arr = (C.c_void_p * 4)(*[None if a is None else a.ctypes.data for a in keep])
    cap = max_faces(edge)
    out = np.zeros(cap, FACE_DTYPE)
    n = lib().vxo_mesh_faces(_p(ids), edge, origin[0], origin[1], origin[2],
                             C.cast(arr, C.c_void_p), _p(out), cap)
    assert n <= cap
    return out[:n].copy()


def rotation(d: int) -> np.ndarray:
    r = np.zeros((3, 3), np.float32)
    lib().vxo_rotation(d, _p(r))
    return r


def instances(faces: np.ndarray) -> np.ndarray:
    out = np.zeros(len(faces), INSTANCE_DTYPE)
    faces = np.ascontiguousarray(faces)
    for i in range(len(faces)):
        lib().vxo_instance_of(C.c_void_p(faces.ctypes.data + i * FACE_DTYPE.itemsize),
                              C.c_void_p(out.ctypes.data + i * INSTANCE_DTYPE.itemsize))
    return out


def vertices(inst: np.ndarray) -> np.ndarray:
    out = np.zeros(4 * len(inst), VERTEX_DTYPE)
    inst = np.ascontiguousarray(inst)
    for i in range(len(inst)):
        lib().vxo_vertices_of(C.c_void_p(inst.ctypes.data + i * INSTANCE_DTYPE.itemsize),
                              C.c_void_p(out.ctypes.data + 4 * i * VERTEX_DTYPE.itemsize))
    return out


def indices(nfaces: int) -> np.ndarray:
    out = np.zeros(6 * nfaces, np.uint32)
    for i in range(nfaces):
        lib().vxo_indices_of(i, C.c_void_p(out.ctypes.data + 24 * i))
    return out


def frustum_planes(m16) -> np.ndarray:
    m = np.ascontiguousarray(np.asarray(m16, np.float32).reshape(16))
    out = np.zeros(24, np.float32)
    lib().vxo_frustum_planes(_p(m), _p(out))
    return out


def frustum_contains(planes, coords, edge: int = 32) -> np.ndarray:
    planes = np.ascontiguousarray(planes, np.float32)
    coords = np.asarray(coords, np.int32).reshape(-1, 3)
    return np.array([lib().vxo_frustum_contains(_p(planes), int(c[0]), int(c[1]), int(c[2]), edge)
                     for c in coords], np.uint8)


def batch_run(t: Terrain, coords: np.ndarray, nb: np.ndarray | None, edge: int, threads: int):
    coords = np.ascontiguousarray(coords, np.int32).reshape(-1, 3)
    n = len(coords)
    nbp = None if nb is None else _p(np.ascontiguousarray(nb, np.int32))
    faces = np.zeros(n, np.int64)
    total = lib().vxo_batch_run(C.byref(t), _p(coords), nbp, n, edge, threads, _p(faces))
    return total, faces


def batch_mesh(t: Terrain, coords: np.ndarray, nb: np.ndarray | None, edge: int, threads: int,
               bases: np.ndarray, total_faces: int):
    """The batch driver keeping its meshes: vertex and index buffers of
    `total_faces` faces with chunk i's faces from bases[i], and the face
    count per chunk.  Raises if a chunk's range falls outside the buffers."""
    coords = np.ascontiguousarray(coords, np.int32).reshape(-1, 3)
    n = len(coords)
    bases = np.ascontiguousarray(bases, np.int64)
    assert len(bases) == n
    nbp = None if nb is None else _p(np.ascontiguousarray(nb, np.int32))
    vert = np.zeros(4 * total_faces, VERTEX_DTYPE)
    idx = np.zeros(6 * total_faces, np.uint32)
    faces = np.zeros(n, np.int64)
    total = lib().vxo_batch_mesh(C.byref(t), _p(coords), nbp, n, edge, threads, _p(bases),
                                 _p(vert), _p(idx), total_faces, _p(faces))
    if total < 0:
        raise ValueError("a chunk's faces fall outside the buffers for these bases")
    return total, faces, vert, idx


def expand(faces: np.ndarray):
    """InstanceRaw records, vertex buffer and chunk-local index buffer."""
    faces = np.ascontiguousarray(faces, FACE_DTYPE)
    n = len(faces)
    inst = np.zeros(n, INSTANCE_DTYPE)
    vert = np.zeros(4 * n, VERTEX_DTYPE)
    idx = np.zeros(6 * n, np.uint32)
    lib().vxo_expand(_p(faces), n, _p(inst), _p(vert), _p(idx))
    return inst, vert, idx


def mesh_buffers(ids, edge, origin, adj=(None, None, None, None)):
    """Full per-chunk mesh in every output format: faces, InstanceRaw records,
    vertex and index buffers (chunk-local indices)."""
    f = mesh_faces(ids, edge, origin, adj)
    inst, vert, idx = expand(f)
    return f, inst, vert, idx
