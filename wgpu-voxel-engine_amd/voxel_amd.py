"""Python host binding of libvoxel_hip.so (include/voxel.h) through ctypes.

Mirrors the reference's module surface so tests read like its call sites
(paths relative to the reference repository):

    ChunkCoord3D                 src/coordinate.rs:41-66
    Chunk.new / create_mesh      src/chunk.rs:21-57
    ChunkMesh                    src/chunk.rs:248-266
    Frustum.new / update / contains   src/frustum_culling.rs:10-45
    Batch (many pool.execute jobs)    src/chunk_builder.rs:80-84

Every call goes to the HIP library; there is no CPU fallback.  Loading fails
loudly when the library is missing, and creating a Context fails when no GPU
is present.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoxel_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "voxel.h")

OK, ERR_INVALID, ERR_NOMEM, ERR_DEVICE, ERR_CAPACITY, ERR_BUDGET, ERR_NOT_READY, ERR_NODEV = (
    0, -1, -2, -3, -4, -5, -6, -7)
TERRAIN_AIR, TERRAIN_SOLID, TERRAIN_HEIGHTMAP, TERRAIN_DENSITY, TERRAIN_REFERENCE = range(5)
FMT_INSTANCES, FMT_VERTICES = 1, 2
FMT_INDICES = 4  # vx_batch_read_mesh selector for the index buffer
BATCH_AUTO, BATCH_FUSED, BATCH_STAGED, BATCH_PIPELINED, BATCH_SPLIT = 0, 1, 2, 3, 4
AIR, GRASS, DIRT = 0, 1, 2

INSTANCE_DTYPE = np.dtype([("m", "<f4", (4, 4)), ("tex", "<u4")])  # InstanceRaw, 68 B
VERTEX_DTYPE = np.dtype([("pos", "<f4", (3,)), ("uv", "<f4", (2,))])  # Vertex, 20 B


class VoxelError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        msg = f"{what} failed: status {status}"
        if _lib is not None:
            msg += f" ({_lib.vx_status_string(status).decode()})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


class Coord(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class TerrainC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint32), ("octaves", C.c_int32),
                ("frequency", C.c_float), ("lacunarity", C.c_float), ("gain", C.c_float),
                ("height_scale", C.c_float), ("height_max", C.c_int32), ("y_scale", C.c_float)]


class CtxDesc(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32), ("hbm_budget", C.c_uint64)]


class FrustumC(C.Structure):
    _fields_ = [("planes", C.c_float * 24)]


class Request(C.Structure):
    _fields_ = [("coord", Coord), ("edge", C.c_int32), ("terrain_id", C.c_int32),
                ("neighbours", C.c_uint32), ("reserved", C.c_uint32), ("arrival_ns", C.c_uint64)]


class Completion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("coord", Coord), ("edge", C.c_int32),
                ("status", C.c_int32), ("cache_hit", C.c_uint32), ("faces", C.c_uint64),
                ("arrival_ns", C.c_uint64), ("done_ns", C.c_uint64),
                ("nb_present", C.c_int32 * 4)]


class ServerDesc(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("max_inflight", C.c_int32), ("formats", C.c_uint32),
                ("flags", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "submitted", "completed", "cache_hits", "coalesced", "generated", "batches", "evictions",
        "failed", "resident_chunks", "resident_bytes", "peak_resident_bytes", "budget_bytes")]


class SimDesc(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("rate_per_s", C.c_double), ("duration_s", C.c_double),
                ("speed_blocks_per_s", C.c_double), ("radius_chunks", C.c_int32),
                ("n_classes", C.c_int32), ("edges", C.c_int32 * 3), ("weights", C.c_float * 3),
                ("terrain_id", C.c_int32 * 3), ("neighbours", C.c_uint32), ("server", ServerDesc)]


class SimResult(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "requests", "completed", "cache_hits", "generated", "batches", "evictions", "faces")] + \
        [(n, C.c_double) for n in (
            "wall_s", "throughput_rps", "generated_per_s", "p50_ms", "p99_ms", "max_ms", "mean_ms",
            "miss_p50_ms", "miss_p99_ms", "miss_max_ms")] + [("peak_resident_bytes", C.c_uint64)]


class TuneResult(C.Structure):
    _fields_ = [("mode", C.c_int32), ("grid", C.c_int32), ("best_ms", C.c_float),
                ("n_candidates", C.c_int32), ("cand_mode", C.c_int32 * 12),
                ("cand_grid", C.c_int32 * 12), ("cand_ms", C.c_float * 12),
                ("ratio", C.c_int32), ("cand_ratio", C.c_int32 * 12),
                ("fill_split", C.c_int32), ("emit_split", C.c_int32),
                ("cand_fill_split", C.c_int32 * 12), ("cand_emit_split", C.c_int32 * 12)]


class WorldDesc(C.Structure):
    _fields_ = [("terrain_id", C.c_int32), ("render_distance", C.c_int32),
                ("result_delay", C.c_int32), ("formats", C.c_uint32),
                ("player_pos", C.c_float * 3)]


class WorldStats(C.Structure):
    _fields_ = [("tick", C.c_uint64), ("chunks", C.c_int32), ("meshes", C.c_int32),
                ("rebuild_queue", C.c_int32), ("in_process", C.c_int32),
                ("jobs_loaded", C.c_uint64), ("jobs_rebuilt", C.c_uint64),
                ("received_chunks", C.c_uint64), ("received_meshes", C.c_uint64),
                ("unloaded", C.c_uint64), ("player_chunk", Coord)]


class BatchReq(C.Structure):
    _fields_ = [("coord", Coord), ("edge", C.c_int32), ("terrain_id", C.c_int32),
                ("nb", C.c_int32 * 4)]


_lib = None
P = C.c_void_p
_SIGS = {
    "vx_abi_version": (C.c_int, []),
    "vx_status_string": (C.c_char_p, [C.c_int]),
    "vx_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "vx_ctx_create": (C.c_int, [C.POINTER(CtxDesc), C.POINTER(P)]),
    "vx_ctx_destroy": (C.c_int, [P]),
    "vx_ctx_synchronize": (C.c_int, [P]),
    "vx_ctx_stream": (C.c_int, [P, C.POINTER(P)]),
    "vx_ctx_memory": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vx_ctx_last_error": (C.c_char_p, [P]),
    "vx_terrain_register": (C.c_int, [P, C.POINTER(TerrainC), C.POINTER(C.c_int32)]),
    "vx_buffer_alloc": (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
    "vx_buffer_free": (C.c_int, [P, P]),
    "vx_buffer_upload": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_buffer_download": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_buffer_copy": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_chunk_new": (C.c_int, [P, C.c_int32, Coord, C.c_int32, C.POINTER(P)]),
    "vx_chunk_from_ids": (C.c_int, [P, Coord, C.c_int32, P, C.POINTER(P)]),
    "vx_chunk_retain": (C.c_int, [P]),
    "vx_chunk_release": (C.c_int, [P]),
    "vx_chunk_info": (C.c_int, [P, C.POINTER(Coord), C.POINTER(C.c_int32)]),
    "vx_chunk_read_ids": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_chunk_read_occupancy": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_chunk_noise": (C.c_int, [P, C.c_int32, Coord, C.c_int32, P, C.c_uint64]),
    "vx_chunk_create_mesh": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(P)]),
    "vx_mesh_retain": (C.c_int, [P]),
    "vx_mesh_release": (C.c_int, [P]),
    "vx_mesh_face_count": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "vx_mesh_read_instances": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_mesh_read_vertices": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_mesh_read_indices": (C.c_int, [P, P, P, C.c_uint64]),
    "vx_mesh_device_buffers": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "vx_frustum_from_matrix": (C.c_int, [P, C.POINTER(FrustumC)]),
    "vx_frustum_contains": (C.c_int, [C.POINTER(FrustumC), Coord, C.c_int32]),
    "vx_frustum_mask": (C.c_int, [P, C.POINTER(FrustumC), P, C.c_int32, C.c_int32, P]),
    "vx_frustum_mask_device": (C.c_int, [P, C.POINTER(FrustumC), P, C.c_int32, C.c_int32, P]),
    "vx_frustum_draw_list": (C.c_int, [P, C.POINTER(FrustumC), P, P, P, C.c_int32, C.c_int32,
                                       C.c_uint32, P, P]),
    "vx_frustum_draw_list_noalloc": (C.c_int, [P, C.POINTER(FrustumC), P, P, P, C.c_int32,
                                               C.c_int32, C.c_uint32, P, P]),
    "vx_batch_draw_list": (C.c_int, [P, P, C.POINTER(FrustumC), C.c_uint32, P, P]),
    "vx_batch_create": (C.c_int, [P, P, C.c_int32, C.c_uint32, C.c_uint64, C.POINTER(P)]),
    "vx_batch_create_ext": (C.c_int, [P, P, C.c_int32, P, C.c_int32, C.c_uint32, C.c_uint64,
                                      C.POINTER(P)]),
    "vx_batch_chunk": (C.c_int, [P, C.c_int32, C.POINTER(P)]),
    "vx_batch_set_cull": (C.c_int, [P, C.POINTER(FrustumC), P, C.c_int32, C.c_int32, P]),
    "vx_batch_destroy": (C.c_int, [P]),
    "vx_batch_run": (C.c_int, [P, P]),
    "vx_batch_set_mode": (C.c_int, [P, C.c_int32]),
    "vx_batch_set_grid": (C.c_int, [P, C.c_int32]),
    "vx_batch_set_roles": (C.c_int, [P, C.c_int32]),
    "vx_batch_set_splits": (C.c_int, [P, C.c_int32, C.c_int32]),
    "vx_batch_get_mode": (C.c_int, [P, C.POINTER(C.c_int32)]),
    "vx_batch_run_per_chunk": (C.c_int, [P, P]),
    "vx_batch_wait": (C.c_int, [P, P, C.POINTER(C.c_uint64)]),
    "vx_batch_chunk_faces": (C.c_int, [P, P, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vx_batch_read_ids": (C.c_int, [P, P, C.c_int32, P, C.c_uint64]),
    "vx_batch_read_mesh": (C.c_int, [P, P, C.c_uint32, P, C.c_uint64]),
    "vx_batch_device_buffers": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "vx_batch_profile": (C.c_int, [P, P, C.c_int32, C.POINTER(C.c_float)]),
    "vx_batch_autotune": (C.c_int, [P, P, C.c_int32, C.POINTER(TuneResult)]),
    "vx_server_create": (C.c_int, [P, C.POINTER(ServerDesc), C.POINTER(P)]),
    "vx_server_destroy": (C.c_int, [P]),
    "vx_server_submit": (C.c_int, [P, P, C.c_int32, P]),
    "vx_server_pump": (C.c_int, [P, C.c_int32]),
    "vx_server_poll": (C.c_int, [P, P, C.c_int32, C.POINTER(C.c_int32)]),
    "vx_server_drain": (C.c_int, [P]),
    "vx_server_lookup": (C.c_int, [P, Coord, C.c_int32, C.c_int32, C.POINTER(P), C.POINTER(P)]),
    "vx_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
    "vx_serve_simulate": (C.c_int, [P, C.POINTER(SimDesc), C.POINTER(SimResult)]),
    "vx_world_create": (C.c_int, [P, C.POINTER(WorldDesc), C.POINTER(P)]),
    "vx_world_destroy": (C.c_int, [P]),
    "vx_world_tick": (C.c_int, [P, P, C.POINTER(FrustumC)]),
    "vx_world_stats_get": (C.c_int, [P, C.POINTER(WorldStats)]),
    "vx_world_list": (C.c_int, [P, C.c_uint32, P, C.c_int32, C.POINTER(C.c_int32)]),
    "vx_world_lookup": (C.c_int, [P, Coord, C.POINTER(P), C.POINTER(P)]),
}


def header_functions() -> list[str]:
    """Names of every function declared in include/voxel.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(vx_[a-z0-9_]+)\s*\(",
                                 text, re.M)))


def lib():
    """Loads libvoxel_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(or python wgpu-voxel-engine_amd/build.py) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int32(0)
    lib().vx_device_count(C.byref(n))
    return n.value


@dataclass(frozen=True)
class Terrain:
    kind: int = TERRAIN_DENSITY
    seed: int = 7
    octaves: int = 4
    frequency: float = 1.0 / 64
    lacunarity: float = 2.0
    gain: float = 0.5
    height_scale: float = 16.0
    height_max: int = 15
    y_scale: float = 1.0 / 128

    def c(self) -> TerrainC:
        return TerrainC(self.kind, self.seed, self.octaves, self.frequency, self.lacunarity,
                        self.gain, self.height_scale, self.height_max, self.y_scale)


class Context:
    """A device context: stream, caching device heap with HBM budget, tables."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = lib()
        h = P()
        st = L.vx_ctx_create(C.byref(CtxDesc(device, 0, hbm_budget)), C.byref(h))
        if st != OK:
            raise VoxelError(st, "vx_ctx_create")
        self.h = h
        self.L = L
        self._terrains: dict[Terrain, int] = {}

    def check(self, st: int, what: str):
        if st != OK:
            raise VoxelError(st, what, self.L.vx_ctx_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.vx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        self.check(self.L.vx_ctx_synchronize(self.h), "vx_ctx_synchronize")

    def memory(self):
        used, budget = C.c_uint64(), C.c_uint64()
        self.check(self.L.vx_ctx_memory(self.h, C.byref(used), C.byref(budget)), "vx_ctx_memory")
        return used.value, budget.value

    def terrain(self, t: Terrain) -> int:
        if t not in self._terrains:
            tid = C.c_int32()
            self.check(self.L.vx_terrain_register(self.h, C.byref(t.c()), C.byref(tid)),
                       "vx_terrain_register")
            self._terrains[t] = tid.value
        return self._terrains[t]

    def noise(self, t: Terrain, coord, edge: int) -> np.ndarray:
        n = edge ** 2 if t.kind == TERRAIN_HEIGHTMAP else edge ** 3
        out = np.zeros(n, np.float32)
        self.check(self.L.vx_chunk_noise(self.h, self.terrain(t), Coord(*coord), edge, _p(out), n),
                   "vx_chunk_noise")
        return out


class DeviceBuffer:
    """A device allocation from the context heap."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx, self.nbytes, self.ptr = ctx, nbytes, P()
        ctx.check(ctx.L.vx_buffer_alloc(ctx.h, nbytes, C.byref(self.ptr)), "vx_buffer_alloc")

    @classmethod
    def of(cls, ctx: Context, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = cls(ctx, arr.nbytes)
        b.upload(arr)
        return b

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(self.ctx.L.vx_buffer_upload(self.ctx.h, self.ptr, _p(arr), arr.nbytes),
                       "vx_buffer_upload")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(self.ctx.L.vx_buffer_download(self.ctx.h, _p(out), self.ptr, out.nbytes),
                       "vx_buffer_download")
        return out

    def free(self):
        if self.ptr:
            self.ctx.L.vx_buffer_free(self.ctx.h, self.ptr)
            self.ptr = P()


class Chunk:
    """Chunk (src/chunk.rs:15-46): device-resident block ids + occupancy."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.h = ctx, handle

    @classmethod
    def new(cls, ctx: Context, terrain: Terrain, coord, edge: int = 32) -> "Chunk":
        h = P()
        ctx.check(ctx.L.vx_chunk_new(ctx.h, ctx.terrain(terrain), Coord(*coord), edge, C.byref(h)),
                  "vx_chunk_new")
        return cls(ctx, h)

    @classmethod
    def from_ids(cls, ctx: Context, coord, edge: int, ids: np.ndarray) -> "Chunk":
        ids = np.ascontiguousarray(ids, np.uint8)
        assert ids.size == edge ** 3
        h = P()
        ctx.check(ctx.L.vx_chunk_from_ids(ctx.h, Coord(*coord), edge, _p(ids), C.byref(h)),
                  "vx_chunk_from_ids")
        return cls(ctx, h)

    def release(self):
        if self.h:
            self.ctx.L.vx_chunk_release(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.release()
        except Exception:
            pass

    @property
    def edge(self) -> int:
        e = C.c_int32()
        self.ctx.L.vx_chunk_info(self.h, None, C.byref(e))
        return e.value

    def ids(self) -> np.ndarray:
        n = self.edge ** 3
        out = np.zeros(n, np.uint8)
        self.ctx.check(self.ctx.L.vx_chunk_read_ids(self.ctx.h, self.h, _p(out), n),
                       "vx_chunk_read_ids")
        return out

    def occupancy(self) -> np.ndarray:
        e = self.edge
        n = e * e * ((e + 31) // 32)
        out = np.zeros(n, np.uint32)
        self.ctx.check(self.ctx.L.vx_chunk_read_occupancy(self.ctx.h, self.h, _p(out), n),
                       "vx_chunk_read_occupancy")
        return out

    def create_mesh(self, adjacent=(None, None, None, None),
                    formats: int = FMT_INSTANCES | FMT_VERTICES) -> "ChunkMesh":
        arr = (P * 4)(*[None if a is None else a.h for a in adjacent])
        h = P()
        self.ctx.check(self.ctx.L.vx_chunk_create_mesh(self.ctx.h, self.h, arr, formats,
                                                       C.byref(h)), "vx_chunk_create_mesh")
        return ChunkMesh(self.ctx, h, formats)


class ChunkMesh:
    """ChunkMesh (src/chunk.rs:248-266): device-resident face records."""

    def __init__(self, ctx: Context, handle, formats: int):
        self.ctx, self.h, self.formats = ctx, handle, formats

    def release(self):
        if self.h:
            self.ctx.L.vx_mesh_release(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.release()
        except Exception:
            pass

    @property
    def faces(self) -> int:
        n = C.c_uint64()
        self.ctx.L.vx_mesh_face_count(self.h, C.byref(n))
        return n.value

    def instances(self) -> np.ndarray:
        out = np.zeros(self.faces, INSTANCE_DTYPE)
        self.ctx.check(self.ctx.L.vx_mesh_read_instances(self.ctx.h, self.h, _p(out), out.nbytes),
                       "vx_mesh_read_instances")
        return out

    def vertices(self) -> np.ndarray:
        out = np.zeros(4 * self.faces, VERTEX_DTYPE)
        self.ctx.check(self.ctx.L.vx_mesh_read_vertices(self.ctx.h, self.h, _p(out), out.nbytes),
                       "vx_mesh_read_vertices")
        return out

    def indices(self) -> np.ndarray:
        out = np.zeros(6 * self.faces, np.uint32)
        self.ctx.check(self.ctx.L.vx_mesh_read_indices(self.ctx.h, self.h, _p(out), out.nbytes),
                       "vx_mesh_read_indices")
        return out


# vx_draw_indexed: one indirect indexed draw (WebGPU drawIndexedIndirect layout)
DRAW_DTYPE = np.dtype([("index_count", "<u4"), ("instance_count", "<u4"), ("first_index", "<u4"),
                       ("base_vertex", "<i4"), ("first_instance", "<u4")])
DRAW_INSTANCED, DRAW_EXPANDED = 1, 2


def _read_draws(ctx: "Context", args: "DeviceBuffer", cnt: "DeviceBuffer"):
    c = cnt.download(np.uint32, 2)
    if c[1]:
        raise VoxelError(ERR_INVALID, "draw list: an offset does not fit 32 bits")
    n = int(c[0])
    if n == 0:
        return np.zeros(0, DRAW_DTYPE)
    return args.download(np.uint8, n * DRAW_DTYPE.itemsize).view(DRAW_DTYPE)


class Frustum:
    """Frustum (src/frustum_culling.rs): planes from a view-projection matrix."""

    def __init__(self, matrix16):
        self.c = FrustumC()
        self.update(matrix16)

    def update(self, matrix16):
        m = np.ascontiguousarray(np.asarray(matrix16, np.float32).reshape(16))
        st = lib().vx_frustum_from_matrix(_p(m), C.byref(self.c))
        if st != OK:
            raise VoxelError(st, "vx_frustum_from_matrix")

    @property
    def planes(self) -> np.ndarray:
        return np.array(self.c.planes[:], np.float32)

    def contains(self, coord, edge: int = 32) -> bool:
        return bool(lib().vx_frustum_contains(C.byref(self.c), Coord(*coord), edge))

    def mask(self, ctx: Context, coords: np.ndarray, edge: int = 32) -> np.ndarray:
        coords = np.ascontiguousarray(np.asarray(coords, np.int32).reshape(-1, 3))
        out = np.zeros(len(coords), np.uint8)
        ctx.check(ctx.L.vx_frustum_mask(ctx.h, C.byref(self.c), _p(coords), len(coords), edge,
                                        _p(out)), "vx_frustum_mask")
        return out

    def draw_list(self, ctx: Context, coords: np.ndarray, faces: np.ndarray, first: np.ndarray,
                  edge: int = 32, mode: int = DRAW_INSTANCED) -> np.ndarray:
        """World::render's culled draw loop as indirect draw arguments (device-built)."""
        coords = np.ascontiguousarray(np.asarray(coords, np.int32).reshape(-1, 3))
        n = len(coords)
        bufs = []
        if n:
            bufs = [DeviceBuffer.of(ctx, coords),
                    DeviceBuffer.of(ctx, np.ascontiguousarray(faces, np.uint32)),
                    DeviceBuffer.of(ctx, np.ascontiguousarray(first, np.uint64))]
        ptrs = [b.ptr for b in bufs] if n else [None, None, None]
        args, cnt = DeviceBuffer(ctx, max(1, n) * DRAW_DTYPE.itemsize), DeviceBuffer(ctx, 8)
        try:
            ctx.check(ctx.L.vx_frustum_draw_list(ctx.h, C.byref(self.c), *ptrs, n, edge, mode,
                                                 args.ptr if n else None, cnt.ptr),
                      "vx_frustum_draw_list")
            return _read_draws(ctx, args, cnt)
        finally:
            for b in bufs + [args, cnt]:
                b.free()


class Batch:
    """Many Chunk::new + create_mesh requests as one graph-replayed launch
    sequence (fill -> count -> scan -> emit)."""

    def __init__(self, ctx: Context, coords, edges, terrain_ids, neighbours=None,
                 formats: int = FMT_VERTICES, face_capacity: int = 0, ext=None):
        coords = np.asarray(coords, np.int32).reshape(-1, 3)
        n = len(coords)
        edges = np.broadcast_to(np.asarray(edges, np.int32), (n,))
        tids = np.broadcast_to(np.asarray(terrain_ids, np.int32), (n,))
        nb = (np.full((n, 4), -1, np.int32) if neighbours is None
              else np.asarray(neighbours, np.int32).reshape(n, 4))
        reqs = (BatchReq * n)()
        for i in range(n):
            reqs[i].coord = Coord(*map(int, coords[i]))
            reqs[i].edge = int(edges[i])
            reqs[i].terrain_id = int(tids[i])
            for k in range(4):
                reqs[i].nb[k] = int(nb[i, k])
        h = P()
        ext = list(ext or [])
        ext_arr = (P * max(1, len(ext)))(*[c.h for c in ext])
        ctx.check(ctx.L.vx_batch_create_ext(ctx.h, C.cast(reqs, P), n, C.cast(ext_arr, P), len(ext),
                                            formats, face_capacity, C.byref(h)),
                  "vx_batch_create_ext")
        self.ctx, self.h, self.n, self.formats = ctx, h, n, formats
        self.edges = np.array(edges)
        self.total = None

    def close(self):
        if self.h:
            self.ctx.L.vx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.close()
        except Exception:
            pass

    def chunk(self, i: int) -> "Chunk":
        h = P()
        self.ctx.check(self.ctx.L.vx_batch_chunk(self.h, i, C.byref(h)), "vx_batch_chunk")
        return Chunk(self.ctx, h)

    def set_cull(self, frustum, coords: "DeviceBuffer", n: int, edge: int, mask: "DeviceBuffer"):
        self.ctx.check(self.ctx.L.vx_batch_set_cull(
            self.h, C.byref(frustum.c) if frustum is not None else None,
            coords.ptr if coords is not None else None, n, edge,
            mask.ptr if mask is not None else None), "vx_batch_set_cull")

    def set_mode(self, mode: int):
        self.ctx.check(self.ctx.L.vx_batch_set_mode(self.h, mode), "vx_batch_set_mode")

    def draw_list(self, frustum: "Frustum", mode: int = DRAW_EXPANDED) -> np.ndarray:
        """Culled indirect draw list over this batch's meshes (after a run)."""
        args, cnt = DeviceBuffer(self.ctx, self.n * DRAW_DTYPE.itemsize), DeviceBuffer(self.ctx, 8)
        try:
            self.ctx.check(self.ctx.L.vx_batch_draw_list(self.ctx.h, self.h, C.byref(frustum.c),
                                                         mode, args.ptr, cnt.ptr),
                           "vx_batch_draw_list")
            return _read_draws(self.ctx, args, cnt)
        finally:
            args.free()
            cnt.free()

    def set_grid(self, grid: int):
        """Workgroups of the fused launch (0 = one per chunk); staged ignores it."""
        self.ctx.check(self.ctx.L.vx_batch_set_grid(self.h, grid), "vx_batch_set_grid")

    def set_roles(self, ratio: int):
        """Pipelined form: one mesher workgroup in `ratio` (2..8; 0 = default)."""
        self.ctx.check(self.ctx.L.vx_batch_set_roles(self.h, ratio), "vx_batch_set_roles")

    def set_splits(self, fill_parts: int, emit_parts: int):
        """Staged form, small batches: fill workgroups per chunk (1, 2, 4, 8)
        and emit workgroups per item (1..4); 0 = the launchers' rule."""
        self.ctx.check(self.ctx.L.vx_batch_set_splits(self.h, fill_parts, emit_parts),
                       "vx_batch_set_splits")

    @property
    def mode(self) -> int:
        m = C.c_int32()
        self.ctx.check(self.ctx.L.vx_batch_get_mode(self.h, C.byref(m)), "vx_batch_get_mode")
        return m.value

    def run(self):
        self.ctx.check(self.ctx.L.vx_batch_run(self.ctx.h, self.h), "vx_batch_run")

    def run_per_chunk(self):
        self.ctx.check(self.ctx.L.vx_batch_run_per_chunk(self.ctx.h, self.h),
                       "vx_batch_run_per_chunk")

    def wait(self) -> int:
        t = C.c_uint64()
        self.ctx.check(self.ctx.L.vx_batch_wait(self.ctx.h, self.h, C.byref(t)), "vx_batch_wait")
        self.total = t.value
        return t.value

    def chunk_faces(self, i: int):
        a, b = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.vx_batch_chunk_faces(self.ctx.h, self.h, i, C.byref(a),
                                                       C.byref(b)), "vx_batch_chunk_faces")
        return a.value, b.value

    def ids(self, i: int) -> np.ndarray:
        n = int(self.edges[i]) ** 3
        out = np.zeros(n, np.uint8)
        self.ctx.check(self.ctx.L.vx_batch_read_ids(self.ctx.h, self.h, i, _p(out), n),
                       "vx_batch_read_ids")
        return out

    def mesh(self, fmt: int) -> np.ndarray:
        F = self.total
        if fmt == FMT_INSTANCES:
            out = np.zeros(F, INSTANCE_DTYPE)
        elif fmt == FMT_VERTICES:
            out = np.zeros(4 * F, VERTEX_DTYPE)
        else:
            out = np.zeros(6 * F, np.uint32)
        self.ctx.check(self.ctx.L.vx_batch_read_mesh(self.ctx.h, self.h, fmt, _p(out), out.nbytes),
                       "vx_batch_read_mesh")
        return out

    def autotune(self, reps: int = 5) -> dict:
        r = TuneResult()
        self.ctx.check(self.ctx.L.vx_batch_autotune(self.ctx.h, self.h, reps, C.byref(r)),
                       "vx_batch_autotune")
        names = {BATCH_FUSED: "fused", BATCH_STAGED: "staged", BATCH_PIPELINED: "pipelined",
                 BATCH_SPLIT: "split"}
        return {"mode": names.get(r.mode, r.mode), "grid": r.grid, "ratio": r.ratio,
                "splits": [r.fill_split, r.emit_split], "best_ms": round(r.best_ms, 4),
                "candidates": [{"mode": names.get(r.cand_mode[i], r.cand_mode[i]),
                                "grid": r.cand_grid[i], "ratio": r.cand_ratio[i],
                                "splits": [r.cand_fill_split[i], r.cand_emit_split[i]],
                                "ms": round(r.cand_ms[i], 4)}
                               for i in range(r.n_candidates)]}

    def profile(self, reps: int = 10) -> np.ndarray:
        ms = (C.c_float * 5)()
        self.ctx.check(self.ctx.L.vx_batch_profile(self.ctx.h, self.h, reps, ms),
                       "vx_batch_profile")
        return np.array(ms[:], np.float64)


NB_NONE, NB_RESIDENT = 0, 1


WORLD_CHUNKS, WORLD_MESHES, WORLD_REBUILD_QUEUE, WORLD_IN_PROCESS = 1, 2, 3, 4


class World:
    """The reference's World + ChunkGenerator stepped tick by tick
    (src/world.rs:14-58, src/chunk_builder.rs:45-352), jobs on the GPU."""

    def __init__(self, ctx: Context, terrain_id: int, player_pos, render_distance: int = 5,
                 result_delay: int = 1, formats: int = FMT_VERTICES):
        d = WorldDesc(terrain_id, render_distance, result_delay, formats,
                      (C.c_float * 3)(*map(float, player_pos)))
        h = P()
        ctx.check(ctx.L.vx_world_create(ctx.h, C.byref(d), C.byref(h)), "vx_world_create")
        self.ctx, self.h, self.formats = ctx, h, formats

    def close(self):
        if self.h:
            self.ctx.L.vx_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.close()
        except Exception:
            pass

    def tick(self, player_pos, frustum: "Frustum"):
        p = np.ascontiguousarray(np.asarray(player_pos, np.float32).reshape(3))
        self.ctx.check(self.ctx.L.vx_world_tick(self.h, _p(p), C.byref(frustum.c)), "vx_world_tick")

    def stats(self) -> dict:
        s = WorldStats()
        self.ctx.check(self.ctx.L.vx_world_stats_get(self.h, C.byref(s)), "vx_world_stats_get")
        out = {n: getattr(s, n) for n, _ in WorldStats._fields_ if n != "player_chunk"}
        out["player_chunk"] = (s.player_chunk.x, s.player_chunk.y, s.player_chunk.z)
        return out

    def list(self, what: int = WORLD_CHUNKS) -> list:
        n = C.c_int32()
        self.ctx.check(self.ctx.L.vx_world_list(self.h, what, None, 0, C.byref(n)), "vx_world_list")
        arr = (Coord * max(1, n.value))()
        self.ctx.check(self.ctx.L.vx_world_list(self.h, what, C.cast(arr, P), n.value, C.byref(n)),
                       "vx_world_list")
        return [(arr[i].x, arr[i].y, arr[i].z) for i in range(n.value)]

    def mesh(self, coord):
        m = P()
        self.ctx.check(self.ctx.L.vx_world_lookup(self.h, Coord(*coord), None, C.byref(m)),
                       "vx_world_lookup")
        return ChunkMesh(self.ctx, m, self.formats) if m else None

    def chunk(self, coord):
        c = P()
        self.ctx.check(self.ctx.L.vx_world_lookup(self.h, Coord(*coord), C.byref(c), None),
                       "vx_world_lookup")
        return Chunk(self.ctx, c) if c else None


class Server:
    """Continuous-batching chunk server with a device-resident LRU cache
    (the reference's ChunkGenerator + World maps, src/chunk_builder.rs:45-352)."""

    def __init__(self, ctx: Context, max_batch: int = 0, max_inflight: int = 0,
                 formats: int = FMT_VERTICES):
        h = P()
        ctx.check(ctx.L.vx_server_create(ctx.h, C.byref(ServerDesc(max_batch, max_inflight,
                                                                    formats, 0)), C.byref(h)),
                  "vx_server_create")
        self.ctx, self.h, self.formats = ctx, h, formats

    def close(self):
        if self.h:
            self.ctx.L.vx_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, reqs) -> list[int]:
        """reqs: iterable of (coord, edge, terrain_id, neighbours)."""
        reqs = list(reqs)
        arr = (Request * len(reqs))()
        for i, (c, e, t, nbp) in enumerate(reqs):
            arr[i].coord = Coord(*c)
            arr[i].edge, arr[i].terrain_id, arr[i].neighbours = e, t, nbp
        tickets = (C.c_uint64 * len(reqs))()
        self.ctx.check(self.ctx.L.vx_server_submit(self.h, C.cast(arr, P), len(reqs),
                                                   C.cast(tickets, P)), "vx_server_submit")
        return list(tickets)

    def pump(self, wait_us: int = 0) -> int:
        st = self.ctx.L.vx_server_pump(self.h, wait_us)
        if st < 0:
            self.ctx.check(st, "vx_server_pump")
        return st

    def drain(self):
        self.ctx.check(self.ctx.L.vx_server_drain(self.h), "vx_server_drain")

    def poll(self) -> list[Completion]:
        out = []
        buf = (Completion * 1024)()
        while True:
            n = C.c_int32()
            self.ctx.check(self.ctx.L.vx_server_poll(self.h, C.cast(buf, P), 1024, C.byref(n)),
                           "vx_server_poll")
            out += [Completion.from_buffer_copy(buf[i]) for i in range(n.value)]
            if n.value < 1024:
                return out

    def lookup(self, coord, edge: int, terrain_id: int):
        ch, me = P(), P()
        st = self.ctx.L.vx_server_lookup(self.h, Coord(*coord), edge, terrain_id, C.byref(ch),
                                         C.byref(me))
        if st == ERR_NOT_READY:
            return None
        self.ctx.check(st, "vx_server_lookup")
        return Chunk(self.ctx, ch), ChunkMesh(self.ctx, me, self.formats)

    def stats(self) -> ServerStats:
        s = ServerStats()
        self.ctx.check(self.ctx.L.vx_server_stats_get(self.h, C.byref(s)), "vx_server_stats_get")
        return s


def simulate(ctx: Context, *, seed=3, rate=2000.0, duration=2.0, speed=20.0, radius=16,
             classes=((16, 1.0, 0), (32, 1.0, 0), (64, 1.0, 0)), neighbours=NB_NONE,
             max_batch=0, max_inflight=0, formats=FMT_VERTICES) -> SimResult:
    """Streaming simulation (camera random walk + Poisson arrivals) through a
    Server; classes = ((edge, weight, terrain_id), ...)."""
    d = SimDesc()
    d.seed, d.rate_per_s, d.duration_s, d.speed_blocks_per_s = seed, rate, duration, speed
    d.radius_chunks, d.n_classes, d.neighbours = radius, len(classes), neighbours
    for i, (e, w, t) in enumerate(classes):
        d.edges[i], d.weights[i], d.terrain_id[i] = e, w, t
    d.server = ServerDesc(max_batch, max_inflight, formats, 0)
    r = SimResult()
    ctx.check(ctx.L.vx_serve_simulate(ctx.h, C.byref(d), C.byref(r)), "vx_serve_simulate")
    return r


def grid_neighbours(nx: int, nz: int) -> np.ndarray:
    """Neighbour snapshot for an nx x nz grid of chunks indexed i = x*nz + z,
    every in-grid neighbour present (-x, +x, -z, +z), -1 outside."""
    nb = np.full((nx * nz, 4), -1, np.int32)
    for x in range(nx):
        for z in range(nz):
            i = x * nz + z
            if x > 0:
                nb[i, 0] = (x - 1) * nz + z
            if x < nx - 1:
                nb[i, 1] = (x + 1) * nz + z
            if z > 0:
                nb[i, 2] = x * nz + z - 1
            if z < nz - 1:
                nb[i, 3] = x * nz + z + 1
    return nb
