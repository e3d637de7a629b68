"""CPU restatement of the reference's World + ChunkGenerator tick
(src/world.rs:14-58, src/chunk_builder.rs:28-352), test infrastructure: the
GPU library's vx_world must produce the same world state tick for tick.

Only tests/ may use this module.  Jobs run synchronously on the CPU oracle
(chunk_fill + mesh_faces); as in the library, a job's result becomes
receivable `result_delay` ticks after submission, in submission order (one
schedule of the reference's thread pool).

State per tick: world.chunks (coord -> ids), world.meshes (coord -> faces),
chunk_load_queue, chunk_rebuild_queue, data_in_process, player.chunk.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np

import oracle as O

EDGE = 32  # CHUNK_I32 (src/world.rs:21-22)


def as_i32(v):
    """Rust's `f as i32`: truncation toward zero, saturating, NaN -> 0."""
    if math.isnan(v):
        return 0
    if v >= 2.0 ** 31:
        return 2 ** 31 - 1
    if v < -2.0 ** 31:
        return -2 ** 31
    return int(v)


def to_chunk_coord(pos):
    """Coord3DF::to_chunk_coord (src/coordinate.rs:33-38): floor, as i32, >> 5
    (positions are f32)."""
    return tuple(as_i32(float(np.floor(np.float32(v)))) >> 5 for v in pos)


class WorldModel:
    def __init__(self, terrain, player_pos, render_distance=5, result_delay=1):
        """terrain: oracle Terrain; player_pos: Player::new's position (world units)."""
        self.t = terrain
        self.rd = render_distance
        self.delay = result_delay
        self.chunks = {}        # world.chunks: coord -> ids (Arc<Chunk>)
        self.meshes = {}        # world.meshes: coord -> faces (ChunkMesh)
        self.load_queue = []    # chunk_load_queue
        self.rebuild_queue = []  # chunk_rebuild_queue
        self.in_process = []    # data_in_process
        self.data_rx = deque()  # (ready_tick, coord, ids, faces)
        self.mesh_rx = deque()  # (ready_tick, coord, faces)
        self.player_pos = tuple(map(float, player_pos))
        self.player_chunk = to_chunk_coord(self.player_pos)  # Player::new (src/player.rs:10-14)
        self.tick_no = 0
        self.jobs_loaded = self.jobs_rebuilt = 0
        self.received_chunks = self.received_meshes = self.unloaded = 0

    # -- jobs ------------------------------------------------------------
    def _mesh(self, ids, pos, adj):
        origin = (pos[0] * EDGE, pos[1] * EDGE, pos[2] * EDGE)
        return O.mesh_faces(ids, EDGE, origin, adj)

    def _adjacent(self, pos):
        """adjacent_chunks (chunk_builder.rs:134-169): -x, +x, -z, +z."""
        x, y, z = pos
        return [self.chunks.get((x - 1, y, z)), self.chunks.get((x + 1, y, z)),
                self.chunks.get((x, y, z - 1)), self.chunks.get((x, y, z + 1))]

    def _submit_load(self, pos):
        """The pool.execute closure of chunk_builder.rs:80-84 / 125-129."""
        adj = self._adjacent(pos)
        ids = O.chunk_fill(self.t, pos[0], pos[1], pos[2], EDGE)
        self.data_rx.append((self.tick_no + self.delay, pos, ids, self._mesh(ids, pos, adj)))
        self.in_process.append(pos)
        self.jobs_loaded += 1

    def _is_loaded(self, pos):
        """is_chunk_loaded (262-271)."""
        return pos in self.chunks or pos in self.load_queue or pos in self.in_process

    # -- build_chunks steps (45-67) ----------------------------------------
    def _load_primary(self):
        """load_primary_chunks (171-212)."""
        now = to_chunk_coord(self.player_pos)
        if now == self.player_chunk:        # is_in_new_chunk_pos (player.rs:24-27)
            return
        self.player_chunk = now
        px, pz = now[0], now[2]
        seq = [(px, 0, pz)]
        r = 1
        for z in range(-r, r + 1):
            seq += [(px + r, 0, pz + z), (px - r, 0, pz + z)]
        for x in range(-r + 1, r):
            seq += [(px + x, 0, pz + r), (px + x, 0, pz - r)]
        for p in seq:                       # load_chunk_directly (115-132)
            if not self._is_loaded(p):
                self._submit_load(p)

    def _enqueue_in_frustum(self, planes):
        """enqueue_chunks_in_frustum (214-253) + enqueue_data (255-261)."""
        px, pz = self.player_chunk[0], self.player_chunk[2]
        if self.rd <= 1:
            return
        cand = []
        for r in range(2, self.rd):
            for z in range(-r, r + 1):
                cand += [(px + r, 0, pz + z), (px - r, 0, pz + z)]
            for x in range(-r + 1, r):
                cand += [(px + x, 0, pz + r), (px + x, 0, pz - r)]
        for p in cand:
            if not self._is_loaded(p) and O.frustum_contains(planes, [p], EDGE)[0]:
                self.load_queue.append(p)

    def _process_loading(self):
        """process_chunk_loading_queue (69-87)."""
        if self.load_queue:
            self._submit_load(self.load_queue.pop(0))

    def _process_rebuild(self):
        """process_rebuild_queue (89-113)."""
        if not self.in_process and self.rebuild_queue:
            pos = self.rebuild_queue.pop(0)
            if pos in self.chunks:
                faces = self._mesh(self.chunks[pos], pos, self._adjacent(pos))
                self.mesh_rx.append((self.tick_no + self.delay, pos, faces))
                self.jobs_rebuilt += 1

    def _rebuild_adjacent(self, pos):
        """rebuild_adjacent_chunks (273-300): +x, -x, +z, -z."""
        if not self.chunks:
            return
        x, y, z = pos
        for q in ((x + 1, y, z), (x - 1, y, z), (x, y, z + 1), (x, y, z - 1)):
            if q in self.chunks:
                self.rebuild_queue.append(q)

    def _update_world(self):
        """update_world (310-332): one data result and one mesh result."""
        if self.data_rx and self.data_rx[0][0] <= self.tick_no:
            _, pos, ids, faces = self.data_rx.popleft()
            self.chunks[pos] = ids
            self.meshes[pos] = faces
            self.in_process = [p for p in self.in_process if p != pos]
            self._rebuild_adjacent(pos)
            self.received_chunks += 1
        if self.mesh_rx and self.mesh_rx[0][0] <= self.tick_no:
            _, pos, faces = self.mesh_rx.popleft()
            self.meshes[pos] = faces
            self.received_meshes += 1

    def _clean_up(self):
        """clean_up_queues (302-308): ChunkCoord3D orders by x, then y, then z."""
        self.load_queue.clear()
        self.rebuild_queue.sort()
        dedup = []
        for p in self.rebuild_queue:
            if not dedup or dedup[-1] != p:
                dedup.append(p)
        self.rebuild_queue = dedup

    def _filter_unseen(self):
        """filter_unseen_chunks (334-352)."""
        cx, cy, cz = self.player_chunk
        rd = self.rd
        for p in list(self.chunks):
            if (p[0] <= rd + cx and p[2] <= rd + cz and p[0] >= -rd + cx and p[2] >= -rd + cz
                    and p[1] <= rd + cy and p[1] >= -rd + cy):
                continue
            self.meshes.pop(p, None)
            if p in self.rebuild_queue:
                self.rebuild_queue.remove(p)   # the first occurrence
            del self.chunks[p]
            self.unloaded += 1

    def tick(self, player_pos, planes):
        """One World::update (src/engine.rs:65-74)."""
        self.player_pos = tuple(map(float, player_pos))
        self._load_primary()
        self._enqueue_in_frustum(planes)
        self._process_loading()
        self._process_rebuild()
        self._update_world()
        self._clean_up()
        self._filter_unseen()
        self.tick_no += 1

    def state(self):
        return {"chunks": sorted(self.chunks), "meshes": sorted(self.meshes),
                "rebuild_queue": list(self.rebuild_queue), "in_process": list(self.in_process),
                "player_chunk": self.player_chunk, "jobs_loaded": self.jobs_loaded,
                "jobs_rebuilt": self.jobs_rebuilt, "received_chunks": self.received_chunks,
                "received_meshes": self.received_meshes, "unloaded": self.unloaded}
