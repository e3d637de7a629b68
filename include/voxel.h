/*
 * voxel.h -- C ABI of libvoxel_hip.so, the MI355X (gfx950) chunk pipeline:
 * noise terrain fill -> visible-face culling and quad emission -> AABB
 * frustum tests, with chunk state and meshes resident in HBM.
 *
 * The entry points mirror the module surface the reference's scheduler uses
 * (paths relative to the reference repository, Blatko1/wgpu-voxel-engine):
 *
 *   vx_chunk_new            Chunk::new                 src/chunk.rs:21-25
 *   vx_chunk_create_mesh    Chunk::create_mesh         src/chunk.rs:48-57
 *                           + ChunkMesh::new           src/chunk.rs:254-266
 *   vx_frustum_from_matrix  Frustum::new / update      src/frustum_culling.rs:10-14,42-45
 *   vx_frustum_contains     Frustum::contains          src/frustum_culling.rs:16-40
 *   vx_frustum_mask         World::render's per-mesh Frustum::contains loop
 *                                                      src/world.rs:34-38
 *   vx_batch_*              many ChunkGenerator jobs at once: the body of
 *                           pool.execute(Chunk::new + create_mesh)
 *                                                      src/chunk_builder.rs:80-84
 *   vx_server_submit /      pool.execute / flume try_recv
 *   vx_server_pump /                                   src/chunk_builder.rs:80-86,310-332
 *   vx_server_poll
 *   vx_server_lookup        world.chunks / world.meshes src/world.rs:14-17
 *   vx_serve_simulate       the World + ChunkGenerator streaming loop
 *                                                      src/world.rs:14-58,
 *                                                      src/chunk_builder.rs:45-352
 *
 * Conventions
 *   - Every call returns an int status (VX_OK = 0, negative = error); nothing
 *     panics or aborts across the ABI (the reference unwrap()s:
 *     src/chunk_builder.rs:83,107,128).
 *   - Chunks and meshes are immutable after construction and reference
 *     counted like the reference's Arc<Chunk> (src/chunk_builder.rs:81,140);
 *     release every handle you own with vx_*_release.
 *   - Calls on one vx_ctx are serialised by the context; any host thread may
 *     call them.  GPU work is queued on the context's stream.
 *   - Coordinates are chunk coordinates (ChunkCoord3D, src/coordinate.rs:41-66);
 *     a chunk of edge N covers world voxels [c*N, c*N + N) on each axis (the
 *     reference fixes N = 32, src/coordinate.rs:60-65).  Calls that build a
 *     chunk return VX_ERR_INVALID unless c*N - 1 and c*N + N fit in int32 on
 *     every axis.
 *   - Block ids are u8: 0 air, 1 grass, 2 dirt (the reference's CubeType
 *     GRASS=0, DIRT=1, AIR=2, src/cube.rs:16-21, shifted so 0 is air).
 *   - Voxel index inside a chunk: x + N*z + N*N*y (src/chunk.rs:38).
 */
#ifndef VOXEL_H
#define VOXEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
/* Everything declared here is the library's exported surface; it is built
 * with hidden visibility otherwise. */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define VX_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------- */
#define VX_OK 0
#define VX_ERR_INVALID (-1)  /* bad argument */
#define VX_ERR_NOMEM (-2)    /* host or device allocation failed */
#define VX_ERR_DEVICE (-3)   /* HIP runtime error */
#define VX_ERR_CAPACITY (-4) /* caller buffer too small; required size reported */
#define VX_ERR_BUDGET (-5)   /* HBM admission budget exceeded */
#define VX_ERR_NOT_READY (-6)
#define VX_ERR_NODEV (-7)    /* no usable GPU */

/* ---- types --------------------------------------------------------------- */
typedef struct vx_ctx vx_ctx;
typedef struct vx_chunk vx_chunk;
typedef struct vx_mesh vx_mesh;
typedef struct vx_batch vx_batch;

typedef struct vx_coord {
    int32_t x, y, z;
} vx_coord;

/* Terrain generator.  kind:
 *   VX_TERRAIN_AIR / VX_TERRAIN_SOLID   every voxel air / grass
 *   VX_TERRAIN_HEIGHTMAP  n = fbm2(wx*f, wz*f); h = floor(height_scale*(n+1)/2)
 *                         clamped to [0, height_max]; voxel solid iff wy <= h
 *   VX_TERRAIN_DENSITY    solid iff fbm3(wx*f, wy*f, wz*f) - wy*y_scale > 0
 *   VX_TERRAIN_REFERENCE  solid iff fbm3(wx*f, wz*f, wy*f) < 0: the reference's
 *                         own rule (noise < 0 -> GRASS, src/chunk.rs:38-42) and
 *                         axis order (world z on the noise's y axis, world y
 *                         on its z axis, src/perlin_noise.rs:6-13); y_scale
 *                         unused.  The reference's parameters are seed 1,
 *                         frequency 0.09, gain 2.5, lacunarity 0.5, 3 octaves
 *                         (src/perlin_noise.rs:14-18), with Perlin's noise in
 *                         place of simdnoise's.  Batches of it run staged.
 * fbm: sum over `octaves` of amp*perlin(p*freq), freq *= lacunarity,
 * amp *= gain, from freq = frequency, amp = 1 (the reference's fbm_3d_offset
 * with_freq/with_gain/with_lacunarity/with_octaves, src/perlin_noise.rs:6-20).
 * perlin = Improved Noise (Perlin 2002) with the permutation shuffled by
 * splitmix64(seed).  Every octave frequency must be >= 2^-20. */
#define VX_TERRAIN_AIR 0
#define VX_TERRAIN_SOLID 1
#define VX_TERRAIN_HEIGHTMAP 2
#define VX_TERRAIN_DENSITY 3
#define VX_TERRAIN_REFERENCE 4

typedef struct vx_terrain {
    int32_t kind;
    uint32_t seed;
    int32_t octaves; /* 1..8 */
    float frequency;
    float lacunarity;
    float gain;
    float height_scale;
    int32_t height_max;
    float y_scale;
} vx_terrain;

/* Mesh output formats (bit mask). */
#define VX_FMT_INSTANCES 1u /* InstanceRaw: f32 4x4 column-major + u32 layer, 68 B
                               (src/renderer/instance.rs:1-6, src/quad.rs:47-54) */
#define VX_FMT_VERTICES 2u  /* 4 Vertex {f32 pos[3], f32 uv[2]} (20 B) + 6 u32
                               chunk-local indices per face: 104 B per face
                               (src/quad.rs:66-89, src/shaders/vertex.vert:24) */

typedef struct vx_ctx_desc {
    int32_t device;        /* HIP device ordinal */
    uint32_t flags;        /* reserved, 0 */
    uint64_t hbm_budget;   /* bytes the context may hold resident; 0 = no limit */
} vx_ctx_desc;

/* ---- library / context --------------------------------------------------- */
int vx_abi_version(void);
const char *vx_status_string(int status);
/* Number of visible HIP devices (0 when none); does not create a context. */
int vx_device_count(int32_t *count);

int vx_ctx_create(const vx_ctx_desc *desc, vx_ctx **out);
int vx_ctx_destroy(vx_ctx *ctx);
int vx_ctx_synchronize(vx_ctx *ctx);
/* The hipStream_t all of this context's work is queued on. */
int vx_ctx_stream(vx_ctx *ctx, void **hip_stream);
/* Device bytes currently held by chunks, meshes and batches of the context. */
int vx_ctx_memory(vx_ctx *ctx, uint64_t *resident_bytes, uint64_t *budget_bytes);
/* Last HIP error text recorded by the context (static storage). */
const char *vx_ctx_last_error(vx_ctx *ctx);

/* Device buffers for callers that keep inputs/outputs resident in HBM (e.g.
 * chunk coordinates and cull masks of vx_frustum_mask_device). Allocated
 * from the context heap and counted against its budget. */
int vx_buffer_alloc(vx_ctx *ctx, uint64_t bytes, void **dev);
int vx_buffer_free(vx_ctx *ctx, void *dev);
int vx_buffer_upload(vx_ctx *ctx, void *dev, const void *host, uint64_t bytes);
int vx_buffer_download(vx_ctx *ctx, void *host, const void *dev, uint64_t bytes);
/* Device-to-device copy by a plain kernel (16-byte aligned, multiple of 16
 * bytes); queued on the context stream. */
int vx_buffer_copy(vx_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Registers a terrain generator; returns its id in *terrain_id. */
int vx_terrain_register(vx_ctx *ctx, const vx_terrain *terrain, int32_t *terrain_id);

/* ---- Chunk (src/chunk.rs:15-46) ------------------------------------------ */
/* Chunk::new: generates the chunk's block ids on the GPU (asynchronous). */
int vx_chunk_new(vx_ctx *ctx, int32_t terrain_id, vx_coord coord, int32_t edge,
                 vx_chunk **out);
/* Builds a chunk from caller-provided block ids (host memory, edge^3 bytes,
 * values 0..2).  The reference builds chunks only from noise; this is the
 * entry point for edited or externally stored chunks. */
int vx_chunk_from_ids(vx_ctx *ctx, vx_coord coord, int32_t edge, const uint8_t *ids,
                      vx_chunk **out);
int vx_chunk_retain(vx_chunk *chunk);
int vx_chunk_release(vx_chunk *chunk);
int vx_chunk_info(const vx_chunk *chunk, vx_coord *coord, int32_t *edge);
/* Copies the block ids (edge^3 bytes) / occupancy rows (edge^2 rows of
 * (edge+31)/32 u32, bit x of row y*edge+z) to host memory; synchronises. */
int vx_chunk_read_ids(vx_ctx *ctx, const vx_chunk *chunk, uint8_t *dst, uint64_t cap);
int vx_chunk_read_occupancy(vx_ctx *ctx, const vx_chunk *chunk, uint32_t *dst,
                            uint64_t cap_words);
/* Raw noise the chunk was generated from (density: edge^3 f32 at the voxel
 * index; heightmap: edge^2 f32 at x + edge*z).  For parity checks. */
int vx_chunk_noise(vx_ctx *ctx, int32_t terrain_id, vx_coord coord, int32_t edge, float *dst,
                   uint64_t cap);

/* ---- ChunkMesh (src/chunk.rs:48-57, 248-266) ----------------------------- */
/* Chunk::create_mesh with the neighbour snapshot adj = {-x, +x, -z, +z}
 * (src/chunk_builder.rs:134-169); NULL = absent neighbour (its boundary
 * faces are emitted, src/chunk.rs:102-108).  Neighbours must have the same
 * edge.  Synchronises once to size the mesh. */
int vx_chunk_create_mesh(vx_ctx *ctx, vx_chunk *chunk, vx_chunk *const adj[4], uint32_t formats,
                         vx_mesh **out);
int vx_mesh_retain(vx_mesh *mesh);
int vx_mesh_release(vx_mesh *mesh);
int vx_mesh_face_count(const vx_mesh *mesh, uint64_t *faces);
/* Copy-out with explicit capacities in bytes; VX_ERR_CAPACITY when short. */
int vx_mesh_read_instances(vx_ctx *ctx, const vx_mesh *mesh, void *dst, uint64_t cap);
int vx_mesh_read_vertices(vx_ctx *ctx, const vx_mesh *mesh, void *dst, uint64_t cap);
int vx_mesh_read_indices(vx_ctx *ctx, const vx_mesh *mesh, void *dst, uint64_t cap);
/* Zero-copy: device pointers of the mesh's records (NULL for formats not
 * built).  Valid while the mesh is retained. */
int vx_mesh_device_buffers(const vx_mesh *mesh, void **instances, void **vertices,
                           void **indices);

/* ---- Frustum (src/frustum_culling.rs) ------------------------------------ */
typedef struct vx_frustum {
    float planes[24]; /* [near far left right top bottom] x (a b c d) */
} vx_frustum;
/* matrix_to_planes: m is the view-projection matrix, column-major (m[c*4+r]). */
int vx_frustum_from_matrix(const float m[16], vx_frustum *out);
/* contains: 1 if some corner of the chunk AABB is inside all six planes. */
int vx_frustum_contains(const vx_frustum *f, vx_coord coord, int32_t edge);
/* The same test for n chunks on the GPU; coords/out are host arrays. */
int vx_frustum_mask(vx_ctx *ctx, const vx_frustum *f, const vx_coord *coords, int32_t n,
                    int32_t edge, uint8_t *out);
/* Device-resident variant: coords_dev (3n i32) and out_dev (n u8) are device
 * pointers; queued on the context stream, no synchronisation. */
int vx_frustum_mask_device(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                           int32_t n, int32_t edge, uint8_t *out_dev);

/* ---- draw lists (src/world.rs:34-38 + ChunkMesh::render, src/chunk.rs:268-277) ----
 * World::render issues ChunkMesh::render's draw_indexed(0..6, 0, 0..instances)
 * for every resident mesh whose chunk passes Frustum::contains.  With the
 * meshes in one arena (mesh i = faces[i] records starting at first[i]) that
 * loop becomes an indirect draw list, built on the GPU in mesh order:
 *   VX_DRAW_INSTANCED  InstanceRaw arena:   {6, faces, 0, 0, first}
 *   VX_DRAW_EXPANDED   vertex/index arena:  {6*faces, 1, 6*first, 4*first, 0}
 * One entry per visible mesh (zero-face meshes give no-op draws, as in the
 * reference).  count_dev[0] = number of draws; count_dev[1] != 0 if some
 * offset did not fit the 32-bit fields.  Asynchronous on the context stream;
 * all pointers are device pointers; args_dev holds n entries. */
typedef struct vx_draw_indexed {
    uint32_t index_count;
    uint32_t instance_count;
    uint32_t first_index;
    int32_t base_vertex;
    uint32_t first_instance;
} vx_draw_indexed;
#define VX_DRAW_INSTANCED 1u
#define VX_DRAW_EXPANDED 2u
int vx_frustum_draw_list(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                         const uint32_t *faces_dev, const uint64_t *first_dev, int32_t n,
                         int32_t edge, uint32_t mode, vx_draw_indexed *args_dev,
                         uint32_t *count_dev);
/* vx_frustum_draw_list that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, a per-frame render loop, harnesses that forbid
 * allocation).  Same arguments, kernels and results as vx_frustum_draw_list;
 * its per-workgroup partial counts live in a fixed 16 KiB device array per
 * GPU, held until the process ends, instead of a context allocation, and it
 * queues no memset.  n may not exceed VX_DRAW_LIST_NOALLOC_MAX
 * (VX_ERR_INVALID).  Every call on one GPU uses that GPU's array: calls that
 * overlap on one GPU (two contexts' streams) race; calls on one context are
 * ordered by its stream. */
#define VX_DRAW_LIST_NOALLOC_MAX 4194304
int vx_frustum_draw_list_noalloc(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                                 const uint32_t *faces_dev, const uint64_t *first_dev, int32_t n,
                                 int32_t edge, uint32_t mode, vx_draw_indexed *args_dev,
                                 uint32_t *count_dev);

/* ---- Batches: many Chunk::new + create_mesh jobs in one launch sequence --- */
/* One request: chunk coord, edge (16, 32 or 64), terrain, and the neighbour
 * snapshot as indices into the same batch (-1 = absent).  A batch holds 1 to
 * VX_BATCH_MAX_CHUNKS requests. */
#define VX_BATCH_MAX_CHUNKS (1 << 26)
typedef struct vx_batch_req {
    vx_coord coord;
    int32_t edge;
    int32_t terrain_id;
    int32_t nb[4]; /* -x, +x, -z, +z */
} vx_batch_req;

int vx_batch_create(vx_ctx *ctx, const vx_batch_req *reqs, int32_t n, uint32_t formats,
                    uint64_t face_capacity, vx_batch **out);
/* As vx_batch_create, with resident neighbours: nb[k] = -2 - e names ext[e]
 * (a chunk of the same edge, e.g. a view from vx_batch_chunk or a server
 * lookup), retained by the batch. */
int vx_batch_create_ext(vx_ctx *ctx, const vx_batch_req *reqs, int32_t n, vx_chunk *const *ext,
                        int32_t n_ext, uint32_t formats, uint64_t face_capacity, vx_batch **out);
/* Chunk i of the batch as a retained chunk handle (a view into the batch's
 * storage, valid after the batch is destroyed). */
int vx_batch_chunk(vx_batch *batch, int32_t i, vx_chunk **out);
/* Frustum job run with every later run of the batch (in the fused form,
 * inside the same launch): mask_dev[i] = contains(coords_dev[3i..3i+2]).
 * Each call uploads new planes (a camera move needs no graph re-capture);
 * n = 0 removes the job.  coords_dev / mask_dev are device pointers. */
int vx_batch_set_cull(vx_batch *batch, const vx_frustum *f, const int32_t *coords_dev, int32_t n,
                      int32_t edge, uint8_t *mask_dev);
int vx_batch_destroy(vx_batch *batch);
/* Execution form of a batch:
 *   VX_BATCH_AUTO    staged, or the form vx_batch_autotune chose for batches
 *                    of this shape (default)
 *   VX_BATCH_FUSED   one persistent kernel: noise -> fill -> cull -> compact per
 *                    chunk, neighbour planes recomputed, output offsets from a
 *                    single-pass device-wide scan.  Eligible when every in-batch
 *                    neighbour is the adjacent chunk with the same terrain.
 *   VX_BATCH_STAGED  four kernels: fill -> count -> scan -> emit
 *   VX_BATCH_PIPELINED  one kernel with fill and mesh roles: one workgroup in
 *                    three (vx_batch_set_roles) meshes chunks in batch order
 *                    (count, single-pass scan, emit) as their fills and
 *                    their neighbours' fills
 *                    publish; the others fill, then mesh once every fill is
 *                    handed out, so noise (VALU) and emission (HBM stores)
 *                    run side by side on every CU; neighbour planes come
 *                    from the neighbours' fills.  Eligible as VX_BATCH_FUSED,
 *                    for edges up to 32.
 *   VX_BATCH_SPLIT   the staged kernels over two halves of the batch on two
 *                    streams of one graph: the first half (and the later
 *                    chunks it borders) is filled, then its count, scan and
 *                    emit run beside the second half's fill; the second
 *                    half's scan continues from the first half's total.
 *                    Edges up to 32; a batch with no in-batch neighbour
 *                    needs 2,048 chunks or more (whole-chunk fills).
 * Every form produces identical bytes. */
#define VX_BATCH_AUTO 0
#define VX_BATCH_FUSED 1
#define VX_BATCH_STAGED 2
#define VX_BATCH_PIPELINED 3
#define VX_BATCH_SPLIT 4
int vx_batch_set_mode(vx_batch *batch, int32_t mode);
/* Workgroups of the persistent launch in VX_BATCH_FUSED / VX_BATCH_PIPELINED
 * (0 = one per chunk; fewer = persistent workgroups that keep taking chunk
 * tickets).  The staged form ignores it, and so does VX_BATCH_AUTO once the
 * autotuner has chosen a shape for the batch. */
int vx_batch_set_grid(vx_batch *batch, int32_t grid);
/* Roles of the VX_BATCH_PIPELINED launch: one workgroup in `ratio` (2..8)
 * meshes from the start, the others fill and then mesh; 0 = the default
 * (one in three).  Other forms ignore it, and so does VX_BATCH_AUTO once the
 * autotuner has chosen a split for the batch. */
int vx_batch_set_roles(vx_batch *batch, int32_t ratio);
/* Tile shape of the VX_BATCH_STAGED form for small batches: workgroups per
 * chunk of the fill (1, 2, 4 or 8; used below 1,024 chunks, where the rule
 * is 2, 4 or 8, and 8 only for 64^3 chunks) and per item of the emit (1..4;
 * used below 1,280 items, where the rule is 2, 3 or 4); 0 = the rule.
 * Larger batches take one workgroup per chunk and per item whatever is set.
 * Other forms ignore it, and so does VX_BATCH_AUTO once the autotuner has
 * chosen a shape for the batch. */
int vx_batch_set_splits(vx_batch *batch, int32_t fill_parts, int32_t emit_parts);
/* The form the next run uses (VX_BATCH_FUSED, VX_BATCH_STAGED or
 * VX_BATCH_PIPELINED). */
int vx_batch_get_mode(vx_batch *batch, int32_t *mode);
/* Queues fill -> count -> scan -> emit on the context stream (replayed from a
 * captured hipGraph after the first run).  Asynchronous. */
int vx_batch_run(vx_ctx *ctx, vx_batch *batch);
/* Waits for the last run; if the mesh arena was too small it is grown and the
 * emit stage re-run.  Reports the batch's total face count. */
int vx_batch_wait(vx_ctx *ctx, vx_batch *batch, uint64_t *total_faces);
/* Per-chunk result: first face (batch-global) and face count. */
int vx_batch_chunk_faces(vx_ctx *ctx, vx_batch *batch, int32_t i, uint64_t *first,
                         uint64_t *count);
int vx_batch_read_ids(vx_ctx *ctx, vx_batch *batch, int32_t i, uint8_t *dst, uint64_t cap);
/* Copies the whole batch's vertex / index / instance buffers (faces in
 * request order; indices are chunk-local). */
int vx_batch_read_mesh(vx_ctx *ctx, vx_batch *batch, uint32_t format, void *dst, uint64_t cap);
int vx_batch_device_buffers(vx_batch *batch, void **instances, void **vertices, void **indices);
/* Draw list over the batch's meshes (vx_frustum_draw_list with the batch's
 * coordinates, face counts and first faces; uniform edge; the mode must
 * match a format the batch writes).  After a run. */
int vx_batch_draw_list(vx_ctx *ctx, vx_batch *batch, const vx_frustum *f, uint32_t mode,
                       vx_draw_indexed *args_dev, uint32_t *count_dev);
/* Times `reps` replays of each stage with HIP events on the context stream.
 * Staged form: ms[0..3] = average fill, count, scan, emit kernel time;
 * fused / pipelined form: ms[0] = the kernel (with its control reset),
 * ms[1..3] = 0;
 * ms[4] = whole run replayed from the graph. */
int vx_batch_profile(vx_ctx *ctx, vx_batch *batch, int32_t reps, float ms[5]);
/* Autotuner: runs the batch once (sizing its mesh arena, so every candidate
 * is timed on the whole output), then times every candidate execution form
 * and launch shape of this batch (staged; for small batches also staged with other tile shapes,
 * vx_batch_set_splits: half and twice the fill's workgroups per chunk, one
 * and four (or two) emit workgroups per item; fused with one workgroup per
 * chunk and with 3 or 4 persistent workgroups per CU; pipelined with one
 * workgroup per chunk and one mesher in three, and with 8 workgroups per CU
 * and one mesher in two, three or four; split in two halves on two streams)
 * over `reps` graph replays each, in
 * two rounds (each candidate's faster round counts), keeps the fastest for
 * the batch, and caches the choice for batches of the same shape
 * (chunk-count bucket, largest edge, uniform or ragged, neighbours).  The
 * batch is left in VX_BATCH_AUTO running the choice; shapes set before with
 * vx_batch_set_grid / _roles / _splits are cleared.  With VX_TUNE_CACHE set
 * to a file path the cache persists across processes.  Batches created later
 * in VX_BATCH_AUTO use a cached choice. */
#define VX_TUNE_MAX_CANDIDATES 12
typedef struct vx_tune_result {
    int32_t mode;  /* chosen VX_BATCH_FUSED / _STAGED / _PIPELINED / _SPLIT */
    int32_t grid;  /* chosen launch shape (vx_batch_set_grid) */
    float best_ms;
    int32_t n_candidates;
    int32_t cand_mode[VX_TUNE_MAX_CANDIDATES];
    int32_t cand_grid[VX_TUNE_MAX_CANDIDATES];
    float cand_ms[VX_TUNE_MAX_CANDIDATES];
    int32_t ratio; /* chosen pipelined roles (vx_batch_set_roles; 0 elsewhere) */
    int32_t cand_ratio[VX_TUNE_MAX_CANDIDATES];
    int32_t fill_split; /* chosen staged tile shape (vx_batch_set_splits; 0 = the rule) */
    int32_t emit_split;
    int32_t cand_fill_split[VX_TUNE_MAX_CANDIDATES];
    int32_t cand_emit_split[VX_TUNE_MAX_CANDIDATES];
} vx_tune_result;
int vx_batch_autotune(vx_ctx *ctx, vx_batch *batch, int32_t reps, vx_tune_result *result);
/* Plain path for comparison: the same work as one launch per chunk per
 * stage (no batching, no graph), chunk order preserved. */
int vx_batch_run_per_chunk(vx_ctx *ctx, vx_batch *batch);

/* ---- Serving: continuous batching of streamed chunk requests ------------ */
/* The reference's ChunkGenerator submits one chunk job per tick to a thread
 * pool and collects one result per tick (src/chunk_builder.rs:69-87,
 * 310-332).  A vx_server instead admits whatever requests are pending into
 * the next device batch as soon as an in-flight slot and HBM budget are
 * free, runs up to max_inflight batches concurrently on separate streams,
 * and keeps chunks + meshes resident in an LRU cache bounded by the
 * context's HBM budget (evicting least-recently-used entries on admission,
 * like filter_unseen_chunks drops chunks outside the box,
 * src/chunk_builder.rs:334-352). */
typedef struct vx_server vx_server;

/* Neighbour policy of a request: mesh standalone (absent neighbours, every
 * boundary face emitted), or snapshot the resident neighbours of the same
 * edge when the request is admitted (adjacent_chunks,
 * src/chunk_builder.rs:134-169). */
#define VX_NB_NONE 0u
#define VX_NB_RESIDENT 1u

typedef struct vx_request {
    vx_coord coord;
    int32_t edge;
    int32_t terrain_id;
    uint32_t neighbours; /* VX_NB_* */
    uint32_t reserved;
    uint64_t arrival_ns; /* CLOCK_MONOTONIC ns of arrival; 0 = at submission */
} vx_request;

typedef struct vx_completion {
    uint64_t ticket;
    vx_coord coord;
    int32_t edge;
    int32_t status;      /* VX_OK, or the error that failed the request */
    uint32_t cache_hit;  /* 1 when served from the resident cache */
    uint64_t faces;
    uint64_t arrival_ns; /* as submitted */
    uint64_t done_ns;    /* CLOCK_MONOTONIC ns at which the result was ready */
    int32_t nb_present[4]; /* neighbour snapshot used: -x, +x, -z, +z */
} vx_completion;

typedef struct vx_server_desc {
    int32_t max_batch;    /* requests per device batch (0 -> 512) */
    int32_t max_inflight; /* batches on the device at once (0 -> 4) */
    uint32_t formats;     /* VX_FMT_* kept resident (0 -> VX_FMT_VERTICES) */
    uint32_t flags;       /* reserved, 0 */
} vx_server_desc;

typedef struct vx_server_stats {
    uint64_t submitted, completed, cache_hits, coalesced, generated, batches, evictions,
        failed;
    uint64_t resident_chunks, resident_bytes, peak_resident_bytes, budget_bytes;
} vx_server_stats;

int vx_server_create(vx_ctx *ctx, const vx_server_desc *desc, vx_server **out);
int vx_server_destroy(vx_server *server);
/* Queues n requests; tickets[i] identifies request i in completions. */
int vx_server_submit(vx_server *server, const vx_request *reqs, int32_t n, uint64_t *tickets);
/* Makes progress without blocking longer than wait_us: retires finished
 * device work, admits pending requests.  Returns the number of completions
 * ready to poll (>= 0) or an error status. */
int vx_server_pump(vx_server *server, int32_t wait_us);
int vx_server_poll(vx_server *server, vx_completion *out, int32_t cap, int32_t *n);
/* Blocks until every submitted request has completed. */
int vx_server_drain(vx_server *server);
/* Resident chunk + mesh for (coord, edge, terrain_id), retained for the
 * caller (either output may be NULL); VX_ERR_NOT_READY when not resident.
 * The cache keys chunks by coordinate, edge and terrain: requests for one
 * coordinate under two terrains are two chunks, and a neighbour snapshot
 * (VX_NB_RESIDENT) only takes neighbours of the request's own terrain.
 * A completed request's entry is not evicted before the second pump after
 * the one (or the submit) that completed it, so a caller that polls and
 * looks up its completions between pumps always finds them; under a budget
 * that binds, later admissions wait for that instead. */
int vx_server_lookup(vx_server *server, vx_coord coord, int32_t edge, int32_t terrain_id,
                     vx_chunk **chunk, vx_mesh **mesh);
int vx_server_stats_get(vx_server *server, vx_server_stats *stats);

/* Streaming simulation: a camera on a seeded random walk, Poisson request
 * arrivals around it within the render radius, mixed chunk edges, served
 * by a vx_server in real time; latency = completion - scheduled arrival. */
typedef struct vx_sim_desc {
    uint64_t seed;
    double rate_per_s;          /* Poisson arrival rate */
    double duration_s;          /* arrivals are generated for this long */
    double speed_blocks_per_s;  /* camera speed */
    int32_t radius_chunks;      /* render radius in 32-voxel chunks */
    int32_t n_classes;          /* chunk size classes (1..3) */
    int32_t edges[3];
    float weights[3];
    int32_t terrain_id[3];
    uint32_t neighbours;        /* VX_NB_* for every request */
    vx_server_desc server;
} vx_sim_desc;

typedef struct vx_sim_result {
    uint64_t requests, completed, cache_hits, generated, batches, evictions, faces;
    double wall_s, throughput_rps, generated_per_s;
    double p50_ms, p99_ms, max_ms, mean_ms;          /* all requests */
    double miss_p50_ms, miss_p99_ms, miss_max_ms;    /* requests that generated */
    uint64_t peak_resident_bytes;
} vx_sim_result;

int vx_serve_simulate(vx_ctx *ctx, const vx_sim_desc *desc, vx_sim_result *result);

/* ---- World + ChunkGenerator, tick for tick -------------------------------
 * The reference's scheduler (ChunkGenerator::build_chunks,
 * src/chunk_builder.rs:45-352) and world maps (World, src/world.rs:14-58)
 * with the same per-tick steps: primary ring of radius 1 when the player
 * enters a new chunk, frustum-tested rings 2..render_distance-1 at y = 0,
 * one queued load and (when nothing is loading) one rebuild submitted per
 * tick, one finished chunk and one rebuilt mesh received per tick, rebuilds
 * queued for the resident neighbours of every arrival, the rebuild queue
 * sorted and de-duplicated, chunks outside the +-render_distance box
 * unloaded.  Jobs (Chunk::new + create_mesh with the resident-neighbour
 * snapshot at submission, or create_mesh of a resident chunk) run on the
 * GPU.  A job's result becomes receivable `result_delay` ticks after
 * submission, in submission order: one of the schedules the reference's
 * thread pool can produce, fixed so that a run is reproducible.  Chunk edge
 * 32 (src/world.rs:21-22). */
typedef struct vx_world vx_world;
typedef struct vx_world_desc {
    int32_t terrain_id;
    int32_t render_distance;  /* RENDER_DISTANCE, 5 in the reference (src/world.rs:19) */
    int32_t result_delay;     /* ticks from a job's submission to its receipt (>= 0) */
    uint32_t formats;         /* VX_FMT_* of the meshes */
    float player_pos[3];      /* Player::new (src/player.rs:10-14): world units */
} vx_world_desc;
typedef struct vx_world_stats {
    uint64_t tick;
    int32_t chunks, meshes, rebuild_queue, in_process;
    uint64_t jobs_loaded, jobs_rebuilt, received_chunks, received_meshes, unloaded;
    vx_coord player_chunk;
} vx_world_stats;
int vx_world_create(vx_ctx *ctx, const vx_world_desc *desc, vx_world **out);
int vx_world_destroy(vx_world *world);
/* One World::update (src/engine.rs:65-74): the player's position this tick
 * (world units; Player::update_pos) and the frustum (Frustum::update). */
int vx_world_tick(vx_world *world, const float player_pos[3], const vx_frustum *f);
int vx_world_stats_get(vx_world *world, vx_world_stats *out);
/* Coordinates of the resident chunks / meshes (sorted by x, y, z), or the
 * rebuild queue / chunks in process (in queue order).  With cap = 0 only
 * *n is set. */
#define VX_WORLD_CHUNKS 1u
#define VX_WORLD_MESHES 2u
#define VX_WORLD_REBUILD_QUEUE 3u
#define VX_WORLD_IN_PROCESS 4u
int vx_world_list(vx_world *world, uint32_t what, vx_coord *out, int32_t cap, int32_t *n);
/* The resident chunk and mesh at a coordinate (retained handles, NULL when
 * absent; pass NULL for the one you do not want). */
int vx_world_lookup(vx_world *world, vx_coord c, vx_chunk **chunk, vx_mesh **mesh);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif /* VOXEL_H */
