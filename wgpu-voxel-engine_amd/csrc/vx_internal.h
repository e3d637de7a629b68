// vx_internal.h -- structures shared by the host runtime (vx_api.cpp) and the
// gfx950 kernels (vx_kernels.hip).  Plain-old-data only: every struct here is
// copied byte-for-byte into device memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vx {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kMaxOctaves = 8;     // fBm octaves supported by the noise tables
constexpr int kDescTile = 4096;    // face descriptors staged in LDS per emit tile
constexpr int kEmitGrid = 1280;    // emit workgroups that fill the GPU (5 per CU)
constexpr int kFillGrid = 1024;    // fill workgroups that fill the GPU (4 per CU)

// Terrain kinds (mirror of vx_terrain.kind in include/voxel.h).
enum : int32_t { kAir = 0, kSolid = 1, kHeightmap = 2, kDensity = 3, kReference = 4 };

// Block ids stored in the u8 voxel array: 0 air, 1 grass, 2 dirt.  The
// reference's CubeType is {GRASS=0, DIRT=1, AIR=2} (src/cube.rs:16-21).
enum : uint32_t { kBlockAir = 0, kBlockGrass = 1, kBlockDirt = 2, kBlockMixed = 0xFFu };

// Output formats (mirror of VX_FMT_* in include/voxel.h).
enum : uint32_t { kFmtInstances = 1u, kFmtVertices = 2u };

// Per-terrain constants, resident in device memory.  The permutation is the
// splitmix64/Fisher-Yates shuffle of 0..255 for the seed, doubled to 512.
struct TerrainDev {
    int32_t kind;
    int32_t octaves;
    float frequency;
    float lacunarity;
    float gain;
    float height_scale;
    int32_t height_max;
    float y_scale;
    int32_t perm[512];
};

// Per-hash gradient coefficients for the column-hoisted noise evaluation,
// indexed [corner k][hash & 15] with corner k = x | row << 1 | col << 2 (row
// = world y, the per-row axis; col = world z).  grad_k = fma(yr, g_row,
// a_x*xr + a_col*zr); the +-1 entries select and sign an operand, the zero
// entries are the signed zero that makes the product vanish exactly (-0
// when the operand is >= 0, +0 when it is < 0), so the sum equals Perlin's
// (+-u) + (+-v) bit for bit.
struct GradLut {
    float4 c[8][16];  // noise at (wx, wy, wz): (a_x, a_col, g_row) = (gx, gz, gy)
    float4 r[8][16];  // noise at (wx, wz, wy), the reference's axes: (gx, gy, gz)
};

// Meshing constants computed on the host from the rotation of each face
// direction (nalgebra from_euler_angles restated) and the translation sign:
//   P[d][r][s][corner] = ((M(r,0)*cx + M(r,1)*cy) + M(r,2)*cz)
//   Mb[d][s][r][c]     = M(r,c) of T*R for a translation component of sign s
//                        (c = 3 unused, zero)
// where s = 1 iff that translation component is negative.  The innermost
// four floats of each are what one face reads for one coordinate r: a
// single 16-byte LDS read.
struct MeshTables {
    alignas(16) float P[6][3][2][4];
    alignas(16) float Mb[6][2][3][4];
    uint32_t tex[3][6];  // [block id][dir] -> texture layer (row 0 unused)
};

// One chunk to fill: terrain -> ids, occupancy rows, boundary planes, meta.
struct FillWork {
    int32_t cx, cy, cz, edge;
    int32_t terrain;       // index into the terrain table
    int32_t pad0;
    uint8_t *ids;          // edge^3
    uint32_t *occ;         // edge^2 rows * W words (W = 1 for edge <= 32, 2 for 64)
    uint32_t *planes;      // 4 planes * edge * W words: x=0, x=N-1, z=0, z=N-1
    uint32_t *meta;        // [0] solid block id (or kBlockMixed)
    float *noise_out;      // optional raw noise (density: edge^3, heightmap: edge^2)
    // Optional (fill_kernel, one workgroup per chunk): the fill also counts
    // the chunk's visible faces, for chunks whose every neighbour is absent
    // or was written before the launch.  cnb[k] = the plane of neighbour k
    // (-x, +x, -z, +z) facing this chunk, null = absent; count_out[0] = the
    // face count; slab_out[0..kSlabs) = a 64^3 chunk's per-slab counts.
    const uint32_t *cnb[4];
    uint32_t *count_out;   // null: the count kernel counts
    uint32_t *slab_out;
};

// One chunk to mesh.  nb[k] points at the neighbour's plane facing this
// chunk (k = -x, +x, -z, +z -> its x=N-1, x=0, z=N-1, z=0 plane) or is null
// when the neighbour is absent (chunk.rs:89-107: absent -> emit).
struct MeshWork {
    int32_t ox, oy, oz, edge;  // world origin of the chunk
    const uint32_t *occ;
    const uint8_t *ids;
    const uint32_t *nb[4];
    uint32_t solid_type;       // block id of every solid voxel, or kBlockMixed
    uint32_t pad0;
};

// Batch-wide mesh outputs.
struct MeshOut {
    float *vertices;      // 20 floats per face (4 x {pos xyz, uv})
    uint32_t *indices;    // 6 per face, chunk-local
    uint32_t *instances;  // 17 dwords per face (InstanceRaw)
    uint64_t capacity;    // faces
    uint32_t formats;
    uint32_t pad0;
};

// Fused noise -> cull -> compact: per-side source of the neighbour plane.
enum : uint32_t { kHaloAbsent = 0u, kHaloResident = 1u, kHaloRecompute = 2u };

// One chunk of the fused kernel: its fill outputs, its mesh description
// (m.nb[k] read for resident sides) and how each neighbour plane is obtained.
// kHaloRecompute is only used when that neighbour is the adjacent chunk of
// the same batch with the same terrain and edge, so the recomputed plane is
// bit-identical to the one the neighbour writes itself.
struct FusedWork {
    FillWork f;
    MeshWork m;
    uint32_t halo[4];
    int32_t nbi[4];  // batch index of each in-batch (kHaloRecompute) neighbour, else -1
};

// Control words of one fused run; zeroed by the reset kernel.
struct FusedCtl {
    uint32_t *ticket;    // next chunk to hand out (processing order = output order)
    uint64_t *status;    // per chunk: flag(2 bits) | face count or inclusive prefix
    uint32_t *counts;    // per chunk face count (for the host)
    uint64_t *bases;     // per chunk first face (for the host)
    uint64_t *total;     // faces of the whole batch
    uint32_t *overflow;  // set when a chunk did not fit the arena
    uint32_t *error;     // set when a look-back wait exceeded its bound
    uint32_t *filled;    // per chunk: its fill has published (pipelined form; may be null)
    uint32_t *mesh_ticket;  // next chunk to mesh (pipelined form; may be null)
};

// Frustum job carried by a fused launch: planes in device memory (updated
// per frame without re-capturing the graph), n chunk coordinates, edge,
// one mask byte per chunk.  n = 0 disables it.
struct CullJob {
    const float *planes;
    const int32_t *coords;
    uint8_t *mask;
    int32_t n;
    int32_t edge;
};

// Six frustum planes [near far left right top bottom] x (a b c d), passed
// to the cull kernel by value.
struct Planes {
    float p[24];
};

// Indirect draw arguments (the layout of vx_draw_indexed in include/voxel.h
// and of WebGPU's drawIndexedIndirect buffer).
struct DrawArgs {
    uint32_t index_count, instance_count, first_index;
    int32_t base_vertex;
    uint32_t first_instance;
};
enum : uint32_t { kDrawInstanced = 1u, kDrawExpanded = 2u };

inline int words_per_row(int edge) { return (edge + 31) / 32; }

// ---- kernel launchers (vx_kernels.hip) --------------------------------
// Schedules: `order` (fill, emit) lists the n items in launch order, or is
// null for 0..n-1.  An entry outside [0, n) makes its workgroup do nothing
// and sets a sticky error word: `error` for the fill, overflow[1] for the
// emit, whose overflow[0] is the arena-overflow flag the scan clears.
// Workgroups per chunk of the fill: one for batches of kFillGrid chunks or
// more; below that 2, 4 or 8 so that the grid still covers about kFillGrid
// workgroups (8 only for batches of at most kFillGrid / 8 chunks that hold
// 64^3 chunks).  `hint` (1, 2, 4 or 8; 0 = this rule) replaces the rule
// for the batches it would split.
inline int fill_parts_for(int32_t n, int32_t max_edge, int32_t hint = 0) {
    if (n >= kFillGrid) return 1;
    if (hint > 0) return hint;
    return 2 * n >= kFillGrid ? 2 : (max_edge > 32 && 8 * n <= kFillGrid) ? 8 : 4;
}
// Workgroups per emit item: one for kEmitGrid items or more; below that up to
// 4, so that the grid still covers about kEmitGrid workgroups.  `hint` (1..4;
// 0 = this rule) replaces the rule for the batches it would split.
inline int emit_parts_for(int32_t items, int32_t hint = 0) {
    if (items >= kEmitGrid) return 1;
    if (hint > 0) return hint;
    const int p = (kEmitGrid + items - 1) / items;
    return p < 4 ? p : 4;
}
// parts_hint: fill_parts_for's hint.
hipError_t launch_fill(const FillWork *work, const int32_t *order, int32_t n, int32_t max_edge,
                       const TerrainDev *terrains, const GradLut *lut, uint32_t *error,
                       int32_t parts_hint, hipStream_t s);
// True iff launch_fill gives each chunk one workgroup (fill_kernel), so that
// FillWork.count_out is honoured for every chunk of the batch (whatever the
// hint).
inline bool fill_whole_chunks(int32_t n) { return n >= kFillGrid; }
// Batches that hold 64^3 chunks emit by items: each 64^3 chunk is kSlabs
// y-slabs of kSlabLayers layers (face counts per slab in slabs[kSlabs * i
// ..], written by the count kernel), any other chunk one item.  With `items`
// (the batch's mesh work) the scan's order lists item codes chunk << 2 |
// slab, at most kSlabs * n of them; launch_emit takes their number.
constexpr int kSlabs = 4;
hipError_t launch_count(const MeshWork *work, int32_t n, int32_t max_edge, uint32_t *counts,
                        uint32_t *slabs, hipStream_t s);
// base_in (device, may be null): faces before these chunks, added to every
// base and to the total (the second half of a split batch).
hipError_t launch_scan(const uint32_t *counts, int32_t n, const MeshWork *items, uint64_t *bases,
                       uint64_t *total, uint32_t *overflow, int32_t *order_out, hipStream_t s,
                       const uint64_t *base_in = nullptr);
hipError_t launch_fill_one(const FillWork &w, const TerrainDev *terrains, const GradLut *lut,
                           hipStream_t s);
hipError_t launch_pack_one(const FillWork &w, hipStream_t s);
hipError_t launch_count_one(const MeshWork &w, uint32_t *count, hipStream_t s);
hipError_t launch_emit_one(const MeshWork &w, uint64_t base, uint32_t faces, MeshOut out,
                           const MeshTables *tables, uint32_t *overflow, hipStream_t s);
// parts_hint: emit_parts_for's hint.
hipError_t launch_emit(const MeshWork *work, const int32_t *order, int32_t n, int32_t n_items,
                       int32_t max_edge, const uint64_t *bases, const uint32_t *counts,
                       const uint32_t *slabs, MeshOut out, const MeshTables *tables,
                       uint32_t *overflow, int32_t parts_hint, hipStream_t s);
// Emit items of a batch: its chunks, or for batches that hold 64^3 chunks
// kSlabs per 64^3 chunk and one per other chunk.
inline int32_t emit_items_of(int32_t n, int32_t n_items, int32_t max_edge) {
    return max_edge <= 32 ? n : n_items;
}
hipError_t launch_fused(const FusedWork *work, int32_t n, int32_t max_edge,
                        const TerrainDev *terrains, const GradLut *lut, const MeshTables *tables,
                        FusedCtl ctl, MeshOut out, const CullJob &cull, int32_t grid,
                        hipStream_t s);
// Pipelined form (edges up to 32): fill and mesh tasks from two ticket
// counters; one workgroup in `ratio` (2..kPipeRatioMax, 0 = kPipeRatio)
// meshes from the start, the others fill and then mesh; mesh m is served
// once fills up to m + lag are out (lag >= the farthest in-batch neighbour
// index minus the chunk's own).  Needs ctl.filled and ctl.mesh_ticket.
constexpr int kPipeRatio = 3;
constexpr int kPipeRatioMax = 8;
hipError_t launch_pipe(const FusedWork *work, int32_t n, int32_t max_edge, int32_t lag,
                       const TerrainDev *terrains, const GradLut *lut, const MeshTables *tables,
                       FusedCtl ctl, MeshOut out, const CullJob &cull, int32_t grid, int32_t ratio,
                       hipStream_t s);
hipError_t launch_cull(const CullJob &job, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t s);
hipError_t launch_frustum(const Planes &planes, const int32_t *coords, int32_t n, int32_t edge,
                          uint8_t *mask, hipStream_t s);
// Workgroups (and block_counts entries) a draw list of n meshes uses.
int32_t draw_blocks(int32_t n);
// The draw-list scratch array of the allocation-free entry point: device
// address of `device`'s copy (looked up once per device, with that device
// current at the first call), and the most meshes it serves.
constexpr int kMaxDevices = 64;
hipError_t draw_scratch_address(int device, uint32_t **ptr);
int32_t draw_scratch_max_meshes();
// Zeroes n <= 64 words with a kernel (capturable; no memset node).
hipError_t launch_clear_words(uint32_t *p, int32_t n, hipStream_t s);
hipError_t launch_draw_list(const Planes &planes, const int32_t *coords, const uint32_t *faces,
                            const uint64_t *first, int32_t n, int32_t edge, uint32_t mode,
                            uint32_t *block_counts, DrawArgs *args, uint32_t *count, hipStream_t s);

}  // namespace vx
