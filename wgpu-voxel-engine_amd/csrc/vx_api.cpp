// vx_api.cpp -- host runtime behind include/voxel.h: contexts, a caching
// device heap with an HBM budget, terrain tables, chunk/mesh handles, batches
// (captured into a hipGraph) and the frustum entry points.
//
// Reference surface mirrored (paths relative to the reference repository):
//   Chunk::new / create_mesh        src/chunk.rs:21-57
//   ChunkMesh::new                  src/chunk.rs:254-266
//   Quad::new / to_raw              src/quad.rs:14-54 (rotation tables below)
//   Frustum::matrix_to_planes/contains  src/frustum_culling.rs:16-85
//   pool.execute per chunk          src/chunk_builder.rs:80-84 (batches)
#include "vx_internal.h"
#include "vx_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

using namespace vx;

// ------------------------------------------------------------------ errors
namespace {

const char *kStatusText[] = {"ok",          "invalid argument", "out of memory",
                             "device error", "capacity too small", "HBM budget exceeded",
                             "not ready",   "no GPU device"};

}  // namespace

#define VX_HIP(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->set_error(e_, #call);                                                    \
            return VX_ERR_DEVICE;                                                           \
        }                                                                                   \
    } while (0)

// ------------------------------------------------------------ device heap
size_t DeviceHeap::size_class(size_t n) {
    if (n < 256) n = 256;
    // classes: 4 steps per power of two -> at most 25% slack
    size_t p = 256;
    while (p < n) p <<= 1;
    const size_t q = p >> 3;
    size_t c = p >> 1;
    while (c < n) c += q;
    return c;
}

hipError_t DeviceHeap::alloc(size_t n, void **out) {
    const size_t c = size_class(n);
    auto it = free_.find(c);
    if (it != free_.end() && !it->second.empty()) {
        *out = it->second.back();
        it->second.pop_back();
        cached_ -= c;
        live_ += c;
        return hipSuccess;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();  // give cached blocks back and retry once
        e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return e;
        }
    }
    sizes_[p] = c;
    live_ += c;
    *out = p;
    return hipSuccess;
}

void DeviceHeap::release_pending(void *p) {
    if (!p) return;
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    live_ -= it->second;
    cached_ += it->second;
    pending_.push_back(p);
}

void DeviceHeap::flush_pending() {
    for (void *p : pending_) free_[sizes_[p]].push_back(p);
    pending_.clear();
}

void DeviceHeap::trim() {
    for (auto &kv : free_) {
        for (void *p : kv.second) {
            (void)hipFree(p);
            sizes_.erase(p);
        }
        kv.second.clear();
    }
    cached_ = 0;
}

DeviceHeap::~DeviceHeap() {
    for (auto &kv : sizes_) (void)hipFree(kv.first);
}

// ------------------------------------------------------------ tables
namespace {

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Seeded permutation of 0..255 (Fisher-Yates driven by splitmix64), doubled.
void make_perm(uint32_t seed, int32_t perm[512]) {
    int32_t p[256];
    for (int i = 0; i < 256; i++) p[i] = i;
    uint64_t s = seed;
    for (int i = 255; i >= 1; i--) {
        const int j = (int)(splitmix64(s) % (uint64_t)(i + 1));
        std::swap(p[i], p[j]);
    }
    for (int i = 0; i < 512; i++) perm[i] = p[i & 255];
}

// Gradient coefficient tables for Perlin's grad(hash, x, y, z) in the
// column-hoisted evaluation: per corner k (bit0 x+1, bit1 row+1, bit2
// col+1; row = world y, col = world z) and hash h, (a_x, a_col, g_row) such
// that fma(row_r, g_row, fma(a_x, xr, a_col*col_r)) == ((h&1)?-u:u) +
// ((h&2)?-v:v) exactly.  Table c samples the noise at (wx, wy, wz) (noise y
// = row, noise z = col); table r at (wx, wz, wy), the reference's axis order
// (src/perlin_noise.rs:6-13: noise y = col, noise z = row).
void make_grad_lut(GradLut &L) {
    for (int k = 0; k < 8; k++) {
        // zero whose product with the corner's relative coordinate is -0:
        // the coordinate is >= 0 on the low side (x - X) and < 0 on the high
        // side (x - X - 1).
        const float zx = (k & 1) ? 0.0f : -0.0f;
        const float zrow = (k & 2) ? 0.0f : -0.0f;
        const float zcol = (k & 4) ? 0.0f : -0.0f;
        for (int h = 0; h < 16; h++) {
            const float su = (h & 1) ? -1.0f : 1.0f, sv = (h & 2) ? -1.0f : 1.0f;
            // Perlin: u = h<8 ? x : y; v = h<4 ? y : (h==12||h==14 ? x : z)
            // coefficient of the noise's x, y and z axes, zero elsewhere
            float g[3];
            const int ui = h < 8 ? 0 : 1;
            const int vi = h < 4 ? 1 : (h == 12 || h == 14 ? 0 : 2);
            g[0] = g[1] = g[2] = 0.0f;
            g[ui] = su;
            g[vi] = sv;
            const bool nz[3] = {ui == 0 || vi == 0, ui == 1 || vi == 1, ui == 2 || vi == 2};
            // table c: noise (x, y, z) = (x, row, col)
            L.c[k][h] = make_float4(nz[0] ? g[0] : zx, nz[2] ? g[2] : zcol, nz[1] ? g[1] : zrow,
                                    0.0f);
            // table r: noise (x, y, z) = (x, col, row)
            L.r[k][h] = make_float4(nz[0] ? g[0] : zx, nz[1] ? g[1] : zcol, nz[2] ? g[2] : zrow,
                                    0.0f);
        }
    }
}

// Rotation of each face direction: Quad::new (src/quad.rs:14-46) calls
// nalgebra 0.29 Rotation3::from_euler_angles(roll, pitch, yaw) with
// UP roll -90, DOWN roll 90, LEFT pitch -90, RIGHT pitch 90, FRONT 0,
// BACK pitch 180 degrees; degrees -> radians as f32::to_radians
// (deg * (PI / 180) in f32); R rows as nalgebra builds them.
void face_rotation(int dir, float R[3][3]) {
    const float k = 3.14159265358979323846f / 180.0f;
    float roll = 0.0f, pitch = 0.0f, yaw = 0.0f;
    switch (dir) {
    case 5: roll = -90.0f * k; break;   // UP
    case 4: roll = 90.0f * k; break;    // DOWN
    case 0: pitch = -90.0f * k; break;  // LEFT
    case 1: pitch = 90.0f * k; break;   // RIGHT
    case 3: break;                      // FRONT
    case 2: pitch = 180.0f * k; break;  // BACK
    }
    const float sr = sinf(roll), cr = cosf(roll);
    const float sp = sinf(pitch), cp = cosf(pitch);
    const float sy = sinf(yaw), cy = cosf(yaw);
    R[0][0] = cy * cp;
    R[0][1] = cy * sp * sr - sy * cr;
    R[0][2] = cy * sp * cr + sy * sr;
    R[1][0] = sy * cp;
    R[1][1] = sy * sp * sr + cy * cr;
    R[1][2] = sy * sp * cr - cy * sr;
    R[2][0] = -sp;
    R[2][1] = cp * sr;
    R[2][2] = cp * cr;
}

// Quad::to_raw (src/quad.rs:47-54): M = T.to_homogeneous() * R.to_homogeneous();
// for rows r < 3 and columns c < 3 the product depends on the translation only
// through the sign of t_r (t_r * 0), so it is tabulated for t_r = +1 and -1.
// The vertex expansion (src/shaders/vertex.vert:24, left to right) then is
// pos[r] = P[d][sign(t_r)][corner][r] + t_r.
void make_mesh_tables(MeshTables &T) {
    static const float corner[4][3] = {
        {-0.5f, 0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}};
    for (int d = 0; d < 6; d++) {
        float R[3][3];
        face_rotation(d, R);
        float H[4][4] = {};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) H[i][j] = R[i][j];
        H[3][3] = 1.0f;
        for (int s = 0; s < 2; s++) {
            const float t = s ? -1.0f : 1.0f;
            for (int r = 0; r < 3; r++) {
                float Trow[4] = {0.0f, 0.0f, 0.0f, t};
                Trow[r] = 1.0f;
                for (int c = 0; c < 3; c++) {
                    float acc = Trow[0] * H[0][c];
                    acc = acc + Trow[1] * H[1][c];
                    acc = acc + Trow[2] * H[2][c];
                    acc = acc + Trow[3] * H[3][c];
                    T.Mb[d][s][r][c] = acc;
                }
                T.Mb[d][s][r][3] = 0.0f;
                for (int q = 0; q < 4; q++) {
                    float acc = T.Mb[d][s][r][0] * corner[q][0];
                    acc = acc + T.Mb[d][s][r][1] * corner[q][1];
                    acc = acc + T.Mb[d][s][r][2] * corner[q][2];
                    T.P[d][r][s][q] = acc;
                }
            }
        }
    }
    // TEXTURE_INDEX_LIST (src/texture.rs:238-243) by block id; per face slot
    // LEFT ti[0] RIGHT ti[1] BACK ti[2] FRONT ti[3] DOWN ti[5] UP ti[4]
    // (src/chunk.rs:85,118,149,182,213,231).
    static const uint32_t list[2][6] = {{0, 0, 0, 0, 2, 1}, {1, 1, 1, 1, 1, 1}};
    static const int slot[6] = {0, 1, 2, 3, 5, 4};
    for (int d = 0; d < 6; d++) {
        T.tex[0][d] = 0;
        T.tex[1][d] = list[0][slot[d]];
        T.tex[2][d] = list[1][slot[d]];
    }
}

bool valid_edge(int e) { return e == 16 || e == 32 || e == 64; }

}  // namespace

// ------------------------------------------------------------ context
void vx_ctx::set_error(hipError_t e, const char *what) {
    (void)hipGetLastError();
    char buf[512];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    last_error = buf;
}

int vx_ctx::dev_alloc(size_t n, void **out) {
    if (heap.has_pending()) {
        // blocks released since the last allocation become reusable once
        // the context stream has run everything queued before their release
        // (other streams of this library, the server's slots, only release
        // blocks after their own work on them has completed)
        if (free_fence) (void)hipEventSynchronize(free_fence);
        heap.flush_pending();
    }
    const size_t c = DeviceHeap::size_class(n);
    if (budget && heap.live() + c > budget) return VX_ERR_BUDGET;
    hipError_t e = heap.alloc(n, out);
    if (e != hipSuccess) {
        set_error(e, "hipMalloc");
        return VX_ERR_NOMEM;
    }
    return VX_OK;
}

// A released block may still be read or written by work queued on the
// context stream (an emit into a mesh whose handle the caller released right
// away, a fill of a chunk released before it ran).  The fence recorded here
// is waited on before the block is handed out again, so an allocation used
// on another stream (a server slot) never overlaps that work.
void vx_ctx::dev_free(void *p) {
    if (!p) return;
    if (free_fence) (void)hipEventRecord(free_fence, stream);
    heap.release_pending(p);
}

int vx_ctx::ensure_scratch(size_t bytes) {
    if (scratch_bytes >= bytes) return VX_OK;
    if (scratch) dev_free(scratch);
    scratch = nullptr;
    scratch_bytes = 0;
    int st = dev_alloc(bytes, &scratch);
    if (st) return st;
    scratch_bytes = DeviceHeap::size_class(bytes);
    return VX_OK;
}

extern "C" int vx_abi_version(void) { return VX_ABI_VERSION; }

extern "C" const char *vx_status_string(int status) {
    if (status > 0 || status < -7) return "unknown status";
    return kStatusText[-status];
}

extern "C" int vx_device_count(int32_t *count) {
    if (!count) return VX_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return VX_OK;
}

extern "C" int vx_ctx_create(const vx_ctx_desc *desc, vx_ctx **out) {
    if (!out) return VX_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return VX_ERR_NODEV;
    }
    const int dev = desc ? desc->device : 0;
    if (dev < 0 || dev >= ndev) return VX_ERR_INVALID;
    if (hipSetDevice(dev) != hipSuccess) return VX_ERR_DEVICE;
    vx_ctx *c = new (std::nothrow) vx_ctx();
    if (!c) return VX_ERR_NOMEM;
    c->device = dev;
    c->budget = desc ? desc->hbm_budget : 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VX_ERR_DEVICE;
    }
    if (hipEventCreateWithFlags(&c->free_fence, hipEventDisableTiming) != hipSuccess) {
        c->free_fence = nullptr;
        vx_ctx_destroy(c);
        return VX_ERR_DEVICE;
    }
    // device address of this GPU's draw-list scratch (the device is current)
    if (hipError_t e = draw_scratch_address(dev, &c->draw_scratch); e != hipSuccess) {
        c->set_error(e, "hipGetSymbolAddress(draw-list scratch)");
        vx_ctx_destroy(c);
        return VX_ERR_DEVICE;
    }
    GradLut lut;
    make_grad_lut(lut);
    MeshTables tabs;
    make_mesh_tables(tabs);
    int st = VX_OK;
    if ((st = c->dev_alloc(sizeof(GradLut), (void **)&c->lut_dev)) ||
        (st = c->dev_alloc(sizeof(MeshTables), (void **)&c->tables_dev)) ||
        (st = c->dev_alloc(kMaxTerrains * sizeof(TerrainDev), (void **)&c->terrains_dev)) ||
        (st = c->dev_alloc(64, (void **)&c->small_dev))) {
        vx_ctx_destroy(c);
        return st;
    }
    if (hipHostMalloc((void **)&c->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        c->pinned = nullptr;
        vx_ctx_destroy(c);
        return VX_ERR_NOMEM;
    }
    if (hipMemcpyAsync(c->lut_dev, &lut, sizeof(lut), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->tables_dev, &tabs, sizeof(tabs), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        vx_ctx_destroy(c);
        return VX_ERR_DEVICE;
    }
    *out = c;
    return VX_OK;
}

extern "C" int vx_ctx_destroy(vx_ctx *ctx) {
    if (!ctx) return VX_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->free_fence) (void)hipEventDestroy(ctx->free_fence);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;  // DeviceHeap frees every device block
    return VX_OK;
}

extern "C" int vx_ctx_synchronize(vx_ctx *ctx) {
    if (!ctx) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_ctx_stream(vx_ctx *ctx, void **s) {
    if (!ctx || !s) return VX_ERR_INVALID;
    *s = (void *)ctx->stream;
    return VX_OK;
}

extern "C" int vx_ctx_memory(vx_ctx *ctx, uint64_t *resident, uint64_t *budget) {
    if (!ctx) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (resident) *resident = ctx->heap.live();
    if (budget) *budget = ctx->budget;
    return VX_OK;
}

extern "C" const char *vx_ctx_last_error(vx_ctx *ctx) {
    if (!ctx) return "no context";
    return ctx->last_error.c_str();
}

extern "C" int vx_buffer_alloc(vx_ctx *ctx, uint64_t bytes, void **dev) {
    if (!ctx || !dev || bytes == 0) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return ctx->dev_alloc(bytes, dev);
}

extern "C" int vx_buffer_free(vx_ctx *ctx, void *dev) {
    if (!ctx || !dev) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // no queued work may still use it
    ctx->dev_free(dev);
    return VX_OK;
}

extern "C" int vx_buffer_upload(vx_ctx *ctx, void *dev, const void *host, uint64_t bytes) {
    if (!ctx || !dev || (!host && bytes)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_buffer_download(vx_ctx *ctx, void *host, const void *dev, uint64_t bytes) {
    if (!ctx || !dev || (!host && bytes)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_buffer_copy(vx_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || !dst || !src || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, launch_copy(dst, src, bytes, ctx->stream));
    return VX_OK;
}

extern "C" int vx_terrain_register(vx_ctx *ctx, const vx_terrain *t, int32_t *id) {
    if (!ctx || !t || !id) return VX_ERR_INVALID;
    if (t->kind < VX_TERRAIN_AIR || t->kind > VX_TERRAIN_REFERENCE) return VX_ERR_INVALID;
    if (t->kind == VX_TERRAIN_HEIGHTMAP || t->kind == VX_TERRAIN_DENSITY ||
        t->kind == VX_TERRAIN_REFERENCE) {
        if (t->octaves < 1 || t->octaves > kMaxOctaves) return VX_ERR_INVALID;
        // every octave frequency must keep x - floor(x) - 1 strictly negative
        float f = t->frequency;
        for (int o = 0; o < t->octaves; o++) {
            if (!(f >= 9.5367431640625e-07f) || !std::isfinite(f)) return VX_ERR_INVALID;
            f = f * t->lacunarity;
        }
        if (!std::isfinite(t->gain)) return VX_ERR_INVALID;
        if (t->kind == VX_TERRAIN_HEIGHTMAP &&
            (!std::isfinite(t->height_scale) || t->height_max < 0))
            return VX_ERR_INVALID;
        if (t->kind == VX_TERRAIN_DENSITY && !std::isfinite(t->y_scale)) return VX_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if ((int)ctx->terrains.size() >= kMaxTerrains) return VX_ERR_NOMEM;
    TerrainDev T{};
    T.kind = t->kind;
    T.octaves = t->octaves;
    T.frequency = t->frequency;
    T.lacunarity = t->lacunarity;
    T.gain = t->gain;
    T.height_scale = t->height_scale;
    T.height_max = t->height_max;
    T.y_scale = t->y_scale;
    make_perm(t->seed, T.perm);
    const int idx = (int)ctx->terrains.size();
    ctx->terrains.push_back(T);
    VX_HIP(ctx, hipMemcpyAsync(ctx->terrains_dev + idx, &ctx->terrains.back(), sizeof(TerrainDev),
                               hipMemcpyHostToDevice, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *id = idx;
    return VX_OK;
}

// ------------------------------------------------------------ chunks
namespace {

struct ChunkLayout {
    size_t ids, occ, planes, meta, total;
};

ChunkLayout chunk_layout(int N) {
    const size_t W = (size_t)words_per_row(N);
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    ChunkLayout L;
    L.ids = 0;
    L.occ = up((size_t)N * N * N);
    L.planes = L.occ + up((size_t)N * N * W * 4);
    L.meta = L.planes + up(4 * (size_t)N * W * 4);
    L.total = L.meta + 256;
    return L;
}

void bind_chunk_ptrs(vx_chunk *c, char *base) {
    const ChunkLayout L = chunk_layout(c->edge);
    c->ids = (uint8_t *)(base + L.ids);
    c->occ = (uint32_t *)(base + L.occ);
    c->planes = (uint32_t *)(base + L.planes);
    c->meta = (uint32_t *)(base + L.meta);
}

FillWork fill_work_of(const vx_chunk *c, int terrain) {
    FillWork w{};
    w.cx = c->coord.x;
    w.cy = c->coord.y;
    w.cz = c->coord.z;
    w.edge = c->edge;
    w.terrain = terrain;
    w.ids = c->ids;
    w.occ = c->occ;
    w.planes = c->planes;
    w.meta = c->meta;
    return w;
}

// Plane of `nb` that faces a chunk on its side k (-x, +x, -z, +z): the
// neighbour's x=N-1, x=0, z=N-1, z=0 plane (src/chunk.rs:89-92,122-123,
// 153-156,186-187).
const uint32_t *facing_plane(const vx_chunk *nb, int k) {
    static const int which[4] = {1, 0, 3, 2};
    const size_t stride = (size_t)nb->edge * words_per_row(nb->edge);
    return nb->planes + which[k] * stride;
}

vx_chunk *new_chunk(vx_ctx *ctx, vx_coord coord, int edge) {
    vx_chunk *c = new (std::nothrow) vx_chunk();
    if (!c) return nullptr;
    c->refs.store(1);
    c->ctx = ctx;
    c->coord = coord;
    c->edge = edge;
    return c;
}

}  // namespace

extern "C" int vx_chunk_new(vx_ctx *ctx, int32_t terrain_id, vx_coord coord, int32_t edge,
                            vx_chunk **out) {
    if (!ctx || !out || !valid_edge(edge) || !coord_in_range(coord, edge)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (terrain_id < 0 || terrain_id >= (int)ctx->terrains.size()) return VX_ERR_INVALID;
    vx_chunk *c = new_chunk(ctx, coord, edge);
    if (!c) return VX_ERR_NOMEM;
    int st = ctx->dev_alloc(chunk_layout(edge).total, &c->block);
    if (st) {
        delete c;
        return st;
    }
    bind_chunk_ptrs(c, (char *)c->block);
    c->solid_type = kBlockGrass;
    hipError_t e = launch_fill_one(fill_work_of(c, terrain_id), ctx->terrains_dev, ctx->lut_dev,
                                   ctx->stream);
    if (e != hipSuccess) {
        ctx->set_error(e, "fill kernel");
        ctx->dev_free(c->block);
        delete c;
        return VX_ERR_DEVICE;
    }
    *out = c;
    return VX_OK;
}

extern "C" int vx_chunk_from_ids(vx_ctx *ctx, vx_coord coord, int32_t edge, const uint8_t *ids,
                                 vx_chunk **out) {
    if (!ctx || !out || !ids || !valid_edge(edge) || !coord_in_range(coord, edge))
        return VX_ERR_INVALID;
    const size_t nv = (size_t)edge * edge * edge;
    for (size_t i = 0; i < nv; i++)
        if (ids[i] > kBlockDirt) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    vx_chunk *c = new_chunk(ctx, coord, edge);
    if (!c) return VX_ERR_NOMEM;
    int st = ctx->dev_alloc(chunk_layout(edge).total, &c->block);
    if (st) {
        delete c;
        return st;
    }
    bind_chunk_ptrs(c, (char *)c->block);
    hipError_t e = hipMemcpyAsync(c->ids, ids, nv, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // caller's buffer is pageable
    if (e == hipSuccess) e = launch_pack_one(fill_work_of(c, 0), ctx->stream);
    uint32_t meta0 = kBlockMixed;
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->pinned, c->meta, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->set_error(e, "vx_chunk_from_ids");
        ctx->dev_free(c->block);
        delete c;
        return VX_ERR_DEVICE;
    }
    memcpy(&meta0, ctx->pinned, 4);
    c->solid_type = meta0;
    *out = c;
    return VX_OK;
}

extern "C" int vx_chunk_retain(vx_chunk *c) {
    if (!c) return VX_ERR_INVALID;
    c->refs.fetch_add(1);
    return VX_OK;
}

void vx_shared_unref_locked(vx_ctx *ctx, vx::SharedBlock *s) {
    if (s && s->refs.fetch_sub(1) == 1) {
        ctx->dev_free(s->ptr);
        delete s;
    }
}

void vx_chunk_destroy_locked(vx_chunk *c) {
    vx_ctx *ctx = c->ctx;
    if (c->block) ctx->dev_free(c->block);
    vx_shared_unref_locked(ctx, c->shared);
    delete c;
}

extern "C" int vx_chunk_release(vx_chunk *c) {
    if (!c) return VX_ERR_INVALID;
    if (c->refs.fetch_sub(1) == 1) {
        std::lock_guard<std::recursive_mutex> g(c->ctx->mu);
        vx_chunk_destroy_locked(c);
    }
    return VX_OK;
}

extern "C" int vx_chunk_info(const vx_chunk *c, vx_coord *coord, int32_t *edge) {
    if (!c) return VX_ERR_INVALID;
    if (coord) *coord = c->coord;
    if (edge) *edge = c->edge;
    return VX_OK;
}

extern "C" int vx_chunk_read_ids(vx_ctx *ctx, const vx_chunk *c, uint8_t *dst, uint64_t cap) {
    if (!ctx || !c || !dst) return VX_ERR_INVALID;
    const uint64_t n = (uint64_t)c->edge * c->edge * c->edge;
    if (cap < n) return VX_ERR_CAPACITY;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipMemcpyAsync(dst, c->ids, n, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_chunk_read_occupancy(vx_ctx *ctx, const vx_chunk *c, uint32_t *dst,
                                       uint64_t cap_words) {
    if (!ctx || !c || !dst) return VX_ERR_INVALID;
    const uint64_t n = (uint64_t)c->edge * c->edge * words_per_row(c->edge);
    if (cap_words < n) return VX_ERR_CAPACITY;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipMemcpyAsync(dst, c->occ, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_chunk_noise(vx_ctx *ctx, int32_t terrain_id, vx_coord coord, int32_t edge,
                              float *dst, uint64_t cap) {
    if (!ctx || !dst || !valid_edge(edge) || !coord_in_range(coord, edge)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (terrain_id < 0 || terrain_id >= (int)ctx->terrains.size()) return VX_ERR_INVALID;
    const int kind = ctx->terrains[terrain_id].kind;
    if (kind != kDensity && kind != kHeightmap && kind != kReference) return VX_ERR_INVALID;
    const uint64_t n = kind == kHeightmap ? (uint64_t)edge * edge : (uint64_t)edge * edge * edge;
    if (cap < n) return VX_ERR_CAPACITY;
    vx_chunk tmp;
    tmp.ctx = ctx;
    tmp.coord = coord;
    tmp.edge = edge;
    const size_t cbytes = chunk_layout(edge).total;
    void *block = nullptr;
    int st = ctx->dev_alloc(cbytes + n * 4, &block);
    if (st) return st;
    bind_chunk_ptrs(&tmp, (char *)block);
    FillWork w = fill_work_of(&tmp, terrain_id);
    w.noise_out = (float *)((char *)block + cbytes);
    hipError_t e = launch_fill_one(w, ctx->terrains_dev, ctx->lut_dev, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst, w.noise_out, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->dev_free(block);
    tmp.block = nullptr;
    if (e != hipSuccess) {
        ctx->set_error(e, "vx_chunk_noise");
        return VX_ERR_DEVICE;
    }
    return VX_OK;
}

// ------------------------------------------------------------ meshes
namespace {

MeshWork mesh_work_of(const vx_chunk *c, vx_chunk *const adj[4]) {
    MeshWork w{};
    w.ox = c->coord.x * c->edge;
    w.oy = c->coord.y * c->edge;
    w.oz = c->coord.z * c->edge;
    w.edge = c->edge;
    w.occ = c->occ;
    w.ids = c->ids;
    for (int k = 0; k < 4; k++) w.nb[k] = (adj && adj[k]) ? facing_plane(adj[k], k) : nullptr;
    w.solid_type = c->solid_type;
    return w;
}

size_t mesh_bytes(uint64_t faces, uint32_t formats, size_t *off_v, size_t *off_i) {
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t b = 0;
    *off_v = *off_i = 0;
    if (formats & kFmtInstances) b = up(faces * 68);
    if (formats & kFmtVertices) {
        *off_v = b;
        b += up(faces * 80);
        *off_i = b;
        b += up(faces * 24);
    }
    return b ? b : 256;
}

vx_mesh *new_mesh(vx_ctx *ctx, uint64_t faces, uint32_t formats) {
    vx_mesh *m = new (std::nothrow) vx_mesh();
    if (!m) return nullptr;
    m->refs.store(1);
    m->ctx = ctx;
    m->faces = faces;
    m->formats = formats;
    return m;
}

}  // namespace

extern "C" int vx_chunk_create_mesh(vx_ctx *ctx, vx_chunk *chunk, vx_chunk *const adj[4],
                                    uint32_t formats, vx_mesh **out) {
    if (!ctx || !chunk || !out) return VX_ERR_INVALID;
    if (formats == 0 || (formats & ~(kFmtInstances | kFmtVertices))) return VX_ERR_INVALID;
    if (adj)
        for (int k = 0; k < 4; k++)
            if (adj[k] && adj[k]->edge != chunk->edge) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    const MeshWork w = mesh_work_of(chunk, adj);
    uint32_t *count_dev = (uint32_t *)ctx->small_dev;
    VX_HIP(ctx, launch_count_one(w, count_dev, ctx->stream));
    VX_HIP(ctx, hipMemcpyAsync(ctx->pinned, count_dev, 4, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t faces;
    memcpy(&faces, ctx->pinned, 4);
    vx_mesh *m = new_mesh(ctx, faces, formats);
    if (!m) return VX_ERR_NOMEM;
    size_t off_v, off_i;
    const size_t bytes = mesh_bytes(faces, formats, &off_v, &off_i);
    int st = ctx->dev_alloc(bytes, &m->block);
    if (st) {
        delete m;
        return st;
    }
    char *b = (char *)m->block;
    if (formats & kFmtInstances) m->instances = (uint32_t *)b;
    if (formats & kFmtVertices) {
        m->vertices = (float *)(b + off_v);
        m->indices = (uint32_t *)(b + off_i);
    }
    if (faces > 0) {
        MeshOut o{};
        o.vertices = m->vertices;
        o.indices = m->indices;
        o.instances = m->instances;
        o.capacity = faces;
        o.formats = formats;
        hipError_t e = launch_emit_one(w, 0, faces, o, ctx->tables_dev, count_dev + 1, ctx->stream);
        if (e != hipSuccess) {
            ctx->set_error(e, "emit kernel");
            ctx->dev_free(m->block);
            delete m;
            return VX_ERR_DEVICE;
        }
    }
    *out = m;
    return VX_OK;
}

extern "C" int vx_mesh_retain(vx_mesh *m) {
    if (!m) return VX_ERR_INVALID;
    m->refs.fetch_add(1);
    return VX_OK;
}

extern "C" int vx_mesh_release(vx_mesh *m) {
    if (!m) return VX_ERR_INVALID;
    if (m->refs.fetch_sub(1) == 1) {
        std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
        if (m->block) m->ctx->dev_free(m->block);
        vx_shared_unref_locked(m->ctx, m->shared);
        delete m;
    }
    return VX_OK;
}

extern "C" int vx_mesh_face_count(const vx_mesh *m, uint64_t *faces) {
    if (!m || !faces) return VX_ERR_INVALID;
    *faces = m->faces;
    return VX_OK;
}

static int mesh_read(vx_ctx *ctx, const vx_mesh *m, const void *src, uint64_t bytes, void *dst,
                     uint64_t cap) {
    if (!ctx || !m || !dst) return VX_ERR_INVALID;
    if (!src && bytes) return VX_ERR_INVALID;
    if (cap < bytes) return VX_ERR_CAPACITY;
    if (!bytes) return VX_OK;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_mesh_read_instances(vx_ctx *ctx, const vx_mesh *m, void *dst, uint64_t cap) {
    if (!m || !(m->formats & kFmtInstances)) return VX_ERR_INVALID;
    return mesh_read(ctx, m, m->instances, m->faces * 68, dst, cap);
}

extern "C" int vx_mesh_read_vertices(vx_ctx *ctx, const vx_mesh *m, void *dst, uint64_t cap) {
    if (!m || !(m->formats & kFmtVertices)) return VX_ERR_INVALID;
    return mesh_read(ctx, m, m->vertices, m->faces * 80, dst, cap);
}

extern "C" int vx_mesh_read_indices(vx_ctx *ctx, const vx_mesh *m, void *dst, uint64_t cap) {
    if (!m || !(m->formats & kFmtVertices)) return VX_ERR_INVALID;
    return mesh_read(ctx, m, m->indices, m->faces * 24, dst, cap);
}

extern "C" int vx_mesh_device_buffers(const vx_mesh *m, void **inst, void **vert, void **idx) {
    if (!m) return VX_ERR_INVALID;
    if (inst) *inst = m->instances;
    if (vert) *vert = m->vertices;
    if (idx) *idx = m->indices;
    return VX_OK;
}

// ------------------------------------------------------------ frustum
// matrix_to_planes (src/frustum_culling.rs:47-85): data[c][r] column-major,
// planes near=r3+r2, far=r3-r2, left=r3+r0, right=r3-r0, top=r3-r1,
// bottom=r3+r1, unnormalised, in the order [near far left right top bottom].
extern "C" int vx_frustum_from_matrix(const float m[16], vx_frustum *out) {
    if (!m || !out) return VX_ERR_INVALID;
    for (int c = 0; c < 4; c++) {
        const float r0 = m[c * 4 + 0], r1 = m[c * 4 + 1], r2 = m[c * 4 + 2], r3 = m[c * 4 + 3];
        out->planes[0 * 4 + c] = r3 + r2;
        out->planes[1 * 4 + c] = r3 - r2;
        out->planes[2 * 4 + c] = r3 + r0;
        out->planes[3 * 4 + c] = r3 - r0;
        out->planes[4 * 4 + c] = r3 - r1;
        out->planes[5 * 4 + c] = r3 + r1;
    }
    return VX_OK;
}

// contains (src/frustum_culling.rs:16-40) for one chunk, on the host.
extern "C" int vx_frustum_contains(const vx_frustum *f, vx_coord c, int32_t edge) {
    if (!f || edge <= 0) return VX_ERR_INVALID;
    // corner positions exact in 64 bits, then rounded to f32 (equal to the
    // reference's i32 product wherever that does not overflow)
    const float px = (float)((int64_t)c.x * edge), py = (float)((int64_t)c.y * edge),
                pz = (float)((int64_t)c.z * edge);
    const float e = (float)edge;
    for (int k = 0; k < 8; k++) {
        const float x = (k & 1) ? px + e : px;
        const float z = (k & 2) ? pz + e : pz;
        const float y = (k & 4) ? py + e : py;
        bool inside = true;
        for (int i = 0; i < 6 && inside; i++) {
            const float *p = f->planes + 4 * i;
            const float dist = p[0] * x + p[1] * y + p[2] * z + p[3];
            if (dist < 0.0f) inside = false;
        }
        if (inside) return 1;
    }
    return 0;
}

extern "C" int vx_frustum_mask_device(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                                      int32_t n, int32_t edge, uint8_t *out_dev) {
    if (!ctx || !f || n < 0 || edge <= 0 || (n && (!coords_dev || !out_dev))) return VX_ERR_INVALID;
    if (!n) return VX_OK;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    Planes P;
    memcpy(P.p, f->planes, sizeof(P.p));
    VX_HIP(ctx, launch_frustum(P, coords_dev, n, edge, out_dev, ctx->stream));
    return VX_OK;
}

int vx_draw_list_locked(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                        const uint32_t *faces_dev, const uint64_t *first_dev, int32_t n,
                        int32_t edge, uint32_t mode, vx_draw_indexed *args_dev,
                        uint32_t *count_dev) {
    const int32_t nb = draw_blocks(n);
    if (nb > ctx->draw_counts_cap) {
        if (ctx->draw_counts) {
            VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // earlier lists may still read it
            ctx->dev_free(ctx->draw_counts);
            ctx->draw_counts = nullptr;
            ctx->draw_counts_cap = 0;
        }
        const int32_t cap = std::max(nb, 256);
        int st = ctx->dev_alloc((size_t)cap * 4, (void **)&ctx->draw_counts);
        if (st) return st;
        ctx->draw_counts_cap = cap;
    }
    Planes P;
    memcpy(P.p, f->planes, sizeof(P.p));
    VX_HIP(ctx, launch_draw_list(P, coords_dev, faces_dev, first_dev, n, edge, mode,
                                 ctx->draw_counts, reinterpret_cast<DrawArgs *>(args_dev),
                                 count_dev, ctx->stream));
    return VX_OK;
}

extern "C" int vx_frustum_draw_list(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                                    const uint32_t *faces_dev, const uint64_t *first_dev,
                                    int32_t n, int32_t edge, uint32_t mode,
                                    vx_draw_indexed *args_dev, uint32_t *count_dev) {
    if (!ctx || !f || n < 0 || edge <= 0 || !count_dev ||
        (mode != VX_DRAW_INSTANCED && mode != VX_DRAW_EXPANDED) ||
        (n && (!coords_dev || !faces_dev || !first_dev || !args_dev)))
        return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!n) {
        VX_HIP(ctx, hipMemsetAsync(count_dev, 0, 8, ctx->stream));
        return VX_OK;
    }
    return vx_draw_list_locked(ctx, f, coords_dev, faces_dev, first_dev, n, edge, mode, args_dev,
                               count_dev);
}

// The allocation-free variant: the per-workgroup counts live in the static
// scratch array of this context's GPU; nothing is allocated or freed and no
// memset is queued, so a caller may capture the call into a HIP graph.
extern "C" int vx_frustum_draw_list_noalloc(vx_ctx *ctx, const vx_frustum *f,
                                            const int32_t *coords_dev, const uint32_t *faces_dev,
                                            const uint64_t *first_dev, int32_t n, int32_t edge,
                                            uint32_t mode, vx_draw_indexed *args_dev,
                                            uint32_t *count_dev) {
    if (!ctx || !f || n < 0 || edge <= 0 || !count_dev ||
        (mode != VX_DRAW_INSTANCED && mode != VX_DRAW_EXPANDED) ||
        (n && (!coords_dev || !faces_dev || !first_dev || !args_dev)))
        return VX_ERR_INVALID;
    if (n > draw_scratch_max_meshes()) {
        ctx->last_error = "vx_frustum_draw_list_noalloc: more meshes than its scratch serves";
        return VX_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!ctx->draw_scratch) return VX_ERR_DEVICE;
    if (!n) {
        VX_HIP(ctx, launch_clear_words(count_dev, 2, ctx->stream));
        return VX_OK;
    }
    Planes P;
    memcpy(P.p, f->planes, sizeof(P.p));
    VX_HIP(ctx, launch_draw_list(P, coords_dev, faces_dev, first_dev, n, edge, mode,
                                 ctx->draw_scratch, reinterpret_cast<DrawArgs *>(args_dev),
                                 count_dev, ctx->stream));
    return VX_OK;
}

extern "C" int vx_frustum_mask(vx_ctx *ctx, const vx_frustum *f, const vx_coord *coords, int32_t n,
                               int32_t edge, uint8_t *out) {
    if (!ctx || !f || n < 0 || edge <= 0 || (n && (!coords || !out))) return VX_ERR_INVALID;
    if (!n) return VX_OK;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    const size_t cb = (size_t)n * 12, cbr = (cb + 255) & ~(size_t)255, ob = (size_t)n;
    int st = ctx->ensure_scratch(cbr + ob);
    if (st) return st;
    int32_t *cd = (int32_t *)ctx->scratch;
    uint8_t *od = (uint8_t *)ctx->scratch + cbr;
    Planes P;
    memcpy(P.p, f->planes, sizeof(P.p));
    VX_HIP(ctx, hipMemcpyAsync(cd, coords, cb, hipMemcpyHostToDevice, ctx->stream));
    VX_HIP(ctx, launch_frustum(P, cd, n, edge, od, ctx->stream));
    VX_HIP(ctx, hipMemcpyAsync(out, od, ob, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}
