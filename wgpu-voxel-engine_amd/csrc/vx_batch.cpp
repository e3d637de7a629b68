// vx_batch.cpp -- batched chunk generation: n requests of Chunk::new +
// Chunk::create_mesh (src/chunk.rs:21-57) run as four kernels over the whole
// batch (fill -> count -> scan -> emit), captured once into a hipGraph and
// replayed.  This replaces the reference's one-job-per-chunk thread pool
// (src/chunk_builder.rs:80-84) and its per-mesh wgpu upload
// (src/chunk.rs:254-266): ids, occupancy and meshes stay in HBM.
//
// Layout in HBM
//   store   per chunk i at chunk_off[i]: ids (N^3 B) | occupancy rows
//           (N^2 * W u32) | 4 boundary planes (4 * N * W u32) | meta (solid block id, u32),
//           each 256-byte aligned.
//   ctl     counts[n] u32 | bases[n] u64 | total u64 | overflow u32 |
//           emit_order[n_items] i32 | slabs[4 n] u32 (batches holding 64^3 chunks,
//           whose emit works on y-slabs: n_items = n + 3 * (number of 64^3 chunks))
//   arena   instances (68 B/face) | vertices (80 B/face) | indices (24 B/face)
//           for `capacity` faces; chunk i's faces are [bases[i], +counts[i]).
#include "vx_host.h"

#include <algorithm>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <numeric>
#include <tuple>

using namespace vx;

#define VX_HIP(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->set_error(e_, #call);                                                    \
            return VX_ERR_DEVICE;                                                           \
        }                                                                                   \
    } while (0)

namespace {

size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
    size_t occ, planes, meta, total;
};

Layout layout_of(int N) {
    const size_t W = (size_t)words_per_row(N);
    Layout L;
    L.occ = up256((size_t)N * N * N);
    L.planes = L.occ + up256((size_t)N * N * W * 4);
    L.meta = L.planes + up256(4 * (size_t)N * W * 4);
    L.total = L.meta + 256;
    return L;
}

// Work estimate used to schedule large chunks first in ragged batches.
double cost_of(const vx_ctx *ctx, const vx_batch_req &r) {
    const double n3 = (double)r.edge * r.edge * r.edge;
    const TerrainDev &T = ctx->terrains[r.terrain_id];
    if (T.kind == kDensity || T.kind == kReference) return n3 * T.octaves;
    if (T.kind == kHeightmap) return (double)r.edge * r.edge * T.octaves * 8.0 + n3 / 8.0;
    return n3 / 16.0;
}

size_t arena_bytes(uint64_t cap, uint32_t formats, size_t *off_v, size_t *off_i) {
    size_t b = 0;
    *off_v = *off_i = 0;
    if (formats & kFmtInstances) b = up256(cap * 68);
    if (formats & kFmtVertices) {
        *off_v = b;
        b += up256(cap * 80);
        *off_i = b;
        b += up256(cap * 24);
    }
    return b ? b : 256;
}

void drop_graph(vx_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

// (Re)allocates the mesh arena for `cap` faces.  The old arena is written by
// queued runs and its pointers are baked into the captured graph, so the
// stream is drained and the graph dropped before it is freed.  On failure
// the batch holds no arena (capacity 0: a later run writes no faces and
// reports overflow) and no graph.
int alloc_arena(vx_batch *b, uint64_t cap) {
    vx_ctx *ctx = b->ctx;
    if (b->arena) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        drop_graph(b);
        if (e != hipSuccess) {
            ctx->set_error(e, "arena reallocation");
            return VX_ERR_DEVICE;
        }
        ctx->dev_free(b->arena);
    }
    b->arena = nullptr;
    size_t off_v, off_i;
    const size_t bytes = arena_bytes(cap, b->formats, &off_v, &off_i);
    int st = ctx->dev_alloc(bytes, &b->arena);
    if (st) {
        b->capacity = 0;
        b->out = MeshOut{};
        return st;
    }
    char *a = (char *)b->arena;
    b->capacity = cap;
    b->out = MeshOut{};
    b->out.capacity = cap;
    b->out.formats = b->formats;
    if (b->formats & kFmtInstances) b->out.instances = (uint32_t *)a;
    if (b->formats & kFmtVertices) {
        b->out.vertices = (float *)(a + off_v);
        b->out.indices = (uint32_t *)(a + off_i);
    }
    return VX_OK;
}

// VX_CHECK_SCHEDULES=1 (a debugging aid): every vx_batch_wait also reads
// the schedules back and checks that each is a permutation of the requests.
bool check_schedules() {
    static const bool on = [] {
        const char *v = getenv("VX_CHECK_SCHEDULES");
        return v && v[0] == '1';
    }();
    return on;
}

int form_of(const vx_batch *b);

bool schedules_ok(vx_batch *b) {
    std::vector<int32_t> h((size_t)std::max(b->n, b->n_items));
    auto fetch = [&](const int32_t *a, int32_t m) {
        if (hipMemcpy(h.data(), a, (size_t)m * 4, hipMemcpyDeviceToHost) == hipSuccess) return true;
        (void)hipGetLastError();
        return false;
    };
    if (b->order_dev && !(fetch(b->order_dev, b->n) && is_schedule(h.data(), b->n))) return false;
    if (!fetch(b->emit_order, b->n_items)) return false;
    if (form_of(b) == VX_BATCH_SPLIT)  // each half's order, indices relative to the half
        return is_schedule(h.data(), b->split_m) &&
               is_schedule(h.data() + b->split_m, b->n - b->split_m);
    if (!b->slabs) return is_schedule(h.data(), b->n);
    std::vector<int32_t> per(b->n);
    for (int i = 0; i < b->n; i++) per[i] = b->reqs[i].edge == 64 ? kSlabs : 1;
    return is_item_schedule(h.data(), b->n_items, per.data(), b->n);
}

// The staged form's workgroups per chunk (fill) and per item (emit): the
// autotuner's choice under VX_BATCH_AUTO, else vx_batch_set_splits' (0 =
// the launchers' own rule).
bool tuned_auto(const vx_batch *b) { return b->mode == VX_BATCH_AUTO && b->tuned; }
int32_t fill_split_of(const vx_batch *b) {
    return tuned_auto(b) ? b->tuned_fill_split : b->fill_split;
}
int32_t emit_split_of(const vx_batch *b) {
    return tuned_auto(b) ? b->tuned_emit_split : b->emit_split;
}

hipError_t enqueue_fill(vx_batch *b) {
    vx_ctx *ctx = b->ctx;
    return launch_fill(b->fill_dev, b->order_dev, b->n, b->max_edge, ctx->terrains_dev,
                       ctx->lut_dev, b->overflow + 1, fill_split_of(b), ctx->stream);
}
hipError_t enqueue_count(vx_batch *b) {
    return launch_count(b->mesh_dev, b->n, b->max_edge, b->counts, b->slabs, b->ctx->stream);
}
hipError_t enqueue_scan(vx_batch *b) {
    return launch_scan(b->counts, b->n, b->slabs ? b->mesh_dev : nullptr, b->bases, b->total,
                       b->overflow, b->emit_order, b->ctx->stream);
}
hipError_t enqueue_emit(vx_batch *b) {
    return launch_emit(b->mesh_dev, b->emit_order, b->n, b->n_items, b->max_edge, b->bases,
                       b->counts, b->slabs, b->out, b->ctx->tables_dev, b->overflow,
                       emit_split_of(b), b->ctx->stream);
}

// The form a run takes.  VX_BATCH_AUTO runs the staged form unless the
// autotuner chose a persistent form (fused or pipelined) for this batch
// shape: the plain fused kernel measured as slow or slower than the staged
// graph at every size, from 64 chunks (0.20 vs 0.21 ms) through a 512-chunk
// frame with its cull job (0.23 vs 0.24 ms) to 4096 chunks (0.76 vs 1.1 ms).
int form_of(const vx_batch *b) {
    int m = b->mode;
    if (m == VX_BATCH_AUTO) {
        if (!b->tuned) return VX_BATCH_STAGED;
        m = b->tuned_mode;
    }
    if (m == VX_BATCH_SPLIT) return b->split_ok ? VX_BATCH_SPLIT : VX_BATCH_STAGED;
    if (!b->fused_ok) return VX_BATCH_STAGED;
    if (m == VX_BATCH_PIPELINED && !b->pipe_ok) return VX_BATCH_STAGED;
    return m == VX_BATCH_FUSED || m == VX_BATCH_PIPELINED ? m : VX_BATCH_STAGED;
}

// A persistent form (one launch per run, the fused or the pipelined kernel).
bool use_fused(const vx_batch *b) {
    const int f = form_of(b);
    return f == VX_BATCH_FUSED || f == VX_BATCH_PIPELINED;
}

int32_t fused_grid_of(const vx_batch *b) {
    return b->mode == VX_BATCH_AUTO && b->tuned ? b->tuned_grid : b->grid;
}

int32_t pipe_ratio_of(const vx_batch *b) {
    return b->mode == VX_BATCH_AUTO && b->tuned ? b->tuned_ratio : b->pipe_ratio;
}

hipError_t enqueue_fused(vx_batch *b) {
    vx_ctx *ctx = b->ctx;
    if (form_of(b) == VX_BATCH_PIPELINED)
        return launch_pipe(b->fused_dev, b->n, b->max_edge, b->pipe_lag, ctx->terrains_dev,
                           ctx->lut_dev,
                           ctx->tables_dev, b->fctl_ptrs, b->out, b->cull, fused_grid_of(b),
                           pipe_ratio_of(b), ctx->stream);
    return launch_fused(b->fused_dev, b->n, b->max_edge, ctx->terrains_dev, ctx->lut_dev,
                        ctx->tables_dev, b->fctl_ptrs, b->out, b->cull, fused_grid_of(b),
                        ctx->stream);
}

// The split form: the staged kernels over two halves on two streams.  The
// context stream fills chunks [0, m2) (the first half [0, m) and the later
// chunks its counts read planes of) and then [m2, n) back to back, so the
// fills stay on one queue; the second stream counts, scans and emits the
// first half as soon as its fill is out, beside the second fill.  The
// context stream then counts [m, n), scans it from the first half's total
// and emits it, and waits for the second stream before the run ends.  Each
// half's emission order holds indices relative to the half.
hipError_t enqueue_split(vx_batch *b) {
    vx_ctx *ctx = b->ctx;
    hipError_t e = hipSuccess;
    if (!ctx->stream2) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    hipStream_t s0 = ctx->stream, s1 = ctx->stream2;
    const int32_t n = b->n, m = b->split_m, m2 = b->split_m2;
    const int32_t fs = fill_split_of(b), es = emit_split_of(b);
    e = launch_cull(b->cull, s0);
    if (e == hipSuccess)
        e = launch_fill(b->fill_dev, nullptr, m2, b->max_edge, ctx->terrains_dev, ctx->lut_dev,
                        b->overflow + 1, fs, s0);
    if (e == hipSuccess) e = hipEventRecord(b->split_ev[0], s0);
    // first half on the second stream
    if (e == hipSuccess) e = hipStreamWaitEvent(s1, b->split_ev[0], 0);
    if (e == hipSuccess && !b->fill_counts)
        e = launch_count(b->mesh_dev, m, b->max_edge, b->counts, nullptr, s1);
    if (e == hipSuccess)
        e = launch_scan(b->counts, m, nullptr, b->bases, b->split_total, b->overflow,
                        b->emit_order, s1);
    if (e == hipSuccess) e = hipEventRecord(b->split_ev[1], s1);
    if (e == hipSuccess)
        e = launch_emit(b->mesh_dev, b->emit_order, m, m, b->max_edge, b->bases, b->counts,
                        nullptr, b->out, ctx->tables_dev, b->overflow, es, s1);
    if (e == hipSuccess) e = hipEventRecord(b->split_ev[2], s1);
    // second half on the context stream
    if (e == hipSuccess)
        e = launch_fill(b->fill_dev + m2, nullptr, n - m2, b->max_edge, ctx->terrains_dev,
                        ctx->lut_dev, b->overflow + 1, fs, s0);
    if (e == hipSuccess && !b->fill_counts)
        e = launch_count(b->mesh_dev + m, n - m, b->max_edge, b->counts + m, nullptr, s0);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->split_ev[1], 0);
    // the overflow word was cleared by the first half's scan: not again here
    if (e == hipSuccess)
        e = launch_scan(b->counts + m, n - m, nullptr, b->bases + m, b->total, nullptr,
                        b->emit_order + m, s0, b->split_total);
    if (e == hipSuccess)
        e = launch_emit(b->mesh_dev + m, b->emit_order + m, n - m, n - m, b->max_edge,
                        b->bases + m, b->counts + m, nullptr, b->out, ctx->tables_dev,
                        b->overflow, es, s0);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->split_ev[2], 0);
    return e;
}

hipError_t enqueue_all(vx_batch *b) {
    if (use_fused(b)) return enqueue_fused(b);
    if (form_of(b) == VX_BATCH_SPLIT) return enqueue_split(b);
    hipError_t e = launch_cull(b->cull, b->ctx->stream);
    if (e == hipSuccess) e = enqueue_fill(b);
    if (e == hipSuccess && !b->fill_counts) e = enqueue_count(b);
    if (e == hipSuccess) e = enqueue_scan(b);
    if (e == hipSuccess) e = enqueue_emit(b);
    return e;
}

int capture_graph(vx_batch *b) {
    vx_ctx *ctx = b->ctx;
    drop_graph(b);
    VX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_all(b);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        ctx->set_error(e != hipSuccess ? e : e2, "batch graph capture");
        return VX_ERR_DEVICE;
    }
    b->graph = g;
    VX_HIP(ctx, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    return VX_OK;
}

void free_batch(vx_batch *b) {
    vx_ctx *ctx = b->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    drop_graph(b);
    vx_shared_unref_locked(ctx, b->store_blk);
    for (vx_chunk *c : b->ext)
        if (c->refs.fetch_sub(1) == 1) vx_chunk_destroy_locked(c);
    ctx->dev_free(b->planes_dev);
    if (b->planes_pin) (void)hipHostFree(b->planes_pin);
    if (b->planes_ev) (void)hipEventDestroy(b->planes_ev);
    ctx->dev_free(b->fill_dev);
    ctx->dev_free(b->mesh_dev);
    ctx->dev_free(b->order_dev);
    ctx->dev_free(b->coords_dev);
    ctx->dev_free(b->ctl);
    ctx->dev_free(b->arena);
    ctx->dev_free(b->fused_dev);
    ctx->dev_free(b->fctl);
    for (hipEvent_t &ev : b->split_ev)
        if (ev) (void)hipEventDestroy(ev);
    delete b;
}

int fetch_results(vx_batch *b) {
    vx_ctx *ctx = b->ctx;
    if (b->results) return VX_OK;
    b->h_counts.resize(b->n);
    b->h_bases.resize(b->n);
    VX_HIP(ctx, hipMemcpyAsync(b->h_counts.data(), b->counts, (size_t)b->n * 4,
                               hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipMemcpyAsync(b->h_bases.data(), b->bases, (size_t)b->n * 8,
                               hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->results = true;
    return VX_OK;
}

// ---- autotuning -------------------------------------------------------
// Shape of a batch for the tuning cache: chunk count bucket (power of two),
// largest edge, uniform or ragged, fused form eligible, has neighbours.
using TuneKey = std::tuple<int, int, int, int, int>;
bool fill_split_ok(int32_t v) { return v == 0 || v == 1 || v == 2 || v == 4 || v == 8; }
bool emit_split_ok(int32_t v) { return v >= 0 && v <= 4; }

struct TuneChoice {
    int32_t mode, grid, ratio;  // ratio: pipelined roles (0 = default)
    int32_t fill_split = 0, emit_split = 0;  // staged workgroups per chunk / item (0 = rule)
};
std::map<TuneKey, TuneChoice> g_tune;
std::mutex g_tune_mu;
bool g_tune_loaded = false;

TuneKey tune_key(const vx_batch *b) {
    int bucket = 0;
    while ((1 << bucket) < b->n) bucket++;
    bool nb = false;
    for (const vx_batch_req &r : b->reqs)
        for (int k = 0; k < 4; k++) nb |= r.nb[k] != -1;
    return TuneKey{bucket, b->max_edge, b->uniform ? 1 : 0, b->fused_ok ? 1 : 0, nb ? 1 : 0};
}

// The cache persists across processes when VX_TUNE_CACHE names a file:
// one line per shape, "bucket max_edge uniform fused_ok has_nb mode grid
// ratio fill_split emit_split".
void tune_load() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = getenv("VX_TUNE_CACHE");
    if (!path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    // line by line: lines of the earlier seven- and eight-field forms (no
    // role ratio, no splits) read with the defaults; a line that is none of
    // the forms, or holds a value no setter accepts, is skipped
    char line[256];
    while (fgets(line, sizeof line, f)) {
        int a, c, d, e, h, m, g, r = 0, fs = 0, es = 0;
        const int k = sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &a, &c, &d, &e, &h, &m, &g, &r,
                             &fs, &es);
        if (k != 7 && k != 8 && k != 10) continue;
        const bool mode_ok = m == VX_BATCH_STAGED || m == VX_BATCH_FUSED ||
                             m == VX_BATCH_PIPELINED || m == VX_BATCH_SPLIT;
        const bool ratio_ok = r == 0 || (r >= 2 && r <= kPipeRatioMax);
        if (mode_ok && ratio_ok && g >= 0 && fill_split_ok(fs) && emit_split_ok(es))
            g_tune[TuneKey{a, c, d, e, h}] = TuneChoice{m, g, r, fs, es};
    }
    fclose(f);
}

void tune_save() {
    const char *path = getenv("VX_TUNE_CACHE");
    if (!path) return;
    FILE *f = fopen(path, "w");
    if (!f) return;
    for (auto &kv : g_tune)
        fprintf(f, "%d %d %d %d %d %d %d %d %d %d\n", std::get<0>(kv.first),
                std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first),
                std::get<4>(kv.first), kv.second.mode, kv.second.grid, kv.second.ratio,
                kv.second.fill_split, kv.second.emit_split);
    fclose(f);
}

}  // namespace

void vx_batch_unref(vx_batch *b) {
    if (b->refs.fetch_sub(1) == 1) free_batch(b);
}

extern "C" int vx_batch_create(vx_ctx *ctx, const vx_batch_req *reqs, int32_t n, uint32_t formats,
                               uint64_t face_capacity, vx_batch **out) {
    return vx_batch_create_ext(ctx, reqs, n, nullptr, 0, formats, face_capacity, out);
}

extern "C" int vx_batch_create_ext(vx_ctx *ctx, const vx_batch_req *reqs, int32_t n,
                                   vx_chunk *const *ext, int32_t n_ext, uint32_t formats,
                                   uint64_t face_capacity, vx_batch **out) {
    if (!ctx || !out || n <= 0 || n > VX_BATCH_MAX_CHUNKS || !reqs || n_ext < 0 || (n_ext && !ext))
        return VX_ERR_INVALID;  // the limit keeps emit item codes (chunk << 2 | slab) in an int
    for (int e = 0; e < n_ext; e++)
        if (!ext[e] || ext[e]->ctx != ctx) return VX_ERR_INVALID;
    if (formats == 0 || (formats & ~(kFmtInstances | kFmtVertices))) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    for (int i = 0; i < n; i++) {
        const vx_batch_req &r = reqs[i];
        if (r.edge != 16 && r.edge != 32 && r.edge != 64) return VX_ERR_INVALID;
        if (!coord_in_range(r.coord, r.edge)) return VX_ERR_INVALID;
        if (r.terrain_id < 0 || r.terrain_id >= (int)ctx->terrains.size()) return VX_ERR_INVALID;
        for (int k = 0; k < 4; k++) {
            const int j = r.nb[k];
            const bool in_batch = j >= 0 && j < n && j != i;
            const bool external = j <= -2 && -2 - j < n_ext;
            if (!(in_batch || j == -1 || external)) return VX_ERR_INVALID;
            if (in_batch && reqs[j].edge != r.edge) return VX_ERR_INVALID;
            if (external && ext[-2 - j]->edge != r.edge) return VX_ERR_INVALID;
        }
    }
    vx_batch *b = new (std::nothrow) vx_batch();
    if (!b) return VX_ERR_NOMEM;
    b->ctx = ctx;
    b->n = n;
    b->formats = formats;
    b->reqs.assign(reqs, reqs + n);
    b->chunk_off.resize(n);
    size_t off = 0;
    int emin = 64;
    for (int i = 0; i < n; i++) {
        b->chunk_off[i] = off;
        off += layout_of(reqs[i].edge).total;
        b->max_edge = std::max(b->max_edge, reqs[i].edge);
        emin = std::min(emin, reqs[i].edge);
    }
    // schedule: identical requests keep launch order; otherwise longest first
    std::vector<double> cost(n);
    for (int i = 0; i < n; i++) cost[i] = cost_of(ctx, reqs[i]);
    b->uniform = emin == b->max_edge &&
                 std::all_of(cost.begin(), cost.end(), [&](double c) { return c == cost[0]; });
    int st = VX_OK;
    auto fail = [&](int s) {
        free_batch(b);
        return s;
    };
    if ((st = ctx->dev_alloc(off, &b->store))) return fail(st);
    b->store_blk = new vx::SharedBlock();
    b->store_blk->ptr = b->store;
    for (int e = 0; e < n_ext; e++) {
        ext[e]->refs.fetch_add(1);
        b->ext.push_back(ext[e]);
    }
    if ((st = ctx->dev_alloc((size_t)n * sizeof(FillWork), (void **)&b->fill_dev))) return fail(st);
    if ((st = ctx->dev_alloc((size_t)n * sizeof(MeshWork), (void **)&b->mesh_dev))) return fail(st);
    const bool slabbed = b->max_edge > 32;
    b->n_items = n;
    if (slabbed)
        for (int i = 0; i < n; i++) b->n_items += reqs[i].edge == 64 ? kSlabs - 1 : 0;
    const size_t ctl_bytes = up256((size_t)n * 4) + up256((size_t)n * 8) + 256 +
                             up256((size_t)b->n_items * 4) +
                             (slabbed ? up256((size_t)kSlabs * n * 4) : 0);
    if ((st = ctx->dev_alloc(ctl_bytes, &b->ctl))) return fail(st);
    b->counts = (uint32_t *)b->ctl;
    b->bases = (uint64_t *)((char *)b->ctl + up256((size_t)n * 4));
    b->total = (uint64_t *)((char *)b->bases + up256((size_t)n * 8));
    b->overflow = (uint32_t *)(b->total + 1);  // [0] arena overflow, [1] schedule error
    b->emit_order = (int32_t *)((char *)b->total + 256);
    if (slabbed) {
        b->slabs = (uint32_t *)((char *)b->emit_order + up256((size_t)b->n_items * 4));
        VX_HIP(ctx, hipMemsetAsync(b->slabs, 0, (size_t)kSlabs * n * 4, ctx->stream));
    }
    VX_HIP(ctx, hipMemsetAsync(b->overflow, 0, 8, ctx->stream));

    std::vector<FillWork> fw(n);
    std::vector<MeshWork> mw(n);
    char *store = (char *)b->store;
    for (int i = 0; i < n; i++) {
        const vx_batch_req &r = reqs[i];
        const Layout L = layout_of(r.edge);
        char *base = store + b->chunk_off[i];
        FillWork &f = fw[i];
        f = FillWork{};
        f.cx = r.coord.x;
        f.cy = r.coord.y;
        f.cz = r.coord.z;
        f.edge = r.edge;
        f.terrain = r.terrain_id;
        f.ids = (uint8_t *)base;
        f.occ = (uint32_t *)(base + L.occ);
        f.planes = (uint32_t *)(base + L.planes);
        f.meta = (uint32_t *)(base + L.meta);
    }
    for (int i = 0; i < n; i++) {
        const vx_batch_req &r = reqs[i];
        MeshWork &m = mw[i];
        m = MeshWork{};
        m.ox = r.coord.x * r.edge;
        m.oy = r.coord.y * r.edge;
        m.oz = r.coord.z * r.edge;
        m.edge = r.edge;
        m.occ = fw[i].occ;
        m.ids = fw[i].ids;
        m.solid_type = kBlockGrass;  // generated chunks hold only grass
        static const int which[4] = {1, 0, 3, 2};  // facing plane of the neighbour
        const size_t plane = (size_t)r.edge * words_per_row(r.edge);
        for (int k = 0; k < 4; k++) {
            const int j = r.nb[k];
            if (j >= 0) m.nb[k] = fw[j].planes + which[k] * plane;
            else if (j <= -2) m.nb[k] = ext[-2 - j]->planes + which[k] * plane;
            else m.nb[k] = nullptr;
        }
    }
    // fused form: every in-batch neighbour must be the adjacent chunk with the
    // The fill counts the faces itself when every chunk has its own
    // workgroup and no neighbour inside the batch (each neighbour plane is
    // absent or resident before the run): the staged form then has no
    // count kernel.
    b->fill_counts = fill_whole_chunks(n);
    for (int i = 0; i < n && b->fill_counts; i++)
        for (int k = 0; k < 4; k++)
            if (reqs[i].nb[k] >= 0) b->fill_counts = false;
    if (b->fill_counts)
        for (int i = 0; i < n; i++) {
            for (int k = 0; k < 4; k++) fw[i].cnb[k] = mw[i].nb[k];
            fw[i].count_out = b->counts + i;
            fw[i].slab_out = b->slabs ? b->slabs + (size_t)kSlabs * i : nullptr;
        }
    // split form (edges up to 32): the first half [0, m) and the later chunks
    // it borders, up to m2, are filled first; a batch whose fill counts needs
    // whole-chunk fills (kFillGrid chunks or more) in both fill launches
    if (b->max_edge <= 32 && n >= 2) {
        const int32_t m = n / 2;
        int32_t m2 = m;
        for (int i = 0; i < m; i++)
            for (int k = 0; k < 4; k++) m2 = std::max(m2, reqs[i].nb[k] + 1);
        b->split_m = m;
        b->split_m2 = m2;
        b->split_ok = !b->fill_counts || (m2 >= kFillGrid && n - m2 >= kFillGrid);
        b->split_total = (uint64_t *)((char *)b->total + 16);  // in the control block's spare words
        for (int k = 0; k < 3 && b->split_ok; k++)
            if (hipEventCreateWithFlags(&b->split_ev[k], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                b->split_ok = false;
            }
    }
    // same terrain, so its facing plane can be recomputed from noise; the
    // reference-rule terrain is generated by the fill kernels only
    static const int ddx[4] = {-1, 1, 0, 0}, ddz[4] = {0, 0, -1, 1};
    b->fused_ok = true;
    for (int i = 0; i < n; i++)
        if (ctx->terrains[reqs[i].terrain_id].kind == kReference) b->fused_ok = false;
    for (int i = 0; i < n && b->fused_ok; i++)
        for (int k = 0; k < 4; k++) {
            const int j = reqs[i].nb[k];
            if (j < 0) continue;
            const vx_batch_req &a = reqs[i], &c = reqs[j];
            if (c.coord.x != a.coord.x + ddx[k] || c.coord.y != a.coord.y ||
                c.coord.z != a.coord.z + ddz[k] || c.terrain_id != a.terrain_id) {
                b->fused_ok = false;
                break;
            }
        }
    // pipelined form: edges up to 32 (its LDS holds four workgroups per CU)
    b->pipe_ok = b->fused_ok && b->max_edge <= 32;
    std::vector<FusedWork> fu;
    if (b->fused_ok) {
        fu.resize(n);
        for (int i = 0; i < n; i++) {
            fu[i].f = fw[i];
            fu[i].m = mw[i];
            for (int k = 0; k < 4; k++) {
                fu[i].halo[k] = reqs[i].nb[k] >= 0    ? kHaloRecompute
                                : reqs[i].nb[k] <= -2 ? kHaloResident
                                                      : kHaloAbsent;
                fu[i].nbi[k] = reqs[i].nb[k] >= 0 ? reqs[i].nb[k] : -1;
            }
        }
        if ((st = ctx->dev_alloc((size_t)n * sizeof(FusedWork), (void **)&b->fused_dev)))
            return fail(st);
        // ticket | error | ... | status[n] (u64) | filled[n] (u32)
        if ((st = ctx->dev_alloc(256 + (size_t)n * 12, &b->fctl))) return fail(st);
        b->fctl_ptrs.ticket = (uint32_t *)b->fctl;
        b->fctl_ptrs.error = (uint32_t *)b->fctl + 1;
        b->fctl_ptrs.status = (uint64_t *)((char *)b->fctl + 256);
        b->fctl_ptrs.filled = (uint32_t *)(b->fctl_ptrs.status + n);
        b->fctl_ptrs.mesh_ticket = (uint32_t *)b->fctl + 2;
        // the farthest in-batch neighbour ahead of a chunk, in batch order
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 4; k++) b->pipe_lag = std::max(b->pipe_lag, reqs[i].nb[k] - i);
        b->fctl_ptrs.counts = b->counts;
        b->fctl_ptrs.bases = b->bases;
        b->fctl_ptrs.total = b->total;
        b->fctl_ptrs.overflow = b->overflow;
        VX_HIP(ctx, hipMemcpyAsync(b->fused_dev, fu.data(), fu.size() * sizeof(FusedWork),
                                   hipMemcpyHostToDevice, ctx->stream));
    }
    {
        // the scan rewrites it every run; start from a valid order regardless
        // (chunk indices, or every emit item's code chunk << 2 | slab)
        std::vector<int32_t> iota_order;
        iota_order.reserve(b->n_items);
        for (int i = 0; i < n; i++) {
            if (!b->slabs) iota_order.push_back(i);
            else
                for (int q = 0; q < (reqs[i].edge == 64 ? kSlabs : 1); q++)
                    iota_order.push_back(i << 2 | q);
        }
        VX_HIP(ctx, hipMemcpyAsync(b->emit_order, iota_order.data(), (size_t)b->n_items * 4,
                                   hipMemcpyHostToDevice, ctx->stream));
        VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // pageable source
    }
    VX_HIP(ctx, hipMemcpyAsync(b->fill_dev, fw.data(), fw.size() * sizeof(FillWork),
                               hipMemcpyHostToDevice, ctx->stream));
    VX_HIP(ctx, hipMemcpyAsync(b->mesh_dev, mw.data(), mw.size() * sizeof(MeshWork),
                               hipMemcpyHostToDevice, ctx->stream));
    if (!b->uniform) {
        std::vector<int32_t> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int c) { return cost[a] > cost[c]; });
        if (!is_schedule(order.data(), n)) {
            ctx->last_error = "batch fill schedule is not a permutation of the requests";
            return fail(VX_ERR_DEVICE);
        }
        if ((st = ctx->dev_alloc((size_t)n * 4, (void **)&b->order_dev))) return fail(st);
        VX_HIP(ctx, hipMemcpyAsync(b->order_dev, order.data(), (size_t)n * 4,
                                   hipMemcpyHostToDevice, ctx->stream));
    }
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // host vectors are pageable
    uint64_t cap = face_capacity;
    if (cap == 0) {
        for (int i = 0; i < n; i++) cap += 4ull * reqs[i].edge * reqs[i].edge;
    }
    if ((st = alloc_arena(b, cap))) return fail(st);
    {
        std::lock_guard<std::mutex> tg(g_tune_mu);
        tune_load();
        auto it = g_tune.find(tune_key(b));
        const int m = it != g_tune.end() ? it->second.mode : VX_BATCH_STAGED;
        if (it != g_tune.end() && (m != VX_BATCH_FUSED || b->fused_ok) &&
            (m != VX_BATCH_PIPELINED || b->pipe_ok) && (m != VX_BATCH_SPLIT || b->split_ok)) {
            b->tuned = true;
            b->tuned_mode = it->second.mode;
            b->tuned_grid = it->second.grid;
            b->tuned_ratio = it->second.ratio;
            b->tuned_fill_split = it->second.fill_split;
            b->tuned_emit_split = it->second.emit_split;
        }
    }
    *out = b;
    return VX_OK;
}

extern "C" int vx_batch_destroy(vx_batch *b) {
    if (!b) return VX_ERR_INVALID;
    vx_ctx *ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    vx_batch_unref(b);
    return VX_OK;
}

extern "C" int vx_batch_set_mode(vx_batch *b, int32_t mode) {
    if (!b || mode < VX_BATCH_AUTO || mode > VX_BATCH_SPLIT) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    if (mode == VX_BATCH_FUSED && !b->fused_ok) return VX_ERR_INVALID;
    if (mode == VX_BATCH_PIPELINED && !b->pipe_ok) return VX_ERR_INVALID;
    if (mode == VX_BATCH_SPLIT && !b->split_ok) return VX_ERR_INVALID;
    if (mode != b->mode) {
        (void)hipStreamSynchronize(b->ctx->stream);
        drop_graph(b);
        b->mode = mode;
    }
    return VX_OK;
}

extern "C" int vx_batch_set_grid(vx_batch *b, int32_t grid) {
    if (!b || grid < 0) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    if (grid != b->grid) {
        (void)hipStreamSynchronize(b->ctx->stream);
        drop_graph(b);
        b->grid = grid;
    }
    return VX_OK;
}

extern "C" int vx_batch_set_roles(vx_batch *b, int32_t ratio) {
    if (!b || (ratio != 0 && (ratio < 2 || ratio > kPipeRatioMax))) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    if (ratio != b->pipe_ratio) {
        (void)hipStreamSynchronize(b->ctx->stream);
        drop_graph(b);
        b->pipe_ratio = ratio;
    }
    return VX_OK;
}

extern "C" int vx_batch_set_splits(vx_batch *b, int32_t fill_parts, int32_t emit_parts) {
    if (!b || !fill_split_ok(fill_parts) || !emit_split_ok(emit_parts)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    if (fill_parts != b->fill_split || emit_parts != b->emit_split) {
        (void)hipStreamSynchronize(b->ctx->stream);
        drop_graph(b);
        b->fill_split = fill_parts;
        b->emit_split = emit_parts;
    }
    return VX_OK;
}

extern "C" int vx_batch_get_mode(vx_batch *b, int32_t *mode) {
    if (!b || !mode) return VX_ERR_INVALID;
    *mode = form_of(b);
    return VX_OK;
}

extern "C" int vx_batch_run(vx_ctx *ctx, vx_batch *b) {
    if (!ctx || !b || b->ctx != ctx) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!b->graph_exec) {
        int st = capture_graph(b);
        if (st) return st;
    }
    VX_HIP(ctx, hipGraphLaunch(b->graph_exec, ctx->stream));
    b->ran = true;
    b->results = false;
    return VX_OK;
}

extern "C" int vx_batch_run_per_chunk(vx_ctx *ctx, vx_batch *b) {
    if (!ctx || !b || b->ctx != ctx) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    const int n = b->n;
    for (int i = 0; i < n; i++)
        VX_HIP(ctx, launch_fill(b->fill_dev + i, nullptr, 1, b->reqs[i].edge, ctx->terrains_dev,
                                ctx->lut_dev, b->overflow + 1, 0, ctx->stream));
    auto slabs_of = [&](int i) { return b->slabs ? b->slabs + (size_t)kSlabs * i : nullptr; };
    for (int i = 0; i < n; i++)
        VX_HIP(ctx, launch_count(b->mesh_dev + i, 1, b->reqs[i].edge, b->counts + i, slabs_of(i),
                                 ctx->stream));
    VX_HIP(ctx, enqueue_scan(b));
    for (int i = 0; i < n; i++)
        VX_HIP(ctx, launch_emit(b->mesh_dev + i, nullptr, 1, 0, b->reqs[i].edge, b->bases + i,
                                b->counts + i, slabs_of(i), b->out, ctx->tables_dev, b->overflow,
                                0, ctx->stream));
    b->ran = true;
    b->results = false;
    return VX_OK;
}

extern "C" int vx_batch_wait(vx_ctx *ctx, vx_batch *b, uint64_t *total_faces) {
    if (!ctx || !b || b->ctx != ctx) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!b->ran) return VX_ERR_NOT_READY;
    VX_HIP(ctx, hipMemcpyAsync(ctx->pinned, b->total, 8, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipMemcpyAsync(ctx->pinned + 8, b->overflow, 8, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t sched_err;
    memcpy(&sched_err, ctx->pinned + 12, 4);
    if (sched_err || (check_schedules() && !schedules_ok(b))) {
        ctx->last_error = "batch: a fill or emit launch met a schedule entry outside the batch";
        return VX_ERR_DEVICE;
    }
    uint32_t ferr = 0;
    if (use_fused(b)) {
        VX_HIP(ctx, hipMemcpyAsync(ctx->pinned + 16, b->fctl_ptrs.error, 4, hipMemcpyDeviceToHost,
                                   ctx->stream));
        VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        memcpy(&ferr, ctx->pinned + 16, 4);
    }
    uint64_t total;
    uint32_t overflow;
    memcpy(&total, ctx->pinned, 8);
    memcpy(&overflow, ctx->pinned + 8, 4);
    if (ferr) {
        ctx->last_error = "fused batch: a look-back wait exceeded its bound";
        return VX_ERR_DEVICE;
    }
    if (overflow || total > b->capacity) {
        // grow the arena to the exact total and re-run only the meshing
        // output: the scan again (same bases from this run's counts; it also
        // writes the emission order, which a fused run never wrote, and clears
        // the overflow word), then the staged emit
        int st = alloc_arena(b, total + total / 16 + 1024);  // also drops the graph
        if (st) {
            // the run's meshes are lost with the old arena: nothing is
            // readable until a later run succeeds
            b->ran = false;
            b->h_total = 0;
            b->results = false;
            return st;
        }
        // a persistent form writes no slab counts: count again first
        if (b->slabs) VX_HIP(ctx, enqueue_count(b));
        VX_HIP(ctx, enqueue_scan(b));
        VX_HIP(ctx, enqueue_emit(b));
        VX_HIP(ctx, hipMemcpyAsync(ctx->pinned + 8, b->overflow, 8, hipMemcpyDeviceToHost,
                                   ctx->stream));
        VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        memcpy(&overflow, ctx->pinned + 8, 4);
        memcpy(&sched_err, ctx->pinned + 12, 4);
        if (sched_err) {
            ctx->last_error = "batch: the re-emit met a schedule entry outside the batch";
            return VX_ERR_DEVICE;
        }
        if (overflow) return VX_ERR_CAPACITY;
    }
    b->h_total = total;
    b->results = false;
    if (total_faces) *total_faces = total;
    return VX_OK;
}

extern "C" int vx_batch_chunk_faces(vx_ctx *ctx, vx_batch *b, int32_t i, uint64_t *first,
                                    uint64_t *count) {
    if (!ctx || !b || i < 0 || i >= b->n) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!b->ran) return VX_ERR_NOT_READY;
    int st = fetch_results(b);
    if (st) return st;
    if (first) *first = b->h_bases[i];
    if (count) *count = b->h_counts[i];
    return VX_OK;
}

extern "C" int vx_batch_read_ids(vx_ctx *ctx, vx_batch *b, int32_t i, uint8_t *dst, uint64_t cap) {
    if (!ctx || !b || !dst || i < 0 || i >= b->n) return VX_ERR_INVALID;
    const uint64_t nv = (uint64_t)b->reqs[i].edge * b->reqs[i].edge * b->reqs[i].edge;
    if (cap < nv) return VX_ERR_CAPACITY;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    VX_HIP(ctx, hipMemcpyAsync(dst, (char *)b->store + b->chunk_off[i], nv, hipMemcpyDeviceToHost,
                               ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_batch_read_mesh(vx_ctx *ctx, vx_batch *b, uint32_t format, void *dst,
                                  uint64_t cap) {
    if (!ctx || !b || !dst) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!b->ran) return VX_ERR_NOT_READY;
    const uint64_t F = b->h_total;
    const void *src;
    uint64_t bytes;
    if (format == kFmtInstances && (b->formats & kFmtInstances)) {
        src = b->out.instances;
        bytes = F * 68;
    } else if (format == kFmtVertices && (b->formats & kFmtVertices)) {
        src = b->out.vertices;
        bytes = F * 80;
    } else if (format == 4u && (b->formats & kFmtVertices)) {  // indices
        src = b->out.indices;
        bytes = F * 24;
    } else {
        return VX_ERR_INVALID;
    }
    if (cap < bytes) return VX_ERR_CAPACITY;
    if (!bytes) return VX_OK;
    VX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VX_OK;
}

extern "C" int vx_batch_chunk(vx_batch *b, int32_t i, vx_chunk **out) {
    if (!b || !out || i < 0 || i >= b->n) return VX_ERR_INVALID;
    vx_ctx *ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    const vx_batch_req &r = b->reqs[i];
    const Layout L = layout_of(r.edge);
    char *c = (char *)b->store + b->chunk_off[i];
    vx_chunk *ch = new (std::nothrow) vx_chunk();
    if (!ch) return VX_ERR_NOMEM;
    ch->ctx = ctx;
    ch->coord = r.coord;
    ch->edge = r.edge;
    ch->shared = b->store_blk;
    b->store_blk->refs.fetch_add(1);
    ch->ids = (uint8_t *)c;
    ch->occ = (uint32_t *)(c + L.occ);
    ch->planes = (uint32_t *)(c + L.planes);
    ch->meta = (uint32_t *)(c + L.meta);
    ch->solid_type = kBlockGrass;
    *out = ch;
    return VX_OK;
}

extern "C" int vx_batch_set_cull(vx_batch *b, const vx_frustum *f, const int32_t *coords_dev,
                                 int32_t n, int32_t edge, uint8_t *mask_dev) {
    if (!b || n < 0 || (n && (!f || !coords_dev || !mask_dev || edge <= 0))) return VX_ERR_INVALID;
    vx_ctx *ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (n == 0) {
        if (b->cull.n != 0) {
            VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // a queued replay may still run
            drop_graph(b);
        }
        b->cull = CullJob{};
        return VX_OK;
    }
    if (!b->planes_dev) {
        int st = ctx->dev_alloc(sizeof(f->planes), (void **)&b->planes_dev);
        if (st) return st;
        VX_HIP(ctx, hipHostMalloc((void **)&b->planes_pin, sizeof(f->planes), hipHostMallocDefault));
        VX_HIP(ctx, hipEventCreateWithFlags(&b->planes_ev, hipEventDisableTiming));
    } else {
        VX_HIP(ctx, hipEventSynchronize(b->planes_ev));  // previous frame's copy has run
    }
    memcpy(b->planes_pin, f->planes, sizeof(f->planes));
    VX_HIP(ctx, hipMemcpyAsync(b->planes_dev, b->planes_pin, sizeof(f->planes),
                               hipMemcpyHostToDevice, ctx->stream));
    VX_HIP(ctx, hipEventRecord(b->planes_ev, ctx->stream));
    CullJob job{b->planes_dev, coords_dev, mask_dev, n, edge};
    if (job.coords != b->cull.coords || job.mask != b->cull.mask || job.n != b->cull.n ||
        job.edge != b->cull.edge) {
        // kernel arguments change (planes alone are read from memory); a
        // queued replay of the old graph may still run
        VX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        drop_graph(b);
    }
    b->cull = job;
    return VX_OK;
}

extern "C" int vx_batch_draw_list(vx_ctx *ctx, vx_batch *b, const vx_frustum *f, uint32_t mode,
                                  vx_draw_indexed *args_dev, uint32_t *count_dev) {
    if (!ctx || !b || b->ctx != ctx || !f || !args_dev || !count_dev) return VX_ERR_INVALID;
    if (!((mode == VX_DRAW_INSTANCED && (b->formats & kFmtInstances)) ||
          (mode == VX_DRAW_EXPANDED && (b->formats & kFmtVertices))))
        return VX_ERR_INVALID;
    for (const vx_batch_req &r : b->reqs)
        if (r.edge != b->max_edge) return VX_ERR_INVALID;  // one AABB edge per list
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!b->ran) return VX_ERR_NOT_READY;
    if (!b->coords_dev) {
        std::vector<int32_t> c(3 * (size_t)b->n);
        for (int32_t i = 0; i < b->n; i++) {
            c[3 * i] = b->reqs[i].coord.x;
            c[3 * i + 1] = b->reqs[i].coord.y;
            c[3 * i + 2] = b->reqs[i].coord.z;
        }
        int st = ctx->dev_alloc(c.size() * 4, (void **)&b->coords_dev);
        if (st) return st;
        VX_HIP(ctx, hipMemcpyAsync(b->coords_dev, c.data(), c.size() * 4, hipMemcpyHostToDevice,
                                   ctx->stream));
        VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // c is pageable and goes out of scope
    }
    return vx_draw_list_locked(ctx, f, b->coords_dev, b->counts, b->bases, b->n, b->max_edge, mode,
                               args_dev, count_dev);
}

extern "C" int vx_batch_device_buffers(vx_batch *b, void **inst, void **vert, void **idx) {
    if (!b) return VX_ERR_INVALID;
    if (inst) *inst = b->out.instances;
    if (vert) *vert = b->out.vertices;
    if (idx) *idx = b->out.indices;
    return VX_OK;
}

extern "C" int vx_batch_autotune(vx_ctx *ctx, vx_batch *b, int32_t reps, vx_tune_result *res) {
    if (!ctx || !b || b->ctx != ctx || reps <= 0) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    {
        // one run in the current form first: it grows the mesh arena if the
        // batch's faces outgrow it, so that no candidate is timed on a run
        // that skips the faces beyond the arena
        uint64_t faces = 0;
        int st0 = vx_batch_run(ctx, b);
        if (!st0) st0 = vx_batch_wait(ctx, b, &faces);
        if (st0) return st0;
    }
    int cus = 1;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    std::vector<TuneChoice> cands;
    cands.push_back({VX_BATCH_STAGED, 0, 0});
    // staged tile shapes for batches small enough that the launchers split
    // chunks (the fill, under kFillGrid chunks) or items (the emit, under
    // kEmitGrid items) over several workgroups: half and twice the fill's
    // rule, and one or four (two when the rule gives four) emit workgroups
    // per item, each beside the other stage's rule
    {
        const int32_t items = emit_items_of(b->n, b->n_items, b->max_edge);
        const int hf = fill_parts_for(b->n, b->max_edge), he = emit_parts_for(items);
        if (b->n < kFillGrid) {
            cands.push_back({VX_BATCH_STAGED, 0, 0, hf / 2, 0});
            if (hf * 2 <= (b->max_edge > 32 ? 8 : 4)) cands.push_back({VX_BATCH_STAGED, 0, 0, hf * 2, 0});
        }
        if (items < kEmitGrid) {
            if (he != 1) cands.push_back({VX_BATCH_STAGED, 0, 0, 0, 1});
            cands.push_back({VX_BATCH_STAGED, 0, 0, 0, he == 4 ? 2 : 4});
        }
    }
    if (b->fused_ok) {
        cands.push_back({VX_BATCH_FUSED, 0, 0});
        for (int per = 3; per <= 4; per++)
            if (cus * per < b->n) cands.push_back({VX_BATCH_FUSED, cus * per, 0});
    }
    if (b->split_ok) cands.push_back({VX_BATCH_SPLIT, 0, 0});
    if (b->pipe_ok) {
        // fill and mesh roles: one workgroup per chunk with one mesher in
        // three; 8 per CU with one mesher in two, three or four
        cands.push_back({VX_BATCH_PIPELINED, 0, 3});
        for (int ratio = 2; ratio <= 4; ratio++)
            if (cus * 8 < b->n) cands.push_back({VX_BATCH_PIPELINED, cus * 8, ratio});
    }
    VX_HIP(ctx, hipStreamSynchronize(ctx->stream));  // a queued replay may still run
    hipEvent_t e0, e1;
    VX_HIP(ctx, hipEventCreate(&e0));
    VX_HIP(ctx, hipEventCreate(&e1));
    vx_tune_result r{};
    float best = 1e30f;
    TuneChoice bestc = cands[0];
    int st = VX_OK;
    // two rounds over the candidates, each keeping its faster round: the
    // first candidates timed would otherwise also carry the GPU's clock ramp
    std::vector<float> tmin(cands.size(), 1e30f);
    for (int round = 0; round < 2 && !st; round++) {
        for (size_t c = 0; c < cands.size() && !st; c++) {
            drop_graph(b);
            b->mode = cands[c].mode;
            b->grid = cands[c].grid;
            b->pipe_ratio = cands[c].ratio;
            b->fill_split = cands[c].fill_split;
            b->emit_split = cands[c].emit_split;
            if ((st = capture_graph(b))) break;
            hipError_t e = hipGraphLaunch(b->graph_exec, ctx->stream);  // warm-up
            if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
            for (int i = 0; i < reps && e == hipSuccess; i++)
                e = hipGraphLaunch(b->graph_exec, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e != hipSuccess) {
                ctx->set_error(e, "vx_batch_autotune");
                st = VX_ERR_DEVICE;
                break;
            }
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            tmin[c] = std::min(tmin[c], t / reps);
        }
    }
    for (size_t c = 0; c < cands.size() && !st; c++) {
        if (c < VX_TUNE_MAX_CANDIDATES) {
            r.cand_mode[c] = cands[c].mode;
            r.cand_grid[c] = cands[c].grid;
            r.cand_ratio[c] = cands[c].ratio;
            r.cand_fill_split[c] = cands[c].fill_split;
            r.cand_emit_split[c] = cands[c].emit_split;
            r.cand_ms[c] = tmin[c];
            r.n_candidates = (int32_t)c + 1;
        }
        if (tmin[c] < best) best = tmin[c], bestc = cands[c];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graph(b);
    b->mode = VX_BATCH_AUTO;
    b->grid = 0;
    b->pipe_ratio = 0;
    b->fill_split = b->emit_split = 0;
    b->tuned = true;
    b->tuned_mode = bestc.mode;
    b->tuned_grid = bestc.grid;
    b->tuned_ratio = bestc.ratio;
    b->tuned_fill_split = bestc.fill_split;
    b->tuned_emit_split = bestc.emit_split;
    b->ran = true;
    b->results = false;
    if (st) return st;
    r.mode = bestc.mode;
    r.grid = bestc.grid;
    r.ratio = bestc.ratio;
    r.fill_split = bestc.fill_split;
    r.emit_split = bestc.emit_split;
    r.best_ms = best;
    {
        std::lock_guard<std::mutex> tg(g_tune_mu);
        g_tune[tune_key(b)] = bestc;
        tune_save();
    }
    if (res) *res = r;
    return VX_OK;
}

extern "C" int vx_batch_profile(vx_ctx *ctx, vx_batch *b, int32_t reps, float ms[5]) {
    if (!ctx || !b || !ms || reps <= 0) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    hipEvent_t ev[6];
    for (int i = 0; i < 6; i++) VX_HIP(ctx, hipEventCreate(&ev[i]));
    double acc[5] = {0, 0, 0, 0, 0};
    int st = VX_OK;
    // one-launch forms and the split form: the whole run as one figure
    const bool whole = use_fused(b) || form_of(b) == VX_BATCH_SPLIT;
    for (int r = 0; r < reps && !st && whole; r++) {
        hipError_t e = hipEventRecord(ev[0], ctx->stream);
        if (e == hipSuccess) e = enqueue_all(b);
        if (e == hipSuccess) e = hipEventRecord(ev[1], ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e != hipSuccess) {
            ctx->set_error(e, "vx_batch_profile");
            st = VX_ERR_DEVICE;
            break;
        }
        float t = 0;
        (void)hipEventElapsedTime(&t, ev[0], ev[1]);
        acc[0] += t;
    }
    for (int r = 0; r < reps && !st && !whole; r++) {
        hipError_t e = hipEventRecord(ev[0], ctx->stream);
        if (e == hipSuccess) e = enqueue_fill(b);
        if (e == hipSuccess) e = hipEventRecord(ev[1], ctx->stream);
        if (e == hipSuccess && !b->fill_counts) e = enqueue_count(b);
        if (e == hipSuccess) e = hipEventRecord(ev[2], ctx->stream);
        if (e == hipSuccess) e = enqueue_scan(b);
        if (e == hipSuccess) e = hipEventRecord(ev[3], ctx->stream);
        if (e == hipSuccess) e = enqueue_emit(b);
        if (e == hipSuccess) e = hipEventRecord(ev[4], ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev[4]);
        if (e != hipSuccess) {
            ctx->set_error(e, "vx_batch_profile");
            st = VX_ERR_DEVICE;
            break;
        }
        for (int k = 0; k < 4; k++) {
            float t = 0;
            (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
            acc[k] += t;
        }
    }
    if (!st) {
        if (!b->graph_exec) st = capture_graph(b);
        if (!st) {
            hipError_t e = hipEventRecord(ev[0], ctx->stream);
            for (int r = 0; r < reps && e == hipSuccess; r++)
                e = hipGraphLaunch(b->graph_exec, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[5], ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(ev[5]);
            if (e != hipSuccess) {
                ctx->set_error(e, "vx_batch_profile graph");
                st = VX_ERR_DEVICE;
            } else {
                float t = 0;
                (void)hipEventElapsedTime(&t, ev[0], ev[5]);
                acc[4] = t;
            }
        }
    }
    for (int i = 0; i < 6; i++) (void)hipEventDestroy(ev[i]);
    if (st) return st;
    for (int k = 0; k < 5; k++) ms[k] = (float)(acc[k] / reps);
    b->ran = true;
    b->results = false;
    return VX_OK;
}
