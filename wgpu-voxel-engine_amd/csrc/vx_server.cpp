// vx_server.cpp -- continuous-batching chunk server with a device-resident
// LRU cache, and the streaming simulation driver.
//
// Reference behaviour restated (paths relative to the reference repository):
//   pool.execute(Chunk::new + create_mesh) per queued chunk
//                                  src/chunk_builder.rs:69-87, 114-132
//   adjacent_chunks snapshot at submission      src/chunk_builder.rs:134-169
//   results collected with try_recv and inserted into World.chunks/meshes
//                                  src/chunk_builder.rs:310-332
//   unload of chunks outside the view box       src/chunk_builder.rs:334-352
// The reference submits one chunk per 7-frame tick; here every pending
// request is admitted into the next device batch as soon as an in-flight
// slot and memory are free, and batches overlap on separate streams.
//
// Life of a batch ("flight"), never blocking the host:
//   admit   pick pending requests, allocate chunk storage (evicting LRU
//           entries to fit the HBM budget), upload descriptors from pinned
//           memory, queue fill -> count -> scan, copy the face total back,
//           record event 1
//   size    event 1 done: allocate the exact mesh arena, queue emit,
//           record event 2
//   retire  event 2 done: results become resident cache entries (chunk and
//           mesh views sharing the flight's two device blocks) and
//           completions are posted
#include <time.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <list>
#include <random>
#include <unordered_map>
#include <vector>

#include "vx_host.h"

using namespace vx;

namespace {

uint64_t now_ns() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

// A resident or in-flight chunk is identified by its coordinate, edge and
// terrain: the same coordinate under two terrains is two different chunks.
struct Key {
    int32_t x, y, z, edge, terrain;
    bool operator==(const Key &o) const {
        return x == o.x && y == o.y && z == o.z && edge == o.edge && terrain == o.terrain;
    }
};

struct KeyHash {
    size_t operator()(const Key &k) const {
        uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.edge + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.terrain * 0xD6E8FEB86659FD93ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

struct ChunkLayoutS {
    size_t occ, planes, meta, total;
};

ChunkLayoutS layout_s(int N) {
    const size_t W = (size_t)words_per_row(N);
    ChunkLayoutS L;
    L.occ = up256((size_t)N * N * N);
    L.planes = L.occ + up256((size_t)N * N * W * 4);
    L.meta = L.planes + up256(4 * (size_t)N * W * 4);
    L.total = L.meta + 256;
    return L;
}

struct Entry {
    vx_chunk *chunk = nullptr;
    vx_mesh *mesh = nullptr;
    std::list<Key>::iterator lru;
    uint64_t born = 0;  // pump epoch of its latest completion (miss or hit)
};

struct Waiter {
    uint64_t ticket;
    uint64_t arrival_ns;
};

struct Flight {
    int slot = -1;
    int phase = 0;  // 1 = fill/count/scan queued, 2 = emit queued
    int n = 0;
    int max_edge = 0;
    std::vector<vx_request> reqs;
    std::vector<Key> keys;
    std::vector<std::vector<Waiter>> waiters;  // per request: tickets (first is the owner)
    std::vector<std::array<int32_t, 4>> nb_present;
    std::vector<vx_chunk *> nb_hold;  // resident neighbours read by this flight
    std::vector<size_t> chunk_off;
    SharedBlock *store = nullptr;     // chunk storage | descriptors | ctl
    SharedBlock *arena = nullptr;     // mesh records
    FillWork *fill_dev = nullptr;
    MeshWork *mesh_dev = nullptr;
    int32_t *order_dev = nullptr;     // fill schedule, then the scan's emission order
    int32_t n_items = 0;              // emit items (kSlabs per 64^3 chunk when slabs is set)
    uint32_t *slabs = nullptr;        // per-slab face counts (flights holding 64^3 chunks)
    uint32_t *counts = nullptr;
    uint64_t *bases = nullptr;
    uint64_t *total = nullptr;
    uint32_t *overflow = nullptr;
    uint64_t faces = 0;
    MeshOut out{};
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev1 = nullptr, ev2 = nullptr;
    unsigned char *pinned = nullptr;  // descriptors out, counts/bases/total back
    size_t pinned_bytes = 0;
    bool busy = false;
};

}  // namespace

struct vx_server {
    vx_ctx *ctx = nullptr;
    int max_batch = 512;
    int max_inflight = 4;
    uint32_t formats = kFmtVertices;
    uint64_t next_ticket = 1;
    std::deque<std::pair<uint64_t, vx_request>> pending;
    std::unordered_map<Key, Entry, KeyHash> cache;
    std::list<Key> lru;  // front = most recently used
    std::unordered_map<Key, std::pair<Flight *, int>, KeyHash> inflight_keys;
    std::vector<Flight *> flights;
    std::vector<Slot> slots;
    std::vector<vx_completion> done;
    vx_server_stats st{};
    // Pump epoch: pump_once advances it.  An entry completed in this epoch
    // or the previous one is not evicted, so a client that polls and looks
    // up its completions between pumps finds them resident (admission of a
    // later flight in the same pump would otherwise evict them first).
    uint64_t epoch = 2;
};

namespace {

void complete(vx_server *S, uint64_t ticket, const vx_request &r, int status, bool hit,
              uint64_t faces, uint64_t arrival, uint64_t t, const int32_t nbp[4]) {
    vx_completion c{};
    c.ticket = ticket;
    c.coord = r.coord;
    c.edge = r.edge;
    c.status = status;
    c.cache_hit = hit ? 1u : 0u;
    c.faces = faces;
    c.arrival_ns = arrival;
    c.done_ns = t;
    for (int k = 0; k < 4; k++) c.nb_present[k] = nbp ? nbp[k] : 0;
    S->done.push_back(c);
    S->st.completed++;
    if (hit) S->st.cache_hits++;
    if (status != VX_OK) S->st.failed++;
}

void update_resident(vx_server *S) {
    S->st.resident_chunks = S->cache.size();
    S->st.resident_bytes = S->ctx->heap.live();
    S->st.peak_resident_bytes = std::max(S->st.peak_resident_bytes, S->st.resident_bytes);
    S->st.budget_bytes = S->ctx->budget;
}

void evict_entry(vx_server *S, const Key &k) {
    auto it = S->cache.find(k);
    if (it == S->cache.end()) return;
    S->lru.erase(it->second.lru);
    vx_chunk *c = it->second.chunk;
    vx_mesh *m = it->second.mesh;
    S->cache.erase(it);
    if (c && c->refs.fetch_sub(1) == 1) vx_chunk_destroy_locked(c);
    if (m && m->refs.fetch_sub(1) == 1) {
        vx_shared_unref_locked(S->ctx, m->shared);
        delete m;
    }
    S->st.evictions++;
}

bool pinned(const vx_server *S, const Entry &e) { return S->epoch - e.born < 2; }

void mark_completed(vx_server *S, Entry &e) { e.born = S->epoch; }

// Allocates `bytes` from the context heap, evicting least-recently-used
// cache entries while the HBM budget refuses it; entries completed in this
// pump or the previous one are kept.  On VX_ERR_BUDGET, *pinned_left tells
// whether kept entries remain (memory a later pump can free).
int alloc_evicting(vx_server *S, size_t bytes, void **out, bool *pinned_left) {
    *pinned_left = false;
    for (;;) {
        int st = S->ctx->dev_alloc(bytes, out);
        if (st != VX_ERR_BUDGET) return st;
        auto it = S->lru.end();
        bool found = false;
        while (it != S->lru.begin()) {
            --it;
            if (!pinned(S, S->cache.at(*it))) {
                found = true;
                break;
            }
        }
        if (!found) {
            *pinned_left = !S->lru.empty();
            return VX_ERR_BUDGET;
        }
        evict_entry(S, *it);
    }
}

SharedBlock *new_shared(void *p) {
    SharedBlock *s = new SharedBlock();
    s->ptr = p;
    return s;
}

const uint32_t *facing_plane_of(const vx_chunk *nb, int k) {
    static const int which[4] = {1, 0, 3, 2};
    return nb->planes + (size_t)which[k] * nb->edge * words_per_row(nb->edge);
}

void fail_flight_requests(vx_server *S, Flight *F, int status) {
    const uint64_t t = now_ns();
    for (int i = 0; i < F->n; i++) {
        for (const Waiter &w : F->waiters[i])
            complete(S, w.ticket, F->reqs[i], status, false, 0, w.arrival_ns, t, nullptr);
        S->inflight_keys.erase(F->keys[i]);
    }
}

void release_flight(vx_server *S, Flight *F) {
    for (vx_chunk *c : F->nb_hold)
        if (c->refs.fetch_sub(1) == 1) vx_chunk_destroy_locked(c);
    F->nb_hold.clear();
    vx_shared_unref_locked(S->ctx, F->store);
    vx_shared_unref_locked(S->ctx, F->arena);
    if (F->slot >= 0) S->slots[F->slot].busy = false;
    S->flights.erase(std::find(S->flights.begin(), S->flights.end(), F));
    delete F;
}

// Admits up to max_batch pending requests into a free slot.
int admit(vx_server *S) {
    vx_ctx *ctx = S->ctx;
    int slot = -1;
    for (int i = 0; i < (int)S->slots.size(); i++)
        if (!S->slots[i].busy) {
            slot = i;
            break;
        }
    if (slot < 0 || S->pending.empty()) return VX_OK;
    Flight *F = new Flight();
    F->slot = slot;
    // take pending requests: cache hits and in-flight duplicates complete or
    // coalesce without device work
    while (!S->pending.empty() && F->n < S->max_batch) {
        auto [ticket, r] = S->pending.front();
        S->pending.pop_front();
        const Key k{r.coord.x, r.coord.y, r.coord.z, r.edge, r.terrain_id};
        auto hit = S->cache.find(k);
        if (hit != S->cache.end()) {
            S->lru.splice(S->lru.begin(), S->lru, hit->second.lru);
            mark_completed(S, hit->second);
            complete(S, ticket, r, VX_OK, true, hit->second.mesh->faces, r.arrival_ns, now_ns(),
                     nullptr);
            continue;
        }
        auto inf = S->inflight_keys.find(k);
        if (inf != S->inflight_keys.end()) {
            inf->second.first->waiters[inf->second.second].push_back({ticket, r.arrival_ns});
            S->st.coalesced++;
            continue;
        }
        S->inflight_keys[k] = {F, F->n};
        F->reqs.push_back(r);
        F->keys.push_back(k);
        F->waiters.push_back({{ticket, r.arrival_ns}});
        F->n++;
    }
    if (F->n == 0) {
        delete F;
        return VX_OK;
    }
    const int n = F->n;
    // device storage: chunks | FillWork[n] | MeshWork[n] | order[n_items] |
    // counts[n] | bases[n] | total, overflow | slabs[kSlabs * n] (64^3 chunks)
    F->chunk_off.resize(n);
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        F->chunk_off[i] = off;
        off += layout_s(F->reqs[i].edge).total;
        F->max_edge = std::max(F->max_edge, F->reqs[i].edge);
    }
    const bool slabbed = F->max_edge > 32;
    F->n_items = n;
    if (slabbed)
        for (int i = 0; i < n; i++) F->n_items += F->reqs[i].edge == 64 ? kSlabs - 1 : 0;
    const size_t off_fill = off;
    const size_t off_mesh = off_fill + up256((size_t)n * sizeof(FillWork));
    const size_t off_order = off_mesh + up256((size_t)n * sizeof(MeshWork));
    const size_t off_counts = off_order + up256((size_t)F->n_items * 4);
    const size_t off_bases = off_counts + up256((size_t)n * 4);
    const size_t off_total = off_bases + up256((size_t)n * 8);
    const size_t off_slabs = off_total + 256;
    const size_t bytes = off_slabs + (slabbed ? up256((size_t)kSlabs * n * 4) : 0);
    void *blk = nullptr;
    bool pinned_left = false;
    int st = alloc_evicting(S, bytes, &blk, &pinned_left);
    if (st != VX_OK) {
        // nothing fits now: put the requests back unless neither a flight
        // nor a later pump can free memory, in which case they fail
        if (S->flights.empty() && !pinned_left) {
            fail_flight_requests(S, F, st);
        } else {
            for (int i = n - 1; i >= 0; i--) {
                S->inflight_keys.erase(F->keys[i]);
                for (auto w = F->waiters[i].rbegin(); w != F->waiters[i].rend(); ++w) {
                    vx_request r = F->reqs[i];
                    r.arrival_ns = w->arrival_ns;
                    S->pending.push_front({w->ticket, r});
                }
            }
        }
        delete F;
        return VX_OK;
    }
    F->store = new_shared(blk);
    char *base = (char *)blk;
    F->fill_dev = (FillWork *)(base + off_fill);
    F->mesh_dev = (MeshWork *)(base + off_mesh);
    F->counts = (uint32_t *)(base + off_counts);
    F->bases = (uint64_t *)(base + off_bases);
    F->total = (uint64_t *)(base + off_total);
    F->overflow = (uint32_t *)(base + off_total + 8);
    F->slabs = slabbed ? (uint32_t *)(base + off_slabs) : nullptr;
    Slot &sl = S->slots[slot];
    FillWork *fw = (FillWork *)sl.pinned;
    MeshWork *mw = (MeshWork *)(sl.pinned + up256((size_t)S->max_batch * sizeof(FillWork)));
    int32_t *ord = (int32_t *)((char *)mw + up256((size_t)S->max_batch * sizeof(MeshWork)));
    F->nb_present.resize(n);
    for (int i = 0; i < n; i++) {
        const vx_request &r = F->reqs[i];
        const ChunkLayoutS L = layout_s(r.edge);
        char *c = base + F->chunk_off[i];
        FillWork &f = fw[i];
        f = FillWork{};
        f.cx = r.coord.x;
        f.cy = r.coord.y;
        f.cz = r.coord.z;
        f.edge = r.edge;
        f.terrain = r.terrain_id;
        f.ids = (uint8_t *)c;
        f.occ = (uint32_t *)(c + L.occ);
        f.planes = (uint32_t *)(c + L.planes);
        f.meta = (uint32_t *)(c + L.meta);
        MeshWork &m = mw[i];
        m = MeshWork{};
        m.ox = r.coord.x * r.edge;
        m.oy = r.coord.y * r.edge;
        m.oz = r.coord.z * r.edge;
        m.edge = r.edge;
        m.occ = f.occ;
        m.ids = f.ids;
        m.solid_type = kBlockGrass;
        static const int dx[4] = {-1, 1, 0, 0}, dz[4] = {0, 0, -1, 1};
        for (int k = 0; k < 4; k++) {
            F->nb_present[i][k] = 0;
            if (r.neighbours != VX_NB_RESIDENT) continue;
            const Key nk{r.coord.x + dx[k], r.coord.y, r.coord.z + dz[k], r.edge, r.terrain_id};
            auto it = S->cache.find(nk);
            if (it == S->cache.end()) continue;
            vx_chunk *nc = it->second.chunk;
            nc->refs.fetch_add(1);
            F->nb_hold.push_back(nc);
            m.nb[k] = facing_plane_of(nc, k);
            F->nb_present[i][k] = 1;
        }
    }
    // longest-first schedule when sizes differ
    bool uniform = true;
    for (int i = 1; i < n; i++) uniform &= F->reqs[i].edge == F->reqs[0].edge;
    hipStream_t s = sl.stream;
    hipError_t e = hipMemcpyAsync(F->fill_dev, fw, (size_t)n * sizeof(FillWork),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(F->mesh_dev, mw, (size_t)n * sizeof(MeshWork), hipMemcpyHostToDevice, s);
    // overflow flag and sticky schedule-error word of this flight
    if (e == hipSuccess) e = hipMemsetAsync(F->overflow, 0, 8, s);
    const char *what = "server admit";
    if (!uniform && e == hipSuccess) {
        for (int i = 0; i < n; i++) ord[i] = i;
        std::stable_sort(ord, ord + n, [&](int a, int b) { return F->reqs[a].edge > F->reqs[b].edge; });
        F->order_dev = (int32_t *)(base + off_order);
        if (!is_schedule(ord, n)) {
            e = hipErrorInvalidValue;
            what = "server admit: fill schedule is not a permutation of the flight";
        } else {
            e = hipMemcpyAsync(F->order_dev, ord, (size_t)n * 4, hipMemcpyHostToDevice, s);
        }
    }
    if (e == hipSuccess)
        e = launch_fill(F->fill_dev, F->order_dev, n, F->max_edge, ctx->terrains_dev, ctx->lut_dev,
                        F->overflow + 1, 0, s);
    if (e == hipSuccess) e = launch_count(F->mesh_dev, n, F->max_edge, F->counts, F->slabs, s);
    // the scan rewrites the order array (the fill has consumed it) as the
    // emission order, largest chunks first (emit items when slabbed)
    F->order_dev = (int32_t *)(base + off_order);
    if (e == hipSuccess)
        e = launch_scan(F->counts, n, F->slabs ? F->mesh_dev : nullptr, F->bases, F->total,
                        F->overflow, F->order_dev, s);
    unsigned char *back = sl.pinned + sl.pinned_bytes - 256;
    // total (8 B) | overflow | schedule error
    if (e == hipSuccess) e = hipMemcpyAsync(back, F->total, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(sl.ev1, s);
    if (e != hipSuccess) {
        ctx->set_error(e, what);
        (void)hipStreamSynchronize(s);
        fail_flight_requests(S, F, VX_ERR_DEVICE);
        F->slot = -1;
        S->flights.push_back(F);
        release_flight(S, F);
        return VX_ERR_DEVICE;
    }
    sl.busy = true;
    F->phase = 1;
    S->flights.push_back(F);
    S->st.batches++;
    S->st.generated += n;
    return VX_OK;
}

int advance(vx_server *S, Flight *F) {
    vx_ctx *ctx = S->ctx;
    Slot &sl = S->slots[F->slot];
    if (F->phase == 1) {
        hipError_t q = hipEventQuery(sl.ev1);
        if (q == hipErrorNotReady) return VX_OK;
        if (q != hipSuccess) {
            ctx->set_error(q, "server event");
            fail_flight_requests(S, F, VX_ERR_DEVICE);
            release_flight(S, F);
            return VX_ERR_DEVICE;
        }
        uint64_t total;
        uint32_t sched_err;
        memcpy(&total, sl.pinned + sl.pinned_bytes - 256, 8);
        memcpy(&sched_err, sl.pinned + sl.pinned_bytes - 256 + 12, 4);
        if (sched_err) {
            ctx->last_error = "server: a fill launch met a schedule entry outside the flight";
            fail_flight_requests(S, F, VX_ERR_DEVICE);
            release_flight(S, F);
            return VX_ERR_DEVICE;
        }
        F->faces = total;
        size_t bytes = 0, off_v = 0, off_i = 0;
        if (S->formats & kFmtInstances) bytes = up256(total * 68);
        if (S->formats & kFmtVertices) {
            off_v = bytes;
            bytes += up256(total * 80);
            off_i = bytes;
            bytes += up256(total * 24);
        }
        if (bytes == 0) bytes = 256;
        void *blk = nullptr;
        bool pinned_left = false;
        int st = alloc_evicting(S, bytes, &blk, &pinned_left);
        // Younger flights wait for older ones to retire and free memory; the
        // oldest flight waits only for entries a later pump can evict (it
        // would otherwise wait on flights waiting on it).
        if (st == VX_ERR_BUDGET && (S->flights.front() != F || pinned_left)) return VX_OK;
        if (st != VX_OK) {
            fail_flight_requests(S, F, st);
            release_flight(S, F);
            return VX_OK;
        }
        F->arena = new_shared(blk);
        F->out = MeshOut{};
        F->out.capacity = total;
        F->out.formats = S->formats;
        if (S->formats & kFmtInstances) F->out.instances = (uint32_t *)blk;
        if (S->formats & kFmtVertices) {
            F->out.vertices = (float *)((char *)blk + off_v);
            F->out.indices = (uint32_t *)((char *)blk + off_i);
        }
        hipError_t e = launch_emit(F->mesh_dev, F->order_dev, F->n, F->n_items, F->max_edge,
                                   F->bases, F->counts, F->slabs, F->out, ctx->tables_dev,
                                   F->overflow, 0, sl.stream);
        unsigned char *back = sl.pinned + sl.pinned_bytes - 256 - up256((size_t)S->max_batch * 12);
        if (e == hipSuccess)
            e = hipMemcpyAsync(back, F->counts, (size_t)F->n * 4, hipMemcpyDeviceToHost, sl.stream);
        if (e == hipSuccess)  // overflow | schedule error after the emit
            e = hipMemcpyAsync(sl.pinned + sl.pinned_bytes - 256 + 16, F->overflow, 8,
                               hipMemcpyDeviceToHost, sl.stream);
        if (e == hipSuccess) e = hipEventRecord(sl.ev2, sl.stream);
        if (e != hipSuccess) {
            ctx->set_error(e, "server emit");
            (void)hipStreamSynchronize(sl.stream);
            fail_flight_requests(S, F, VX_ERR_DEVICE);
            release_flight(S, F);
            return VX_ERR_DEVICE;
        }
        F->phase = 2;
        return VX_OK;
    }
    hipError_t q = hipEventQuery(sl.ev2);
    if (q == hipErrorNotReady) return VX_OK;
    const uint64_t t = now_ns();
    if (q != hipSuccess) {
        ctx->set_error(q, "server event");
        fail_flight_requests(S, F, VX_ERR_DEVICE);
        release_flight(S, F);
        return VX_ERR_DEVICE;
    }
    {
        uint32_t flags[2];
        memcpy(flags, sl.pinned + sl.pinned_bytes - 256 + 16, 8);
        if (flags[1] || flags[0]) {  // the arena is sized exactly: no overflow either
            ctx->last_error = flags[1] ? "server: an emit launch met a schedule entry outside the flight"
                                       : "server: mesh arena overflow";
            fail_flight_requests(S, F, VX_ERR_DEVICE);
            release_flight(S, F);
            return VX_ERR_DEVICE;
        }
    }
    const uint32_t *counts =
        (const uint32_t *)(sl.pinned + sl.pinned_bytes - 256 - up256((size_t)S->max_batch * 12));
    uint64_t first = 0;
    for (int i = 0; i < F->n; i++) {
        const vx_request &r = F->reqs[i];
        const ChunkLayoutS L = layout_s(r.edge);
        char *c = (char *)F->store->ptr + F->chunk_off[i];
        vx_chunk *ch = new vx_chunk();
        ch->ctx = ctx;
        ch->coord = r.coord;
        ch->edge = r.edge;
        ch->shared = F->store;
        F->store->refs.fetch_add(1);
        ch->ids = (uint8_t *)c;
        ch->occ = (uint32_t *)(c + L.occ);
        ch->planes = (uint32_t *)(c + L.planes);
        ch->meta = (uint32_t *)(c + L.meta);
        ch->solid_type = kBlockGrass;
        vx_mesh *m = new vx_mesh();
        m->ctx = ctx;
        m->faces = counts[i];
        m->formats = S->formats;
        m->shared = F->arena;
        F->arena->refs.fetch_add(1);
        if (S->formats & kFmtInstances) m->instances = F->out.instances + first * 17;
        if (S->formats & kFmtVertices) {
            m->vertices = F->out.vertices + first * 20;
            m->indices = F->out.indices + first * 6;
        }
        first += counts[i];
        const Key &k = F->keys[i];
        S->lru.push_front(k);
        Entry en;
        en.chunk = ch;
        en.mesh = m;
        en.lru = S->lru.begin();
        en.born = S->epoch;
        S->cache[k] = en;
        S->inflight_keys.erase(k);
        for (const Waiter &w : F->waiters[i])
            complete(S, w.ticket, r, VX_OK, false, counts[i], w.arrival_ns, t, F->nb_present[i].data());
    }
    release_flight(S, F);
    update_resident(S);
    return VX_OK;
}

int pump_once(vx_server *S) {
    int rc = VX_OK;
    S->epoch++;
    // retire / advance in admission order
    std::vector<Flight *> fl = S->flights;
    for (Flight *F : fl) {
        int st = advance(S, F);
        if (st && rc == VX_OK) rc = st;
    }
    while (!S->pending.empty()) {
        size_t before = S->pending.size();
        int busy = 0;
        for (const Slot &s : S->slots) busy += s.busy;
        if (busy == (int)S->slots.size()) break;
        int st = admit(S);
        if (st && rc == VX_OK) rc = st;
        if (S->pending.size() == before) break;
    }
    update_resident(S);
    return rc;
}

}  // namespace

extern "C" int vx_server_create(vx_ctx *ctx, const vx_server_desc *d, vx_server **out) {
    if (!ctx || !out) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    vx_server *S = new (std::nothrow) vx_server();
    if (!S) return VX_ERR_NOMEM;
    S->ctx = ctx;
    if (d) {
        if (d->max_batch < 0 || d->max_inflight < 0 || d->max_batch > 65536 || d->max_inflight > 16 ||
            (d->formats & ~(kFmtInstances | kFmtVertices))) {
            delete S;
            return VX_ERR_INVALID;
        }
        if (d->max_batch) S->max_batch = d->max_batch;
        if (d->max_inflight) S->max_inflight = d->max_inflight;
        if (d->formats) S->formats = d->formats;
    }
    S->slots.resize(S->max_inflight);
    const size_t pin = up256((size_t)S->max_batch * sizeof(FillWork)) +
                       up256((size_t)S->max_batch * sizeof(MeshWork)) +
                       up256((size_t)S->max_batch * 4) + up256((size_t)S->max_batch * 12) + 256;
    for (Slot &sl : S->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev1, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev2, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&sl.pinned, pin, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            vx_server_destroy(S);
            return VX_ERR_DEVICE;
        }
        sl.pinned_bytes = pin;
    }
    *out = S;
    return VX_OK;
}

extern "C" int vx_server_destroy(vx_server *S) {
    if (!S) return VX_ERR_INVALID;
    vx_ctx *ctx = S->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    for (Slot &sl : S->slots)
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    while (!S->flights.empty()) {
        Flight *F = S->flights.back();
        fail_flight_requests(S, F, VX_ERR_NOT_READY);
        release_flight(S, F);
    }
    while (!S->lru.empty()) evict_entry(S, S->lru.back());
    for (Slot &sl : S->slots) {
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.ev1) (void)hipEventDestroy(sl.ev1);
        if (sl.ev2) (void)hipEventDestroy(sl.ev2);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete S;
    return VX_OK;
}

extern "C" int vx_server_submit(vx_server *S, const vx_request *reqs, int32_t n, uint64_t *tickets) {
    if (!S || n < 0 || (n && !reqs)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(S->ctx->mu);
    for (int i = 0; i < n; i++) {
        const vx_request &r = reqs[i];
        if ((r.edge != 16 && r.edge != 32 && r.edge != 64) || r.terrain_id < 0 ||
            r.terrain_id >= (int)S->ctx->terrains.size() || r.neighbours > VX_NB_RESIDENT ||
            !coord_in_range(r.coord, r.edge))
            return VX_ERR_INVALID;
    }
    const uint64_t t = now_ns();
    for (int i = 0; i < n; i++) {
        vx_request r = reqs[i];
        if (r.arrival_ns == 0) r.arrival_ns = t;
        const uint64_t ticket = S->next_ticket++;
        if (tickets) tickets[i] = ticket;
        S->st.submitted++;
        const Key k{r.coord.x, r.coord.y, r.coord.z, r.edge, r.terrain_id};
        auto hit = S->cache.find(k);
        if (hit != S->cache.end()) {  // resident: served without device work
            S->lru.splice(S->lru.begin(), S->lru, hit->second.lru);
            mark_completed(S, hit->second);
            complete(S, ticket, r, VX_OK, true, hit->second.mesh->faces, r.arrival_ns, now_ns(),
                     nullptr);
            continue;
        }
        auto inf = S->inflight_keys.find(k);
        if (inf != S->inflight_keys.end()) {  // same chunk already on the device
            inf->second.first->waiters[inf->second.second].push_back({ticket, r.arrival_ns});
            S->st.coalesced++;
            continue;
        }
        S->pending.push_back({ticket, r});
    }
    return VX_OK;
}

extern "C" int vx_server_pump(vx_server *S, int32_t wait_us) {
    if (!S) return VX_ERR_INVALID;
    const uint64_t until = now_ns() + (uint64_t)std::max(0, wait_us) * 1000ull;
    for (;;) {
        int rc;
        {
            std::lock_guard<std::recursive_mutex> g(S->ctx->mu);
            rc = pump_once(S);
            if (rc != VX_OK) return rc;
            if (!S->done.empty() || now_ns() >= until) return (int)S->done.size();
        }
        struct timespec ts = {0, 20000};
        nanosleep(&ts, nullptr);
    }
}

extern "C" int vx_server_poll(vx_server *S, vx_completion *out, int32_t cap, int32_t *n) {
    if (!S || cap < 0 || (cap && !out) || !n) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(S->ctx->mu);
    const int k = std::min<int>(cap, (int)S->done.size());
    std::copy(S->done.begin(), S->done.begin() + k, out);
    S->done.erase(S->done.begin(), S->done.begin() + k);
    *n = k;
    return VX_OK;
}

extern "C" int vx_server_drain(vx_server *S) {
    if (!S) return VX_ERR_INVALID;
    for (;;) {
        {
            std::lock_guard<std::recursive_mutex> g(S->ctx->mu);
            int rc = pump_once(S);
            if (rc != VX_OK) return rc;
            if (S->pending.empty() && S->flights.empty()) return VX_OK;
        }
        struct timespec ts = {0, 20000};
        nanosleep(&ts, nullptr);
    }
}

extern "C" int vx_server_lookup(vx_server *S, vx_coord c, int32_t edge, int32_t terrain_id,
                                vx_chunk **chunk, vx_mesh **mesh) {
    if (!S) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(S->ctx->mu);
    auto it = S->cache.find(Key{c.x, c.y, c.z, edge, terrain_id});
    if (it == S->cache.end()) return VX_ERR_NOT_READY;
    S->lru.splice(S->lru.begin(), S->lru, it->second.lru);
    if (chunk) {
        it->second.chunk->refs.fetch_add(1);
        *chunk = it->second.chunk;
    }
    if (mesh) {
        it->second.mesh->refs.fetch_add(1);
        *mesh = it->second.mesh;
    }
    return VX_OK;
}

extern "C" int vx_server_stats_get(vx_server *S, vx_server_stats *out) {
    if (!S || !out) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(S->ctx->mu);
    update_resident(S);
    *out = S->st;
    return VX_OK;
}

// ------------------------------------------------------------ simulation
extern "C" int vx_serve_simulate(vx_ctx *ctx, const vx_sim_desc *d, vx_sim_result *res) {
    if (!ctx || !d || !res) return VX_ERR_INVALID;
    if (!(d->rate_per_s > 0) || !(d->duration_s > 0) || d->n_classes < 1 || d->n_classes > 3 ||
        d->radius_chunks < 1)
        return VX_ERR_INVALID;
    vx_server *S = nullptr;
    int st = vx_server_create(ctx, &d->server, &S);
    if (st) return st;
    std::mt19937_64 rng(d->seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> G(0.0, 1.0);
    double wsum = 0;
    for (int c = 0; c < d->n_classes; c++) wsum += d->weights[c];
    const double R = d->radius_chunks * 32.0;
    double cx = 0, cz = 0, heading = 0;
    const uint64_t t0 = now_ns();
    const uint64_t t_end = t0 + (uint64_t)(d->duration_s * 1e9);
    double t_next = 0;  // seconds after t0 of the next arrival
    uint64_t t_cam = t0;
    std::vector<double> lat, lat_miss;
    lat.reserve((size_t)(d->rate_per_s * d->duration_s * 1.2) + 16);
    std::vector<vx_completion> comp(4096);
    uint64_t requests = 0, faces = 0;
    int rc = VX_OK;
    auto collect = [&]() {
        int32_t got = 0;
        do {
            vx_server_poll(S, comp.data(), (int32_t)comp.size(), &got);
            for (int i = 0; i < got; i++) {
                const vx_completion &c = comp[i];
                const double ms = (double)(c.done_ns - c.arrival_ns) * 1e-6;
                lat.push_back(ms);
                if (!c.cache_hit) lat_miss.push_back(ms);
                faces += c.cache_hit ? 0 : c.faces;
                if (c.status != VX_OK && rc == VX_OK) rc = c.status;
            }
        } while (got == (int32_t)comp.size());
    };
    std::vector<vx_request> batch;
    for (;;) {
        const uint64_t now = now_ns();
        // camera: constant speed, heading on a Gaussian random walk (1 rad/sqrt(s))
        const double dt = (double)(now - t_cam) * 1e-9;
        t_cam = now;
        heading += G(rng) * std::sqrt(dt);
        cx += d->speed_blocks_per_s * dt * std::cos(heading);
        cz += d->speed_blocks_per_s * dt * std::sin(heading);
        batch.clear();
        while (t0 + (uint64_t)(t_next * 1e9) <= now && t0 + (uint64_t)(t_next * 1e9) < t_end) {
            const double r = R * std::sqrt(U(rng)), a = 6.283185307179586 * U(rng);
            const double px = cx + r * std::cos(a), pz = cz + r * std::sin(a);
            double pick = U(rng) * wsum;
            int cls = 0;
            while (cls < d->n_classes - 1 && pick >= d->weights[cls]) pick -= d->weights[cls++];
            const int e = d->edges[cls];
            vx_request q{};
            q.coord = {(int32_t)std::floor(px / e), 0, (int32_t)std::floor(pz / e)};
            q.edge = e;
            q.terrain_id = d->terrain_id[cls];
            q.neighbours = d->neighbours;
            q.arrival_ns = t0 + (uint64_t)(t_next * 1e9);
            batch.push_back(q);
            t_next += -std::log(1.0 - U(rng)) / d->rate_per_s;
        }
        if (!batch.empty()) {
            st = vx_server_submit(S, batch.data(), (int32_t)batch.size(), nullptr);
            if (st) {
                rc = st;
                break;
            }
            requests += batch.size();
        }
        st = vx_server_pump(S, 0);
        if (st < 0) {
            rc = st;
            break;
        }
        collect();
        if (now >= t_end) break;
    }
    if (rc == VX_OK) {
        st = vx_server_drain(S);
        if (st) rc = st;
        collect();
    }
    const double wall = (double)(now_ns() - t0) * 1e-9;
    vx_server_stats ss{};
    vx_server_stats_get(S, &ss);
    vx_server_destroy(S);
    auto pct = [](std::vector<double> &v, double p) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        const size_t i = std::min(v.size() - 1, (size_t)std::ceil(p * (double)v.size()) - 1);
        return v[i];
    };
    std::memset(res, 0, sizeof(*res));
    res->requests = requests;
    res->completed = lat.size();
    res->cache_hits = ss.cache_hits;
    res->generated = ss.generated;
    res->batches = ss.batches;
    res->evictions = ss.evictions;
    res->faces = faces;
    res->wall_s = wall;
    res->throughput_rps = (double)lat.size() / wall;
    res->generated_per_s = (double)ss.generated / wall;
    double sum = 0;
    for (double v : lat) sum += v;
    res->mean_ms = lat.empty() ? 0 : sum / (double)lat.size();
    res->p50_ms = pct(lat, 0.50);
    res->p99_ms = pct(lat, 0.99);
    res->max_ms = lat.empty() ? 0 : lat.back();
    res->miss_p50_ms = pct(lat_miss, 0.50);
    res->miss_p99_ms = pct(lat_miss, 0.99);
    res->miss_max_ms = lat_miss.empty() ? 0 : lat_miss.back();
    res->peak_resident_bytes = ss.peak_resident_bytes;
    return rc;
}
