// vx_world.cpp -- the reference's tick-driven chunk scheduler and world
// (ChunkGenerator, src/chunk_builder.rs:28-352; World, src/world.rs:14-58)
// with its jobs -- Chunk::new + create_mesh, or create_mesh of a resident
// chunk -- run on the GPU.  The host state machine follows the reference
// step for step so that a run reproduces the reference's observable world
// (which chunks and meshes are resident, each mesh built from the neighbour
// snapshot the reference would have used):
//
//   build_chunks (45-67): load_primary_chunks -> enqueue_chunks_in_frustum ->
//   process_chunk_loading_queue -> process_rebuild_queue -> update_world ->
//   clean_up_queues -> filter_unseen_chunks
//
// The reference's thread pool delivers results whenever a worker finishes;
// here a job's result becomes receivable `result_delay` ticks after it was
// submitted, in submission order -- one of the schedules the reference can
// produce, fixed so that runs are reproducible.
#include "vx_host.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <map>
#include <tuple>
#include <vector>

using namespace vx;

namespace {

constexpr int32_t kWorldEdge = 32;  // CHUNK_I32 (src/world.rs:21-22)

using Key = std::tuple<int32_t, int32_t, int32_t>;  // ChunkCoord3D's derived Ord: x, y, z
Key key_of(vx_coord c) { return Key{c.x, c.y, c.z}; }
bool same(vx_coord a, vx_coord b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
bool less(vx_coord a, vx_coord b) { return key_of(a) < key_of(b); }

// Rust's `f as i32`: truncation, saturating at the i32 range, NaN -> 0.
int32_t as_i32(float f) {
    if (std::isnan(f)) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

// Coord3DF::to_chunk_coord (src/coordinate.rs:33-38): floor, as i32, >> 5.
vx_coord chunk_of(const float p[3]) {
    return vx_coord{as_i32(std::floor(p[0])) >> 5, as_i32(std::floor(p[1])) >> 5,
                    as_i32(std::floor(p[2])) >> 5};
}

struct Job {
    vx_coord pos{0, 0, 0};
    uint64_t ready = 0;         // first tick whose update_world may receive it
    vx_chunk *chunk = nullptr;  // load jobs: the new chunk (owned)
    vx_mesh *mesh = nullptr;    // owned
};

}  // namespace

struct vx_world {
    vx_ctx *ctx = nullptr;
    vx_world_desc desc{};
    std::map<Key, vx_chunk *> chunks;  // world.chunks (retained handles)
    std::map<Key, vx_mesh *> meshes;   // world.meshes (retained handles)
    std::vector<vx_coord> load_queue, rebuild_queue, in_process;
    std::deque<Job> data_channel, mesh_channel;
    vx_coord player_chunk{0, 0, 0};
    float player_pos[3] = {0, 0, 0};
    vx_world_stats stats{};
};

namespace {

bool contains(const std::vector<vx_coord> &v, vx_coord c) {
    return std::any_of(v.begin(), v.end(), [&](vx_coord q) { return same(q, c); });
}

vx_chunk *chunk_at(vx_world *w, vx_coord c) {
    auto it = w->chunks.find(key_of(c));
    return it == w->chunks.end() ? nullptr : it->second;
}

// adjacent_chunks (134-169): the resident neighbours -x, +x, -z, +z now.
void adjacent(vx_world *w, vx_coord p, vx_chunk *adj[4]) {
    adj[0] = chunk_at(w, vx_coord{p.x - 1, p.y, p.z});
    adj[1] = chunk_at(w, vx_coord{p.x + 1, p.y, p.z});
    adj[2] = chunk_at(w, vx_coord{p.x, p.y, p.z - 1});
    adj[3] = chunk_at(w, vx_coord{p.x, p.y, p.z + 1});
}

// is_chunk_loaded (262-271)
bool is_loaded(vx_world *w, vx_coord p) {
    return w->chunks.count(key_of(p)) || contains(w->load_queue, p) || contains(w->in_process, p);
}

// The job closure of process_chunk_loading_queue / load_chunk_directly
// (80-84, 125-129): Chunk::new(pos) then create_mesh with the snapshot taken
// at submission.  The GPU work is queued now; the result waits in the
// channel until its tick.
int submit_load(vx_world *w, vx_coord pos) {
    vx_chunk *adj[4];
    adjacent(w, pos, adj);
    Job j;
    j.pos = pos;
    j.ready = w->stats.tick + (uint64_t)w->desc.result_delay;
    int st = vx_chunk_new(w->ctx, w->desc.terrain_id, pos, kWorldEdge, &j.chunk);
    if (st) return st;
    st = vx_chunk_create_mesh(w->ctx, j.chunk, adj, w->desc.formats, &j.mesh);
    if (st) {
        vx_chunk_release(j.chunk);
        return st;
    }
    w->data_channel.push_back(j);
    w->in_process.push_back(pos);
    w->stats.jobs_loaded++;
    return VX_OK;
}

// load_chunk_directly (115-132)
int load_directly(vx_world *w, vx_coord pos) {
    return is_loaded(w, pos) ? VX_OK : submit_load(w, pos);
}

// load_primary_chunks (171-212): on entering a new chunk, the player's
// column at y = 0 and the ring of radius 1 around it.
int load_primary(vx_world *w) {
    const vx_coord now = chunk_of(w->player_pos);
    if (same(now, w->player_chunk)) return VX_OK;  // is_in_new_chunk_pos (src/player.rs:24-27)
    w->player_chunk = now;
    const int32_t px = now.x, pz = now.z, r = 1;
    int st = load_directly(w, vx_coord{px, 0, pz});
    for (int32_t z = -r; z <= r && !st; z++) {
        st = load_directly(w, vx_coord{px + r, 0, pz + z});
        if (!st) st = load_directly(w, vx_coord{px - r, 0, pz + z});
    }
    for (int32_t x = -r + 1; x < r && !st; x++) {
        st = load_directly(w, vx_coord{px + x, 0, pz + r});
        if (!st) st = load_directly(w, vx_coord{px + x, 0, pz - r});
    }
    return st;
}

// enqueue_data (255-261)
int enqueue(vx_world *w, vx_coord p, const vx_frustum *f) {
    if (is_loaded(w, p)) return VX_OK;
    const int in = vx_frustum_contains(f, p, kWorldEdge);
    if (in < 0) return in;
    if (in) w->load_queue.push_back(p);
    return VX_OK;
}

// enqueue_chunks_in_frustum (214-253): rings 2 .. RENDER_DISTANCE-1 at y = 0.
int enqueue_in_frustum(vx_world *w, const vx_frustum *f) {
    const int32_t px = w->player_chunk.x, pz = w->player_chunk.z, RD = w->desc.render_distance;
    int st = VX_OK;
    if (RD > 1) {
        for (int32_t r = 2; r < RD && !st; r++) {
            for (int32_t z = -r; z <= r && !st; z++) {
                st = enqueue(w, vx_coord{px + r, 0, pz + z}, f);
                if (!st) st = enqueue(w, vx_coord{px - r, 0, pz + z}, f);
            }
            for (int32_t x = -r + 1; x < r && !st; x++) {
                st = enqueue(w, vx_coord{px + x, 0, pz + r}, f);
                if (!st) st = enqueue(w, vx_coord{px + x, 0, pz - r}, f);
            }
        }
    }
    return st;
}

// process_rebuild_queue (89-113): only while nothing is loading.
int process_rebuild(vx_world *w) {
    if (!w->in_process.empty() || w->rebuild_queue.empty()) return VX_OK;
    const vx_coord pos = w->rebuild_queue.front();
    w->rebuild_queue.erase(w->rebuild_queue.begin());
    vx_chunk *c = chunk_at(w, pos);
    if (!c) return VX_OK;
    vx_chunk *adj[4];
    adjacent(w, pos, adj);
    Job j;
    j.pos = pos;
    j.ready = w->stats.tick + (uint64_t)w->desc.result_delay;
    int st = vx_chunk_create_mesh(w->ctx, c, adj, w->desc.formats, &j.mesh);
    if (st) return st;
    w->mesh_channel.push_back(j);
    w->stats.jobs_rebuilt++;
    return VX_OK;
}

void put_mesh(vx_world *w, vx_coord pos, vx_mesh *m) {
    auto it = w->meshes.find(key_of(pos));
    if (it != w->meshes.end()) {
        vx_mesh_release(it->second);
        it->second = m;
    } else {
        w->meshes.emplace(key_of(pos), m);
    }
}

// rebuild_adjacent_chunks (273-300): order +x, -x, +z, -z.
void rebuild_adjacent(vx_world *w, vx_coord p) {
    if (w->chunks.empty()) return;
    const vx_coord n[4] = {{p.x + 1, p.y, p.z}, {p.x - 1, p.y, p.z}, {p.x, p.y, p.z + 1},
                           {p.x, p.y, p.z - 1}};
    for (const vx_coord &q : n)
        if (w->chunks.count(key_of(q))) w->rebuild_queue.push_back(q);
}

// update_world (310-332): at most one finished chunk and one rebuilt mesh.
void update_world(vx_world *w) {
    if (!w->data_channel.empty() && w->data_channel.front().ready <= w->stats.tick) {
        Job j = w->data_channel.front();
        w->data_channel.pop_front();
        auto it = w->chunks.find(key_of(j.pos));
        if (it != w->chunks.end()) {
            vx_chunk_release(it->second);
            it->second = j.chunk;
        } else {
            w->chunks.emplace(key_of(j.pos), j.chunk);
        }
        put_mesh(w, j.pos, j.mesh);
        w->in_process.erase(std::remove_if(w->in_process.begin(), w->in_process.end(),
                                           [&](vx_coord q) { return same(q, j.pos); }),
                            w->in_process.end());
        rebuild_adjacent(w, j.pos);
        w->stats.received_chunks++;
    }
    if (!w->mesh_channel.empty() && w->mesh_channel.front().ready <= w->stats.tick) {
        Job j = w->mesh_channel.front();
        w->mesh_channel.pop_front();
        put_mesh(w, j.pos, j.mesh);  // inserted even if the chunk was unloaded meanwhile
        w->stats.received_meshes++;
    }
}

// clean_up_queues (302-308)
void clean_up(vx_world *w) {
    w->load_queue.clear();
    std::sort(w->rebuild_queue.begin(), w->rebuild_queue.end(), less);
    w->rebuild_queue.erase(std::unique(w->rebuild_queue.begin(), w->rebuild_queue.end(), same),
                           w->rebuild_queue.end());
}

// filter_unseen_chunks (334-352): keep the box of +-RENDER_DISTANCE chunks
// around the player on all three axes.
void filter_unseen(vx_world *w) {
    const vx_coord pc = w->player_chunk;
    const int32_t RD = w->desc.render_distance;
    for (auto it = w->chunks.begin(); it != w->chunks.end();) {
        const vx_coord p{std::get<0>(it->first), std::get<1>(it->first), std::get<2>(it->first)};
        if (p.x <= RD + pc.x && p.z <= RD + pc.z && p.x >= -RD + pc.x && p.z >= -RD + pc.z &&
            p.y <= RD + pc.y && p.y >= -RD + pc.y) {
            ++it;
            continue;
        }
        auto m = w->meshes.find(it->first);
        if (m != w->meshes.end()) {
            vx_mesh_release(m->second);
            w->meshes.erase(m);
        }
        auto q = std::find_if(w->rebuild_queue.begin(), w->rebuild_queue.end(),
                              [&](vx_coord c) { return same(c, p); });
        if (q != w->rebuild_queue.end()) w->rebuild_queue.erase(q);
        vx_chunk_release(it->second);
        it = w->chunks.erase(it);
        w->stats.unloaded++;
    }
}

void release_all(vx_world *w) {
    for (auto &kv : w->chunks) vx_chunk_release(kv.second);
    for (auto &kv : w->meshes) vx_mesh_release(kv.second);
    for (Job &j : w->data_channel) {
        if (j.chunk) vx_chunk_release(j.chunk);
        if (j.mesh) vx_mesh_release(j.mesh);
    }
    for (Job &j : w->mesh_channel)
        if (j.mesh) vx_mesh_release(j.mesh);
    w->chunks.clear();
    w->meshes.clear();
    w->data_channel.clear();
    w->mesh_channel.clear();
}

}  // namespace

extern "C" int vx_world_create(vx_ctx *ctx, const vx_world_desc *desc, vx_world **out) {
    if (!ctx || !desc || !out || desc->render_distance < 0 || desc->result_delay < 0 ||
        desc->formats == 0 || (desc->formats & ~(kFmtInstances | kFmtVertices)))
        return VX_ERR_INVALID;
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(desc->player_pos[k])) return VX_ERR_INVALID;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        if (desc->terrain_id < 0 || desc->terrain_id >= (int)ctx->terrains.size())
            return VX_ERR_INVALID;
    }
    vx_world *w = new (std::nothrow) vx_world();
    if (!w) return VX_ERR_NOMEM;
    w->ctx = ctx;
    w->desc = *desc;
    for (int k = 0; k < 3; k++) w->player_pos[k] = desc->player_pos[k];
    w->player_chunk = chunk_of(w->player_pos);  // Player::new (src/player.rs:10-14)
    w->stats.player_chunk = w->player_chunk;
    *out = w;
    return VX_OK;
}

extern "C" int vx_world_destroy(vx_world *w) {
    if (!w) return VX_ERR_INVALID;
    {
        std::lock_guard<std::recursive_mutex> g(w->ctx->mu);
        release_all(w);
    }
    delete w;
    return VX_OK;
}

extern "C" int vx_world_tick(vx_world *w, const float player_pos[3], const vx_frustum *f) {
    if (!w || !player_pos || !f) return VX_ERR_INVALID;
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(player_pos[k])) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(w->ctx->mu);
    for (int k = 0; k < 3; k++) w->player_pos[k] = player_pos[k];  // Player::update_pos
    int st = load_primary(w);
    if (!st) st = enqueue_in_frustum(w, f);
    if (!st && !w->load_queue.empty()) {  // process_chunk_loading_queue (69-87)
        const vx_coord pos = w->load_queue.front();
        w->load_queue.erase(w->load_queue.begin());
        st = submit_load(w, pos);
    }
    if (!st) st = process_rebuild(w);
    if (st) return st;
    update_world(w);
    clean_up(w);
    filter_unseen(w);
    w->stats.tick++;
    w->stats.player_chunk = w->player_chunk;
    return VX_OK;
}

extern "C" int vx_world_stats_get(vx_world *w, vx_world_stats *out) {
    if (!w || !out) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(w->ctx->mu);
    *out = w->stats;
    out->chunks = (int32_t)w->chunks.size();
    out->meshes = (int32_t)w->meshes.size();
    out->rebuild_queue = (int32_t)w->rebuild_queue.size();
    out->in_process = (int32_t)w->in_process.size();
    return VX_OK;
}

extern "C" int vx_world_list(vx_world *w, uint32_t what, vx_coord *out, int32_t cap, int32_t *n) {
    if (!w || !n || cap < 0 || (cap && !out)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(w->ctx->mu);
    std::vector<vx_coord> v;
    if (what == VX_WORLD_CHUNKS)
        for (auto &kv : w->chunks)
            v.push_back(vx_coord{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)});
    else if (what == VX_WORLD_MESHES)
        for (auto &kv : w->meshes)
            v.push_back(vx_coord{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)});
    else if (what == VX_WORLD_REBUILD_QUEUE)
        v = w->rebuild_queue;
    else if (what == VX_WORLD_IN_PROCESS)
        v = w->in_process;
    else
        return VX_ERR_INVALID;
    *n = (int32_t)v.size();
    if (cap < (int32_t)v.size()) return cap == 0 ? VX_OK : VX_ERR_CAPACITY;
    std::copy(v.begin(), v.end(), out);
    return VX_OK;
}

extern "C" int vx_world_lookup(vx_world *w, vx_coord c, vx_chunk **chunk, vx_mesh **mesh) {
    if (!w || (!chunk && !mesh)) return VX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(w->ctx->mu);
    if (chunk) {
        auto it = w->chunks.find(key_of(c));
        *chunk = it == w->chunks.end() ? nullptr : it->second;
        if (*chunk) vx_chunk_retain(*chunk);
    }
    if (mesh) {
        auto it = w->meshes.find(key_of(c));
        *mesh = it == w->meshes.end() ? nullptr : it->second;
        if (*mesh) vx_mesh_retain(*mesh);
    }
    return VX_OK;
}
