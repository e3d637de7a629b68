// Host-side memory check of libvoxel_hip: drives every subsystem through the
// C ABI -- contexts, chunks and meshes (retained and released in every
// order), batches in every form and launch shape with arena growth, the
// autotuner and its cache file, chunks that outlive their batch, the server under an HBM budget (eviction, destruction with work
// in flight, handles that outlive it), the world tick loop, frustum
// masks and draw lists -- so that AddressSanitizer and UBSan, compiled into
// the host code only, see its allocations, reference counts and copies.
// The kernels run normally.  Exit status 0 and "host check ok" on success.
#include <sanitizer/lsan_interface.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "voxel.h"

#define CHECK(call)                                                                         \
    do {                                                                                    \
        const int st_ = (call);                                                             \
        if (st_ != VX_OK) {                                                                 \
            std::fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #call, st_,  \
                         vx_status_string(st_));                                            \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

#define EXPECT(cond)                                                                        \
    do {                                                                                    \
        if (!(cond)) {                                                                      \
            std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond);        \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

static vx_terrain density(uint32_t seed) {
    vx_terrain t{};
    t.kind = VX_TERRAIN_DENSITY;
    t.seed = seed;
    t.octaves = 4;
    t.frequency = 1.0f / 64;
    t.lacunarity = 2.0f;
    t.gain = 0.5f;
    t.y_scale = 1.0f / 128;
    return t;
}

// A perspective x look-at matrix (column-major) looking along +x from (eye).
static void camera(float eye_x, float m[16]) {
    const float f = 1.0f / std::tan(0.6f), a = 16.0f / 9.0f, n = 0.1f, fa = 1024.0f;
    float p[16] = {f / a, 0, 0, 0, 0, f, 0, 0, 0, 0, (fa + n) / (n - fa), -1, 0, 0,
                   2 * fa * n / (n - fa), 0};
    // view: forward +x, up +y -> right = +z; rows of the rotation (column-major)
    float v[16] = {0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, -40.0f, eye_x, 1};
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            float s = 0;
            for (int k = 0; k < 4; k++) s += p[k * 4 + r] * v[c * 4 + k];
            m[c * 4 + r] = s;
        }
}

static int chunks_and_meshes(vx_ctx *ctx, int32_t tid) {
    vx_chunk *c[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < 3; i++) CHECK(vx_chunk_new(ctx, tid, vx_coord{i, 0, 0}, 32, &c[i]));
    vx_chunk *const adj[4] = {c[0], c[2], nullptr, nullptr};
    vx_mesh *m = nullptr;
    CHECK(vx_chunk_create_mesh(ctx, c[1], adj, VX_FMT_VERTICES | VX_FMT_INSTANCES, &m));
    for (int i = 0; i < 3; i++) CHECK(vx_chunk_release(c[i]));  // the mesh owns its block
    uint64_t faces = 0;
    CHECK(vx_mesh_face_count(m, &faces));
    EXPECT(faces > 0);
    std::vector<unsigned char> buf(faces * 80);
    CHECK(vx_mesh_read_vertices(ctx, m, buf.data(), buf.size()));
    CHECK(vx_mesh_read_indices(ctx, m, buf.data(), faces * 24));
    CHECK(vx_mesh_read_instances(ctx, m, buf.data(), faces * 68));
    EXPECT(vx_mesh_read_vertices(ctx, m, buf.data(), faces * 80 - 1) == VX_ERR_CAPACITY);
    CHECK(vx_mesh_retain(m));
    CHECK(vx_mesh_release(m));
    CHECK(vx_mesh_release(m));
    // a chunk built from caller ids, meshed without neighbours
    std::vector<uint8_t> ids(16 * 16 * 16, 0);
    for (size_t i = 0; i < ids.size(); i += 3) ids[i] = 2;
    vx_chunk *u = nullptr;
    CHECK(vx_chunk_from_ids(ctx, vx_coord{-1, 2, 3}, 16, ids.data(), &u));
    vx_chunk *const none[4] = {nullptr, nullptr, nullptr, nullptr};
    CHECK(vx_chunk_create_mesh(ctx, u, none, VX_FMT_INSTANCES, &m));
    std::vector<uint8_t> back(ids.size());
    CHECK(vx_chunk_read_ids(ctx, u, back.data(), back.size()));
    EXPECT(back == ids);
    CHECK(vx_mesh_release(m));
    CHECK(vx_chunk_release(u));
    return 0;
}

static int batches(vx_ctx *ctx, int32_t tid) {
    const int side = 8, n = side * side;
    std::vector<vx_batch_req> reqs(n);
    for (int x = 0; x < side; x++)
        for (int z = 0; z < side; z++) {
            vx_batch_req &r = reqs[x * side + z];
            r.coord = vx_coord{x, 0, z};
            r.edge = 32;
            r.terrain_id = tid;
            r.nb[0] = x > 0 ? (x - 1) * side + z : -1;
            r.nb[1] = x < side - 1 ? (x + 1) * side + z : -1;
            r.nb[2] = z > 0 ? x * side + z - 1 : -1;
            r.nb[3] = z < side - 1 ? x * side + z + 1 : -1;
        }
    vx_batch *b = nullptr;
    // 1000 faces: far too small, so the first wait grows the arena
    CHECK(vx_batch_create(ctx, reqs.data(), n, VX_FMT_VERTICES | VX_FMT_INSTANCES, 1000, &b));
    uint64_t t_staged = 0, t_fused = 0;
    CHECK(vx_batch_set_mode(b, VX_BATCH_STAGED));
    CHECK(vx_batch_run(ctx, b));
    CHECK(vx_batch_wait(ctx, b, &t_staged));
    CHECK(vx_batch_set_mode(b, VX_BATCH_FUSED));
    CHECK(vx_batch_run(ctx, b));
    CHECK(vx_batch_wait(ctx, b, &t_fused));
    EXPECT(t_staged == t_fused && t_staged > 1000);
    // launch shapes: pipelined roles and grid, staged tile shapes, and the
    // autotuner over all of them (it writes the VX_TUNE_CACHE file)
    uint64_t t_shape = 0;
    CHECK(vx_batch_set_mode(b, VX_BATCH_PIPELINED));
    CHECK(vx_batch_set_grid(b, 7));
    CHECK(vx_batch_set_roles(b, 4));
    CHECK(vx_batch_run(ctx, b));
    CHECK(vx_batch_wait(ctx, b, &t_shape));
    EXPECT(t_shape == t_staged);
    CHECK(vx_batch_set_mode(b, VX_BATCH_STAGED));
    CHECK(vx_batch_set_splits(b, 8, 1));
    CHECK(vx_batch_run(ctx, b));
    CHECK(vx_batch_wait(ctx, b, &t_shape));
    EXPECT(t_shape == t_staged);
    EXPECT(vx_batch_set_splits(b, 3, 0) == VX_ERR_INVALID);
    EXPECT(vx_batch_set_roles(b, 1) == VX_ERR_INVALID);
    CHECK(vx_batch_set_mode(b, VX_BATCH_AUTO));
    vx_tune_result tr{};
    CHECK(vx_batch_autotune(ctx, b, 1, &tr));
    EXPECT(tr.n_candidates >= 5 && tr.best_ms > 0.0f);
    CHECK(vx_batch_run(ctx, b));
    CHECK(vx_batch_wait(ctx, b, &t_shape));
    EXPECT(t_shape == t_staged);
    uint64_t first = 0, count = 0;
    CHECK(vx_batch_chunk_faces(ctx, b, n - 1, &first, &count));
    EXPECT(first + count == t_staged);
    std::vector<unsigned char> all(t_staged * 80);
    CHECK(vx_batch_read_mesh(ctx, b, VX_FMT_VERTICES, all.data(), all.size()));
    vx_chunk *kept = nullptr;
    CHECK(vx_batch_chunk(b, 9, &kept));
    CHECK(vx_batch_destroy(b));  // the chunk view keeps the batch's store alive
    std::vector<uint8_t> ids(32 * 32 * 32);
    CHECK(vx_chunk_read_ids(ctx, kept, ids.data(), ids.size()));
    CHECK(vx_chunk_release(kept));
    return 0;
}

static int frustum_and_draws(vx_ctx *ctx, int32_t tid) {
    float m[16];
    camera(-20.0f, m);
    vx_frustum f{};
    CHECK(vx_frustum_from_matrix(m, &f));
    const int n = 256;
    std::vector<vx_coord> coords(n);
    for (int i = 0; i < n; i++) coords[i] = vx_coord{i % 16, 0, i / 16 - 8};
    std::vector<uint8_t> mask(n);
    CHECK(vx_frustum_mask(ctx, &f, coords.data(), n, 32, mask.data()));
    int host = 0, dev = 0;
    for (int i = 0; i < n; i++) {
        host += vx_frustum_contains(&f, coords[i], 32) == 1;
        dev += mask[i] != 0;
    }
    EXPECT(host == dev);
    std::vector<vx_batch_req> reqs(16);
    for (int i = 0; i < 16; i++) {
        reqs[i].coord = vx_coord{i, 0, 0};
        reqs[i].edge = 32;
        reqs[i].terrain_id = tid;
        for (int k = 0; k < 4; k++) reqs[i].nb[k] = -1;
    }
    vx_batch *b = nullptr;
    CHECK(vx_batch_create(ctx, reqs.data(), 16, VX_FMT_VERTICES, 0, &b));
    CHECK(vx_batch_run(ctx, b));
    CHECK(vx_batch_wait(ctx, b, nullptr));
    void *args = nullptr, *cnt = nullptr;
    CHECK(vx_buffer_alloc(ctx, 16 * sizeof(vx_draw_indexed), &args));
    CHECK(vx_buffer_alloc(ctx, 8, &cnt));
    CHECK(vx_batch_draw_list(ctx, b, &f, VX_DRAW_EXPANDED, (vx_draw_indexed *)args,
                             (uint32_t *)cnt));
    uint32_t count[2] = {0, 0};
    CHECK(vx_buffer_download(ctx, count, cnt, 8));
    EXPECT(count[0] <= 16);
    CHECK(vx_buffer_free(ctx, args));
    CHECK(vx_buffer_free(ctx, cnt));
    CHECK(vx_batch_destroy(b));
    return 0;
}

static int server(int32_t tid_seed) {
    vx_ctx *ctx = nullptr;
    vx_ctx_desc d{0, 0, 48ull << 20};
    CHECK(vx_ctx_create(&d, &ctx));
    vx_terrain t = density(tid_seed);
    int32_t tid = 0;
    CHECK(vx_terrain_register(ctx, &t, &tid));
    vx_server_desc sd{16, 3, VX_FMT_VERTICES, 0};
    vx_server *s = nullptr;
    CHECK(vx_server_create(ctx, &sd, &s));
    std::vector<vx_request> reqs;
    for (int i = 0; i < 240; i++) {
        vx_request r{};
        r.coord = vx_coord{i % 20, 0, i / 20};
        r.edge = i % 3 == 0 ? 16 : 32;
        r.terrain_id = tid;
        r.neighbours = i % 2 ? VX_NB_RESIDENT : VX_NB_NONE;
        reqs.push_back(r);
    }
    std::vector<uint64_t> tickets(reqs.size());
    CHECK(vx_server_submit(s, reqs.data(), (int32_t)reqs.size(), tickets.data()));
    CHECK(vx_server_drain(s));
    std::vector<vx_completion> done(512);
    int32_t got = 0;
    CHECK(vx_server_poll(s, done.data(), (int32_t)done.size(), &got));
    EXPECT(got == (int32_t)reqs.size());
    for (int i = 0; i < got; i++) EXPECT(done[i].status == VX_OK);
    vx_server_stats st{};
    CHECK(vx_server_stats_get(s, &st));
    EXPECT(st.evictions > 0);
    vx_chunk *c = nullptr;
    vx_mesh *m = nullptr;
    CHECK(vx_server_lookup(s, reqs.back().coord, reqs.back().edge, reqs.back().terrain_id, &c, &m));
    // a 64^3 request and 120 more queued, left in flight at destruction
    vx_request big{};
    big.coord = vx_coord{100, 0, 100};
    big.edge = 64;
    big.terrain_id = tid;
    CHECK(vx_server_submit(s, &big, 1, tickets.data()));
    CHECK(vx_server_submit(s, reqs.data(), 120, tickets.data()));
    (void)vx_server_pump(s, 0);
    CHECK(vx_server_destroy(s));
    std::vector<uint8_t> ids(32 * 32 * 32);
    CHECK(vx_chunk_read_ids(ctx, c, ids.data(), ids.size()));  // outlives the server
    CHECK(vx_chunk_release(c));
    CHECK(vx_mesh_release(m));
    CHECK(vx_ctx_destroy(ctx));
    return 0;
}

static int world(vx_ctx *ctx, int32_t tid) {
    vx_world_desc wd{};
    wd.terrain_id = tid;
    wd.render_distance = 4;
    wd.result_delay = 1;
    wd.formats = VX_FMT_VERTICES;
    wd.player_pos[0] = 16;
    wd.player_pos[1] = 40;
    wd.player_pos[2] = 16;
    vx_world *w = nullptr;
    CHECK(vx_world_create(ctx, &wd, &w));
    float pos[3] = {16, 40, 16};
    for (int k = 0; k < 60; k++) {
        if (k < 20) pos[0] += 7.0f;
        if (k == 40) pos[2] += 300.0f;
        float m[16];
        camera(pos[0] - 30.0f, m);
        vx_frustum f{};
        CHECK(vx_frustum_from_matrix(m, &f));
        CHECK(vx_world_tick(w, pos, &f));
    }
    vx_world_stats st{};
    CHECK(vx_world_stats_get(w, &st));
    EXPECT(st.jobs_loaded > 0);
    std::vector<vx_coord> list(1024);
    int32_t n = 0;
    CHECK(vx_world_list(w, VX_WORLD_CHUNKS, list.data(), (int32_t)list.size(), &n));
    if (n > 0) {
        vx_chunk *c = nullptr;
        vx_mesh *m = nullptr;
        CHECK(vx_world_lookup(w, list[0], &c, &m));
        if (c) CHECK(vx_chunk_release(c));
        if (m) CHECK(vx_mesh_release(m));
    }
    CHECK(vx_world_destroy(w));
    return 0;
}

int main() {
    int32_t ndev = 0;
    CHECK(vx_device_count(&ndev));
    if (ndev < 1) {
        std::fprintf(stderr, "no HIP device\n");
        return 2;
    }
    // the autotuner's cache file, read when the first batch is created: an
    // entry of each form and lines of every accepted and refused shape
    char cache[] = "/tmp/vx_host_check_tune_XXXXXX";
    const int fd = mkstemp(cache);
    if (fd < 0) return 1;
    const char lines[] =
        "not an entry\n"
        "6 32 1 1 1 1 0\n"
        "7 32 1 1 1 3 0 4\n"
        "8 32 1 1 1 2 0 0 2 3\n"
        "9 32 1 1 1 2 0 0 3 0\n"
        "10 32 1 1 1 3 0 1\n";
    if (write(fd, lines, sizeof lines - 1) != (ssize_t)(sizeof lines - 1)) return 1;
    close(fd);
    setenv("VX_TUNE_CACHE", cache, 1);
    vx_ctx *ctx = nullptr;
    vx_ctx_desc d{0, 0, 0};
    CHECK(vx_ctx_create(&d, &ctx));
    vx_terrain t = density(7);
    int32_t tid = 0;
    CHECK(vx_terrain_register(ctx, &t, &tid));
    if (chunks_and_meshes(ctx, tid)) return 1;
    std::printf("chunks and meshes ok\n");
    std::fflush(stdout);
    if (batches(ctx, tid)) return 1;
    std::printf("batches ok\n");
    std::fflush(stdout);
    if (frustum_and_draws(ctx, tid)) return 1;
    std::printf("frustum and draw lists ok\n");
    std::fflush(stdout);
    if (world(ctx, tid)) return 1;
    std::printf("world ok\n");
    std::fflush(stdout);
    CHECK(vx_ctx_destroy(ctx));
    if (server(5)) return 1;
    std::printf("server ok\n");
    std::fflush(stdout);
    // Leak check now, with every library object released, then leave without
    // running the HIP runtime's own static teardown: the sanitizer runtime's
    // device-memory bookkeeping cannot release blocks once that teardown has
    // unloaded the HSA runtime, which is no property of this library.
    unlink(cache);
    const int leaks = __lsan_do_recoverable_leak_check();
    if (leaks) {
        std::fprintf(stderr, "leak check failed\n");
        std::fflush(stderr);
        _exit(1);
    }
    std::printf("host check ok\n");
    std::fflush(stdout);
    std::fflush(stderr);
    _exit(0);
}
