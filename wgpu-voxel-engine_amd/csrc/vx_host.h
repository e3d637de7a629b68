// vx_host.h -- host-side object layouts behind the opaque handles of
// include/voxel.h (context, chunk, mesh, batch) and the caching device heap.
#pragma once

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/voxel.h"
#include "vx_internal.h"
#include "vx_schedule.h"

namespace vx {

constexpr int kMaxTerrains = 64;

// Caching device allocator: blocks are rounded to size classes (4 per power
// of two, >= 256 B) and recycled through per-class free lists, so the chunk
// and mesh churn of a streaming world never reaches hipMalloc/hipFree after
// warm-up.  live() counts bytes handed out (what the HBM budget limits).
class DeviceHeap {
  public:
    static size_t size_class(size_t n);
    hipError_t alloc(size_t n, void **out);
    // Released blocks wait here until the caller has seen every stream
    // that may still use them pass the release (flush_pending).
    void release_pending(void *p);
    bool has_pending() const { return !pending_.empty(); }
    void flush_pending();
    void trim();
    uint64_t live() const { return live_; }
    uint64_t cached() const { return cached_; }
    ~DeviceHeap();

  private:
    std::map<size_t, std::vector<void *>> free_;
    std::unordered_map<void *, size_t> sizes_;
    std::vector<void *> pending_;
    uint64_t live_ = 0, cached_ = 0;
};

// A device block shared by several handles (the chunk views and mesh views
// of one admitted batch).  The last release returns it to the heap; callers
// hold the context lock.
struct SharedBlock {
    std::atomic<int> refs{1};
    void *ptr = nullptr;
};

// A chunk at `c` with this edge is accepted when its voxels and the one-voxel
// halo around them (the neighbour planes the fused kernel recomputes) have
// i32 world coordinates: c * edge - 1 >= INT32_MIN and c * edge + edge <=
// INT32_MAX on every axis.  The kernels then do all world-coordinate
// arithmetic in 32 bits without overflow.
inline bool coord_in_range(vx_coord c, int32_t edge) {
    const int64_t v[3] = {c.x, c.y, c.z};
    for (int64_t a : v) {
        const int64_t o = a * edge;
        if (o - 1 < INT32_MIN || o + edge > INT32_MAX) return false;
    }
    return true;
}

}  // namespace vx

struct vx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // second branch of split-batch runs (made on first use)
    std::recursive_mutex mu;
    uint64_t budget = 0;
    std::string last_error;
    vx::DeviceHeap heap;
    std::vector<vx::TerrainDev> terrains;
    vx::TerrainDev *terrains_dev = nullptr;
    vx::GradLut *lut_dev = nullptr;
    vx::MeshTables *tables_dev = nullptr;
    void *small_dev = nullptr;  // [0] face count, [1] overflow flag of the per-chunk path
    void *scratch = nullptr;    // staging for host-array entry points
    size_t scratch_bytes = 0;
    unsigned char *pinned = nullptr;  // 4 KiB pinned host staging for readbacks
    hipEvent_t free_fence = nullptr;  // recorded on `stream` at every dev_free
    uint32_t *draw_counts = nullptr;  // per-workgroup counts of draw-list passes
    int32_t draw_counts_cap = 0;
    uint32_t *draw_scratch = nullptr; // this device's static draw-list scratch

    void set_error(hipError_t e, const char *what);
    int dev_alloc(size_t n, void **out);
    void dev_free(void *p);
    int ensure_scratch(size_t bytes);
};

struct vx_batch;

struct vx_chunk {
    std::atomic<int> refs{1};
    vx_ctx *ctx = nullptr;
    vx_coord coord{0, 0, 0};
    int32_t edge = 0;
    void *block = nullptr;              // owned device block (null for a view)
    vx::SharedBlock *shared = nullptr;  // block this view points into (retained)
    uint8_t *ids = nullptr;
    uint32_t *occ = nullptr;
    uint32_t *planes = nullptr;
    uint32_t *meta = nullptr;
    uint32_t solid_type = vx::kBlockGrass;
};

struct vx_mesh {
    std::atomic<int> refs{1};
    vx_ctx *ctx = nullptr;
    uint64_t faces = 0;
    uint32_t formats = 0;
    void *block = nullptr;              // owned device block (null for a view)
    vx::SharedBlock *shared = nullptr;  // block this view points into (retained)
    float *vertices = nullptr;
    uint32_t *indices = nullptr;
    uint32_t *instances = nullptr;
};

struct vx_batch {
    std::atomic<int> refs{1};
    vx_ctx *ctx = nullptr;
    int32_t n = 0;
    int32_t max_edge = 0;
    uint32_t formats = 0;
    bool uniform = true;
    std::vector<vx_batch_req> reqs;
    std::vector<size_t> chunk_off;  // byte offset of chunk i in `store`
    void *store = nullptr;          // ids | occ | planes | meta per chunk
    vx::SharedBlock *store_blk = nullptr;  // owns `store`; chunk views retain it
    std::vector<vx_chunk *> ext;    // resident neighbours named by requests (retained)
    vx::CullJob cull{};             // optional frustum job run with the batch
    float *planes_dev = nullptr;    // device copy of the cull planes
    float *planes_pin = nullptr;    // pinned staging for per-frame planes
    hipEvent_t planes_ev = nullptr; // the staging copy has been consumed
    vx::FillWork *fill_dev = nullptr;
    vx::MeshWork *mesh_dev = nullptr;
    int32_t *order_dev = nullptr;   // longest-first schedule (null when uniform)
    int32_t *coords_dev = nullptr;  // chunk coordinates (draw lists), made on first use
    void *ctl = nullptr;            // counts[n] | bases[n] | total | overflow | emit_order[n_items]
                                    // | slabs[kSlabs * n] (batches holding 64^3 chunks)
    int32_t *emit_order = nullptr;  // emission order written by the scan (largest first)
    int32_t n_items = 0;            // emit items: n, or with slabs kSlabs per 64^3 chunk
    uint32_t *slabs = nullptr;      // per-slab face counts of 64^3 chunks (null when none)
    uint32_t *counts = nullptr;
    uint64_t *bases = nullptr;
    uint64_t *total = nullptr;
    uint32_t *overflow = nullptr;
    uint64_t capacity = 0;          // faces the mesh arena holds
    void *arena = nullptr;
    vx::MeshOut out{};
    int32_t mode = 0;               // VX_BATCH_* requested
    int32_t grid = 0;               // fused workgroups (0 = one per chunk)
    bool tuned = false;             // an autotuned choice applies in VX_BATCH_AUTO
    int32_t tuned_mode = 0, tuned_grid = 0, tuned_ratio = 0, tuned_fill_split = 0,
            tuned_emit_split = 0;
    int32_t pipe_ratio = 0;         // pipelined roles: one mesher in pipe_ratio (0 = default)
    int32_t fill_split = 0;         // staged fill workgroups per chunk (0 = the launcher's rule)
    int32_t emit_split = 0;         // staged emit workgroups per item (0 = the launcher's rule)
    bool fused_ok = false;          // fused form eligible
    bool pipe_ok = false;           // pipelined form eligible (fused_ok, edges <= 32)
    bool fill_counts = false;       // the staged fill also counts (no count kernel)
    bool split_ok = false;          // VX_BATCH_SPLIT eligible
    int32_t split_m = 0;            // split form: chunks [0, split_m) are the first half
    int32_t split_m2 = 0;           // its first fill covers [0, split_m2) (the half and the
                                    // later chunks it borders)
    uint64_t *split_total = nullptr;   // faces of the first half (device)
    hipEvent_t split_ev[3] = {};    // fill, scan of the first half; the second half's end
    int32_t pipe_lag = 0;           // farthest in-batch neighbour index ahead of a chunk
    vx::FusedWork *fused_dev = nullptr;
    void *fctl = nullptr;           // ticket | error | status[n]
    vx::FusedCtl fctl_ptrs{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool ran = false;
    bool results = false;           // host copies below are current
    uint64_t h_total = 0;
    std::vector<uint32_t> h_counts;
    std::vector<uint64_t> h_bases;
};

void vx_batch_unref(vx_batch *b);
// Frustum-culled indirect draw list over n meshes (callers hold the context lock).
int vx_draw_list_locked(vx_ctx *ctx, const vx_frustum *f, const int32_t *coords_dev,
                        const uint32_t *faces_dev, const uint64_t *first_dev, int32_t n,
                        int32_t edge, uint32_t mode, vx_draw_indexed *args_dev,
                        uint32_t *count_dev);
void vx_chunk_destroy_locked(vx_chunk *c);
void vx_shared_unref_locked(vx_ctx *ctx, vx::SharedBlock *s);
