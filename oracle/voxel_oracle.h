/*
 * voxel_oracle.h -- CPU restatement of the reference chunk pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: a plain-C restatement
 * of the reference's noise -> fill -> cull/emit -> frustum path.  Only the
 * test suite (tests/), __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it, and only as the checker / CPU baseline.  The product
 * library (libvoxel_hip.so) never links, loads or calls anything here.
 *
 * Reference files restated (paths relative to the reference repository):
 *   src/chunk.rs:27-46        generate_terrain (fill loop, y-z-x order)
 *   src/chunk.rs:59-245       cull_unseen_triangles (face test + order)
 *   src/cube.rs:16-21         CubeType
 *   src/texture.rs:238-243    TEXTURE_INDEX_LIST
 *   src/quad.rs:14-54         Quad::new / Quad::to_raw
 *   src/quad.rs:66-89         VERTICES / INDICES
 *   src/shaders/vertex.vert:24  instance * vec4(pos, 1)
 *   src/renderer/instance.rs:1-6, src/renderer/vertex.rs:1-6  record layouts
 *   src/frustum_culling.rs:16-85  contains / matrix_to_planes
 *   src/coordinate.rs:60-65   chunk -> world (x edge)
 *
 * Parity status per function (what pins it):
 *   vxo_noise3 / vxo_noise3d  Perlin's published Improved Noise (2002);
 *                             pinned by its known answer
 *                             noise(3.14,42,7)=0.13691995878400012 with
 *                             Perlin's own permutation (f64 instantiation of
 *                             the same source text), lattice zeros, and an
 *                             independent NumPy f32 implementation.
 *   vxo_noise2                2-D form (8 gradients); pinned by lattice zeros,
 *                             the near-lattice linearisation g.(dx,dy), and
 *                             the independent NumPy f32 implementation.
 *   vxo_perm                  splitmix64 + Fisher-Yates; pinned by
 *                             splitmix64's published first outputs.
 *   vxo_chunk_fill            config rules; pinned by closed forms (all-air,
 *                             all-solid, heightmap monotone columns).  The
 *                             reference rule (chunk.rs:38-42, noise axes of
 *                             perlin_noise.rs:6-13) is pinned against the
 *                             density rule: its noise block is the density
 *                             noise block with y and z transposed, bit for
 *                             bit, and a voxel is solid iff that value < 0.
 *   vxo_mesh_faces            chunk.rs:59-245; pinned by closed forms
 *                             (single voxel, all-solid with/without
 *                             neighbours, checkerboard) and an exhaustive
 *                             N=2 brute force written differently in Python.
 *   vxo_rotation/vxo_instance quad.rs:14-54 via nalgebra 0.29.0
 *                             Rotation3::from_euler_angles and Matrix4 mul.
 *                             Invariants pinned (orthonormal, det 1, face
 *                             normal = test direction); the exact bits of the
 *                             ~4.4e-8 terms and signed zeros against the Rust
 *                             binary are PARITY UNPINNED (nalgebra absent).
 *   vxo_vertices              vertex.vert:24 evaluated left to right without
 *                             FMA; pinned by the exact-corner invariant.
 *   vxo_frustum_*             frustum_culling.rs:16-85; pinned by closed-form
 *                             cameras and an f64 brute force away from ties.
 *
 * Build: gcc -O2 -ffp-contract=off (no fast-math).  See oracle/Makefile.
 */
#ifndef VOXEL_ORACLE_H
#define VOXEL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Block ids (u8).  The reference's CubeType is {GRASS=0, DIRT=1, AIR=2}
 * (cube.rs:16-21); the configs store 0 = air, 1 = grass, so ids are shifted:
 * TEXTURE_INDEX_LIST[cube_type] becomes table[id - 1]. */
enum { VXO_AIR = 0, VXO_GRASS = 1, VXO_DIRT = 2 };

/* Face directions in the reference's per-voxel emission order
 * (chunk.rs:77-240): LEFT(-x) RIGHT(+x) BACK(-z) FRONT(+z) DOWN(-y) UP(+y). */
enum { VXO_LEFT = 0, VXO_RIGHT = 1, VXO_BACK = 2, VXO_FRONT = 3, VXO_DOWN = 4, VXO_UP = 5 };

enum {
    VXO_TERRAIN_AIR = 0,       /* every voxel air */
    VXO_TERRAIN_SOLID = 1,     /* every voxel grass */
    VXO_TERRAIN_HEIGHTMAP = 2, /* 2-D fBm heightmap column rule */
    VXO_TERRAIN_DENSITY = 3,   /* 3-D fBm density rule */
    VXO_TERRAIN_REFERENCE = 4  /* the reference's rule: fbm3(x, z, y) < 0 */
};

typedef struct vxo_terrain {
    int32_t kind;
    uint32_t seed;
    int32_t octaves;
    float frequency;    /* base frequency f0 */
    float lacunarity;
    float gain;
    float height_scale; /* heightmap: h = floor(height_scale*(n+1)/2) */
    int32_t height_max; /* heightmap: h clamped to [0, height_max] */
    float y_scale;      /* density: solid iff fbm3 - wy*y_scale > 0 */
} vxo_terrain;

typedef struct vxo_face {
    int32_t x, y, z; /* voxel world position */
    uint32_t dir;    /* VXO_LEFT.. VXO_UP */
    uint32_t tex;    /* texture layer */
} vxo_face;

typedef struct vxo_instance { /* InstanceRaw, renderer/instance.rs:1-6, 68 B */
    float m[4][4];            /* column-major: m[col][row] */
    uint32_t tex;
} vxo_instance;

typedef struct vxo_vertex { /* Vertex, renderer/vertex.rs:1-6, 20 B */
    float pos[3];
    float uv[2];
} vxo_vertex;

/* ---- noise ------------------------------------------------------------ */
uint64_t vxo_splitmix64_next(uint64_t *state);
void vxo_perm(uint64_t seed, uint8_t perm[512]);
void vxo_perm_from_table(const uint8_t table[256], uint8_t perm[512]);
float vxo_noise3(const uint8_t *perm, float x, float y, float z);
double vxo_noise3d(const uint8_t *perm, double x, double y, double z);
float vxo_noise2(const uint8_t *perm, float x, float y);
float vxo_fbm3(const uint8_t *perm, float x, float y, float z, int octaves, float f0,
               float lacunarity, float gain);
float vxo_fbm2(const uint8_t *perm, float x, float y, int octaves, float f0, float lacunarity,
               float gain);
/* 8 consecutive x samples of vxo_fbm3 at (x0+i, y, z), i = 0..7: AVX2 when the
 * host supports it, otherwise scalar; bit-identical to vxo_fbm3 either way. */
void vxo_fbm3_row8(const uint8_t *perm, int32_t x0, int32_t y, int32_t z, int octaves, float f0,
                   float lacunarity, float gain, float out[8]);
int vxo_have_avx2(void);

/* ---- chunk fill (chunk.rs:27-46) -------------------------------------- */
/* ids: edge^3 bytes, index x + edge*z + edge*edge*y. */
void vxo_chunk_fill(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz, int32_t edge,
                    uint8_t *ids);
/* Same result; noise rows evaluated 8-wide (CPU baseline path). */
void vxo_chunk_fill_fast(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz,
                         int32_t edge, uint8_t *ids);
/* Raw noise the fill rule thresholds: density -> edge^3 values of fbm3 at the
 * voxel index; heightmap -> edge^2 values of fbm2 at x + edge*z. */
void vxo_noise_block(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz, int32_t edge,
                     float *out);
/* Occupancy rows: words_per_row = (edge+31)/32 u32 per (y,z) row, row index
 * y*edge + z, bit x%32 of word x/32 set iff voxel solid. */
void vxo_occupancy(const uint8_t *ids, int32_t edge, uint32_t *occ);

/* ---- cull + emit (chunk.rs:59-245) ------------------------------------ */
/* adj[0..3] = neighbour ids at -x, +x, -z, +z (NULL = absent).  Writes up to
 * cap faces, returns the total number of visible faces. */
int64_t vxo_mesh_faces(const uint8_t *ids, int32_t edge, int32_t ox, int32_t oy, int32_t oz,
                       const uint8_t *const adj[4], vxo_face *out, int64_t cap);
uint32_t vxo_texture_layer(uint8_t block, uint32_t dir);

/* ---- Quad -> InstanceRaw -> vertices/indices -------------------------- */
void vxo_rotation(uint32_t dir, float r[3][3]);
void vxo_instance_of(const vxo_face *f, vxo_instance *out);
void vxo_vertices_of(const vxo_instance *inst, vxo_vertex out[4]);
void vxo_indices_of(int64_t local_face, uint32_t out[6]);
/* Expands n faces into any of: InstanceRaw records, 4 vertices per face,
 * 6 chunk-local indices per face (NULL outputs are skipped). */
void vxo_expand(const vxo_face *faces, int64_t n, vxo_instance *inst, vxo_vertex *vert,
                uint32_t *idx);

/* ---- frustum (frustum_culling.rs:16-85) ------------------------------- */
/* m: 16 floats column-major (m[c*4 + r]); planes: [near far left right top
 * bottom] x (a b c d). */
void vxo_frustum_planes(const float m[16], float planes[24]);
int vxo_frustum_contains(const float planes[24], int32_t cx, int32_t cy, int32_t cz,
                         int32_t edge);

/* ---- multithreaded CPU driver (CPU baseline) -------------------------- */
/* Chunk i at coords[3i..3i+2], edge `edge`; neighbours nb[4i+k] = index of
 * the k-th neighbour (-x,+x,-z,+z) in this batch or -1.  Phase 1 fills every
 * chunk, phase 2 meshes every chunk into vertices+indices (4 and 6 per face),
 * one chunk per task on `threads` threads.  Returns the total face count and
 * writes per-chunk counts to faces_out (may be NULL).  Outputs are
 * discarded after being written (the baseline measures the work, not the
 * keeping). */
int64_t vxo_batch_run(const vxo_terrain *t, const int32_t *coords, const int32_t *nb, int32_t n,
                      int32_t edge, int threads, int64_t *faces_out);
/* The same batch, keeping the meshes: chunk i's 4 vertices and 6 chunk-local
   indices per face written from face bases[i] of vert_out / idx_out, which
   hold cap_faces faces.  Returns the total faces, or -1 when a chunk's range
   falls outside [0, cap_faces) (nothing is written for that chunk). */
int64_t vxo_batch_mesh(const vxo_terrain *t, const int32_t *coords, const int32_t *nb, int32_t n,
                       int32_t edge, int threads, const int64_t *bases, vxo_vertex *vert_out,
                       uint32_t *idx_out, int64_t cap_faces, int64_t *faces_out);

#ifdef __cplusplus
}
#endif
#endif
