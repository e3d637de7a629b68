// vx_kernels.hip -- gfx950 kernels of the chunk pipeline.
//
//   fill    noise -> u8 block ids, packed occupancy rows, boundary planes
//           (src/perlin_noise.rs:5-22 + src/chunk.rs:27-46, config noise)
//   count   visible faces per chunk from occupancy bitmasks
//           (src/chunk.rs:59-245, counting only)
//   scan    device-wide exclusive scan of the per-chunk face counts
//   emit    faces in the reference's order -> vertex / index / InstanceRaw
//           buffers (src/chunk.rs:59-245, src/quad.rs:14-89,
//           src/shaders/vertex.vert:24)
//   frustum AABB-vs-6-planes cull mask (src/frustum_culling.rs:16-40)
//
// Exactness: every floating-point expression is the oracle's expression in
// the same order, compiled with -ffp-contract=off.  The only explicit FMAs
// are fma(a, b, c) where a*b is exact (b in {+-1, +-0}), which round to
// exactly the oracle's separate multiply and add.
#include "vx_internal.h"

namespace vx {

// ---------------------------------------------------------------- helpers
template <int N> struct RowType { using T = uint32_t; };
template <> struct RowType<64> { using T = uint64_t; };

__device__ __forceinline__ float fade(float t) {
    return ((t * t) * t) * (t * (t * 6.0f - 15.0f) + 10.0f);
}
__device__ __forceinline__ float lerp(float t, float a, float b) { return a + t * (b - a); }

__device__ __forceinline__ int popc(uint32_t v) { return __popc(v); }
__device__ __forceinline__ int popc(uint64_t v) { return __popcll(v); }
__device__ __forceinline__ int ctz(uint32_t v) { return __ffs(v) - 1; }
__device__ __forceinline__ int ctz(uint64_t v) { return __ffsll((unsigned long long)v) - 1; }

// Exclusive scan of one u32 per thread over a kBlock-thread workgroup.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave,
                                                         uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, sum = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) {
        uint32_t s = s_wave[i];
        if (i < wave) wbase += s;
        sum += s;
    }
    __syncthreads();
    *total = sum;
    return wbase + incl - v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *s_wave) {
    uint32_t total;
    block_exclusive_scan(v, s_wave, &total);
    return total;
}

// 2-D improved-noise gradient: x+y, -x+y, x-y, -x-y, x, -x, y, -y.
__device__ __forceinline__ float grad2(int hash, float x, float y) {
    const int h = hash & 7;
    const float sx = (h & 1) ? -x : x;
    if (h < 4) return sx + ((h & 2) ? -y : y);
    if (h < 6) return sx;
    return (h & 1) ? -y : y;
}

__device__ float noise2(const int *p, float x, float y) {
    const float fx = floorf(x), fy = floorf(y);
    const int X = (int)fx & 255, Y = (int)fy & 255;
    x = x - fx;
    y = y - fy;
    const float u = fade(x), v = fade(y);
    const int A = p[X] + Y, B = p[X + 1] + Y;
    const float x1 = x - 1.0f, y1 = y - 1.0f;
    return lerp(v, lerp(u, grad2(p[A], x, y), grad2(p[B], x1, y)),
                lerp(u, grad2(p[A + 1], x, y1), grad2(p[B + 1], x1, y1)));
}

// --------------------------------------------------------------- fill
// LDS layout of one fill workgroup, sized for the largest edge it handles.
template <int MAXN> struct FillShared {
    int perm[512];
    float4 lut[8][16];
    float4 ytab[kMaxOctaves * MAXN];   // per (octave, y): fy, fy-1, fade(fy), Y cell
    uint32_t occ[MAXN * MAXN * (MAXN == 64 ? 2 : 1)];
    uint32_t red[kBlock / 64];
};

// Writes the occupancy row(s) produced by one ballot: for N <= 32 the wave
// covers 64/N consecutive z rows at one y; for N = 64 one row of 2 words.
template <int N>
__device__ __forceinline__ void store_ballot_rows(uint32_t *s_occ, uint64_t b, int y, int z0,
                                                  int lane) {
    if (N == 64) {
        if (lane < 2) s_occ[(y * N + z0) * 2 + lane] = lane ? (uint32_t)(b >> 32) : (uint32_t)b;
    } else {
        constexpr int ZPW = 64 / N;
        constexpr uint32_t MASK = N >= 32 ? 0xFFFFFFFFu : ((1u << (N & 31)) - 1u);
        if (lane < ZPW) s_occ[y * N + z0 + lane] = (uint32_t)(b >> (lane * N)) & MASK;
    }
}

template <int N, int MAXN>
__device__ void fill_rows_density(const FillWork &w, const TerrainDev &T, FillShared<MAXN> &S) {
    constexpr int ZPW = 64 / N;
    constexpr int PASSES = N * N / kBlock;
    constexpr int SEG = N < 32 ? N : 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = lane % N;
    const int ox = w.cx * N, oy = w.cy * N, oz = w.cz * N;
    const int octaves = T.octaves;
    for (int p = 0; p < PASSES; p++) {
        const int z0 = (p * (kBlock / 64) + wave) * ZPW;
        const int z = z0 + lane / N;
        for (int y0 = 0; y0 < N; y0 += SEG) {
            float sum[SEG];
#pragma unroll
            for (int i = 0; i < SEG; i++) sum[i] = 0.0f;
            float f = T.frequency, amp = 1.0f;
            for (int o = 0; o < octaves; o++) {
                const float X = (float)(ox + x) * f;
                const float fxl = floorf(X);
                const float fx = X - fxl, fx1 = fx - 1.0f, u = fade(fx);
                const int Xi = (int)fxl & 255;
                const float Z = (float)(oz + z) * f;
                const float fzl = floorf(Z);
                const float fz = Z - fzl, fz1 = fz - 1.0f, wz = fade(fz);
                const int Zi = (int)fzl & 255;
                const int pX0 = S.perm[Xi], pX1 = S.perm[Xi + 1];
                float c[8], gy[8];
                int yprev = -1;
                const float4 *yt = S.ytab + o * N + y0;
                // Runtime loop on purpose: sum[] stays in VGPRs through uniform
                // register indexing, while a fully unrolled loop with the
                // conditional rehash makes the compiler keep every path live.
#pragma unroll 1
                for (int yy = 0; yy < SEG; yy++) {
                    const float4 t = yt[yy];
                    const int Yi = __builtin_amdgcn_readfirstlane(__float_as_int(t.w));
                    if (Yi != yprev) {  // new lattice cell along y: rehash 8 corners
                        yprev = Yi;
                        const int A = pX0 + Yi, B = pX1 + Yi;
                        const int AA = S.perm[A] + Zi, AB = S.perm[A + 1] + Zi;
                        const int BA = S.perm[B] + Zi, BB = S.perm[B + 1] + Zi;
                        const int h[8] = {S.perm[AA],     S.perm[BA],     S.perm[AB],
                                          S.perm[BB],     S.perm[AA + 1], S.perm[BA + 1],
                                          S.perm[AB + 1], S.perm[BB + 1]};
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            const float4 L = S.lut[k][h[k] & 15];
                            const float xr = (k & 1) ? fx1 : fx;
                            const float zr = (k & 4) ? fz1 : fz;
                            c[k] = fmaf(L.x, xr, L.y * zr);
                            gy[k] = L.z;
                        }
                    }
                    const float g0 = fmaf(t.x, gy[0], c[0]), g1 = fmaf(t.x, gy[1], c[1]);
                    const float g2 = fmaf(t.y, gy[2], c[2]), g3 = fmaf(t.y, gy[3], c[3]);
                    const float g4 = fmaf(t.x, gy[4], c[4]), g5 = fmaf(t.x, gy[5], c[5]);
                    const float g6 = fmaf(t.y, gy[6], c[6]), g7 = fmaf(t.y, gy[7], c[7]);
                    const float n = lerp(wz, lerp(t.z, lerp(u, g0, g1), lerp(u, g2, g3)),
                                         lerp(t.z, lerp(u, g4, g5), lerp(u, g6, g7)));
                    sum[yy] = sum[yy] + amp * n;
                }
                f = f * T.lacunarity;
                amp = amp * T.gain;
            }
#pragma unroll
            for (int yy = 0; yy < SEG; yy++) {
                const int y = y0 + yy;
                const float thr = (float)(oy + y) * T.y_scale;
                const bool solid = (sum[yy] - thr) > 0.0f;
                store_ballot_rows<N>(S.occ, __ballot(solid), y, z0, lane);
                if (w.noise_out) w.noise_out[(size_t)x + (size_t)N * z + (size_t)N * N * y] = sum[yy];
            }
        }
    }
}

template <int N, int MAXN>
__device__ void fill_rows_heightmap(const FillWork &w, const TerrainDev &T, FillShared<MAXN> &S) {
    constexpr int ZPW = 64 / N;
    constexpr int PASSES = N * N / kBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = lane % N;
    const int ox = w.cx * N, oy = w.cy * N, oz = w.cz * N;
    for (int p = 0; p < PASSES; p++) {
        const int z0 = (p * (kBlock / 64) + wave) * ZPW;
        const int z = z0 + lane / N;
        const float wx = (float)(ox + x), wzf = (float)(oz + z);
        float sum = 0.0f, amp = 1.0f, f = T.frequency;
        for (int o = 0; o < T.octaves; o++) {
            sum = sum + amp * noise2(S.perm, wx * f, wzf * f);
            f = f * T.lacunarity;
            amp = amp * T.gain;
        }
        const float hf = floorf((T.height_scale * (sum + 1.0f)) / 2.0f);
        int h;
        if (!(hf >= 0.0f)) h = 0;
        else if (hf > (float)T.height_max) h = T.height_max;
        else h = (int)hf;
        if (w.noise_out) w.noise_out[(size_t)x + (size_t)N * z] = sum;
        for (int y = 0; y < N; y++) store_ballot_rows<N>(S.occ, __ballot(oy + y <= h), y, z0, lane);
    }
}

// Writes a chunk whose occupancy rows are complete in S.occ: occupancy and
// boundary planes always, u8 ids (all solid voxels = kBlockGrass) when
// write_ids, and meta = {solid block id, solid voxel count}.
template <int N, int MAXN>
__device__ void write_chunk_outputs(const FillWork &w, FillShared<MAXN> &S, bool write_ids,
                                    uint32_t solid_type) {
    constexpr int W = N == 64 ? 2 : 1;
    constexpr int ROWS = N * N;
    const int tid = threadIdx.x;
    uint32_t solid = 0;
    for (int i = tid; i < ROWS * W; i += kBlock) {
        const uint32_t v = S.occ[i];
        w.occ[i] = v;
        solid += __popc(v);
    }
    // boundary planes: [0] x=0, [1] x=N-1 (bit z per y); [2] z=0, [3] z=N-1 (row bits)
    for (int i = tid; i < 4 * N * W; i += kBlock) {
        const int pl = i / (N * W), rem = i - pl * N * W, y = rem / W, j = rem - y * W;
        uint32_t v = 0;
        if (pl < 2) {
            const int xb = pl == 0 ? 0 : N - 1;
            const int zlim = N < 32 ? N : 32;
            for (int zz = 0; zz < zlim; zz++) {
                const int z = j * 32 + zz;
                const uint32_t word = S.occ[(y * N + z) * W + (xb >> 5)];
                v |= ((word >> (xb & 31)) & 1u) << zz;
            }
        } else {
            const int z = pl == 2 ? 0 : N - 1;
            v = S.occ[(y * N + z) * W + j];
        }
        w.planes[i] = v;
    }
    if (write_ids) {
        // u8 block ids, 16 voxels per 16-byte store
        constexpr int PER_ROW = N / 16;
        uint4 *ids4 = reinterpret_cast<uint4 *>(w.ids);
        for (int q = tid; q < ROWS * PER_ROW; q += kBlock) {
            const int r = q / PER_ROW, part = q - r * PER_ROW;
            const uint32_t word = S.occ[r * W + (part >> 1)];
            const uint32_t bits = (word >> ((part & 1) * 16)) & 0xFFFFu;
            uint32_t o[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t nib = (bits >> (4 * k)) & 0xFu;
                o[k] = ((nib * 0x00204081u) & 0x01010101u) * kBlockGrass;
            }
            ids4[q] = make_uint4(o[0], o[1], o[2], o[3]);
        }
    }
    const uint32_t total = block_sum(solid, S.red);
    if (tid == 0) {
        w.meta[0] = solid_type;
        w.meta[1] = total;
    }
}

template <int N, int MAXN>
__device__ void fill_chunk(const FillWork &w, const TerrainDev &T, FillShared<MAXN> &S) {
    constexpr int W = N == 64 ? 2 : 1;
    constexpr int ROWS = N * N;
    const int tid = threadIdx.x;
    if (T.kind == kDensity) {
        for (int i = tid; i < T.octaves * N; i += kBlock) {
            const int o = i / N, y = i - o * N;
            float f = T.frequency;
            for (int k = 0; k < o; k++) f = f * T.lacunarity;
            const float Y = (float)(w.cy * N + y) * f;
            const float fyl = floorf(Y);
            const float fy = Y - fyl;
            S.ytab[i] = make_float4(fy, fy - 1.0f, fade(fy), __int_as_float((int)fyl & 255));
        }
        __syncthreads();
        fill_rows_density<N, MAXN>(w, T, S);
    } else if (T.kind == kHeightmap) {
        fill_rows_heightmap<N, MAXN>(w, T, S);
    } else {
        const uint32_t full = T.kind == kSolid ? (N >= 32 ? 0xFFFFFFFFu : ((1u << (N & 31)) - 1u)) : 0u;
        for (int i = tid; i < ROWS * W; i += kBlock) S.occ[i] = full;
    }
    __syncthreads();
    write_chunk_outputs<N, MAXN>(w, S, true, kBlockGrass);
}

// Packs caller-provided u8 ids (already in w.ids) into occupancy rows and
// boundary planes; meta[0] = the single solid block id, or kBlockMixed.
template <int N, int MAXN>
__device__ void pack_chunk(const FillWork &w, FillShared<MAXN> &S) {
    constexpr int W = N == 64 ? 2 : 1;
    __shared__ uint32_t s_types;
    if (threadIdx.x == 0) s_types = 0;
    __syncthreads();
    uint32_t types = 0;
    for (int i = threadIdx.x; i < N * N * W; i += kBlock) {
        const int r = i / W, j = i - r * W;
        const int xs = j * 32, xe = min(N, xs + 32);
        const uint8_t *row = w.ids + (size_t)r * N;  // row r = y*N + z -> ids[x + N*z + N*N*y]
        uint32_t bits = 0;
        for (int x = xs; x < xe; x++) {
            const uint32_t b = row[x];
            if (b != kBlockAir) bits |= 1u << (x - xs);
            types |= 1u << (b & 7);
        }
        S.occ[i] = bits;
    }
    atomicOr(&s_types, types);
    __syncthreads();
    const uint32_t solid_types = s_types & ~1u;
    uint32_t st = kBlockMixed;
    if (solid_types == 0) st = kBlockGrass;  // no solid voxel: type unused
    else if ((solid_types & (solid_types - 1)) == 0) st = (uint32_t)(__ffs(solid_types) - 1);
    write_chunk_outputs<N, MAXN>(w, S, false, st);
}

template <int MAXN>
__device__ __forceinline__ void fill_item(const FillWork &w, const TerrainDev *terrains,
                                          const GradLut *lut, FillShared<MAXN> &S) {
    const TerrainDev &T = terrains[w.terrain];
    for (int i = threadIdx.x; i < 512; i += kBlock) S.perm[i] = T.perm[i];
    for (int i = threadIdx.x; i < 8 * 16; i += kBlock) S.lut[i >> 4][i & 15] = lut->c[i >> 4][i & 15];
    __syncthreads();
    switch (w.edge) {
    case 16: fill_chunk<16, MAXN>(w, T, S); break;
    case 32: if (MAXN >= 32) fill_chunk<(MAXN >= 32 ? 32 : 16), MAXN>(w, T, S); break;
    case 64: if (MAXN >= 64) fill_chunk<(MAXN >= 64 ? 64 : 16), MAXN>(w, T, S); break;
    default: break;
    }
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) fill_kernel(const FillWork *__restrict__ work,
                                                      const int32_t *__restrict__ order,
                                                      const TerrainDev *__restrict__ terrains,
                                                      const GradLut *__restrict__ lut) {
    __shared__ FillShared<MAXN> S;
    const int item = order ? order[blockIdx.x] : (int)blockIdx.x;
    fill_item<MAXN>(work[item], terrains, lut, S);
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) fill_one_kernel(FillWork w,
                                                          const TerrainDev *__restrict__ terrains,
                                                          const GradLut *__restrict__ lut) {
    __shared__ FillShared<MAXN> S;
    fill_item<MAXN>(w, terrains, lut, S);
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) pack_one_kernel(FillWork w) {
    __shared__ FillShared<MAXN> S;
    switch (w.edge) {
    case 16: pack_chunk<16, MAXN>(w, S); break;
    case 32: if (MAXN >= 32) pack_chunk<(MAXN >= 32 ? 32 : 16), MAXN>(w, S); break;
    case 64: if (MAXN >= 64) pack_chunk<(MAXN >= 64 ? 64 : 16), MAXN>(w, S); break;
    default: break;
    }
}

// --------------------------------------------------------------- mesh
template <int MAXN> struct MeshShared {
    typename RowType<MAXN>::T occ[MAXN * MAXN];
    typename RowType<MAXN>::T nb[4][MAXN];
    uint32_t desc[kDescTile];
    float P[6][2][4][3];
    float Mb[6][2][3][3];
    uint32_t tex[3][6];
    uint32_t red[kBlock / 64];
};

template <int N, int MAXN>
__device__ __forceinline__ void load_rows(const MeshWork &w, MeshShared<MAXN> &S) {
    using R = typename RowType<MAXN>::T;
    constexpr int W = N == 64 ? 2 : 1;
    const int tid = threadIdx.x;
    for (int r = tid; r < N * N; r += kBlock) {
        R v;
        if (W == 2) v = (R)w.occ[2 * r] | ((R)w.occ[2 * r + 1] << (W == 2 ? 32 : 0));
        else v = (R)w.occ[r];
        S.occ[r] = v;
    }
    for (int i = tid; i < 4 * N; i += kBlock) {
        const int k = i / N, y = i - k * N;
        R v = 0;
        if (w.nb[k]) {
            if (W == 2) v = (R)w.nb[k][2 * y] | ((R)w.nb[k][2 * y + 1] << (W == 2 ? 32 : 0));
            else v = (R)w.nb[k][y];
        }
        S.nb[k][y] = v;
    }
}

// Six face masks of row (y, z), bit x set iff voxel x emits that face
// (chunk.rs:77-240): neighbour cell air, or outside the chunk with the
// neighbour chunk absent, or outside vertically.
template <int N, int MAXN>
__device__ __forceinline__ void face_masks(const MeshShared<MAXN> &S, int y, int z,
                                           typename RowType<MAXN>::T m[6]) {
    using R = typename RowType<MAXN>::T;
    const R self = S.occ[y * N + z];
    const R lb = (S.nb[0][y] >> z) & (R)1;
    const R rb = (S.nb[1][y] >> z) & (R)1;
    m[0] = self & ~((self << 1) | lb);
    m[1] = self & ~((self >> 1) | (rb << (N - 1)));
    const R back = z > 0 ? S.occ[y * N + z - 1] : S.nb[2][y];
    const R front = z < N - 1 ? S.occ[y * N + z + 1] : S.nb[3][y];
    m[2] = self & ~back;
    m[3] = self & ~front;
    const R down = y > 0 ? S.occ[(y - 1) * N + z] : (R)0;
    const R up = y < N - 1 ? S.occ[(y + 1) * N + z] : (R)0;
    m[4] = self & ~down;
    m[5] = self & ~up;
}

template <int N, int MAXN>
__device__ void count_chunk(const MeshWork &w, MeshShared<MAXN> &S, uint32_t *count_out) {
    using R = typename RowType<MAXN>::T;
    load_rows<N, MAXN>(w, S);
    __syncthreads();
    uint32_t c = 0;
    for (int r = threadIdx.x; r < N * N; r += kBlock) {
        R m[6];
        face_masks<N, MAXN>(S, r / N, r % N, m);
#pragma unroll
        for (int d = 0; d < 6; d++) c += popc(m[d]);
    }
    const uint32_t total = block_sum(c, S.red);
    if (threadIdx.x == 0) *count_out = total;
}

template <int MAXN>
__device__ __forceinline__ void count_item(const MeshWork &w, MeshShared<MAXN> &S, uint32_t *out) {
    switch (w.edge) {
    case 16: count_chunk<16, MAXN>(w, S, out); break;
    case 32: if (MAXN >= 32) count_chunk<(MAXN >= 32 ? 32 : 16), MAXN>(w, S, out); break;
    case 64: if (MAXN >= 64) count_chunk<(MAXN >= 64 ? 64 : 16), MAXN>(w, S, out); break;
    default: break;
    }
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) count_kernel(const MeshWork *__restrict__ work,
                                                       uint32_t *__restrict__ counts) {
    __shared__ MeshShared<MAXN> S;
    count_item<MAXN>(work[blockIdx.x], S, counts + blockIdx.x);
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) count_one_kernel(MeshWork w, uint32_t *__restrict__ count) {
    __shared__ MeshShared<MAXN> S;
    count_item<MAXN>(w, S, count);
}

// Device-wide exclusive scan of per-chunk face counts (one workgroup of
// 1024 threads, each owning a contiguous run of chunks).
__global__ void __launch_bounds__(1024) scan_kernel(const uint32_t *__restrict__ counts, int32_t n,
                                                    uint64_t *__restrict__ bases,
                                                    uint64_t *__restrict__ total,
                                                    uint32_t *__restrict__ overflow) {
    __shared__ uint64_t s_wave[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (n + 1023) / 1024;
    const int b = tid * per, e = min(n, b + per);
    uint64_t local = 0;
    for (int i = b; i < e; i++) local += counts[i];
    uint64_t incl = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t wbase = 0, sum = 0;
    for (int i = 0; i < 16; i++) {
        if (i < wave) wbase += s_wave[i];
        sum += s_wave[i];
    }
    uint64_t run = wbase + incl - local;
    for (int i = b; i < e; i++) {
        bases[i] = run;
        run += counts[i];
    }
    if (tid == 0) {
        *total = sum;
        if (overflow) *overflow = 0u;  // the emit stage that follows sets it
    }
}

// Emits one chunk's faces.  Each thread owns RPT consecutive rows; a block
// scan of the per-thread counts gives every row its first face index, so the
// face order is (y, z, x, direction) exactly as chunk.rs:65-240 pushes them.
// Faces are staged as 32-bit descriptors (row | x << 12 | dir << 18) in LDS
// tiles of kDescTile, then one lane per face writes its records.
template <int N, int MAXN>
__device__ void emit_chunk(const MeshWork &w, uint64_t base, uint32_t F, const MeshOut &out,
                           MeshShared<MAXN> &S) {
    using R = typename RowType<MAXN>::T;
    constexpr int RPT = (N * N) / kBlock;
    const int tid = threadIdx.x;
    load_rows<N, MAXN>(w, S);
    __syncthreads();
    const int r0 = tid * RPT;
    uint32_t rc[RPT];
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        R m[6];
        const int r = r0 + i;
        face_masks<N, MAXN>(S, r / N, r % N, m);
        uint32_t c = 0;
#pragma unroll
        for (int d = 0; d < 6; d++) c += popc(m[d]);
        rc[i] = c;
        mine += c;
    }
    uint32_t total;
    const uint32_t first = block_exclusive_scan(mine, S.red, &total);
    (void)total;  // equals F (count kernel); F is authoritative for sizing

    const bool want_v = (out.formats & kFmtVertices) != 0;
    const bool want_i = (out.formats & kFmtInstances) != 0;
    for (uint32_t t0 = 0; t0 < F; t0 += kDescTile) {
        const uint32_t t1 = min(F, t0 + (uint32_t)kDescTile);
        uint32_t f = first;
        if (f < t1 && f + mine > t0) {
#pragma unroll 1
            for (int i = 0; i < RPT; i++) {
                const int r = r0 + i;
                if (f + rc[i] > t0 && f < t1) {
                    R m[6];
                    face_masks<N, MAXN>(S, r / N, r % N, m);
                    R any = m[0] | m[1] | m[2] | m[3] | m[4] | m[5];
                    uint32_t g = f;
                    while (any && g < t1) {
                        const int x = ctz(any);
                        any &= any - 1;
#pragma unroll
                        for (int d = 0; d < 6; d++) {
                            if ((m[d] >> x) & 1) {
                                if (g >= t0 && g < t1)
                                    S.desc[g - t0] = (uint32_t)r | ((uint32_t)x << 12) | ((uint32_t)d << 18);
                                g++;
                            }
                        }
                    }
                }
                f += rc[i];
            }
        }
        __syncthreads();
        for (uint32_t q = tid; q < t1 - t0; q += kBlock) {
            const uint32_t dsc = S.desc[q];
            const int r = dsc & 0xFFF, x = (dsc >> 12) & 0x3F, d = (dsc >> 18) & 7;
            const int y = r / N, z = r - y * N;
            const int wx = w.ox + x, wy = w.oy + y, wz = w.oz + z;
            const float t[3] = {(float)wx, (float)wy, (float)wz};
            const int sg[3] = {wx < 0, wy < 0, wz < 0};
            uint32_t blk = w.solid_type;
            if (blk == kBlockMixed) blk = w.ids[(size_t)x + (size_t)N * z + (size_t)N * N * y];
            const uint32_t tex = S.tex[blk][d];
            const uint64_t face = base + t0 + q;
            if (want_v) {
                float v[20];
#pragma unroll
                for (int c = 0; c < 4; c++) {
#pragma unroll
                    for (int k = 0; k < 3; k++) v[c * 5 + k] = S.P[d][sg[k]][c][k] + t[k];
                    v[c * 5 + 3] = (c & 1) ? 1.0f : 0.0f;
                    v[c * 5 + 4] = (c & 2) ? 1.0f : 0.0f;
                }
                float4 *dst = reinterpret_cast<float4 *>(out.vertices + face * 20);
#pragma unroll
                for (int k = 0; k < 5; k++)
                    dst[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
            }
            if (want_i) {
                uint32_t *dst = out.instances + face * 17;
#pragma unroll
                for (int c = 0; c < 3; c++) {
#pragma unroll
                    for (int k = 0; k < 3; k++) dst[c * 4 + k] = __float_as_uint(S.Mb[d][sg[k]][k][c]);
                    dst[c * 4 + 3] = 0u;  // M(3,c) = +0
                }
#pragma unroll
                for (int k = 0; k < 3; k++) dst[12 + k] = __float_as_uint(t[k]);
                dst[15] = 0x3F800000u;  // 1.0f
                dst[16] = tex;
            }
        }
        __syncthreads();
    }
    if (out.formats & kFmtVertices) {
        // quad::INDICES = [0,2,1,3,1,2] + 4*local face; two indices per store
        uint2 *dst = reinterpret_cast<uint2 *>(out.indices + base * 6);
        for (uint32_t q = tid; q < 3 * F; q += kBlock) {
            const uint32_t face = q / 3, part = q - face * 3;
            const uint32_t b4 = 4 * face;
            uint2 v;
            if (part == 0) v = make_uint2(b4 + 0, b4 + 2);
            else if (part == 1) v = make_uint2(b4 + 1, b4 + 3);
            else v = make_uint2(b4 + 1, b4 + 2);
            dst[q] = v;
        }
    }
}

template <int MAXN>
__device__ __forceinline__ void emit_item(const MeshWork &w, uint64_t base, uint32_t F,
                                          const MeshOut &out, const MeshTables *tables,
                                          uint32_t *overflow, MeshShared<MAXN> &S) {
    if (base + F > out.capacity) {
        if (threadIdx.x == 0) atomicOr(overflow, 1u);
        return;
    }
    for (int i = threadIdx.x; i < 6 * 2 * 4 * 3; i += kBlock) (&S.P[0][0][0][0])[i] = (&tables->P[0][0][0][0])[i];
    for (int i = threadIdx.x; i < 6 * 2 * 3 * 3; i += kBlock) (&S.Mb[0][0][0][0])[i] = (&tables->Mb[0][0][0][0])[i];
    for (int i = threadIdx.x; i < 18; i += kBlock) (&S.tex[0][0])[i] = (&tables->tex[0][0])[i];
    switch (w.edge) {
    case 16: emit_chunk<16, MAXN>(w, base, F, out, S); break;
    case 32: if (MAXN >= 32) emit_chunk<(MAXN >= 32 ? 32 : 16), MAXN>(w, base, F, out, S); break;
    case 64: if (MAXN >= 64) emit_chunk<(MAXN >= 64 ? 64 : 16), MAXN>(w, base, F, out, S); break;
    default: break;
    }
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) emit_kernel(const MeshWork *__restrict__ work,
                                                      const int32_t *__restrict__ order,
                                                      const uint64_t *__restrict__ bases,
                                                      const uint32_t *__restrict__ counts,
                                                      MeshOut out,
                                                      const MeshTables *__restrict__ tables,
                                                      uint32_t *__restrict__ overflow) {
    __shared__ MeshShared<MAXN> S;
    const int item = order ? order[blockIdx.x] : (int)blockIdx.x;
    emit_item<MAXN>(work[item], bases[item], counts[item], out, tables, overflow, S);
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) emit_one_kernel(MeshWork w, uint64_t base, uint32_t F,
                                                          MeshOut out,
                                                          const MeshTables *__restrict__ tables,
                                                          uint32_t *__restrict__ overflow) {
    __shared__ MeshShared<MAXN> S;
    emit_item<MAXN>(w, base, F, out, tables, overflow, S);
}

// --------------------------------------------------------------- frustum
// One thread per chunk AABB: some corner inside all six planes -> 1
// (frustum_culling.rs:16-40; dist = ((a*x + b*y) + c*z) + d, NaN passes).
__global__ void __launch_bounds__(kBlock) frustum_kernel(Planes planes,
                                                         const int32_t *__restrict__ coords,
                                                         int32_t n, int32_t edge,
                                                         uint8_t *__restrict__ mask) {
    const float *pl = planes.p;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float px = (float)(coords[3 * i] * edge), py = (float)(coords[3 * i + 1] * edge),
                pz = (float)(coords[3 * i + 2] * edge);
    const float e = (float)edge;
    const float xs[2] = {px, px + e}, ys[2] = {py, py + e}, zs[2] = {pz, pz + e};
    uint32_t any = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const float cx = xs[c & 1], cy = ys[(c >> 2) & 1], cz = zs[(c >> 1) & 1];
        uint32_t inside = 1;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const float dist = pl[4 * k] * cx + pl[4 * k + 1] * cy + pl[4 * k + 2] * cz + pl[4 * k + 3];
            inside &= !(dist < 0.0f);
        }
        any |= inside;
    }
    mask[i] = (uint8_t)any;
}

// --------------------------------------------------------------- launchers
hipError_t launch_fill(const FillWork *work, const int32_t *order, int32_t n, int32_t max_edge,
                       const TerrainDev *terrains, const GradLut *lut, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_edge <= 16) fill_kernel<16><<<n, kBlock, 0, s>>>(work, order, terrains, lut);
    else if (max_edge <= 32) fill_kernel<32><<<n, kBlock, 0, s>>>(work, order, terrains, lut);
    else fill_kernel<64><<<n, kBlock, 0, s>>>(work, order, terrains, lut);
    return hipGetLastError();
}

hipError_t launch_count(const MeshWork *work, int32_t n, int32_t max_edge, uint32_t *counts,
                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_edge <= 16) count_kernel<16><<<n, kBlock, 0, s>>>(work, counts);
    else if (max_edge <= 32) count_kernel<32><<<n, kBlock, 0, s>>>(work, counts);
    else count_kernel<64><<<n, kBlock, 0, s>>>(work, counts);
    return hipGetLastError();
}

hipError_t launch_scan(const uint32_t *counts, int32_t n, uint64_t *bases, uint64_t *total,
                       uint32_t *overflow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    scan_kernel<<<1, 1024, 0, s>>>(counts, n, bases, total, overflow);
    return hipGetLastError();
}

hipError_t launch_fill_one(const FillWork &w, const TerrainDev *terrains, const GradLut *lut,
                           hipStream_t s) {
    if (w.edge <= 16) fill_one_kernel<16><<<1, kBlock, 0, s>>>(w, terrains, lut);
    else if (w.edge <= 32) fill_one_kernel<32><<<1, kBlock, 0, s>>>(w, terrains, lut);
    else fill_one_kernel<64><<<1, kBlock, 0, s>>>(w, terrains, lut);
    return hipGetLastError();
}

hipError_t launch_pack_one(const FillWork &w, hipStream_t s) {
    if (w.edge <= 16) pack_one_kernel<16><<<1, kBlock, 0, s>>>(w);
    else if (w.edge <= 32) pack_one_kernel<32><<<1, kBlock, 0, s>>>(w);
    else pack_one_kernel<64><<<1, kBlock, 0, s>>>(w);
    return hipGetLastError();
}

hipError_t launch_count_one(const MeshWork &w, uint32_t *count, hipStream_t s) {
    if (w.edge <= 16) count_one_kernel<16><<<1, kBlock, 0, s>>>(w, count);
    else if (w.edge <= 32) count_one_kernel<32><<<1, kBlock, 0, s>>>(w, count);
    else count_one_kernel<64><<<1, kBlock, 0, s>>>(w, count);
    return hipGetLastError();
}

hipError_t launch_emit_one(const MeshWork &w, uint64_t base, uint32_t faces, MeshOut out,
                           const MeshTables *tables, uint32_t *overflow, hipStream_t s) {
    if (w.edge <= 16) emit_one_kernel<16><<<1, kBlock, 0, s>>>(w, base, faces, out, tables, overflow);
    else if (w.edge <= 32) emit_one_kernel<32><<<1, kBlock, 0, s>>>(w, base, faces, out, tables, overflow);
    else emit_one_kernel<64><<<1, kBlock, 0, s>>>(w, base, faces, out, tables, overflow);
    return hipGetLastError();
}

hipError_t launch_emit(const MeshWork *work, const int32_t *order, int32_t n, int32_t max_edge,
                       const uint64_t *bases, const uint32_t *counts, MeshOut out,
                       const MeshTables *tables, uint32_t *overflow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_edge <= 16)
        emit_kernel<16><<<n, kBlock, 0, s>>>(work, order, bases, counts, out, tables, overflow);
    else if (max_edge <= 32)
        emit_kernel<32><<<n, kBlock, 0, s>>>(work, order, bases, counts, out, tables, overflow);
    else
        emit_kernel<64><<<n, kBlock, 0, s>>>(work, order, bases, counts, out, tables, overflow);
    return hipGetLastError();
}

hipError_t launch_frustum(const Planes &planes, const int32_t *coords, int32_t n, int32_t edge,
                          uint8_t *mask, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    frustum_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(planes, coords, n, edge, mask);
    return hipGetLastError();
}

}  // namespace vx
