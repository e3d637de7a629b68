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

#include <mutex>
#include <type_traits>

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

// Exclusive scan of one u32 per thread over a BLOCK-thread workgroup.
template <int BLOCK = kBlock>
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
    for (int i = 0; i < BLOCK / 64; i++) {
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

// Lattice cell mod 256 of a floored coordinate.  Every f32 of magnitude
// >= 2^31 is a multiple of 256, so its cell index is 0; the test also keeps
// the conversion to int in range (and maps NaN, from an infinite coordinate,
// to 0).
__device__ __forceinline__ int cell8(float fl) {
    return fabsf(fl) < 2147483648.0f ? ((int)fl & 255) : 0;
}

__device__ float noise2(const int *p, float x, float y) {
    const float fx = floorf(x), fy = floorf(y);
    const int X = cell8(fx), Y = cell8(fy);
    x = x - fx;
    y = y - fy;
    const float u = fade(x), v = fade(y);
    const int A = p[X] + Y, B = p[X + 1] + Y;
    const float x1 = x - 1.0f, y1 = y - 1.0f;
    return lerp(v, lerp(u, grad2(p[A], x, y), grad2(p[B], x1, y)),
                lerp(u, grad2(p[A + 1], x, y1), grad2(p[B + 1], x1, y1)));
}

// --------------------------------------------------------------- fill
// Occupancy rows live in LDS as RowType<MAXN> (u64 when MAXN = 64), i.e. at
// a stride of RS u32 words; the boundary planes of the four neighbours the
// same way.  The fill, the mesh and the fused kernel share this layout.
template <int MAXN> struct RowStride { static constexpr int v = MAXN == 64 ? 2 : 1; };

// LDS layout of one fill workgroup, sized for the largest edge it handles.
template <int MAXN> struct FillShared {
    int perm[512];
    int plut[512];                     // (perm[i] & 15) * 16: byte offset of the gradient in a lut row
    float4 lut[8][16];
    float4 ytab[kMaxOctaves * MAXN];   // per (octave, y): fy, fy-1, fade(fy), Y cell
    uint32_t ychg[MAXN];               // per y: bit o set iff octave o's Y cell differs from y-1
    typename RowType<MAXN>::T occ[MAXN * MAXN];
    typename RowType<MAXN>::T nb[4][MAXN];
    uint32_t red[kBlock / 64];
};

// Columns are evaluated in "lines" of N lanes.  Lines 0..N-1 are the chunk's
// own rows (columns x = 0..N-1 at z = line); lines N..N+n-1 are recomputed
// neighbour planes: side -x (world x = ox-1, lanes along z), +x (x = ox+N),
// -z (world z = oz-1, lanes along x), +z (z = oz+N).  A ballot over a wave
// then yields occupancy rows or plane words directly.
struct Halo {
    int n;
    uint32_t sides;  // side of halo line i in bits [2i, 2i+2)
    __device__ __forceinline__ int side(int i) const { return (int)((sides >> (2 * i)) & 3u); }
};

template <int N>
__device__ __forceinline__ void line_column(int li, int lane, const Halo &H, int ox, int oz,
                                            int &wx, int &wz) {
    const int p = lane % N;
    if (li < N) {
        wx = ox + p;
        wz = oz + li;
        return;
    }
    const int k = H.side(min(li - N, 3));
    if (k == 0) { wx = ox - 1; wz = oz + p; }
    else if (k == 1) { wx = ox + N; wz = oz + p; }
    else if (k == 2) { wx = ox + p; wz = oz - 1; }
    else { wx = ox + p; wz = oz + N; }
}

// Rows of one line group, one per lane, gathered from the per-height
// ballots: lane k keeps the row at height k % N of line g * L + k / N (the
// ballot's bits [k / N * N, +N)).  Keeping costs a compare and a select per
// height; the rows then go to LDS in one store per lane (put_rows) instead of
// one two-lane store per height.
template <int N> struct RowKeep {
    uint32_t w0 = 0u, w1 = 0u;  // the row (w1: bits 32..63 at N = 64)
};

template <int N>
__device__ __forceinline__ void keep_row(RowKeep<N> &K, uint64_t b, int y, int lane) {
    if constexpr (N == 64) {
        if (lane == y) {
            K.w0 = (uint32_t)b;
            K.w1 = (uint32_t)(b >> 32);
        }
    } else {
        constexpr uint32_t MASK = N >= 32 ? 0xFFFFFFFFu : ((1u << (N & 31)) - 1u);
        if ((lane & (N - 1)) == y) K.w0 = (uint32_t)(b >> (lane & ~(N - 1))) & MASK;
    }
}

// Stores the kept rows of heights [y_lo, y_hi) of line group g.
template <int N, int MAXN, class SH>
__device__ __forceinline__ void put_rows(SH &S, const Halo &H, const RowKeep<N> &K, int g,
                                         int lane, int y_lo, int y_hi) {
    constexpr int RS = RowStride<MAXN>::v;
    constexpr int L = 64 / N;
    const int y = lane & (N - 1);
    if (y < y_lo || y >= y_hi) return;
    const int li = g * L + lane / N;
    uint32_t *dst = nullptr;
    if (li < N) dst = reinterpret_cast<uint32_t *>(S.occ) + (y * N + li) * RS;
    else if (li - N < H.n) dst = reinterpret_cast<uint32_t *>(S.nb[H.side(li - N)]) + y * RS;
    if (!dst) return;
    dst[0] = K.w0;
    if (RS == 2) dst[1] = N == 64 ? K.w1 : 0u;
}

// The 8 corner hashes of a lattice cell for the column-hoisted evaluation,
// indexed by corner k = x | row << 1 | col << 2 (row: world y, the per-row
// axis; col: world z, the column's second axis).  px0/px1 = perm[X], perm[X+1].
// Density rule, noise at (wx, wy, wz): Perlin's hash p[p[p[X]+Y]+Z] takes
// the row cell at its second level and the column cell at its third.
// Reference rule, noise at (wx, wz, wy) (src/perlin_noise.rs:6-13): the
// column cell enters at the second level and the row cell at the third.
template <bool REF, class SH>
__device__ __forceinline__ void corner_hashes(const SH &S, int px0, int px1, int yrow, int zcol,
                                              int h[8]) {
    if (!REF) {
        const int A = px0 + yrow, B = px1 + yrow;
        const int AA = S.perm[A] + zcol, AB = S.perm[A + 1] + zcol;
        const int BA = S.perm[B] + zcol, BB = S.perm[B + 1] + zcol;
        h[0] = S.perm[AA]; h[1] = S.perm[BA]; h[2] = S.perm[AB]; h[3] = S.perm[BB];
        h[4] = S.perm[AA + 1]; h[5] = S.perm[BA + 1]; h[6] = S.perm[AB + 1]; h[7] = S.perm[BB + 1];
    } else {
        const int A = px0 + zcol, B = px1 + zcol;
        const int AA = S.perm[A] + yrow, AB = S.perm[A + 1] + yrow;
        const int BA = S.perm[B] + yrow, BB = S.perm[B + 1] + yrow;
        h[0] = S.perm[AA]; h[1] = S.perm[BA]; h[2] = S.perm[AA + 1]; h[3] = S.perm[BA + 1];
        h[4] = S.perm[AB]; h[5] = S.perm[BB]; h[6] = S.perm[AB + 1]; h[7] = S.perm[BB + 1];
    }
}

// corner_hashes' last level read from S.plut: the byte offset of each
// corner's gradient within its lut row, (hash & 15) * 16.
template <bool REF, class SH>
__device__ __forceinline__ void corner_offsets(const SH &S, int px0, int px1, int yrow, int zcol,
                                               int h[8]) {
    if (!REF) {
        const int A = px0 + yrow, B = px1 + yrow;
        const int AA = S.perm[A] + zcol, AB = S.perm[A + 1] + zcol;
        const int BA = S.perm[B] + zcol, BB = S.perm[B + 1] + zcol;
        h[0] = S.plut[AA]; h[1] = S.plut[BA]; h[2] = S.plut[AB]; h[3] = S.plut[BB];
        h[4] = S.plut[AA + 1]; h[5] = S.plut[BA + 1]; h[6] = S.plut[AB + 1]; h[7] = S.plut[BB + 1];
    } else {
        const int A = px0 + zcol, B = px1 + zcol;
        const int AA = S.perm[A] + yrow, AB = S.perm[A + 1] + yrow;
        const int BA = S.perm[B] + yrow, BB = S.perm[B + 1] + yrow;
        h[0] = S.plut[AA]; h[1] = S.plut[BA]; h[2] = S.plut[AA + 1]; h[3] = S.plut[BA + 1];
        h[4] = S.plut[AB]; h[5] = S.plut[BB]; h[6] = S.plut[AB + 1]; h[7] = S.plut[BB + 1];
    }
}

// Trilinear blend of the 8 corner gradients (indexed as in corner_hashes)
// in Perlin's order: x innermost, then the noise's y axis, then its z axis.
// u: x fade; wz: column (world z) fade; wy: row (world y) fade.
template <bool REF>
__device__ __forceinline__ float blend8(float u, float wz, float wy, float g0, float g1, float g2,
                                        float g3, float g4, float g5, float g6, float g7) {
    if (!REF)  // noise y = world y (row), noise z = world z (column)
        return lerp(wz, lerp(wy, lerp(u, g0, g1), lerp(u, g2, g3)),
                    lerp(wy, lerp(u, g4, g5), lerp(u, g6, g7)));
    // noise y = world z (column), noise z = world y (row)
    return lerp(wy, lerp(wz, lerp(u, g0, g1), lerp(u, g4, g5)),
                lerp(wz, lerp(u, g2, g3), lerp(u, g6, g7)));
}

// The fill rule on a voxel's fBm sum: density terrains are solid iff
// fbm3(wx*f, wy*f, wz*f) - wy*y_scale > 0; the reference rule
// (src/chunk.rs:38-42) is solid iff fbm3(wx*f, wz*f, wy*f) < 0.
template <bool REF>
__device__ __forceinline__ bool solid_of(float sum, const TerrainDev &T, int wy) {
    if (REF) return sum < 0.0f;
    const float thr = (float)wy * T.y_scale;
    return (sum - thr) > 0.0f;
}

// 3-D fBm density for every line: the column's x/z lattice terms are hoisted
// per octave, the y terms come from S.ytab, the 8 corner gradients are
// rehashed only when the y lattice cell changes.  Voxel is solid iff
// fbm3(wx*f, wy*f, wz*f) - wy*y_scale > 0 (REF: fbm3(wx*f, wz*f, wy*f) < 0;
// oracle: vxo_chunk_fill).
// Work items are (line group, y segment of SEG rows) pairs dealt round-robin
// to the waves, so that the halo lines do not leave some waves a whole extra
// line group (N = 32 with four halo lines: 36 items of 16 rows, 9 per wave).
// g_first > 0 skips the first line groups (halo_only: g_first = N / L, the
// neighbour planes alone, with the chunk's own rows left untouched).
template <int N, int MAXN, bool DBG, int SEG, bool REF, class SH>
__device__ void density_lines(const TerrainDev &T, int ox, int oy, int oz, const Halo &H, SH &S,
                              float *noise_out, int g_first = 0) {
    constexpr int L = 64 / N;
    constexpr int NSEG = N / SEG;
    static_assert(N % SEG == 0, "segments tile the column");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int G = (N + H.n + L - 1) / L;
    const int octaves = T.octaves;
    for (int item = g_first * NSEG + wave; item < G * NSEG; item += kBlock / 64) {
        const int g = item / NSEG;
        const int y0 = (item - g * NSEG) * SEG;
        const int li = g * L + lane / N;
        int wxi, wzi;
        line_column<N>(li, lane, H, ox, oz, wxi, wzi);
        {
            float sum[SEG];
#pragma unroll
            for (int i = 0; i < SEG; i++) sum[i] = 0.0f;
            float f = T.frequency, amp = 1.0f;
            for (int o = 0; o < octaves; o++) {
                const float X = (float)wxi * f;
                const float fxl = floorf(X);
                const float fx = X - fxl, fx1 = fx - 1.0f, u = fade(fx);
                const int Xi = cell8(fxl);
                const float Z = (float)wzi * f;
                const float fzl = floorf(Z);
                const float fz = Z - fzl, fz1 = fz - 1.0f, wz = fade(fz);
                const int Zi = cell8(fzl);
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
                        int h[8];
                        corner_hashes<REF>(S, pX0, pX1, Yi, Zi, h);
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            const float4 Lk = S.lut[k][h[k] & 15];
                            const float xr = (k & 1) ? fx1 : fx;
                            const float zr = (k & 4) ? fz1 : fz;
                            c[k] = fmaf(Lk.x, xr, Lk.y * zr);
                            gy[k] = Lk.z;
                        }
                    }
                    const float g0 = fmaf(t.x, gy[0], c[0]), g1 = fmaf(t.x, gy[1], c[1]);
                    const float g2 = fmaf(t.y, gy[2], c[2]), g3 = fmaf(t.y, gy[3], c[3]);
                    const float g4 = fmaf(t.x, gy[4], c[4]), g5 = fmaf(t.x, gy[5], c[5]);
                    const float g6 = fmaf(t.y, gy[6], c[6]), g7 = fmaf(t.y, gy[7], c[7]);
                    const float n = blend8<REF>(u, wz, t.z, g0, g1, g2, g3, g4, g5, g6, g7);
                    sum[yy] = sum[yy] + amp * n;
                }
                f = f * T.lacunarity;
                amp = amp * T.gain;
            }
            RowKeep<N> K;
#pragma unroll 1
            for (int yy = 0; yy < SEG; yy++) {
                const int y = y0 + yy;
                const bool solid = solid_of<REF>(sum[yy], T, oy + y);
                keep_row<N>(K, __ballot(solid), y, lane);
                if (DBG && li < N)
                    noise_out[(size_t)(wxi - ox) + (size_t)N * li + (size_t)N * N * y] = sum[yy];
            }
            put_rows<N, MAXN>(S, H, K, g, lane, y0, y0 + SEG);
        }
    }
}

// The same density with the octave loop inside the row loop, for OCT
// octaves known at compile time: the fBm sum of a voxel is one register (no
// indexed register moves), the occupancy ballot follows each row, and the
// per-octave rehash is taken from the row's change mask (one uniform LDS
// word per row instead of one readfirstlane per octave).  Per-octave fade
// weights of the column and the 8 corner coefficients stay in registers;
// the column's lattice cells and fractions are recomputed in the rare
// rehash.  Octaves are summed in the same order as the oracle.
template <int N, int MAXN, bool DBG, int SEG, int OCT, bool REF, class SH>
__device__ void density_lines_oct(const TerrainDev &T, int ox, int oy, int oz, const Halo &H,
                                  SH &S, float *noise_out, int seg_only = -1, int g_first = 0) {
    constexpr int L = 64 / N;
    constexpr int NSEG = N / SEG;
    static_assert(N % SEG == 0, "segments tile the column");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int G = (N + H.n + L - 1) / L;
    // seg_only >= 0: only rows [seg_only * SEG, +SEG) (a chunk split over
    // several workgroups)
    const int nseg = seg_only >= 0 ? 1 : NSEG;
    float fo[OCT], ao[OCT];
    {
        float f = T.frequency, a = 1.0f;
#pragma unroll
        for (int o = 0; o < OCT; o++) {
            // uniform: the octave frequencies (used per item, not per row) go
            // to SGPRs, which leaves their VGPRs to the row loop
            fo[o] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(f)));
            ao[o] = a;
            f = f * T.lacunarity;
            a = a * T.gain;
        }
    }
    for (int item = g_first * nseg + wave; item < G * nseg; item += kBlock / 64) {
        const int g = item / nseg;
        const int y0 = (seg_only >= 0 ? seg_only : item - g * nseg) * SEG;
        const int li = g * L + lane / N;
        int wxi, wzi;
        // Column terms are computed per item, not hoisted: at edge 32 a lane's
        // x is the same for every item, and hoisted x-lattice terms would be
        // kept in spilled registers.
        int lane_i = lane;
        asm volatile("" : "+v"(lane_i));
        line_column<N>(li, lane_i, H, ox, oz, wxi, wzi);
        float u[OCT], wz[OCT], c[OCT][8], gy[OCT][8];
#pragma unroll
        for (int o = 0; o < OCT; o++) {
            const float X = (float)wxi * fo[o];
            u[o] = fade(X - floorf(X));
            const float Z = (float)wzi * fo[o];
            wz[o] = fade(Z - floorf(Z));
        }
        // The item's cell-change masks in SGPRs: bit r of ychm[o] is set iff
        // octave o enters a new y cell at row y0 + r (one LDS read per lane
        // and a ballot per octave, instead of an LDS read per row).
        uint64_t ychm[OCT];
        {
            const uint32_t yc = lane < SEG ? S.ychg[y0 + lane] : 0u;
#pragma unroll
            for (int o = 0; o < OCT; o++) ychm[o] = __ballot((yc >> o) & 1u);
        }
        RowKeep<N> K;
#pragma unroll 1
        for (int y = y0; y < y0 + SEG; y++) {
            uint32_t chg = 0;
#pragma unroll
            for (int o = 0; o < OCT; o++) chg |= (uint32_t)((ychm[o] >> (y - y0)) & 1ull) << o;
            if (y == y0) chg = (1u << OCT) - 1u;
            // Rehash the octaves entering a new y cell (rare: a uniform loop
            // over the set bits, so the rehash code exists once).
#pragma unroll 1
            for (; chg; chg &= chg - 1u) {
                const int o = __builtin_ctz(chg);
                const int Yi = __builtin_amdgcn_readfirstlane(__float_as_int(S.ytab[o * N + y].w));
                float f = T.frequency;
                for (int k = 0; k < o; k++) f = f * T.lacunarity;
                const float X = (float)wxi * f, Z = (float)wzi * f;
                const float fxl = floorf(X), fzl = floorf(Z);
                const float fx = X - fxl, fz = Z - fzl;
                const float fx1 = fx - 1.0f, fz1 = fz - 1.0f;
                // the column's x and z cells, recomputed here rather than
                // kept per octave in registers (12 VGPRs at four octaves)
                const int Xi = cell8(fxl), zi = cell8(fzl);
                const int px0 = S.perm[Xi], px1 = S.perm[Xi + 1];
                int h[8];
                corner_offsets<REF>(S, px0, px1, Yi, zi, h);
                float cn[8], gn[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const float4 Lk = *reinterpret_cast<const float4 *>(
                        reinterpret_cast<const char *>(&S.lut[k][0]) + h[k]);
                    const float xr = (k & 1) ? fx1 : fx;
                    const float zr = (k & 4) ? fz1 : fz;
                    cn[k] = fmaf(Lk.x, xr, Lk.y * zr);
                    gn[k] = Lk.z;
                }
                // Store into octave o's registers through a uniform branch per
                // octave with plain moves.  Written as selects, every
                // coefficient of every octave costs a v_cndmask per rehash
                // (64 at four octaves, each reading its mask from SGPRs at
                // half issue rate); the asm keeps the branches from being
                // turned back into selects.  (One copy of the rehash per
                // octave, writing its registers in place, raised the
                // register count past four waves per SIMD and spilled.)
#pragma unroll
                for (int q = 0; q < OCT; q++) {
                    if (o == q) {
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            asm volatile("v_mov_b32 %0, %1" : "=v"(c[q][k]) : "v"(cn[k]));
                            asm volatile("v_mov_b32 %0, %1" : "=v"(gy[q][k]) : "v"(gn[k]));
                        }
                    }
                }
            }
            float sum = 0.0f;
#pragma unroll
            for (int o = 0; o < OCT; o++) {
                const float4 t = S.ytab[o * N + y];
                const float g0 = fmaf(t.x, gy[o][0], c[o][0]), g1 = fmaf(t.x, gy[o][1], c[o][1]);
                const float g2 = fmaf(t.y, gy[o][2], c[o][2]), g3 = fmaf(t.y, gy[o][3], c[o][3]);
                const float g4 = fmaf(t.x, gy[o][4], c[o][4]), g5 = fmaf(t.x, gy[o][5], c[o][5]);
                const float g6 = fmaf(t.y, gy[o][6], c[o][6]), g7 = fmaf(t.y, gy[o][7], c[o][7]);
                const float n = blend8<REF>(u[o], wz[o], t.z, g0, g1, g2, g3, g4, g5, g6, g7);
                sum = sum + ao[o] * n;
            }
            keep_row<N>(K, __ballot(solid_of<REF>(sum, T, oy + y)), y, lane);
            if (DBG && li < N)
                noise_out[(size_t)(wxi - ox) + (size_t)N * li + (size_t)N * N * y] = sum;
        }
        put_rows<N, MAXN>(S, H, K, g, lane, y0, y0 + SEG);
    }
}

// 2-D fBm heightmap for every line: h = floor(height_scale*(n+1)/2) clamped
// to [0, height_max]; voxel solid iff wy <= h.
template <int N, int MAXN, bool DBG, class SH>
__device__ void heightmap_lines(const TerrainDev &T, int ox, int oy, int oz, const Halo &H,
                                SH &S, float *noise_out, int g_first = 0) {
    constexpr int L = 64 / N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int G = (N + H.n + L - 1) / L;
    for (int g = g_first + wave; g < G; g += kBlock / 64) {
        const int li = g * L + lane / N;
        int wxi, wzi;
        line_column<N>(li, lane, H, ox, oz, wxi, wzi);
        const float wx = (float)wxi, wzf = (float)wzi;
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
        if (DBG && li < N) noise_out[(size_t)(wxi - ox) + (size_t)N * li] = sum;
        RowKeep<N> K;
        for (int y = 0; y < N; y++) keep_row<N>(K, __ballot(oy + y <= h), y, lane);
        put_rows<N, MAXN>(S, H, K, g, lane, 0, N);
    }
}

// Per-(octave, y) lattice terms of the chunk's world rows.
template <int N, class SH>
__device__ __forceinline__ void build_ytab(const TerrainDev &T, int oy, SH &S) {
    for (int i = threadIdx.x; i < T.octaves * N; i += kBlock) {
        const int o = i / N, y = i - o * N;
        float f = T.frequency;
        for (int k = 0; k < o; k++) f = f * T.lacunarity;
        const float Y = (float)(oy + y) * f;
        const float fyl = floorf(Y);
        const float fy = Y - fyl;
        S.ytab[i] = make_float4(fy, fy - 1.0f, fade(fy), __int_as_float(cell8(fyl)));
    }
    // Rows where an octave enters a new lattice cell (the hash uses the cell
    // index mod 256, which is what is compared); row 0 rehashes everything.
    for (int y = threadIdx.x; y < N; y += kBlock) {
        uint32_t m = 0;
        float f = T.frequency;
        for (int o = 0; o < T.octaves; o++) {
            const int a = cell8(floorf((float)(oy + y) * f));
            const int b = cell8(floorf((float)(oy + y - 1) * f));
            if (y == 0 || a != b) m |= 1u << o;
            f = f * T.lacunarity;
        }
        S.ychg[y] = m;
    }
}

// Fills S.occ (and the recomputed halo planes) for terrain T.
// OCT_INNER: the octave-inner density loop for 1..4 octaves (the fill
// kernels); the fused kernel keeps the octave-outer loop, which needs fewer
// registers next to its emit phase.
template <int N, int MAXN, bool DBG, int SEG, bool REF, class SH>
__device__ __forceinline__ void density_rows(const TerrainDev &T, int ox, int oy, int oz,
                                             const Halo &H, SH &S, float *noise_out) {
    build_ytab<N>(T, oy, S);
    __syncthreads();
    switch (T.octaves) {
    case 1: density_lines_oct<N, MAXN, DBG, SEG, 1, REF>(T, ox, oy, oz, H, S, noise_out); break;
    case 2: density_lines_oct<N, MAXN, DBG, SEG, 2, REF>(T, ox, oy, oz, H, S, noise_out); break;
    case 3: density_lines_oct<N, MAXN, DBG, SEG, 3, REF>(T, ox, oy, oz, H, S, noise_out); break;
    case 4: density_lines_oct<N, MAXN, DBG, SEG, 4, REF>(T, ox, oy, oz, H, S, noise_out); break;
    default: density_lines<N, MAXN, DBG, SEG, REF>(T, ox, oy, oz, H, S, noise_out); break;
    }
}

// The reference rule (kReference) is generated by the fill kernels only; a
// batch holding such a terrain never takes the fused form.
template <int N, int MAXN, bool DBG, int SEG, bool OCT_INNER, class SH>
__device__ void generate_rows(const TerrainDev &T, int ox, int oy, int oz, const Halo &H, SH &S,
                              float *noise_out) {
    using R = typename RowType<MAXN>::T;
    if constexpr (!OCT_INNER) {
        if (T.kind == kDensity) {
            build_ytab<N>(T, oy, S);
            __syncthreads();
            density_lines<N, MAXN, DBG, SEG, false>(T, ox, oy, oz, H, S, noise_out);
            return;
        }
    } else {
        if (T.kind == kDensity) {
            density_rows<N, MAXN, DBG, SEG, false>(T, ox, oy, oz, H, S, noise_out);
            return;
        }
        if (T.kind == kReference) {
            density_rows<N, MAXN, DBG, SEG, true>(T, ox, oy, oz, H, S, noise_out);
            return;
        }
    }
    if (T.kind == kHeightmap) {
        heightmap_lines<N, MAXN, DBG>(T, ox, oy, oz, H, S, noise_out);
    } else {
        const R full = T.kind == kSolid ? (N == 64 ? (R)~(R)0 : (R)((N >= 32 ? 0xFFFFFFFFu
                                                                             : ((1u << (N & 31)) - 1u))))
                                        : (R)0;
        for (int i = threadIdx.x; i < N * N; i += kBlock) S.occ[i] = full;
        for (int i = threadIdx.x; i < H.n * N; i += kBlock) S.nb[H.side(i / N)][i % N] = full;
    }
}

// Writes the layers y_lo <= y < y_hi of a chunk whose occupancy rows for
// them are in S.occ: occupancy and boundary planes always, u8 ids (all solid
// voxels = kBlockGrass) when write_ids; meta[0] = solid block id from the
// workgroup that writes layer 0.  COHERENT: the occupancy rows and boundary
// planes, which other workgroups of the same launch read, are written with
// agent-scope stores (coherent across the XCDs' L2 caches without a cache
// write-back; a plain store may stay in the writing XCD's L2).
template <int N, int MAXN, bool COHERENT = false, class SH>
__device__ void write_chunk_outputs(const FillWork &w, SH &S, bool write_ids, uint32_t solid_type,
                                    int y_lo = 0, int y_hi = N) {
    constexpr int W = N == 64 ? 2 : 1;
    constexpr int RS = RowStride<MAXN>::v;
    const uint32_t *occ = reinterpret_cast<const uint32_t *>(S.occ);
    const int tid = threadIdx.x;
    const int span = y_hi - y_lo;
    for (int i = y_lo * N * W + tid; i < y_hi * N * W; i += kBlock) {
        const int r = i / W, j = i - r * W;
        if (COHERENT)
            __hip_atomic_store(w.occ + i, occ[r * RS + j], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            w.occ[i] = occ[r * RS + j];
    }
    // boundary planes: [0] x=0, [1] x=N-1 (bit z per y); [2] z=0, [3] z=N-1 (row bits)
    for (int k = tid; k < 4 * span * W; k += kBlock) {
        const int pl = k / (span * W), rem = k - pl * span * W, y = y_lo + rem / W;
        const int j = rem - (y - y_lo) * W, i = pl * N * W + y * W + j;
        uint32_t v = 0;
        if (pl < 2) {
            const int xb = pl == 0 ? 0 : N - 1;
            const int zlim = N < 32 ? N : 32;
            for (int zz = 0; zz < zlim; zz++) {
                const int z = j * 32 + zz;
                const uint32_t word = occ[(y * N + z) * RS + (xb >> 5)];
                v |= ((word >> (xb & 31)) & 1u) << zz;
            }
        } else {
            const int z = pl == 2 ? 0 : N - 1;
            v = occ[(y * N + z) * RS + j];
        }
        if (COHERENT)
            __hip_atomic_store(w.planes + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            w.planes[i] = v;
    }
    if (write_ids) {
        // u8 block ids, 16 voxels per 16-byte store
        constexpr int PER_ROW = N / 16;
        uint4 *ids4 = reinterpret_cast<uint4 *>(w.ids);
        for (int q = y_lo * N * PER_ROW + tid; q < y_hi * N * PER_ROW; q += kBlock) {
            const int r = q / PER_ROW, part = q - r * PER_ROW;
            const uint32_t word = occ[r * RS + (part >> 1)];
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
    if (tid == 0 && y_lo == 0) w.meta[0] = solid_type;
}

// Rows [part * N / parts, +N / parts) of chunk `w` through the octave-inner
// density loop, then their outputs (one of `parts` workgroups on the chunk).
template <int N, int MAXN, int PARTS, int OCT, bool REF>
__device__ void fill_part(const FillWork &w, const TerrainDev &T, FillShared<MAXN> &S, int part) {
    const Halo H{0, 0u};
    build_ytab<N>(T, w.cy * N, S);
    __syncthreads();
    density_lines_oct<N, MAXN, false, N / PARTS, OCT, REF>(T, w.cx * N, w.cy * N, w.cz * N, H, S,
                                                           nullptr, part);
    __syncthreads();
    write_chunk_outputs<N, MAXN>(w, S, true, kBlockGrass, part * (N / PARTS),
                                 (part + 1) * (N / PARTS));
}

template <int N, int MAXN, class SH>
__device__ void count_loaded(SH &S, uint32_t *count_out, uint32_t *slabs);

// With w.count_out the fill also counts the chunk's faces from the rows it
// holds in LDS and the neighbour planes w.cnb (all written before the
// launch), which spares the batch its count kernel.
template <int N, int MAXN, bool DBG>
__device__ __forceinline__ void fill_chunk(const FillWork &w, const TerrainDev &T, FillShared<MAXN> &S) {
    const Halo H{0, 0u};
    generate_rows<N, MAXN, DBG, (N < 32 ? N : 32), true>(T, w.cx * N, w.cy * N, w.cz * N, H, S,
                                                         w.noise_out);
    __syncthreads();
    write_chunk_outputs<N, MAXN>(w, S, true, kBlockGrass);
    if (w.count_out) {
        using R = typename RowType<MAXN>::T;
        constexpr int W = N == 64 ? 2 : 1;
        for (int i = threadIdx.x; i < 4 * N; i += kBlock) {  // S.nb: no halo lines here
            const int k = i / N, y = i - k * N;
            const uint32_t *p = w.cnb[k];
            R v = 0;
            if (p) {
                if (W == 2) v = (R)p[2 * y] | ((R)p[2 * y + 1] << (W == 2 ? 32 : 0));
                else v = (R)p[y];
            }
            S.nb[k][y] = v;
        }
        __syncthreads();
        count_loaded<N, MAXN>(S, w.count_out, w.slab_out);
    }
}

template <int N, int MAXN, bool REF>
__device__ __forceinline__ void fill_parts(const FillWork &w, const TerrainDev &T,
                                           FillShared<MAXN> &S, int part, int parts) {
    if (parts == 8) {
        switch (T.octaves) {
        case 1: fill_part<N, MAXN, (N == 64 ? 8 : 4), 1, REF>(w, T, S, part); break;
        case 2: fill_part<N, MAXN, (N == 64 ? 8 : 4), 2, REF>(w, T, S, part); break;
        case 3: fill_part<N, MAXN, (N == 64 ? 8 : 4), 3, REF>(w, T, S, part); break;
        default: fill_part<N, MAXN, (N == 64 ? 8 : 4), 4, REF>(w, T, S, part); break;
        }
    } else if (parts == 2) {
        switch (T.octaves) {
        case 1: fill_part<N, MAXN, 2, 1, REF>(w, T, S, part); break;
        case 2: fill_part<N, MAXN, 2, 2, REF>(w, T, S, part); break;
        case 3: fill_part<N, MAXN, 2, 3, REF>(w, T, S, part); break;
        default: fill_part<N, MAXN, 2, 4, REF>(w, T, S, part); break;
        }
    } else {
        switch (T.octaves) {
        case 1: fill_part<N, MAXN, 4, 1, REF>(w, T, S, part); break;
        case 2: fill_part<N, MAXN, 4, 2, REF>(w, T, S, part); break;
        case 3: fill_part<N, MAXN, 4, 3, REF>(w, T, S, part); break;
        default: fill_part<N, MAXN, 4, 4, REF>(w, T, S, part); break;
        }
    }
}

// One of `parts` (2, 4 or 8) workgroups on a chunk: density terrains (either
// rule) with up to 4 octaves split their rows among the parts (8 only at
// edge 64; a smaller chunk uses the first 4); any other terrain is filled
// whole by part 0 while the other parts return.
template <int N, int MAXN>
__device__ __forceinline__ void fill_chunk_split(const FillWork &w, const TerrainDev &T, FillShared<MAXN> &S,
                                 int part, int parts) {
    if ((T.kind == kDensity || T.kind == kReference) && T.octaves <= 4) {
        if (parts == 8 && N < 64) {
            if (part >= 4) return;
            parts = 4;
        }
        if (T.kind == kReference) fill_parts<N, MAXN, true>(w, T, S, part, parts);
        else fill_parts<N, MAXN, false>(w, T, S, part, parts);
        return;
    }
    if (part == 0) fill_chunk<N, MAXN, false>(w, T, S);
}

// Packs caller-provided u8 ids (already in w.ids) into occupancy rows and
// boundary planes; meta[0] = the single solid block id, or kBlockMixed.
template <int N, int MAXN>
__device__ void pack_chunk(const FillWork &w, FillShared<MAXN> &S) {
    constexpr int W = N == 64 ? 2 : 1;
    constexpr int RS = RowStride<MAXN>::v;
    __shared__ uint32_t s_types;
    if (threadIdx.x == 0) s_types = 0;
    __syncthreads();
    uint32_t *occ = reinterpret_cast<uint32_t *>(S.occ);
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
        occ[r * RS + j] = bits;
        if (RS == 2 && W == 1) occ[r * RS + 1] = 0u;
    }
    atomicOr(&s_types, types);
    __syncthreads();
    const uint32_t solid_types = s_types & ~1u;
    uint32_t st = kBlockMixed;
    if (solid_types == 0) st = kBlockGrass;  // no solid voxel: type unused
    else if ((solid_types & (solid_types - 1)) == 0) st = (uint32_t)(__ffs(solid_types) - 1);
    write_chunk_outputs<N, MAXN>(w, S, false, st);
}

// SPLIT: the workgroup is part `part` of `parts` on this chunk.
template <int MAXN, bool SPLIT = false>
__device__ __forceinline__ void fill_item(const FillWork &w, const TerrainDev *terrains,
                                          const GradLut *lut, FillShared<MAXN> &S, int part = 0,
                                          int parts = 1) {
    const TerrainDev &T = terrains[w.terrain];
    const float4(*L)[16] = T.kind == kReference ? lut->r : lut->c;
    for (int i = threadIdx.x; i < 512; i += kBlock) {
        const int p = T.perm[i];
        S.perm[i] = p;
        S.plut[i] = (p & 15) * 16;
    }
    for (int i = threadIdx.x; i < 8 * 16; i += kBlock) S.lut[i >> 4][i & 15] = L[i >> 4][i & 15];
    __syncthreads();
    if (w.noise_out) {  // parity-check path: also stores the raw noise
        if (SPLIT && part > 0) return;
        switch (w.edge) {
        case 16: fill_chunk<16, MAXN, true>(w, T, S); break;
        case 32: if (MAXN >= 32) fill_chunk<(MAXN >= 32 ? 32 : 16), MAXN, true>(w, T, S); break;
        case 64: if (MAXN >= 64) fill_chunk<(MAXN >= 64 ? 64 : 16), MAXN, true>(w, T, S); break;
        default: break;
        }
        return;
    }
    if (SPLIT) {
        switch (w.edge) {
        case 16: fill_chunk_split<16, MAXN>(w, T, S, part, parts); break;
        case 32: if (MAXN >= 32) fill_chunk_split<(MAXN >= 32 ? 32 : 16), MAXN>(w, T, S, part, parts); break;
        case 64: if (MAXN >= 64) fill_chunk_split<(MAXN >= 64 ? 64 : 16), MAXN>(w, T, S, part, parts); break;
        default: break;
        }
        return;
    }
    switch (w.edge) {
    case 16: fill_chunk<16, MAXN, false>(w, T, S); break;
    case 32: if (MAXN >= 32) fill_chunk<(MAXN >= 32 ? 32 : 16), MAXN, false>(w, T, S); break;
    case 64: if (MAXN >= 64) fill_chunk<(MAXN >= 64 ? 64 : 16), MAXN, false>(w, T, S); break;
    default: break;
    }
}

// Four waves per SIMD (at most 128 VGPRs) for edges up to 32: left to
// itself the allocator takes one register more and drops to three waves,
// about 15% slower.  Edge-64 batches stay at three waves.
// A schedule entry outside [0, n) (a corrupted order array) sets the sticky
// error word and the workgroup does nothing: no access is made through it.
__device__ __forceinline__ bool order_entry_ok(int item, int n, uint32_t *error) {
    if ((unsigned)item < (unsigned)n) return true;
    if (threadIdx.x == 0) atomicOr(error, 1u);
    return false;
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
fill_kernel(const FillWork *__restrict__ work, const int32_t *__restrict__ order, int32_t n,
            const TerrainDev *__restrict__ terrains, const GradLut *__restrict__ lut,
            uint32_t *__restrict__ error) {
    __shared__ FillShared<MAXN> S;
    const int item = order ? order[blockIdx.x] : (int)blockIdx.x;
    if (!order_entry_ok(item, n, error)) return;
    fill_item<MAXN>(work[item], terrains, lut, S);
}

// Small batches: `parts` workgroups per chunk (a kernel of its own so that
// the row-split path does not add to the registers of fill_kernel).
template <int MAXN>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
fill_split_kernel(const FillWork *__restrict__ work, const int32_t *__restrict__ order, int32_t n,
                  const TerrainDev *__restrict__ terrains, const GradLut *__restrict__ lut,
                  int parts, uint32_t *__restrict__ error) {
    __shared__ FillShared<MAXN> S;
    const int k = (int)blockIdx.x / parts, part = (int)blockIdx.x - k * parts;
    const int item = order ? order[k] : k;
    if (!order_entry_ok(item, n, error)) return;
    fill_item<MAXN, true>(work[item], terrains, lut, S, part, parts);
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
// Per-wave staging of 64 face records (80 B of vertices or 68 B of
// InstanceRaw each) so that global stores are lane-contiguous: a record per
// lane written as 16-B pieces at an 80-B stride reaches about half the
// HBM write rate of contiguous 16-B lanes.
constexpr int kStageWords = 32 * 20;  // 32 records per wave pass: 2.5 KB

// Rows of a descriptor tile with more than kHeavyFaces faces, walked by a
// whole wave (one lane per voxel of the row) instead of one lane: the rows
// at a chunk face whose neighbour is absent, or the bottom layer, hold a
// face at nearly every voxel, and a lane walking 64 of them serially holds
// its wave (and the workgroup's next barrier) for 64 steps while its other
// lanes walk rows of two or three faces.  At most tile / (kHeavyFaces + 1)
// + 2 rows of a tile are heavy; a row beyond the list's room is walked by
// its lane.
constexpr int kHeavyFaces = 24;
constexpr int kHeavyMax = kDescTile / 8;
struct HeavyRows {
    uint16_t row[kHeavyMax];
    uint32_t n;
};

// LDS of an emit workgroup of BLOCK threads on a whole chunk (the batch
// emit uses it up to MAXN = 32; a 64^3 chunk's 32 KB of u64 rows and 16 KB
// of row firsts would allow only two workgroups per CU, so batches emit
// those by slabs, and only the single-chunk emit takes MAXN = 64).
template <int MAXN, int BLOCK = kBlock> struct MeshShared {
    typename RowType<MAXN>::T occ[MAXN * MAXN];
    typename RowType<MAXN>::T nb[4][MAXN];
    uint32_t desc[kDescTile / 2];
    // first face of every row, for the interleaved descriptor walk (4 KB at
    // MAXN = 32: 28 KB in all, 5 workgroups per CU; 16 KB at MAXN = 64)
    uint32_t rowfirst[MAXN >= 32 ? MAXN * MAXN : 1];
    alignas(16) float P[6][3][2][4];
    alignas(16) float Mb[6][2][3][4];
    uint32_t tex[3][6];
    uint32_t red[BLOCK / 64];
    HeavyRows heavy;
    float4 stage[BLOCK / 64][kStageWords / 4];
};

// LDS of one y-slab of a 64^3 chunk (the emit of batches that hold 64^3
// chunks works on slabs of kSlabLayers layers, whose face counts the count
// kernel writes): occupancy rows of layers ylo - 1 .. ylo + kSlabLayers
// (zero outside the chunk), the four neighbour planes' rows of the slab's
// layers, and the emit's descriptor tile, row firsts and staging.  33 KB, so
// four 256-thread workgroups per CU instead of the two that a whole 64^3
// chunk's 32 KB of rows and 16 KB of row firsts allow.
constexpr int kSlabLayers = 16;
static_assert(kSlabs * kSlabLayers == 64, "slabs tile a 64^3 chunk");
struct SlabShared {
    uint64_t occ[(kSlabLayers + 2) * 64];
    uint64_t nb[4][kSlabLayers];
    uint32_t desc[kDescTile / 2];
    uint32_t rowfirst[kSlabLayers * 64];
    alignas(16) float P[6][3][2][4];
    alignas(16) float Mb[6][2][3][4];
    uint32_t tex[3][6];
    uint32_t red[kBlock / 64];
    HeavyRows heavy;
    float4 stage[kBlock / 64][kStageWords / 4];
};

// Rows an emit workgroup scans: the whole chunk, or one slab.
template <class SH, int N> struct RowsOf { static constexpr int v = N * N; };
template <int N> struct RowsOf<SlabShared, N> { static constexpr int v = kSlabLayers * N; };

// Orders the LDS accesses of one wave: staging written by some lanes is read
// by others of the same wave (LDS executes one wave's operations in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int N, int MAXN, int BLOCK = kBlock, class SH>
__device__ __forceinline__ void load_rows(const MeshWork &w, SH &S) {
    using R = typename RowType<MAXN>::T;
    constexpr int W = N == 64 ? 2 : 1;
    const int tid = threadIdx.x;
    for (int r = tid; r < N * N; r += BLOCK) {
        R v;
        if (W == 2) v = (R)w.occ[2 * r] | ((R)w.occ[2 * r + 1] << (W == 2 ? 32 : 0));
        else v = (R)w.occ[r];
        S.occ[r] = v;
    }
    for (int i = tid; i < 4 * N; i += BLOCK) {
        const int k = i / N, y = i - k * N;
        R v = 0;
        if (w.nb[k]) {
            if (W == 2) v = (R)w.nb[k][2 * y] | ((R)w.nb[k][2 * y + 1] << (W == 2 ? 32 : 0));
            else v = (R)w.nb[k][y];
        }
        S.nb[k][y] = v;
    }
}

// face_masks for local layer yl of a slab held in SlabShared (its halo
// layers hold the rows above and below, zero outside the chunk).
__device__ __forceinline__ void slab_face_masks(const SlabShared &S, int yl, int z, uint64_t m[6]) {
    const int b = (yl + 1) * 64 + z;
    const uint64_t self = S.occ[b];
    const uint64_t lb = (S.nb[0][yl] >> z) & 1ull;
    const uint64_t rb = (S.nb[1][yl] >> z) & 1ull;
    m[0] = self & ~((self << 1) | lb);
    m[1] = self & ~((self >> 1) | (rb << 63));
    const uint64_t back = z > 0 ? S.occ[b - 1] : S.nb[2][yl];
    const uint64_t front = z < 63 ? S.occ[b + 1] : S.nb[3][yl];
    m[2] = self & ~back;
    m[3] = self & ~front;
    m[4] = self & ~S.occ[b - 64];
    m[5] = self & ~S.occ[b + 64];
}

// Six face masks of row (y, z), bit x set iff voxel x emits that face
// (chunk.rs:77-240): neighbour cell air, or outside the chunk with the
// neighbour chunk absent, or outside vertically.
template <int N, int MAXN, class SH>
__device__ __forceinline__ void face_masks(const SH &S, int y, int z,
                                           typename RowType<MAXN>::T m[6]) {
    using R = typename RowType<MAXN>::T;
    if constexpr (std::is_same<SH, SlabShared>::value) {
        static_assert(N == 64 && MAXN == 64, "slabs are layers of 64^3 chunks");
        slab_face_masks(S, y, z, m);
    } else {
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
}

// Loads slab `ylo / kSlabLayers` of a 64^3 chunk (rows are u32 pairs, low
// word first) with its halo layers and neighbour-plane rows.
__device__ __forceinline__ void load_slab(const MeshWork &w, SlabShared &S, int ylo) {
    const uint64_t *occ = reinterpret_cast<const uint64_t *>(w.occ);
    for (int i = threadIdx.x; i < (kSlabLayers + 2) * 64; i += kBlock) {
        const int y = ylo - 1 + (i >> 6);
        S.occ[i] = (y >= 0 && y < 64) ? occ[y * 64 + (i & 63)] : 0ull;
    }
    for (int i = threadIdx.x; i < 4 * kSlabLayers; i += kBlock) {
        const int k = i / kSlabLayers, yl = i - k * kSlabLayers;
        const uint64_t *p = reinterpret_cast<const uint64_t *>(w.nb[k]);
        S.nb[k][yl] = p ? p[ylo + yl] : 0ull;
    }
}

// Visible faces of the chunk whose rows and neighbour planes are in LDS.
template <int N, int MAXN, class SH>
__device__ uint32_t count_rows(SH &S) {
    using R = typename RowType<MAXN>::T;
    uint32_t c = 0;
    for (int r = threadIdx.x; r < N * N; r += kBlock) {
        R m[6];
        face_masks<N, MAXN>(S, r / N, r % N, m);
#pragma unroll
        for (int d = 0; d < 6; d++) c += popc(m[d]);
    }
    return block_sum(c, S.red);
}

// LDS of the count kernel: rows and neighbour planes only.
template <int MAXN> struct CountShared {
    typename RowType<MAXN>::T occ[MAXN * MAXN];
    typename RowType<MAXN>::T nb[4][MAXN];
    uint32_t red[kBlock / 64];
};

// Counts the faces of the chunk whose rows and neighbour planes are in
// LDS (S.occ, S.nb).  A 64^3 chunk's count also goes out per slab
// (slabs[0..kSlabs), when given): thread t's rows t + kBlock * i lie in
// slab i / (kSlabLayers * 64 / kBlock).
template <int N, int MAXN, class SH>
__device__ void count_loaded(SH &S, uint32_t *count_out, uint32_t *slabs) {
    if constexpr (N == 64) {
        using R = typename RowType<MAXN>::T;
        constexpr int PER = kSlabLayers * 64 / kBlock;  // a thread's rows per slab
        // per-slab sums: wave reductions, then one barrier
        __shared__ uint32_t s_part[kBlock / 64][kSlabs];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll 1
        for (int s = 0; s < kSlabs; s++) {
            uint32_t c = 0;
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int r = threadIdx.x + kBlock * (s * PER + i);
                R m[6];
                face_masks<N, MAXN>(S, r / N, r % N, m);
#pragma unroll
                for (int d = 0; d < 6; d++) c += popc(m[d]);
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
            if (lane == 0) s_part[wave][s] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t total = 0;
            for (int s = 0; s < kSlabs; s++) {
                uint32_t c = 0;
                for (int k = 0; k < kBlock / 64; k++) c += s_part[k][s];
                if (slabs) slabs[s] = c;
                total += c;
            }
            *count_out = total;
        }
    } else {
        const uint32_t total = count_rows<N, MAXN>(S);
        if (threadIdx.x == 0) *count_out = total;
    }
}

template <int N, int MAXN>
__device__ void count_chunk(const MeshWork &w, CountShared<MAXN> &S, uint32_t *count_out,
                            uint32_t *slabs) {
    load_rows<N, MAXN>(w, S);
    __syncthreads();
    count_loaded<N, MAXN>(S, count_out, slabs);
}

template <int MAXN>
__device__ __forceinline__ void count_item(const MeshWork &w, CountShared<MAXN> &S, uint32_t *out,
                                           uint32_t *slabs) {
    switch (w.edge) {
    case 16: count_chunk<16, MAXN>(w, S, out, slabs); break;
    case 32: if (MAXN >= 32) count_chunk<(MAXN >= 32 ? 32 : 16), MAXN>(w, S, out, slabs); break;
    case 64: if (MAXN >= 64) count_chunk<(MAXN >= 64 ? 64 : 16), MAXN>(w, S, out, slabs); break;
    default: break;
    }
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) count_kernel(const MeshWork *__restrict__ work,
                                                       uint32_t *__restrict__ counts,
                                                       uint32_t *__restrict__ slabs) {
    __shared__ CountShared<MAXN> S;
    count_item<MAXN>(work[blockIdx.x], S, counts + blockIdx.x,
                     slabs ? slabs + (size_t)kSlabs * blockIdx.x : nullptr);
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) count_one_kernel(MeshWork w, uint32_t *__restrict__ count) {
    __shared__ CountShared<MAXN> S;
    count_item<MAXN>(w, S, count, nullptr);
}

// Emission order: chunks by descending face count (buckets of 16 faces,
// counts >= 16368 share the first bucket), so the emit grid starts with the
// largest chunks and ends with small ones instead of ending on a long one.
// Order inside a bucket is arbitrary; outputs depend only on the bases.
constexpr int kOrderBuckets = 1024;
__device__ __forceinline__ int order_bucket(uint32_t c) {
    return kOrderBuckets - 1 - (int)min(c >> 4, (uint32_t)(kOrderBuckets - 1));
}

// Device-wide exclusive scan of per-chunk face counts (one workgroup of
// 1024 threads, each owning a contiguous run of chunks); with order_out it
// also writes the emission order (a counting sort by face count).  With
// `items` (the batch's mesh work) the order lists emit items as codes
// chunk << 2 | slab: kSlabs per 64^3 chunk, one (slab 0) per other chunk.
__global__ void __launch_bounds__(1024) scan_kernel(const uint32_t *__restrict__ counts, int32_t n,
                                                    const MeshWork *__restrict__ items,
                                                    uint64_t *__restrict__ bases,
                                                    uint64_t *__restrict__ total,
                                                    uint32_t *__restrict__ overflow,
                                                    int32_t *__restrict__ order_out,
                                                    const uint64_t *__restrict__ base_in) {
    __shared__ uint64_t s_wave[16];
    __shared__ uint32_t s_hist[kOrderBuckets];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (n + 1023) / 1024;
    const int b = tid * per, e = min(n, b + per);
    // a thread's first kCached counts stay in registers for the three passes
    // (all of them up to 8,192 chunks); the rest are read again
    constexpr int kCached = 8;
    uint32_t cv[kCached];
#pragma unroll
    for (int j = 0; j < kCached; j++) cv[j] = b + j < e ? counts[b + j] : 0u;
    auto items_of = [&](int i) -> uint32_t {
        return items && items[i].edge == 64 ? (uint32_t)kSlabs : 1u;
    };
    // emit items per cached chunk, loaded with the counts
    uint32_t iv[kCached];
#pragma unroll
    for (int j = 0; j < kCached; j++) iv[j] = b + j < e && order_out ? items_of(b + j) : 1u;
    const int ec = min(e, b + kCached);
    uint64_t local = 0;
#pragma unroll
    for (int j = 0; j < kCached; j++) local += cv[j];
    for (int i = b + kCached; i < e; i++) local += counts[i];
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
    const uint64_t off = base_in ? *base_in : 0ull;  // faces before these chunks
    uint64_t run = off + wbase + incl - local;
#pragma unroll
    for (int j = 0; j < kCached; j++) {
        if (b + j < ec) bases[b + j] = run;
        run += cv[j];
    }
    for (int i = b + kCached; i < e; i++) {
        bases[i] = run;
        run += counts[i];
    }
    if (tid == 0) {
        *total = off + sum;
        if (overflow) *overflow = 0u;  // the emit stage that follows sets it (not overflow[1])
    }
    if (!order_out) return;
    static_assert(kOrderBuckets == 1024, "one bucket per thread");
    s_hist[tid] = 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCached; j++)
        if (b + j < ec) atomicAdd(&s_hist[order_bucket(cv[j])], iv[j]);
    for (int i = b + kCached; i < e; i++) atomicAdd(&s_hist[order_bucket(counts[i])], items_of(i));
    __syncthreads();
    const uint32_t cnt = s_hist[tid];
    uint32_t cincl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(cincl, d, 64);
        if (lane >= d) cincl += t;
    }
    __shared__ uint32_t s_w32[16];
    if (lane == 63) s_w32[wave] = cincl;
    __syncthreads();
    uint32_t cbase = 0;
    for (int i = 0; i < wave; i++) cbase += s_w32[i];
    s_hist[tid] = cbase + cincl - cnt;
    __syncthreads();
    auto place = [&](int i, uint32_t c, uint32_t k) {
        const uint32_t pos = atomicAdd(&s_hist[order_bucket(c)], k);
        if (!items) {
            order_out[pos] = i;
            return;
        }
        for (uint32_t q = 0; q < k; q++) order_out[pos + q] = (i << 2) | (int)q;
    };
#pragma unroll
    for (int j = 0; j < kCached; j++)
        if (b + j < ec) place(b + j, cv[j], iv[j]);
    for (int i = b + kCached; i < e; i++) place(i, counts[i], items_of(i));
}

// Descriptors (row | x << 12 | dir << 18) of row r's faces that fall in the
// tile [t0, t1), the row's first face being f, written to S.desc[face - t0]
// in (x, direction) order.
template <int N, int MAXN, class SH>
__device__ __forceinline__ void walk_row(SH &S, int r, uint32_t f, uint32_t t0, uint32_t t1) {
    using R = typename RowType<MAXN>::T;
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

// The faces of row r at voxel x (one lane of a wave walking the row): the
// row's faces before x (in (x, direction) order) come from the masks below
// x, then x's own faces follow in direction order.
template <int N, int MAXN, class SH>
__device__ __forceinline__ void walk_row_at(SH &S, int r, int x, uint32_t f, uint32_t t0,
                                            uint32_t t1) {
    using R = typename RowType<MAXN>::T;
    R m[6];
    face_masks<N, MAXN>(S, r / N, r % N, m);
    const R below = ((R)1 << x) - (R)1;
    uint32_t g = f;
#pragma unroll
    for (int d = 0; d < 6; d++) g += popc(m[d] & below);
#pragma unroll
    for (int d = 0; d < 6; d++) {
        if ((m[d] >> x) & 1) {
            if (g >= t0 && g < t1)
                S.desc[g - t0] = (uint32_t)r | ((uint32_t)x << 12) | ((uint32_t)d << 18);
            g++;
        }
    }
}

// Emits one chunk's faces.  Each thread owns RPT consecutive rows; a block
// scan of the per-thread counts gives every row its first face index, so the
// face order is (y, z, x, direction) exactly as chunk.rs:65-240 pushes them.
// Faces are staged as 32-bit descriptors in LDS tiles (the size of S.desc),
// then one lane per face writes its records.  INTERLEAVE (the emit kernels
// for MAXN >= 32 and the fused kernel for MAXN = 32, whatever the chunk's
// edge; S.rowfirst holds MAXN^2 entries): the rows' first faces go to
// S.rowfirst and the
// descriptor walk deals rows to threads round-robin, since a tile's faces
// often come from a band of consecutive rows (a surface) that consecutive
// ownership (4 or 16 rows per thread) would leave to a handful of threads.
// `parts` workgroups may share a chunk (small batches): part p emits tiles
// p, p + parts, ... and the matching share of the index stores; each part
// repeats the row scan.  With SlabShared the rows are those of one slab,
// layers ylo .. ylo + kSlabLayers - 1, whose first face is face `off` of
// the chunk (indices are chunk-local).
template <int N, int MAXN, bool INTERLEAVE, int BLOCK = kBlock, bool HEAVY = INTERLEAVE,
          class SH>
__device__ void emit_rows(const MeshWork &w, uint64_t base, uint32_t F, const MeshOut &out,
                          SH &S, int part = 0, int parts = 1, int ylo = 0, uint32_t off = 0) {
    using R = typename RowType<MAXN>::T;
    constexpr int ROWS = RowsOf<SH, N>::v;
    // rows per thread (rounded up: with more threads than rows, some
    // threads own none)
    constexpr int RPT = (ROWS + BLOCK - 1) / BLOCK;
    const int tid = threadIdx.x;
    const bool want_v = (out.formats & kFmtVertices) != 0;
    const bool want_i = (out.formats & kFmtInstances) != 0;
    if (want_v) {
        // Indices first: pure arithmetic of the face count, so their stores
        // drain while the rows are scanned and the face descriptors built.
        // quad::INDICES = [0,2,1,3,1,2] + 4*local face; two indices per store
        uint2 *dst = reinterpret_cast<uint2 *>(out.indices + base * 6);
        for (uint32_t q = (uint32_t)(part * BLOCK + tid); q < 3 * F; q += (uint32_t)(parts * BLOCK)) {
            const uint32_t face = q / 3, part = q - face * 3;
            const uint32_t b4 = 4 * (face + off);
            uint2 v;
            if (part == 0) v = make_uint2(b4 + 0, b4 + 2);
            else if (part == 1) v = make_uint2(b4 + 1, b4 + 3);
            else v = make_uint2(b4 + 1, b4 + 2);
            dst[q] = v;
        }
    }
    const int r0 = tid * RPT;
    uint32_t rc[RPT];
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        R m[6];
        const int r = r0 + i;
        uint32_t c = 0;
        if (ROWS % BLOCK == 0 || r < ROWS) {
            face_masks<N, MAXN>(S, r / N, r % N, m);
#pragma unroll
            for (int d = 0; d < 6; d++) c += popc(m[d]);
        }
        rc[i] = c;
        mine += c;
    }
    uint32_t total;
    const uint32_t first = block_exclusive_scan<BLOCK>(mine, S.red, &total);
    (void)total;  // equals F (count kernel); F is authoritative for sizing
    if constexpr (INTERLEAVE) {
        uint32_t f = first;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            if (ROWS % BLOCK == 0 || r0 + i < ROWS) S.rowfirst[r0 + i] = f;
            f += rc[i];
        }
        if constexpr (HEAVY)
            if (tid == 0) S.heavy.n = 0u;
        __syncthreads();
    }

    constexpr uint32_t TILE = sizeof(S.desc) / sizeof(S.desc[0]);
    for (uint32_t t0 = (uint32_t)part * TILE; t0 < F; t0 += (uint32_t)parts * TILE) {
        const uint32_t t1 = min(F, t0 + TILE);
        if constexpr (INTERLEAVE && !HEAVY) {
#pragma unroll 1
            for (int r = tid; r < ROWS; r += BLOCK) {
                const uint32_t f = S.rowfirst[r];
                const uint32_t e = r + 1 < ROWS ? S.rowfirst[r + 1] : F;
                if (e > t0 && f < t1) walk_row<N, MAXN>(S, r, f, t0, t1);
            }
        } else if constexpr (INTERLEAVE) {
#pragma unroll 1
            for (int r = tid; r < ROWS; r += BLOCK) {
                const uint32_t f = S.rowfirst[r];
                const uint32_t e = r + 1 < ROWS ? S.rowfirst[r + 1] : F;
                if (e > t0 && f < t1) {
                    uint32_t slot = kHeavyMax;
                    if (e - f > (uint32_t)kHeavyFaces) slot = atomicAdd(&S.heavy.n, 1u);
                    if (slot < (uint32_t)kHeavyMax) S.heavy.row[slot] = (uint16_t)r;
                    else walk_row<N, MAXN>(S, r, f, t0, t1);
                }
            }
            __syncthreads();
            // heavy rows: 64 / N rows per wave pass, lane = voxel
            {
                constexpr int RPP = 64 / N;
                const int lane = tid & 63, wave = tid >> 6, sub = lane / N, x = lane - sub * N;
                const uint32_t nh = min(S.heavy.n, (uint32_t)kHeavyMax);
#pragma unroll 1
                for (uint32_t h = (uint32_t)(wave * RPP + sub); h < nh; h += (uint32_t)(BLOCK / 64 * RPP)) {
                    const int r = S.heavy.row[h];
                    walk_row_at<N, MAXN>(S, r, x, S.rowfirst[r], t0, t1);
                }
            }
        } else {
            uint32_t f = first;
            if (f < t1 && f + mine > t0) {
#pragma unroll 1
                for (int i = 0; i < RPT; i++) {
                    if (rc[i] && f + rc[i] > t0 && f < t1) walk_row<N, MAXN>(S, r0 + i, f, t0, t1);
                    f += rc[i];
                }
            }
        }
        __syncthreads();
        if constexpr (HEAVY)
            if (tid == 0) S.heavy.n = 0u;  // read by every wave before the barrier above
        // Each wave takes 64 consecutive faces of the tile: lane i builds face
        // q0 + i in registers, stages it in the wave's LDS slot, then the wave
        // writes the 64 records as contiguous 16-B (vertices) or 4-B
        // (instances) lanes.
        const int lane = tid & 63, wave = tid >> 6;
        const uint32_t nt = t1 - t0;
        float4 *st4 = S.stage[wave];
        uint32_t *stw = reinterpret_cast<uint32_t *>(S.stage[wave]);
        for (uint32_t q0 = 64u * wave; q0 < nt; q0 += BLOCK) {
            const uint32_t q = q0 + lane;
            const bool valid = q < nt;
            const uint32_t nv = min(64u, nt - q0);
            const uint32_t dsc = valid ? S.desc[q] : 0u;
            const int r = dsc & 0xFFF, x = (dsc >> 12) & 0x3F, d = (dsc >> 18) & 7;
            const int yl = r / N, z = r - yl * N, y = ylo + yl;
            const int wx = w.ox + x, wy = w.oy + y, wz = w.oz + z;
            const float t[3] = {(float)wx, (float)wy, (float)wz};
            const int sg[3] = {wx < 0, wy < 0, wz < 0};
            const uint64_t face0 = base + t0 + q0;
            // the 64 records go out in two passes of 32 (lanes 0-31, then
            // 32-63), each staged in the wave's 2.5 KB slot
            const int half = lane >> 5, hl = lane & 31;
            if (want_v) {
                float v[20];
#pragma unroll
                for (int k = 0; k < 3; k++) {  // one 16-B LDS read per coordinate
                    const float4 pk = *reinterpret_cast<const float4 *>(&S.P[d][k][sg[k]][0]);
                    v[0 * 5 + k] = pk.x + t[k];
                    v[1 * 5 + k] = pk.y + t[k];
                    v[2 * 5 + k] = pk.z + t[k];
                    v[3 * 5 + k] = pk.w + t[k];
                }
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    v[c * 5 + 3] = (c & 1) ? 1.0f : 0.0f;
                    v[c * 5 + 4] = (c & 2) ? 1.0f : 0.0f;
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (nv <= 32u * h) break;
                    const uint32_t hn = min(32u, nv - 32u * h);
                    if (valid && half == h) {
#pragma unroll
                        for (int k = 0; k < 5; k++)
                            st4[hl * 5 + k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
                    }
                    wave_sync();
                    float4 *dst = reinterpret_cast<float4 *>(out.vertices + (face0 + 32u * h) * 20);
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const uint32_t i = 64u * k + lane;
                        if (i < 5u * hn) dst[i] = st4[i];
                    }
                    wave_sync();
                }
            }
            if (want_i) {
                // A chunk of one block type needs no load.  Otherwise the id
                // is used right where it is loaded (clamped to a table row;
                // ids are at most kBlockDirt, checked at the boundary), so
                // no load is pending at the loop's back edge, where the
                // compiler would wait for all of the group's stores.
                uint32_t blk = w.solid_type;
                if (blk == kBlockMixed) {  // uniform
                    blk = valid ? (uint32_t)w.ids[(size_t)x + (size_t)N * z + (size_t)N * N * y]
                                : (uint32_t)kBlockGrass;
                    blk = min(blk, (uint32_t)kBlockDirt);
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (nv <= 32u * h) break;
                    const uint32_t hn = min(32u, nv - 32u * h);
                    if (valid && half == h) {
                        uint32_t *rec = stw + hl * 17;
#pragma unroll
                        for (int k = 0; k < 3; k++) {  // row k of T*R for its sign
                            const float4 mk = *reinterpret_cast<const float4 *>(&S.Mb[d][sg[k]][k][0]);
                            rec[0 * 4 + k] = __float_as_uint(mk.x);
                            rec[1 * 4 + k] = __float_as_uint(mk.y);
                            rec[2 * 4 + k] = __float_as_uint(mk.z);
                        }
#pragma unroll
                        for (int c = 0; c < 3; c++) rec[c * 4 + 3] = 0u;  // M(3,c) = +0
#pragma unroll
                        for (int k = 0; k < 3; k++) rec[12 + k] = __float_as_uint(t[k]);
                        rec[15] = 0x3F800000u;  // 1.0f
                        rec[16] = S.tex[blk][d];
                    }
                    wave_sync();
                    uint32_t *dst = out.instances + (face0 + 32u * h) * 17;
#pragma unroll
                    for (int k = 0; k < 9; k++) {
                        const uint32_t i = 64u * k + lane;
                        if (i < 17u * hn) dst[i] = stw[i];
                    }
                    wave_sync();
                }
            }
        }
        __syncthreads();
    }
}

template <int N, int MAXN, int BLOCK>
__device__ void emit_chunk(const MeshWork &w, uint64_t base, uint32_t F, const MeshOut &out,
                           MeshShared<MAXN, BLOCK> &S, int part, int parts) {
    load_rows<N, MAXN, BLOCK>(w, S);
    __syncthreads();
    emit_rows<N, MAXN, (MAXN >= 32), BLOCK>(w, base, F, out, S, part, parts);
}

template <int MAXN, int BLOCK = kBlock>
__device__ __forceinline__ void emit_item(const MeshWork &w, uint64_t base, uint32_t F,
                                          const MeshOut &out, const MeshTables *tables,
                                          uint32_t *overflow, MeshShared<MAXN, BLOCK> &S, int part,
                                          int parts) {
    if (base + F > out.capacity) {
        if (threadIdx.x == 0) atomicOr(overflow, 1u);
        return;
    }
    if (F == 0) return;  // an air chunk (or one with no visible face)
    for (int i = threadIdx.x; i < 6 * 3 * 2 * 4; i += BLOCK) (&S.P[0][0][0][0])[i] = (&tables->P[0][0][0][0])[i];
    for (int i = threadIdx.x; i < 6 * 2 * 3 * 4; i += BLOCK) (&S.Mb[0][0][0][0])[i] = (&tables->Mb[0][0][0][0])[i];
    for (int i = threadIdx.x; i < 18; i += BLOCK) (&S.tex[0][0])[i] = (&tables->tex[0][0])[i];
    switch (w.edge) {
    case 16: emit_chunk<16, MAXN, BLOCK>(w, base, F, out, S, part, parts); break;
    case 32: if (MAXN >= 32) emit_chunk<(MAXN >= 32 ? 32 : 16), MAXN, BLOCK>(w, base, F, out, S, part, parts); break;
    case 64: if (MAXN >= 64) emit_chunk<(MAXN >= 64 ? 64 : 16), MAXN, BLOCK>(w, base, F, out, S, part, parts); break;
    default: break;
    }
}

// Register cap: five waves per SIMD (96 VGPRs), matching the five
// workgroups per CU that the 29 KB of LDS allows at MAXN = 32, so more
// workgroups overlap their descriptor and store phases.
template <int MAXN, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5, 8)))
emit_kernel(const MeshWork *__restrict__ work, const int32_t *__restrict__ order,
            const uint64_t *__restrict__ bases, const uint32_t *__restrict__ counts, MeshOut out,
            const MeshTables *__restrict__ tables, uint32_t *__restrict__ overflow, int parts,
            int32_t n) {
    __shared__ MeshShared<MAXN, BLOCK> S;
    const int k = (int)blockIdx.x / parts, part = (int)blockIdx.x - k * parts;
    const int item = order ? order[k] : k;
    if (!order_entry_ok(item, n, overflow + 1)) return;
    emit_item<MAXN, BLOCK>(work[item], bases[item], counts[item], out, tables, overflow, S, part,
                           parts);
}

// One slab of a 64^3 chunk: its faces start `off` faces into the chunk,
// off = the face counts of the slabs below it (sc = the chunk's slab counts).
__device__ void emit_slab_item(const MeshWork &w, uint64_t chunk_base, const uint32_t *sc, int slab,
                               const MeshOut &out, const MeshTables *tables, uint32_t *overflow,
                               SlabShared &S, int part, int parts) {
    uint32_t off = 0;
    for (int q = 0; q < slab; q++) off += sc[q];
    const uint32_t F = sc[slab];
    const uint64_t base = chunk_base + off;
    if (base + F > out.capacity) {
        if (threadIdx.x == 0) atomicOr(overflow, 1u);
        return;
    }
    if (F == 0) return;
    for (int i = threadIdx.x; i < 6 * 3 * 2 * 4; i += kBlock) (&S.P[0][0][0][0])[i] = (&tables->P[0][0][0][0])[i];
    for (int i = threadIdx.x; i < 6 * 2 * 3 * 4; i += kBlock) (&S.Mb[0][0][0][0])[i] = (&tables->Mb[0][0][0][0])[i];
    for (int i = threadIdx.x; i < 18; i += kBlock) (&S.tex[0][0])[i] = (&tables->tex[0][0])[i];
    load_slab(w, S, slab * kSlabLayers);
    __syncthreads();
    emit_rows<64, 64, true, kBlock>(w, base, F, out, S, part, parts, slab * kSlabLayers, off);
}

// Emit of a batch that holds 64^3 chunks, over emit items (codes chunk << 2
// | slab, the scan's order; with no order, code = workgroup / parts): a
// slab of a 64^3 chunk, or a whole smaller chunk (slab 0; a code with
// another slab for such a chunk sets the schedule error word).  Both kinds
// share the workgroup's LDS, 33 KB, four workgroups per CU.
union EmitShared64 {
    MeshShared<32, kBlock> small;
    SlabShared slab;
};

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
emit_slab_kernel(const MeshWork *__restrict__ work, const int32_t *__restrict__ order,
                 const uint64_t *__restrict__ bases, const uint32_t *__restrict__ counts,
                 const uint32_t *__restrict__ slabs, MeshOut out,
                 const MeshTables *__restrict__ tables, uint32_t *__restrict__ overflow, int parts,
                 int32_t n) {
    __shared__ EmitShared64 S;
    const int k = (int)blockIdx.x / parts, part = (int)blockIdx.x - k * parts;
    const int code = order ? order[k] : k;
    const int item = code >> 2, slab = code & 3;
    if (!order_entry_ok(item, n, overflow + 1)) return;
    const MeshWork &w = work[item];
    if (w.edge == 64) {
        emit_slab_item(w, bases[item], slabs + (size_t)kSlabs * item, slab, out, tables, overflow,
                       S.slab, part, parts);
    } else if (slab == 0) {
        emit_item<32, kBlock>(w, bases[item], counts[item], out, tables, overflow, S.small, part,
                              parts);
    } else if (order && threadIdx.x == 0) {
        atomicOr(overflow + 1, 1u);
    }
}

template <int MAXN>
__global__ void __launch_bounds__(kBlock) emit_one_kernel(MeshWork w, uint64_t base, uint32_t F,
                                                          MeshOut out,
                                                          const MeshTables *__restrict__ tables,
                                                          uint32_t *__restrict__ overflow) {
    __shared__ MeshShared<MAXN> S;
    emit_item<MAXN>(w, base, F, out, tables, overflow, S, 0, 1);
}

// --------------------------------------------------------------- fused
// One persistent kernel runs noise -> fill -> cull -> compact for a whole
// batch.  Workgroups take chunks in batch order from a ticket counter; each
// chunk recomputes the neighbour planes it needs (a 4N-column halo, 12.5% of
// the noise work at N = 32) instead of waiting for its neighbours, counts
// its faces, and obtains its first output face with a decoupled look-back
// over the chunks before it (Merrill & Garland, "Single-pass Parallel Prefix
// Scan with Decoupled Look-back", 2016).  A workgroup only waits on tickets
// handed out before its own, i.e. on running workgroups, so the wait always
// ends; it is also bounded (kSpinLimit) and reports through ctl.error.
template <int MAXN> struct FusedShared {
    int perm[512];
    float4 lut[8][16];
    float4 ytab[kMaxOctaves * MAXN];
    uint32_t ychg[MAXN];
    typename RowType<MAXN>::T occ[MAXN * MAXN];
    typename RowType<MAXN>::T nb[4][MAXN];
    uint32_t desc[kDescTile];
    uint32_t rowfirst[MAXN == 32 ? MAXN * MAXN : 1];  // interleaved descriptor walk
    alignas(16) float P[6][3][2][4];
    alignas(16) float Mb[6][2][3][4];
    uint32_t tex[3][6];
    uint32_t red[kBlock / 64];
    float4 stage[kBlock / 64][kStageWords / 4];
    int item;
    int terrain;
    uint64_t base;
};

constexpr uint64_t kFlagAgg = 1ull << 62, kFlagIncl = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 24;

// A status word carries its flag and its value together and no workgroup
// reads anything else another workgroup wrote during the launch, so the
// look-back needs coherence of the word only: relaxed agent-scope atomics.
// Release/acquire would order every earlier store too, which on this part
// means writing back the issuing XCD's whole L2 per publish (the emit phase
// keeps it full of vertex data) and serialises the prefix chain behind it.
__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load(uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back (Merrill & Garland 2016) over the chunk tickets: ticket
// t publishes its face count F, sums its predecessors' counts (64 per round
// trip, stopping at the nearest inclusive prefix), publishes its inclusive
// prefix and leaves its first face in S.base.  Waits are only on earlier
// tickets (held by running workgroups) and bounded (kSpinLimit polls, then
// ctl.error).  Wave 0 of the workgroup runs it; S.base is valid after the
// caller's next barrier.
template <class SH>
__device__ void publish_and_look_back(const FusedCtl &ctl, int t, int n, uint32_t F, SH &S) {
    const int tid = threadIdx.x;
    if (tid < 64) {  // wave 0: publish, then look back 64 predecessors per round trip
        const int lane = tid;
        uint64_t prefix = 0;
        if (t > 0) {
            if (lane == 0) status_store(ctl.status + t, kFlagAgg | F);
            int hi = t - 1;
            uint32_t spins = 0;
            for (;;) {
                const int j = hi - lane;
                const uint64_t st = j >= 0 ? status_load(ctl.status + j) : kFlagIncl;
                const uint64_t flag = st & ~kValMask;
                if (__ballot(flag == 0)) {  // some predecessor has not published yet
                    if (++spins > kSpinLimit) {
                        if (lane == 0) atomicOr(ctl.error, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                // Flags only grow (0 -> aggregate -> inclusive) and an inclusive
                // prefix implies every earlier chunk has published, so the
                // nearest inclusive entry ends the walk.
                const uint64_t incl = __ballot(flag == kFlagIncl);
                const int stop = incl ? __ffsll((unsigned long long)incl) - 1 : 63;
                uint64_t v = lane <= stop ? (st & kValMask) : 0ull;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                prefix += v;
                if (incl) break;
                hi -= 64;
            }
        }
        if (lane == 0) {
            status_store(ctl.status + t, kFlagIncl | (prefix + F));
            ctl.counts[t] = F;
            ctl.bases[t] = prefix;
            if (t == n - 1) *ctl.total = prefix + F;
            S.base = prefix;
        }
    }
}

template <int N, int MAXN>
__device__ void fused_chunk(const FusedWork &W, const TerrainDev &T, int t, int n,
                            const FusedCtl &ctl, const MeshOut &out, FusedShared<MAXN> &S) {
    using R = typename RowType<MAXN>::T;
    constexpr int WW = N == 64 ? 2 : 1;
    const int tid = threadIdx.x;
    const FillWork &w = W.f;
    Halo H{0, 0u};
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (W.halo[k] == kHaloRecompute) H.sides |= (uint32_t)k << (2 * H.n++);
    for (int i = tid; i < 4 * N; i += kBlock) {
        const int k = i / N, y = i - k * N;
        if (W.halo[k] == kHaloRecompute) continue;
        R v = 0;
        const uint32_t *p = W.m.nb[k];
        if (W.halo[k] == kHaloResident && p) {
            if (WW == 2) v = (R)p[2 * y] | ((R)p[2 * y + 1] << (WW == 2 ? 32 : 0));
            else v = (R)p[y];
        }
        S.nb[k][y] = v;
    }
    // Halo lines make N + 4 lines: shorter y segments keep the waves even.
    generate_rows<N, MAXN, false, (N == 64 ? 32 : N / 2), false>(T, w.cx * N, w.cy * N, w.cz * N,
                                                                 H, S, nullptr);
    __syncthreads();
    write_chunk_outputs<N, MAXN>(w, S, true, kBlockGrass);
    const uint32_t F = count_rows<N, MAXN>(S);
    publish_and_look_back(ctl, t, n, F, S);
    __syncthreads();
    const uint64_t base = S.base;
    if (base + F > out.capacity) {
        if (tid == 0) atomicOr(ctl.overflow, 1u);
    } else {
        emit_rows<N, MAXN, (MAXN == 32), kBlock, false>(W.m, base, F, out, S);
    }
    __syncthreads();
}

// Frustum test of resident chunk AABBs folded into the fused launch: every
// workgroup culls a grid-stride slice before it takes chunk tickets
// (frustum_culling.rs:16-40; dist = ((a*x + b*y) + c*z) + d, NaN passes).
__device__ __forceinline__ uint8_t aabb_visible(const float *pl, int cx, int cy, int cz, int edge) {
    // exact 64-bit products rounded to f32: equal to the 32-bit product
    // wherever that does not overflow, defined for every coordinate
    const float px = (float)((long long)cx * edge), py = (float)((long long)cy * edge),
                pz = (float)((long long)cz * edge);
    const float e = (float)edge;
    const float xs[2] = {px, px + e}, ys[2] = {py, py + e}, zs[2] = {pz, pz + e};
    uint32_t any = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const float x = xs[c & 1], y = ys[(c >> 2) & 1], z = zs[(c >> 1) & 1];
        uint32_t inside = 1;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const float dist = pl[4 * k] * x + pl[4 * k + 1] * y + pl[4 * k + 2] * z + pl[4 * k + 3];
            inside &= !(dist < 0.0f);
        }
        any |= inside;
    }
    return (uint8_t)any;
}

__device__ __forceinline__ void cull_slice(const CullJob &job) {
    float pl[24];
#pragma unroll
    for (int i = 0; i < 24; i++) pl[i] = job.planes[i];
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < job.n; i += gridDim.x * kBlock)
        job.mask[i] = aabb_visible(pl, job.coords[3 * i], job.coords[3 * i + 1],
                                   job.coords[3 * i + 2], job.edge);
}

// Register cap: three waves per SIMD (the noise and emit phases together
// otherwise need ~185 VGPRs at MAXN = 32 and leave two).
template <int MAXN>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8)))
fused_kernel(const FusedWork *__restrict__ work, int32_t n,
                                                       const TerrainDev *__restrict__ terrains,
                                                       const GradLut *__restrict__ lut,
                                                       const MeshTables *__restrict__ tables,
                                                       FusedCtl ctl, MeshOut out,
                                                       CullJob cull) {
    __shared__ FusedShared<MAXN> S;
    const int tid = threadIdx.x;
    if (cull.n > 0) cull_slice(cull);
    for (int i = tid; i < 8 * 16; i += kBlock) S.lut[i >> 4][i & 15] = lut->c[i >> 4][i & 15];
    for (int i = tid; i < 6 * 3 * 2 * 4; i += kBlock) (&S.P[0][0][0][0])[i] = (&tables->P[0][0][0][0])[i];
    for (int i = tid; i < 6 * 2 * 3 * 4; i += kBlock) (&S.Mb[0][0][0][0])[i] = (&tables->Mb[0][0][0][0])[i];
    for (int i = tid; i < 18; i += kBlock) (&S.tex[0][0])[i] = (&tables->tex[0][0])[i];
    if (tid == 0) S.terrain = -1;
    for (;;) {
        if (tid == 0) S.item = (int)atomicAdd(ctl.ticket, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(S.item);  // uniform: descriptor loads go scalar
        if (t >= n) break;
        const FusedWork &W = work[t];
        const int terrain = W.f.terrain;
        if (terrain != S.terrain) {
            for (int i = tid; i < 512; i += kBlock) S.perm[i] = terrains[terrain].perm[i];
            __syncthreads();
            if (tid == 0) S.terrain = terrain;
        }
        __syncthreads();
        const TerrainDev &T = terrains[terrain];
        switch (W.f.edge) {
        case 16: fused_chunk<16, MAXN>(W, T, t, n, ctl, out, S); break;
        case 32: if (MAXN >= 32) fused_chunk<(MAXN >= 32 ? 32 : 16), MAXN>(W, T, t, n, ctl, out, S); break;
        case 64: if (MAXN >= 64) fused_chunk<(MAXN >= 64 ? 64 : 16), MAXN>(W, T, t, n, ctl, out, S); break;
        default: break;
        }
    }
}

// Zeroes the control words of a fused run (a kernel rather than a memset so
// the graph that replays the run holds kernel nodes only).
__global__ void __launch_bounds__(kBlock) fused_reset_kernel(FusedCtl ctl, int32_t n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) ctl.status[i] = 0ull;
    if (i < n && ctl.filled) ctl.filled[i] = 0u;
    if (i == 0) {
        *ctl.ticket = 0u;
        if (ctl.mesh_ticket) *ctl.mesh_ticket = 0u;
        *ctl.overflow = 0u;
        *ctl.error = 0u;
        *ctl.total = 0ull;
    }
}

// ------------------------------------------------- pipelined persistent kernel
// One launch for the whole batch, like fused_kernel, with fills and meshes
// as separate tasks taken from two ticket counters:
//
//   fill ticket t:  noise -> ids, rows, planes of chunk t; publish filled[t]
//   mesh ticket m:  rows of m and its neighbours' planes from memory, count,
//                   look-back for the first face, emit
//
// Tickets are taken by atomic add (a compare-and-swap claim of ready
// tickets serialised the workgroups and ran 80 times slower).  One
// workgroup in `ratio` meshes from the start; the others fill, and mesh once
// the fill tickets run out.  So on every CU some workgroups run the
// VALU-bound noise while others run the store-bound emit, at the fill's
// register budget (4 waves per SIMD) and with no halo noise.  Mesh m reads
// the fills of chunks up to m + lag (the farthest in-batch neighbour).
// Waits end: fills never wait, so every fill ticket is handed out and every
// fill publishes while fill workgroups run (the grid is at most what is
// resident at once, and a prefix of it holds both roles); a mesh waits only
// for fills and, in the look-back, for earlier mesh tickets, whose holders
// are running and wait only on earlier tickets; all waits are bounded
// (kSpinLimit polls, then ctl.error).  Rows and planes cross
// workgroups through agent-scope loads and stores (coherent across the XCDs'
// L2 caches without a cache write-back), the stores completed before the
// flag is published.
template <int MAXN> struct PipeShared {
    int perm[512];
    int plut[512];
    float4 lut[8][16];
    float4 ytab[kMaxOctaves * MAXN];
    uint32_t ychg[MAXN];
    typename RowType<MAXN>::T occ[MAXN * MAXN];
    typename RowType<MAXN>::T nb[4][MAXN];
    uint32_t desc[kDescTile / 2];
    uint32_t rowfirst[MAXN * MAXN];  // interleaved descriptor walk
    alignas(16) float P[6][3][2][4];
    alignas(16) float Mb[6][2][3][4];
    uint32_t tex[3][6];
    uint32_t red[kBlock / 64];
    float4 stage[kBlock / 64][kStageWords / 4];
    int item;
    int mitem;
    int terrain;
    uint64_t base;
};

template <int N, int MAXN>
__device__ void pipe_fill(const FillWork &w, const TerrainDev &T, uint32_t *filled,
                          PipeShared<MAXN> &S) {
    const Halo H{0, 0u};
    generate_rows<N, MAXN, false, (N < 32 ? N : 32), true>(T, w.cx * N, w.cy * N, w.cz * N, H, S,
                                                           nullptr);
    __syncthreads();
    write_chunk_outputs<N, MAXN, true>(w, S, true, kBlockGrass);
    __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
    __syncthreads();                // ... and every thread's
    if (threadIdx.x == 0) __hip_atomic_store(filled, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits (bounded) until chunk j's fill has published.
__device__ __forceinline__ void pipe_wait_filled(const FusedCtl &ctl, int j) {
    uint32_t spins = 0;
    while (!__hip_atomic_load(ctl.filled + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        if (++spins > kSpinLimit) {
            atomicOr(ctl.error, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ uint32_t coherent_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int N, int MAXN>
__device__ void pipe_mesh(const FusedWork &W, int m, int n, const FusedCtl &ctl,
                          const MeshOut &out, PipeShared<MAXN> &S) {
    using R = typename RowType<MAXN>::T;
    constexpr int WW = N == 64 ? 2 : 1;
    const int tid = threadIdx.x;
    if (tid == 0) {  // the chunk's own fill and its in-batch neighbours' fills
        pipe_wait_filled(ctl, m);
        for (int k = 0; k < 4; k++)
            if (W.halo[k] == kHaloRecompute) pipe_wait_filled(ctl, W.nbi[k]);
    }
    __syncthreads();
    for (int r = tid; r < N * N; r += kBlock) {
        R v;
        if (WW == 2)
            v = (R)coherent_load(W.f.occ + 2 * r) |
                ((R)coherent_load(W.f.occ + 2 * r + 1) << (WW == 2 ? 32 : 0));
        else
            v = (R)coherent_load(W.f.occ + r);
        S.occ[r] = v;
    }
    for (int i = tid; i < 4 * N; i += kBlock) {
        const int k = i / N, y = i - k * N;
        R v = 0;
        const uint32_t *p = W.m.nb[k];
        if (W.halo[k] == kHaloRecompute) {  // in this batch: written during this launch
            if (WW == 2)
                v = (R)coherent_load(p + 2 * y) |
                    ((R)coherent_load(p + 2 * y + 1) << (WW == 2 ? 32 : 0));
            else
                v = (R)coherent_load(p + y);
        } else if (W.halo[k] == kHaloResident && p) {  // written before this launch
            if (WW == 2) v = (R)p[2 * y] | ((R)p[2 * y + 1] << (WW == 2 ? 32 : 0));
            else v = (R)p[y];
        }
        S.nb[k][y] = v;
    }
    __syncthreads();
    const uint32_t F = count_rows<N, MAXN>(S);
    publish_and_look_back(ctl, m, n, F, S);
    __syncthreads();
    const uint64_t base = S.base;
    if (base + F > out.capacity) {
        if (tid == 0) atomicOr(ctl.overflow, 1u);
    } else {
        emit_rows<N, MAXN, true, kBlock, false>(W.m, base, F, out, S);
    }
    __syncthreads();
}


// Register cap: four waves per SIMD, the fill kernel's budget.  Roles:
// workgroup b meshes from the start when b % ratio == ratio - 1; the others
// fill until the fill tickets run out, then mesh.  The two loops are
// separate so that neither holds the other's values (one loop taking both
// kinds of task kept both sets live and spilled).
template <int MAXN>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
pipe_kernel(const FusedWork *__restrict__ work, int32_t n, int32_t lag,
            const TerrainDev *__restrict__ terrains, const GradLut *__restrict__ lut,
            const MeshTables *__restrict__ tables, FusedCtl ctl, MeshOut out, CullJob cull,
            int32_t ratio) {
    static_assert(MAXN <= 32, "64^3 chunks take the fused or staged form");
    __shared__ PipeShared<MAXN> S;
    const int tid = threadIdx.x;
    if (cull.n > 0) cull_slice(cull);
    for (int i = tid; i < 8 * 16; i += kBlock) S.lut[i >> 4][i & 15] = lut->c[i >> 4][i & 15];
    for (int i = tid; i < 6 * 3 * 2 * 4; i += kBlock) (&S.P[0][0][0][0])[i] = (&tables->P[0][0][0][0])[i];
    for (int i = tid; i < 6 * 2 * 3 * 4; i += kBlock) (&S.Mb[0][0][0][0])[i] = (&tables->Mb[0][0][0][0])[i];
    for (int i = tid; i < 18; i += kBlock) (&S.tex[0][0])[i] = (&tables->tex[0][0])[i];
    if (tid == 0) S.terrain = -1;
    const bool mesher = ratio > 1 && (int)(blockIdx.x % (uint32_t)ratio) == ratio - 1;
    if (!mesher) {
#pragma unroll 1
        for (;;) {
            __syncthreads();
            if (tid == 0) S.item = (int)atomicAdd(ctl.ticket, 1u);
            __syncthreads();
            const int t = __builtin_amdgcn_readfirstlane(S.item);
            if (t >= n) break;
            const FusedWork &W = work[t];
            const int terrain = W.f.terrain;
            if (terrain != S.terrain) {
                for (int i = tid; i < 512; i += kBlock) {
                    const int p = terrains[terrain].perm[i];
                    S.perm[i] = p;
                    S.plut[i] = (p & 15) * 16;
                }
                __syncthreads();
                if (tid == 0) S.terrain = terrain;
            }
            const TerrainDev &T = terrains[terrain];
            if (W.f.edge == 16) pipe_fill<16, MAXN>(W.f, T, ctl.filled + t, S);
            else if (MAXN >= 32) pipe_fill<(MAXN >= 32 ? 32 : 16), MAXN>(W.f, T, ctl.filled + t, S);
        }
    }
    // Meshes in ticket order.  Mesh m needs the fills of m and of its
    // in-batch neighbours (at most m + lag): the workgroup first waits until
    // fill tickets that far have been handed out, then for those fills.
#pragma unroll 1
    for (;;) {
        __syncthreads();
        if (tid == 0) {
            const int mm = (int)atomicAdd(ctl.mesh_ticket, 1u);
            S.mitem = mm;
            if (mm < n) {
                const uint32_t need = (uint32_t)min((int64_t)n, (int64_t)mm + lag + 1);
                uint32_t spins = 0;
                while (__hip_atomic_load(ctl.ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       need) {
                    if (++spins > kSpinLimit) {
                        atomicOr(ctl.error, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
        }
        __syncthreads();
        const int m = __builtin_amdgcn_readfirstlane(S.mitem);
        if (m >= n) break;
        const FusedWork &W = work[m];
        if (W.f.edge == 16) pipe_mesh<16, MAXN>(W, m, n, ctl, out, S);
        else if (MAXN >= 32) pipe_mesh<(MAXN >= 32 ? 32 : 16), MAXN>(W, m, n, ctl, out, S);
    }
}

// --------------------------------------------------------------- frustum
// One thread per chunk AABB: some corner inside all six planes -> 1
// (frustum_culling.rs:16-40; dist = ((a*x + b*y) + c*z) + d, NaN passes).
__global__ void __launch_bounds__(kBlock) frustum_kernel(Planes planes,
                                                         const int32_t *__restrict__ coords,
                                                         int32_t n, int32_t edge,
                                                         uint8_t *__restrict__ mask) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    mask[i] = aabb_visible(planes.p, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2], edge);
}

// --------------------------------------------------------------- draw lists
// World::render (world.rs:34-38) calls ChunkMesh::render (chunk.rs:268-277)
// -- draw_indexed(0..6, 0, 0..instances) on the mesh's own instance buffer --
// for every resident mesh whose chunk passes Frustum::contains.  Here the
// meshes live in one arena (mesh i = faces[i] records from first[i]) and
// the draws become indirect arguments: one pass counts visible meshes per
// workgroup, a second gives each workgroup its offset (the sum of the counts
// before it) and writes the arguments in mesh order.  No atomics; the order
// is deterministic.
constexpr int kDrawPer = 4;                     // meshes per thread
constexpr int kDrawBlock = kBlock * kDrawPer;   // meshes per workgroup

__global__ void __launch_bounds__(kBlock) draw_count_kernel(Planes planes,
                                                            const int32_t *__restrict__ coords,
                                                            int32_t n, int32_t edge,
                                                            uint32_t *__restrict__ block_counts,
                                                            uint32_t *__restrict__ count) {
    __shared__ uint32_t red[kBlock / 64];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kDrawPer; k++) {
        const int i = blockIdx.x * kDrawBlock + k * kBlock + threadIdx.x;
        if (i < n) c += aabb_visible(planes.p, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2], edge);
    }
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        block_counts[blockIdx.x] = c;
        if (blockIdx.x == 0) count[1] = 0u;  // range flag, set by the write pass
    }
}

__global__ void __launch_bounds__(kBlock) draw_write_kernel(Planes planes,
                                                            const int32_t *__restrict__ coords,
                                                            const uint32_t *__restrict__ faces,
                                                            const uint64_t *__restrict__ first,
                                                            int32_t n, int32_t edge, uint32_t mode,
                                                            const uint32_t *__restrict__ block_counts,
                                                            DrawArgs *__restrict__ args,
                                                            uint32_t *__restrict__ count) {
    __shared__ uint32_t red[kBlock / 64];
    const int tid = threadIdx.x;
    uint32_t before = 0;
    for (int j = tid; j < (int)blockIdx.x; j += kBlock) before += block_counts[j];
    const uint32_t off = block_sum(before, red);
    const int i0 = blockIdx.x * kDrawBlock + tid * kDrawPer;  // contiguous per thread: order kept
    uint32_t vis = 0, mine = 0;
#pragma unroll
    for (int k = 0; k < kDrawPer; k++) {
        const int i = i0 + k;
        const uint32_t v = i < n ? aabb_visible(planes.p, coords[3 * i], coords[3 * i + 1],
                                                coords[3 * i + 2], edge)
                                 : 0u;
        vis |= v << k;
        mine += v;
    }
    uint32_t total;
    uint32_t pos = off + block_exclusive_scan(mine, red, &total);
    bool range = true;
#pragma unroll
    for (int k = 0; k < kDrawPer; k++) {
        if (!((vis >> k) & 1u)) continue;
        const uint64_t F = faces[i0 + k], f0 = first[i0 + k];
        DrawArgs a;
        if (mode == kDrawExpanded) {  // chunk-local indices: base_vertex moves them to the arena
            range &= 6 * (f0 + F) <= 0xFFFFFFFFull && 4 * (f0 + F) <= 0x7FFFFFFFull;
            a = DrawArgs{(uint32_t)(6 * F), 1u, (uint32_t)(6 * f0), (int32_t)(4 * f0), 0u};
        } else {                      // ChunkMesh::render: 6 indices, F instances
            range &= f0 + F <= 0xFFFFFFFFull;
            a = DrawArgs{6u, (uint32_t)F, 0u, 0, (uint32_t)f0};
        }
        args[pos++] = a;
    }
    if (!range) atomicOr(count + 1, 1u);
    if (blockIdx.x == gridDim.x - 1 && tid == 0) count[0] = off + total;
}

// --------------------------------------------------------------- launchers
hipError_t launch_fill(const FillWork *work, const int32_t *order, int32_t n, int32_t max_edge,
                       const TerrainDev *terrains, const GradLut *lut, uint32_t *error,
                       int32_t parts_hint, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (parts_hint != 0 && parts_hint != 1 && parts_hint != 2 && parts_hint != 4 && parts_hint != 8)
        return hipErrorInvalidValue;
    // small batches: several workgroups per chunk (density terrains with up
    // to 4 octaves split their rows; other chunks are filled by the first of
    // their workgroups; a chunk under 64^3 given 8 uses the first 4)
    const int parts = fill_parts_for(n, max_edge, parts_hint);
    if (parts == 1) {
        if (max_edge <= 16) fill_kernel<16><<<n, kBlock, 0, s>>>(work, order, n, terrains, lut, error);
        else if (max_edge <= 32) fill_kernel<32><<<n, kBlock, 0, s>>>(work, order, n, terrains, lut, error);
        else fill_kernel<64><<<n, kBlock, 0, s>>>(work, order, n, terrains, lut, error);
        return hipGetLastError();
    }
    const int grid = n * parts;
    if (max_edge <= 16)
        fill_split_kernel<16><<<grid, kBlock, 0, s>>>(work, order, n, terrains, lut, parts, error);
    else if (max_edge <= 32)
        fill_split_kernel<32><<<grid, kBlock, 0, s>>>(work, order, n, terrains, lut, parts, error);
    else
        fill_split_kernel<64><<<grid, kBlock, 0, s>>>(work, order, n, terrains, lut, parts, error);
    return hipGetLastError();
}

hipError_t launch_count(const MeshWork *work, int32_t n, int32_t max_edge, uint32_t *counts,
                        uint32_t *slabs, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_edge <= 16) count_kernel<16><<<n, kBlock, 0, s>>>(work, counts, nullptr);
    else if (max_edge <= 32) count_kernel<32><<<n, kBlock, 0, s>>>(work, counts, nullptr);
    else count_kernel<64><<<n, kBlock, 0, s>>>(work, counts, slabs);
    return hipGetLastError();
}

hipError_t launch_scan(const uint32_t *counts, int32_t n, const MeshWork *items, uint64_t *bases,
                       uint64_t *total, uint32_t *overflow, int32_t *order_out, hipStream_t s,
                       const uint64_t *base_in) {
    if (n <= 0) return hipSuccess;
    scan_kernel<<<1, 1024, 0, s>>>(counts, n, items, bases, total, overflow, order_out, base_in);
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

hipError_t launch_emit(const MeshWork *work, const int32_t *order, int32_t n, int32_t n_items,
                       int32_t max_edge, const uint64_t *bases, const uint32_t *counts,
                       const uint32_t *slabs, MeshOut out, const MeshTables *tables,
                       uint32_t *overflow, int32_t parts_hint, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (parts_hint < 0 || parts_hint > 4) return hipErrorInvalidValue;
    // items: chunks, or for batches that hold 64^3 chunks the scan's emit
    // items (kSlabs codes per chunk without an order)
    const int items = max_edge <= 32 ? n : order ? n_items : kSlabs * n;
    if (items <= 0) return hipErrorInvalidValue;
    // small batches: several workgroups per item, so the grid still covers
    // about a full GPU of workgroups and an item's tiles are emitted side by
    // side
    const int parts = emit_parts_for(items, parts_hint);
    const int grid = items * parts;
    if (max_edge <= 16)
        emit_kernel<16, kBlock><<<grid, kBlock, 0, s>>>(work, order, bases, counts, out, tables,
                                                        overflow, parts, n);
    else if (max_edge <= 32)
        emit_kernel<32, kBlock><<<grid, kBlock, 0, s>>>(work, order, bases, counts, out, tables,
                                                        overflow, parts, n);
    else
        emit_slab_kernel<<<grid, kBlock, 0, s>>>(work, order, bases, counts, slabs, out, tables,
                                                 overflow, parts, n);
    return hipGetLastError();
}

// Grid of the fused kernel: by default one workgroup per chunk.  Workgroups
// that stay resident keep taking tickets; ones dispatched after the tickets
// run out exit at once, so the grid size only bounds the parallelism (the
// autotuner also tries persistent grids of a few workgroups per CU).
static int fused_grid(int n, int grid) { return grid > 0 && grid < n ? grid : n; }

hipError_t launch_fused(const FusedWork *work, int32_t n, int32_t max_edge,
                        const TerrainDev *terrains, const GradLut *lut, const MeshTables *tables,
                        FusedCtl ctl, MeshOut out, const CullJob &cull, int32_t grid,
                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    fused_reset_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(ctl, n);
    if (max_edge <= 16)
        fused_kernel<16><<<fused_grid(n, grid), kBlock, 0, s>>>(work, n, terrains, lut, tables, ctl, out, cull);
    else if (max_edge <= 32)
        fused_kernel<32><<<fused_grid(n, grid), kBlock, 0, s>>>(work, n, terrains, lut, tables, ctl, out, cull);
    else
        fused_kernel<64><<<fused_grid(n, grid), kBlock, 0, s>>>(work, n, terrains, lut, tables, ctl, out, cull);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) cull_kernel(CullJob job) { cull_slice(job); }

// Plain device copy, 16 bytes per lane per step (also the reference point
// for calibrating HBM byte counters: it reads and writes exactly `n16` x 16 B).
__global__ void __launch_bounds__(kBlock) copy_kernel(uint4 *__restrict__ dst,
                                                      const uint4 *__restrict__ src, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t s) {
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return hipSuccess;
    const uint64_t blocks = (n16 + kBlock - 1) / kBlock;
    copy_kernel<<<(unsigned)(blocks < 8192 ? blocks : 8192), kBlock, 0, s>>>(
        (uint4 *)dst, (const uint4 *)src, n16);
    return hipGetLastError();
}

hipError_t launch_pipe(const FusedWork *work, int32_t n, int32_t max_edge, int32_t lag,
                       const TerrainDev *terrains, const GradLut *lut, const MeshTables *tables,
                       FusedCtl ctl, MeshOut out, const CullJob &cull, int32_t grid, int32_t ratio,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_edge > 32 || !ctl.filled || !ctl.mesh_ticket || lag < 0) return hipErrorInvalidValue;
    if (ratio == 0) ratio = kPipeRatio;
    if (ratio < 2 || ratio > kPipeRatioMax) return hipErrorInvalidValue;
    fused_reset_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(ctl, n);
    if (max_edge <= 16)
        pipe_kernel<16><<<fused_grid(n, grid), kBlock, 0, s>>>(work, n, lag, terrains, lut, tables,
                                                               ctl, out, cull, ratio);
    else
        pipe_kernel<32><<<fused_grid(n, grid), kBlock, 0, s>>>(work, n, lag, terrains, lut, tables,
                                                               ctl, out, cull, ratio);
    return hipGetLastError();
}

hipError_t launch_cull(const CullJob &job, hipStream_t s) {
    if (job.n <= 0) return hipSuccess;
    cull_kernel<<<(job.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(job);
    return hipGetLastError();
}

hipError_t launch_frustum(const Planes &planes, const int32_t *coords, int32_t n, int32_t edge,
                          uint8_t *mask, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    frustum_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(planes, coords, n, edge, mask);
    return hipGetLastError();
}

int32_t draw_blocks(int32_t n) { return (n + kDrawBlock - 1) / kDrawBlock; }

// Scratch of the draw-list entry point that allocates nothing: the count
// pass's per-workgroup counts, one u32 per kDrawBlock meshes.  One copy per
// GPU, held until the process ends; every call overwrites the entries it
// uses before reading them.  Host code never passes the array's name to a
// kernel: it is a host-side stand-in, and the device address differs per
// GPU (draw_scratch_address).
constexpr int32_t kDrawScratchBlocks = 4096;
static __device__ __align__(256) uint32_t scratch[kDrawScratchBlocks];

int32_t draw_scratch_max_meshes() { return kDrawScratchBlocks * kDrawBlock; }

hipError_t draw_scratch_address(int device, uint32_t **ptr) {
    static std::mutex mu;
    static uint32_t *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr[device]) {
        // the caller has made `device` current: this is that GPU's copy
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<uint32_t *>(p);
    }
    *ptr = addr[device];
    return hipSuccess;
}

__global__ void __launch_bounds__(64) clear_words_kernel(uint32_t *__restrict__ p, int32_t n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

hipError_t launch_clear_words(uint32_t *p, int32_t n, hipStream_t s) {
    if (n <= 0 || n > 64) return hipErrorInvalidValue;
    clear_words_kernel<<<1, 64, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_draw_list(const Planes &planes, const int32_t *coords, const uint32_t *faces,
                            const uint64_t *first, int32_t n, int32_t edge, uint32_t mode,
                            uint32_t *block_counts, DrawArgs *args, uint32_t *count, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int32_t nb = draw_blocks(n);
    draw_count_kernel<<<nb, kBlock, 0, s>>>(planes, coords, n, edge, block_counts, count);
    draw_write_kernel<<<nb, kBlock, 0, s>>>(planes, coords, faces, first, n, edge, mode, block_counts,
                                           args, count);
    return hipGetLastError();
}

}  // namespace vx
