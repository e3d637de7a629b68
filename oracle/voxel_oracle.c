/*
 * voxel_oracle.c -- CPU restatement of the reference chunk pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline); never linked into
 * the product library.  Each function cites the reference lines it follows;
 * the header lists what pins each one.
 *
 * Float rules: every expression is evaluated in IEEE binary32, one rounding
 * per operation, left to right as written, no FMA (-ffp-contract=off), no
 * fast-math.  Rust does not contract either, so this matches the reference's
 * own arithmetic wherever the reference's inputs are reproduced.
 */
#include "voxel_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ------------------------------------------------------------------------ */
/* Seeded permutation: splitmix64 (Steele, Lea, Flood 2014; Vigna's public    */
/* reference code) driving a Fisher-Yates shuffle from i = 255 down to 1.     */
/* ------------------------------------------------------------------------ */
uint64_t vxo_splitmix64_next(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void vxo_perm(uint64_t seed, uint8_t perm[512]) {
    uint8_t p[256];
    for (int i = 0; i < 256; i++) p[i] = (uint8_t)i;
    uint64_t s = seed;
    for (int i = 255; i >= 1; i--) {
        uint64_t j = vxo_splitmix64_next(&s) % (uint64_t)(i + 1);
        uint8_t t = p[i];
        p[i] = p[j];
        p[j] = t;
    }
    vxo_perm_from_table(p, perm);
}

void vxo_perm_from_table(const uint8_t table[256], uint8_t perm[512]) {
    for (int i = 0; i < 256; i++) {
        perm[i] = table[i];
        perm[256 + i] = table[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Improved Noise (K. Perlin, "Improving Noise", SIGGRAPH 2002, and his      */
/* reference ImprovedNoise.java).  One source text, instantiated for float   */
/* (the pipeline) and double (to check Perlin's published known answer).     */
/* Lattice cell mod 256 of a floored coordinate: (int)fl & 255 while the
 * conversion is in range, else 0 -- every f32 of magnitude >= 2^31 is a
 * multiple of 256 (and NaN, from an infinite coordinate, maps to 0 too; the
 * AVX2 path's _mm256_cvttps_epi32 gives 0x80000000 there, also 0 mod 256). */
#define VXO_CELL8(fl) (fabs((double)(fl)) < 2147483648.0 ? ((int)(fl) & 255) : 0)

/* ------------------------------------------------------------------------ */
#define VXO_DEFINE_NOISE3(NAME, T, FLOOR)                                                   \
    static inline T NAME##_fade(T t) { return ((t * t) * t) * (t * (t * (T)6 - (T)15) + (T)10); } \
    static inline T NAME##_lerp(T t, T a, T b) { return a + t * (b - a); }                  \
    static inline T NAME##_grad(int hash, T x, T y, T z) {                                  \
        int h = hash & 15;                                                                  \
        T u = h < 8 ? x : y;                                                                \
        T v = h < 4 ? y : (h == 12 || h == 14 ? x : z);                                     \
        return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);                           \
    }                                                                                       \
    static T NAME(const uint8_t *p, T x, T y, T z) {                                        \
        T fx = FLOOR(x), fy = FLOOR(y), fz = FLOOR(z);                                      \
        int X = VXO_CELL8(fx), Y = VXO_CELL8(fy), Z = VXO_CELL8(fz);                        \
        x = x - fx;                                                                         \
        y = y - fy;                                                                         \
        z = z - fz;                                                                         \
        T u = NAME##_fade(x), v = NAME##_fade(y), w = NAME##_fade(z);                       \
        int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;                                 \
        int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;                             \
        T x1 = x - (T)1, y1 = y - (T)1, z1 = z - (T)1;                                      \
        return NAME##_lerp(                                                                 \
            w,                                                                              \
            NAME##_lerp(v, NAME##_lerp(u, NAME##_grad(p[AA], x, y, z), NAME##_grad(p[BA], x1, y, z)), \
                        NAME##_lerp(u, NAME##_grad(p[AB], x, y1, z), NAME##_grad(p[BB], x1, y1, z))), \
            NAME##_lerp(v,                                                                  \
                        NAME##_lerp(u, NAME##_grad(p[AA + 1], x, y, z1),                    \
                                    NAME##_grad(p[BA + 1], x1, y, z1)),                     \
                        NAME##_lerp(u, NAME##_grad(p[AB + 1], x, y1, z1),                   \
                                    NAME##_grad(p[BB + 1], x1, y1, z1))));                  \
    }

VXO_DEFINE_NOISE3(noise3f, float, floorf)
VXO_DEFINE_NOISE3(noise3d, double, floor)

float vxo_noise3(const uint8_t *perm, float x, float y, float z) { return noise3f(perm, x, y, z); }
double vxo_noise3d(const uint8_t *perm, double x, double y, double z) {
    return noise3d(perm, x, y, z);
}

/* 2-D form of the same construction: quintic fade, hash p[p[X]+Y], eight
 * gradient directions (+-1,+-1), (+-1,0), (0,+-1). */
static inline float grad2f(int hash, float x, float y) {
    switch (hash & 7) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x;
    case 5: return -x;
    case 6: return y;
    default: return -y;
    }
}

float vxo_noise2(const uint8_t *p, float x, float y) {
    float fx = floorf(x), fy = floorf(y);
    int X = VXO_CELL8(fx), Y = VXO_CELL8(fy);
    x = x - fx;
    y = y - fy;
    float u = noise3f_fade(x), v = noise3f_fade(y);
    int A = p[X] + Y, B = p[X + 1] + Y;
    float x1 = x - 1.0f, y1 = y - 1.0f;
    return noise3f_lerp(v, noise3f_lerp(u, grad2f(p[A], x, y), grad2f(p[B], x1, y)),
                        noise3f_lerp(u, grad2f(p[A + 1], x, y1), grad2f(p[B + 1], x1, y1)));
}

/* fBm: sum = sum + amp*noise(p*freq); freq *= lacunarity; amp *= gain. */
float vxo_fbm3(const uint8_t *perm, float x, float y, float z, int octaves, float f0,
               float lacunarity, float gain) {
    float sum = 0.0f, amp = 1.0f, f = f0;
    for (int o = 0; o < octaves; o++) {
        sum = sum + amp * noise3f(perm, x * f, y * f, z * f);
        f = f * lacunarity;
        amp = amp * gain;
    }
    return sum;
}

float vxo_fbm2(const uint8_t *perm, float x, float y, int octaves, float f0, float lacunarity,
               float gain) {
    float sum = 0.0f, amp = 1.0f, f = f0;
    for (int o = 0; o < octaves; o++) {
        sum = sum + amp * vxo_noise2(perm, x * f, y * f);
        f = f * lacunarity;
        amp = amp * gain;
    }
    return sum;
}

/* ---- 8-wide AVX2 form of vxo_fbm3 (CPU baseline; the reference's noise is
 * AVX2 SIMD, perlin_noise.rs:20).  Same operations in the same order. ---- */
#if defined(__x86_64__)
__attribute__((target("avx2"))) static inline __m256 v_fade(__m256 t) {
    __m256 t3 = _mm256_mul_ps(_mm256_mul_ps(t, t), t);
    __m256 in = _mm256_sub_ps(_mm256_mul_ps(t, _mm256_set1_ps(6.0f)), _mm256_set1_ps(15.0f));
    in = _mm256_add_ps(_mm256_mul_ps(t, in), _mm256_set1_ps(10.0f));
    return _mm256_mul_ps(t3, in);
}
__attribute__((target("avx2"))) static inline __m256 v_lerp(__m256 t, __m256 a, __m256 b) {
    return _mm256_add_ps(a, _mm256_mul_ps(t, _mm256_sub_ps(b, a)));
}
__attribute__((target("avx2"))) static inline __m256 v_grad(__m256i hash, __m256 x, __m256 y,
                                                            __m256 z) {
    __m256i h = _mm256_and_si256(hash, _mm256_set1_epi32(15));
    __m256 lt8 = _mm256_castsi256_ps(_mm256_cmpgt_epi32(_mm256_set1_epi32(8), h));
    __m256 lt4 = _mm256_castsi256_ps(_mm256_cmpgt_epi32(_mm256_set1_epi32(4), h));
    __m256 is12or14 = _mm256_castsi256_ps(
        _mm256_or_si256(_mm256_cmpeq_epi32(h, _mm256_set1_epi32(12)),
                        _mm256_cmpeq_epi32(h, _mm256_set1_epi32(14))));
    __m256 u = _mm256_blendv_ps(y, x, lt8);
    __m256 v = _mm256_blendv_ps(_mm256_blendv_ps(z, x, is12or14), y, lt4);
    __m256i s1 = _mm256_slli_epi32(_mm256_and_si256(h, _mm256_set1_epi32(1)), 31);
    __m256i s2 = _mm256_slli_epi32(_mm256_and_si256(h, _mm256_set1_epi32(2)), 30);
    u = _mm256_xor_ps(u, _mm256_castsi256_ps(s1));
    v = _mm256_xor_ps(v, _mm256_castsi256_ps(s2));
    return _mm256_add_ps(u, v);
}
__attribute__((target("avx2"))) static __m256 v_noise3(const int32_t *p, __m256 x, __m256 y,
                                                       __m256 z) {
    __m256 fx = _mm256_floor_ps(x), fy = _mm256_floor_ps(y), fz = _mm256_floor_ps(z);
    __m256i m255 = _mm256_set1_epi32(255), one = _mm256_set1_epi32(1);
    __m256i X = _mm256_and_si256(_mm256_cvttps_epi32(fx), m255);
    __m256i Y = _mm256_and_si256(_mm256_cvttps_epi32(fy), m255);
    __m256i Z = _mm256_and_si256(_mm256_cvttps_epi32(fz), m255);
    x = _mm256_sub_ps(x, fx);
    y = _mm256_sub_ps(y, fy);
    z = _mm256_sub_ps(z, fz);
    __m256 u = v_fade(x), v = v_fade(y), w = v_fade(z);
    __m256i A = _mm256_add_epi32(_mm256_i32gather_epi32(p, X, 4), Y);
    __m256i AA = _mm256_add_epi32(_mm256_i32gather_epi32(p, A, 4), Z);
    __m256i AB = _mm256_add_epi32(_mm256_i32gather_epi32(p, _mm256_add_epi32(A, one), 4), Z);
    __m256i B = _mm256_add_epi32(_mm256_i32gather_epi32(p, _mm256_add_epi32(X, one), 4), Y);
    __m256i BA = _mm256_add_epi32(_mm256_i32gather_epi32(p, B, 4), Z);
    __m256i BB = _mm256_add_epi32(_mm256_i32gather_epi32(p, _mm256_add_epi32(B, one), 4), Z);
    __m256 c1 = _mm256_set1_ps(1.0f);
    __m256 x1 = _mm256_sub_ps(x, c1), y1 = _mm256_sub_ps(y, c1), z1 = _mm256_sub_ps(z, c1);
#define G(idx) _mm256_i32gather_epi32(p, (idx), 4)
#define G1(idx) _mm256_i32gather_epi32(p, _mm256_add_epi32((idx), one), 4)
    __m256 l1 = v_lerp(u, v_grad(G(AA), x, y, z), v_grad(G(BA), x1, y, z));
    __m256 l2 = v_lerp(u, v_grad(G(AB), x, y1, z), v_grad(G(BB), x1, y1, z));
    __m256 l3 = v_lerp(u, v_grad(G1(AA), x, y, z1), v_grad(G1(BA), x1, y, z1));
    __m256 l4 = v_lerp(u, v_grad(G1(AB), x, y1, z1), v_grad(G1(BB), x1, y1, z1));
#undef G
#undef G1
    return v_lerp(w, v_lerp(v, l1, l2), v_lerp(v, l3, l4));
}
__attribute__((target("avx2"))) static void fbm3_row8_avx2(const int32_t *p32, int32_t x0,
                                                           int32_t y, int32_t z, int octaves,
                                                           float f0, float lac, float gain,
                                                           float out[8]) {
    __m256 xs = _mm256_cvtepi32_ps(
        _mm256_add_epi32(_mm256_set1_epi32(x0), _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7)));
    __m256 ys = _mm256_set1_ps((float)y), zs = _mm256_set1_ps((float)z);
    __m256 sum = _mm256_setzero_ps();
    float amp = 1.0f, f = f0;
    for (int o = 0; o < octaves; o++) {
        __m256 vf = _mm256_set1_ps(f);
        __m256 n = v_noise3(p32, _mm256_mul_ps(xs, vf), _mm256_mul_ps(ys, vf),
                            _mm256_mul_ps(zs, vf));
        sum = _mm256_add_ps(sum, _mm256_mul_ps(_mm256_set1_ps(amp), n));
        f = f * lac;
        amp = amp * gain;
    }
    _mm256_storeu_ps(out, sum);
}
#endif

int vxo_have_avx2(void) {
#if defined(__x86_64__)
    return __builtin_cpu_supports("avx2") ? 1 : 0;
#else
    return 0;
#endif
}

static void fbm3_row8_with(const uint8_t *perm, const int32_t *p32, int32_t x0, int32_t y,
                           int32_t z, int octaves, float f0, float lac, float gain, float out[8]) {
#if defined(__x86_64__)
    if (p32 && vxo_have_avx2()) {
        fbm3_row8_avx2(p32, x0, y, z, octaves, f0, lac, gain, out);
        return;
    }
#endif
    (void)p32;
    for (int i = 0; i < 8; i++)
        out[i] = vxo_fbm3(perm, (float)(x0 + i), (float)y, (float)z, octaves, f0, lac, gain);
}

void vxo_fbm3_row8(const uint8_t *perm, int32_t x0, int32_t y, int32_t z, int octaves, float f0,
                   float lacunarity, float gain, float out[8]) {
    int32_t p32[512];
    for (int i = 0; i < 512; i++) p32[i] = perm[i];
    fbm3_row8_with(perm, p32, x0, y, z, octaves, f0, lacunarity, gain, out);
}

/* ------------------------------------------------------------------------ */
/* Fill (chunk.rs:27-46): loop y, z, x; one byte per voxel at x+N*z+N*N*y.   */
/* The reference's rule is `noise < 0 -> GRASS`; the configured rules are:   */
/*   heightmap: n = fbm2(wx*f, wz*f); h = floor(height_scale*(n+1)/2),       */
/*              clamped to [0, height_max]; solid iff wy <= h               */
/*   density:   solid iff fbm3(wx*f, wy*f, wz*f) - wy*y_scale > 0            */
/*   reference: solid iff fbm3(wx*f, wz*f, wy*f) < 0 -- the reference's own  */
/*              rule (chunk.rs:38-42) and axis order: perlin_noise.rs:6-13   */
/*              passes world z as the noise's y offset and world y as its z  */
/*              offset, so noise[x + 32z + 1024y] is sampled at (x, z, y)    */
/* with (wx, wy, wz) = chunk*edge + local (coordinate.rs:60-65 generalised). */
/* ------------------------------------------------------------------------ */
static inline int32_t heightmap_h(const vxo_terrain *t, const uint8_t *perm, int32_t wx,
                                  int32_t wz) {
    float n = vxo_fbm2(perm, (float)wx, (float)wz, t->octaves, t->frequency, t->lacunarity,
                       t->gain);
    float hf = floorf((t->height_scale * (n + 1.0f)) / 2.0f);
    int32_t h;
    if (!(hf >= 0.0f)) h = 0; /* also catches NaN */
    else if (hf > (float)t->height_max) h = t->height_max;
    else h = (int32_t)hf;
    return h;
}

static void chunk_fill_impl(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz, int32_t N,
                            uint8_t *ids, int fast) {
    const int32_t ox = cx * N, oy = cy * N, oz = cz * N;
    const size_t NN = (size_t)N * (size_t)N;
    uint8_t perm[512];
    int32_t p32[512];
    if (t->kind == VXO_TERRAIN_AIR) {
        memset(ids, VXO_AIR, NN * (size_t)N);
        return;
    }
    if (t->kind == VXO_TERRAIN_SOLID) {
        memset(ids, VXO_GRASS, NN * (size_t)N);
        return;
    }
    vxo_perm(t->seed, perm);
    for (int i = 0; i < 512; i++) p32[i] = perm[i];
    if (t->kind == VXO_TERRAIN_HEIGHTMAP) {
        int32_t *h = (int32_t *)malloc(NN * sizeof(int32_t));
        for (int32_t z = 0; z < N; z++)
            for (int32_t x = 0; x < N; x++) h[(size_t)z * N + x] = heightmap_h(t, perm, ox + x, oz + z);
        for (int32_t y = 0; y < N; y++)
            for (int32_t z = 0; z < N; z++)
                for (int32_t x = 0; x < N; x++)
                    ids[(size_t)x + (size_t)N * z + NN * y] =
                        (oy + y) <= h[(size_t)z * N + x] ? VXO_GRASS : VXO_AIR;
        free(h);
        return;
    }
    if (t->kind == VXO_TERRAIN_REFERENCE) {
        for (int32_t y = 0; y < N; y++) {
            const int32_t wy = oy + y;
            for (int32_t z = 0; z < N; z++) {
                const int32_t wz = oz + z;
                int32_t x = 0;
                if (fast) {
                    for (; x + 8 <= N; x += 8) {
                        float d[8];
                        fbm3_row8_with(perm, p32, ox + x, wz, wy, t->octaves, t->frequency,
                                       t->lacunarity, t->gain, d);
                        for (int i = 0; i < 8; i++)
                            ids[(size_t)(x + i) + (size_t)N * z + NN * y] =
                                d[i] < 0.0f ? VXO_GRASS : VXO_AIR;
                    }
                }
                for (; x < N; x++) {
                    float d = vxo_fbm3(perm, (float)(ox + x), (float)wz, (float)wy, t->octaves,
                                       t->frequency, t->lacunarity, t->gain);
                    ids[(size_t)x + (size_t)N * z + NN * y] = d < 0.0f ? VXO_GRASS : VXO_AIR;
                }
            }
        }
        return;
    }
    /* density */
    for (int32_t y = 0; y < N; y++) {
        const int32_t wy = oy + y;
        const float thr = (float)wy * t->y_scale;
        for (int32_t z = 0; z < N; z++) {
            const int32_t wz = oz + z;
            int32_t x = 0;
            if (fast) {
                for (; x + 8 <= N; x += 8) {
                    float d[8];
                    fbm3_row8_with(perm, p32, ox + x, wy, wz, t->octaves, t->frequency,
                                   t->lacunarity, t->gain, d);
                    for (int i = 0; i < 8; i++)
                        ids[(size_t)(x + i) + (size_t)N * z + NN * y] =
                            (d[i] - thr > 0.0f) ? VXO_GRASS : VXO_AIR;
                }
            }
            for (; x < N; x++) {
                float d = vxo_fbm3(perm, (float)(ox + x), (float)wy, (float)wz, t->octaves,
                                   t->frequency, t->lacunarity, t->gain);
                ids[(size_t)x + (size_t)N * z + NN * y] = (d - thr > 0.0f) ? VXO_GRASS : VXO_AIR;
            }
        }
    }
}

void vxo_chunk_fill(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz, int32_t edge,
                    uint8_t *ids) {
    chunk_fill_impl(t, cx, cy, cz, edge, ids, 0);
}

void vxo_chunk_fill_fast(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz,
                         int32_t edge, uint8_t *ids) {
    chunk_fill_impl(t, cx, cy, cz, edge, ids, 1);
}

void vxo_noise_block(const vxo_terrain *t, int32_t cx, int32_t cy, int32_t cz, int32_t N,
                     float *out) {
    uint8_t perm[512];
    vxo_perm(t->seed, perm);
    const int32_t ox = cx * N, oy = cy * N, oz = cz * N;
    if (t->kind == VXO_TERRAIN_HEIGHTMAP) {
        for (int32_t z = 0; z < N; z++)
            for (int32_t x = 0; x < N; x++)
                out[(size_t)x + (size_t)N * z] =
                    vxo_fbm2(perm, (float)(ox + x), (float)(oz + z), t->octaves, t->frequency,
                             t->lacunarity, t->gain);
        return;
    }
    const int ref = t->kind == VXO_TERRAIN_REFERENCE; /* noise axes (x, z, y) */
    for (int32_t y = 0; y < N; y++)
        for (int32_t z = 0; z < N; z++)
            for (int32_t x = 0; x < N; x++) {
                const float wx = (float)(ox + x), wy = (float)(oy + y), wz = (float)(oz + z);
                out[(size_t)x + (size_t)N * z + (size_t)N * N * y] =
                    vxo_fbm3(perm, wx, ref ? wz : wy, ref ? wy : wz, t->octaves, t->frequency,
                             t->lacunarity, t->gain);
            }
}

void vxo_occupancy(const uint8_t *ids, int32_t N, uint32_t *occ) {
    const int32_t W = (N + 31) / 32;
    for (int32_t y = 0; y < N; y++)
        for (int32_t z = 0; z < N; z++) {
            uint32_t *row = occ + ((size_t)y * N + z) * W;
            for (int32_t w = 0; w < W; w++) row[w] = 0;
            for (int32_t x = 0; x < N; x++)
                if (ids[(size_t)x + (size_t)N * z + (size_t)N * N * y] != VXO_AIR)
                    row[x >> 5] |= 1u << (x & 31);
        }
}

/* ------------------------------------------------------------------------ */
/* Texture layers: TEXTURE_INDEX_LIST (texture.rs:238-243) GRASS [0,0,0,0,2,1] */
/* DIRT [1;6]; slot per face k: LEFT ti[0], RIGHT ti[1], BACK ti[2], FRONT     */
/* ti[3], DOWN ti[5], UP ti[4] (chunk.rs:85,118,149,182,213,231).              */
/* ------------------------------------------------------------------------ */
static const uint32_t kTexList[2][6] = {{0, 0, 0, 0, 2, 1}, {1, 1, 1, 1, 1, 1}};
static const uint32_t kTexSlot[6] = {0, 1, 2, 3, 5, 4};

uint32_t vxo_texture_layer(uint8_t block, uint32_t dir) {
    return kTexList[block - 1][kTexSlot[dir]];
}

/* ------------------------------------------------------------------------ */
/* cull_unseen_triangles (chunk.rs:59-245), restated branch for branch.       */
/* ------------------------------------------------------------------------ */
int64_t vxo_mesh_faces(const uint8_t *ids, int32_t N, int32_t ox, int32_t oy, int32_t oz,
                       const uint8_t *const adj[4], vxo_face *out, int64_t cap) {
    const size_t NN = (size_t)N * (size_t)N;
#define CELL(a, X, Y, Z) ((a)[(size_t)(X) + (size_t)N * (size_t)(Z) + NN * (size_t)(Y)])
    int64_t n = 0;
#define EMIT(D)                                                                              \
    do {                                                                                     \
        if (n < cap) {                                                                       \
            out[n].x = px;                                                                   \
            out[n].y = py;                                                                   \
            out[n].z = pz;                                                                   \
            out[n].dir = (D);                                                                \
            out[n].tex = vxo_texture_layer(c, (D));                                          \
        }                                                                                    \
        n++;                                                                                 \
    } while (0)
    for (int32_t y = 0; y < N; y++) {
        const int32_t py = y + oy;
        for (int32_t z = 0; z < N; z++) {
            const int32_t pz = z + oz;
            for (int32_t x = 0; x < N; x++) {
                const uint8_t c = CELL(ids, x, y, z);
                if (c == VXO_AIR) continue; /* chunk.rs:71-73 */
                const int32_t px = x + ox;
                /* LEFT, chunk.rs:77-109 */
                if (x > 0) {
                    if (CELL(ids, x - 1, y, z) == VXO_AIR) EMIT(VXO_LEFT);
                } else if (adj && adj[0]) {
                    if (CELL(adj[0], N - 1, y, z) == VXO_AIR) EMIT(VXO_LEFT);
                } else {
                    EMIT(VXO_LEFT);
                }
                /* RIGHT, chunk.rs:110-140 */
                if (x < N - 1) {
                    if (CELL(ids, x + 1, y, z) == VXO_AIR) EMIT(VXO_RIGHT);
                } else if (adj && adj[1]) {
                    if (CELL(adj[1], 0, y, z) == VXO_AIR) EMIT(VXO_RIGHT);
                } else {
                    EMIT(VXO_RIGHT);
                }
                /* BACK, chunk.rs:141-173 */
                if (z > 0) {
                    if (CELL(ids, x, y, z - 1) == VXO_AIR) EMIT(VXO_BACK);
                } else if (adj && adj[2]) {
                    if (CELL(adj[2], x, y, N - 1) == VXO_AIR) EMIT(VXO_BACK);
                } else {
                    EMIT(VXO_BACK);
                }
                /* FRONT, chunk.rs:174-204 */
                if (z < N - 1) {
                    if (CELL(ids, x, y, z + 1) == VXO_AIR) EMIT(VXO_FRONT);
                } else if (adj && adj[3]) {
                    if (CELL(adj[3], x, y, 0) == VXO_AIR) EMIT(VXO_FRONT);
                } else {
                    EMIT(VXO_FRONT);
                }
                /* DOWN, chunk.rs:205-222: the bottom layer always emits */
                if (y > 0) {
                    if (CELL(ids, x, y - 1, z) == VXO_AIR) EMIT(VXO_DOWN);
                } else {
                    EMIT(VXO_DOWN);
                }
                /* UP, chunk.rs:223-240: the top layer always emits */
                if (y < N - 1) {
                    if (CELL(ids, x, y + 1, z) == VXO_AIR) EMIT(VXO_UP);
                } else {
                    EMIT(VXO_UP);
                }
            }
        }
    }
#undef EMIT
#undef CELL
    return n;
}

/* ------------------------------------------------------------------------ */
/* Quad::new (quad.rs:14-46): Rotation3::from_euler_angles(roll, pitch, yaw) */
/* with UP roll -90, DOWN roll 90, LEFT pitch -90, RIGHT pitch 90, FRONT 0,  */
/* BACK pitch 180 degrees.  Degrees -> radians as Rust's f32::to_radians:    */
/* deg * (PI_f32 / 180).  nalgebra 0.29.0 from_euler_angles builds           */
/*   [cy*cp, cy*sp*sr - sy*cr, cy*sp*cr + sy*sr]                             */
/*   [sy*cp, sy*sp*sr + cy*cr, sy*sp*cr - cy*sr]                             */
/*   [-sp,   cp*sr,            cp*cr          ]                              */
/* with (s, c) = sin_cos of roll r, pitch p, yaw y (libm sinf/cosf).         */
/* ------------------------------------------------------------------------ */
void vxo_rotation(uint32_t dir, float r[3][3]) {
    const float pi_f = 3.14159265358979323846f;
    const float k = pi_f / 180.0f;
    float roll = 0.0f, pitch = 0.0f, yaw = 0.0f;
    switch (dir) {
    case VXO_UP: roll = -90.0f * k; break;
    case VXO_DOWN: roll = 90.0f * k; break;
    case VXO_LEFT: pitch = -90.0f * k; break;
    case VXO_RIGHT: pitch = 90.0f * k; break;
    case VXO_FRONT: break;
    case VXO_BACK: pitch = 180.0f * k; break;
    default: break;
    }
    const float sr = sinf(roll), cr = cosf(roll);
    const float sp = sinf(pitch), cp = cosf(pitch);
    const float sy = sinf(yaw), cy = cosf(yaw);
    r[0][0] = cy * cp;
    r[0][1] = cy * sp * sr - sy * cr;
    r[0][2] = cy * sp * cr + sy * sr;
    r[1][0] = sy * cp;
    r[1][1] = sy * sp * sr + cy * cr;
    r[1][2] = sy * sp * cr - cy * sr;
    r[2][0] = -sp;
    r[2][1] = cp * sr;
    r[2][2] = cp * cr;
}

/* Quad::to_raw (quad.rs:47-54): T.to_homogeneous() * R.to_homogeneous(),
 * element (i,j) = sum_k T[i][k]*H[k][j] for k = 0..3, stored column-major. */
void vxo_instance_of(const vxo_face *f, vxo_instance *out) {
    float R[3][3];
    vxo_rotation(f->dir, R);
    float T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    T[0][3] = (float)f->x;
    T[1][3] = (float)f->y;
    T[2][3] = (float)f->z;
    float H[4][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 1}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H[i][j] = R[i][j];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = T[i][0] * H[0][j];
            acc = acc + T[i][1] * H[1][j];
            acc = acc + T[i][2] * H[2][j];
            acc = acc + T[i][3] * H[3][j];
            out->m[j][i] = acc;
        }
    out->tex = f->tex;
}

/* quad::VERTICES (quad.rs:66-87), expanded as vertex.vert:24 does:
 * pos[r] = ((m[0][r]*cx + m[1][r]*cy) + m[2][r]*cz) + m[3][r]*1. */
static const float kCorner[4][3] = {
    {-0.5f, 0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}};
static const float kUV[4][2] = {{0.0f, 0.0f}, {1.0f, 0.0f}, {0.0f, 1.0f}, {1.0f, 1.0f}};

void vxo_vertices_of(const vxo_instance *inst, vxo_vertex out[4]) {
    for (int c = 0; c < 4; c++) {
        for (int r = 0; r < 3; r++) {
            float acc = inst->m[0][r] * kCorner[c][0];
            acc = acc + inst->m[1][r] * kCorner[c][1];
            acc = acc + inst->m[2][r] * kCorner[c][2];
            acc = acc + inst->m[3][r] * 1.0f;
            out[c].pos[r] = acc;
        }
        out[c].uv[0] = kUV[c][0];
        out[c].uv[1] = kUV[c][1];
    }
}

/* quad::INDICES = [0, 2, 1, 3, 1, 2] (quad.rs:89), offset by 4 per face. */
void vxo_indices_of(int64_t local_face, uint32_t out[6]) {
    static const uint32_t pat[6] = {0, 2, 1, 3, 1, 2};
    for (int i = 0; i < 6; i++) out[i] = (uint32_t)(4 * local_face) + pat[i];
}

void vxo_expand(const vxo_face *faces, int64_t n, vxo_instance *inst, vxo_vertex *vert,
                uint32_t *idx) {
    for (int64_t f = 0; f < n; f++) {
        vxo_instance r;
        vxo_instance_of(&faces[f], &r);
        if (inst) inst[f] = r;
        if (vert) vxo_vertices_of(&r, vert + 4 * f);
        if (idx) vxo_indices_of(f, idx + 6 * f);
    }
}

/* ------------------------------------------------------------------------ */
/* Frustum (frustum_culling.rs:47-85): data[c][r] is nalgebra's column-major */
/* storage; planes left=r3+r0, right=r3-r0, top=r3-r1, bottom=r3+r1,         */
/* near=r3+r2, far=r3-r2, unnormalised, returned [near far left right top    */
/* bottom].                                                                  */
/* ------------------------------------------------------------------------ */
void vxo_frustum_planes(const float m[16], float planes[24]) {
#define D(c, r) m[(c) * 4 + (r)]
    for (int c = 0; c < 4; c++) {
        planes[0 * 4 + c] = D(c, 3) + D(c, 2);  /* near */
        planes[1 * 4 + c] = D(c, 3) - D(c, 2);  /* far */
        planes[2 * 4 + c] = D(c, 3) + D(c, 0);  /* left */
        planes[3 * 4 + c] = D(c, 3) - D(c, 0);  /* right */
        planes[4 * 4 + c] = D(c, 3) - D(c, 1);  /* top */
        planes[5 * 4 + c] = D(c, 3) + D(c, 1);  /* bottom */
    }
#undef D
}

/* contains (frustum_culling.rs:16-40): true iff some corner of the chunk
 * AABB has dist = ((a*x + b*y) + c*z) + d >= 0 (or NaN) for all six planes. */
int vxo_frustum_contains(const float planes[24], int32_t cx, int32_t cy, int32_t cz,
                         int32_t edge) {
    /* exact 64-bit products, then f32 (the i32 product where it does not overflow) */
    const float px = (float)((int64_t)cx * edge), py = (float)((int64_t)cy * edge),
                pz = (float)((int64_t)cz * edge);
    const float e = (float)edge;
    float cs[8][3] = {{px, py, pz},         {px + e, py, pz},         {px, py, pz + e},
                      {px + e, py, pz + e}, {px, py + e, pz},         {px + e, py + e, pz},
                      {px, py + e, pz + e}, {px + e, py + e, pz + e}};
    for (int p = 0; p < 8; p++) {
        int inside = 1;
        for (int i = 0; i < 6; i++) {
            const float *pl = planes + 4 * i;
            float dist = pl[0] * cs[p][0] + pl[1] * cs[p][1] + pl[2] * cs[p][2] + pl[3];
            if (dist < 0.0f) {
                inside = 0;
                break;
            }
        }
        if (inside) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Multithreaded batch driver: one chunk per task on a pool of threads, as   */
/* the reference's uvth pool runs Chunk::new + create_mesh per chunk          */
/* (chunk_builder.rs:80-84).                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const vxo_terrain *t;
    const int32_t *coords, *nb;
    int32_t n, edge;
    uint8_t *ids; /* n * edge^3 */
    atomic_int next;
    int phase;
    atomic_llong total;
    int64_t *faces_out;
    /* optional mesh output: chunk i's vertices and indices at face bases[i]
       of vert_out / idx_out (cap_faces faces); a chunk that would not fit
       sets overflow instead */
    const int64_t *bases;
    vxo_vertex *vert_out;
    uint32_t *idx_out;
    int64_t cap_faces;
    atomic_int overflow;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *J = (batch_job *)arg;
    const int32_t N = J->edge;
    const size_t vox = (size_t)N * N * N;
    const int64_t fmax = 3ll * N * N * (N + 1);
    vxo_face *faces = NULL;
    uint32_t *occ = NULL;
    if (J->phase == 1) {
        faces = (vxo_face *)malloc((size_t)fmax * sizeof(vxo_face));
    } else {
        occ = (uint32_t *)malloc((size_t)N * N * ((N + 31) / 32) * sizeof(uint32_t));
    }
    for (;;) {
        int i = atomic_fetch_add(&J->next, 1);
        if (i >= J->n) break;
        uint8_t *ids = J->ids + (size_t)i * vox;
        const int32_t cx = J->coords[3 * i], cy = J->coords[3 * i + 1], cz = J->coords[3 * i + 2];
        if (J->phase == 0) {
            vxo_chunk_fill_fast(J->t, cx, cy, cz, N, ids);
            vxo_occupancy(ids, N, occ);
        } else {
            const uint8_t *adj[4];
            for (int k = 0; k < 4; k++) {
                int32_t j = J->nb ? J->nb[4 * i + k] : -1;
                adj[k] = j >= 0 ? J->ids + (size_t)j * vox : NULL;
            }
            int64_t nf = vxo_mesh_faces(ids, N, cx * N, cy * N, cz * N, adj, faces, fmax);
            const int keep = J->vert_out != NULL;
            if (keep && (J->bases[i] < 0 || J->bases[i] + nf > J->cap_faces)) {
                atomic_store(&J->overflow, 1);
                if (J->faces_out) J->faces_out[i] = nf;
                atomic_fetch_add(&J->total, (long long)nf);
                continue;
            }
            vxo_vertex *vb = keep ? J->vert_out + 4 * J->bases[i]
                                  : (vxo_vertex *)malloc((size_t)(nf > 0 ? nf : 1) * 4 * sizeof(vxo_vertex));
            uint32_t *ib = keep ? J->idx_out + 6 * J->bases[i]
                                : (uint32_t *)malloc((size_t)(nf > 0 ? nf : 1) * 6 * sizeof(uint32_t));
            for (int64_t f = 0; f < nf; f++) {
                vxo_instance inst;
                vxo_instance_of(&faces[f], &inst);
                vxo_vertices_of(&inst, vb + 4 * f);
                vxo_indices_of(f, ib + 6 * f);
            }
            /* keep the writes observable so they are not optimised away */
            volatile uint32_t sink = nf > 0 ? ib[6 * nf - 1] : 0;
            (void)sink;
            if (!keep) {
                free(vb);
                free(ib);
            }
            if (J->faces_out) J->faces_out[i] = nf;
            atomic_fetch_add(&J->total, (long long)nf);
        }
    }
    free(faces);
    free(occ);
    return NULL;
}

static int64_t batch_drive(const vxo_terrain *t, const int32_t *coords, const int32_t *nb,
                           int32_t n, int32_t edge, int threads, int64_t *faces_out,
                           const int64_t *bases, vxo_vertex *vert_out, uint32_t *idx_out,
                           int64_t cap_faces) {
    if (n <= 0) return 0;
    if (threads < 1) threads = 1;
    batch_job J;
    J.t = t;
    J.coords = coords;
    J.nb = nb;
    J.n = n;
    J.edge = edge;
    J.ids = (uint8_t *)malloc((size_t)n * edge * edge * edge);
    J.faces_out = faces_out;
    J.bases = bases;
    J.vert_out = vert_out;
    J.idx_out = idx_out;
    J.cap_faces = cap_faces;
    atomic_init(&J.overflow, 0);
    atomic_init(&J.total, 0);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int phase = 0; phase < 2; phase++) {
        J.phase = phase;
        atomic_init(&J.next, 0);
        for (int k = 0; k < threads; k++) pthread_create(&th[k], NULL, batch_worker, &J);
        for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    }
    free(th);
    free(J.ids);
    if (atomic_load(&J.overflow)) return -1;
    return (int64_t)atomic_load(&J.total);
}

int64_t vxo_batch_run(const vxo_terrain *t, const int32_t *coords, const int32_t *nb, int32_t n,
                      int32_t edge, int threads, int64_t *faces_out) {
    return batch_drive(t, coords, nb, n, edge, threads, faces_out, NULL, NULL, NULL, 0);
}

int64_t vxo_batch_mesh(const vxo_terrain *t, const int32_t *coords, const int32_t *nb, int32_t n,
                       int32_t edge, int threads, const int64_t *bases, vxo_vertex *vert_out,
                       uint32_t *idx_out, int64_t cap_faces, int64_t *faces_out) {
    if (!bases || !vert_out || !idx_out) return -1;
    return batch_drive(t, coords, nb, n, edge, threads, faces_out, bases, vert_out, idx_out,
                       cap_faces);
}
