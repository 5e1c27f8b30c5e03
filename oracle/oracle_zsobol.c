/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's Z-Sobol sampler, used as
 * the checker by tests/ and the smoke check; never shipped and never on the product path.
 *
 * Restates Sources/libgonzales/Sampler/ZSobolSampler.swift and the table generator:
 *   sobolSample ............ ZSobolSampler.swift:6-31   (XOR of columns, <= 52 bits, Owen, x 2^-32,
 *                                                         min with 1 - 1e-6)
 *   reverseBits32 .......... ZSobolSampler.swift:34-41
 *   FastOwenScrambler ...... ZSobolSampler.swift:44-67
 *   encodeMorton2 .......... ZSobolSampler.swift:71-88
 *   mixBits ................ ZSobolSampler.swift:91-101  (32-bit result: the second seed of every
 *                                                         get2D/get3D pair is 0)
 *   init / startPixelSample  ZSobolSampler.swift:114-136
 *   getSampleIndex ......... ZSobolSampler.swift:138-166
 *   getSampleHash .......... ZSobolSampler.swift:188-199
 *   get1D / get2D / get3D .. ZSobolSampler.swift:203-280 (hash and index from the pre-wrap dim)
 *   nextPowerOfTwo ......... ZSobolSampler.swift:283-295
 *   generateSobolMatrices .. SobolGenerator/main.swift:11-76 (1024 dims x 52 columns)
 *
 * Pinning (tests/test_oracle_zsobol.py): the reference's own KATs (SobolSamplerTests.swift:9-75),
 * the matrix table against scipy's independent Joe-Kuo implementation (all 1024 x 52 columns),
 * dimension 0 = van der Corput, (0, m)-stratification of unscrambled and Owen-scrambled
 * dimensions, bijectivity of getSampleIndex over a full resolution, ranges.
 * Parity unpinned beyond the restatement: the hash constants of mixBits and of the Owen
 * scrambler (no reference test holds their values).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { O_NDIMS = 1024, O_NCOLS = 52 };

uint32_t orc_reverse_bits32(uint32_t v) {
        v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
        v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
        v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
        v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
        return (v >> 16) | (v << 16);
}

uint32_t orc_owen(uint32_t v, uint32_t seed) {
        v = orc_reverse_bits32(v);
        v ^= v * 0x3d20adeau;
        v += seed;
        v *= (seed >> 16) | 1u;
        v ^= v * 0x05526c56u;
        v ^= v * 0x53a22864u;
        return orc_reverse_bits32(v);
}

uint64_t orc_encode_morton2(uint32_t x, uint32_t y) {
        uint64_t a = x, b = y;
        a = (a | (a << 16)) & 0x0000FFFF0000FFFFull;
        a = (a | (a << 8)) & 0x00FF00FF00FF00FFull;
        a = (a | (a << 4)) & 0x0F0F0F0F0F0F0F0Full;
        a = (a | (a << 2)) & 0x3333333333333333ull;
        a = (a | (a << 1)) & 0x5555555555555555ull;
        b = (b | (b << 16)) & 0x0000FFFF0000FFFFull;
        b = (b | (b << 8)) & 0x00FF00FF00FF00FFull;
        b = (b | (b << 4)) & 0x0F0F0F0F0F0F0F0Full;
        b = (b | (b << 2)) & 0x3333333333333333ull;
        b = (b | (b << 1)) & 0x5555555555555555ull;
        return a | (b << 1);
}

uint32_t orc_mix_bits(uint64_t v) {
        uint32_t x = (uint32_t)(v & 0xFFFFFFFFu);
        x ^= (uint32_t)(v >> 32);
        x ^= x >> 16;
        x *= 0x85ebca77u;
        x ^= x >> 13;
        x *= 0xc2b2ae35u;
        x ^= x >> 16;
        return x;
}

int64_t orc_next_pow2(int64_t v) {
        v -= 1;
        v |= v >> 1;
        v |= v >> 2;
        v |= v >> 4;
        v |= v >> 8;
        v |= v >> 16;
        return v + 1;
}

/* SobolGenerator/main.swift:11-76. Reads the Joe-Kuo direction-number text ("d s a m_1 ...";
 * header and blank lines skipped) and fills out[1024 * 52]. Returns 0, or -1 on a read error. */
int orc_sobol_matrices(const char *path, uint32_t *out) {
        FILE *f = fopen(path, "r");
        if (!f) return -1;
        memset(out, 0, sizeof(uint32_t) * O_NDIMS * O_NCOLS);
        for (int i = 0; i < 32; ++i) out[i] = 1u << (31 - i);
        char line[4096];
        int dim = 1;
        while (dim < O_NDIMS && fgets(line, sizeof line, f)) {
                char *p = line;
                while (*p == ' ' || *p == '\t') ++p;
                if (*p == 'd' || *p == '\n' || *p == '\r' || *p == 0) continue;
                long vals[64];
                int nv = 0;
                char *end;
                for (;;) {
                        long x = strtol(p, &end, 10);
                        if (end == p) break;
                        if (nv < 64) vals[nv++] = x;
                        p = end;
                }
                if (nv < 3) continue;
                int s = (int)vals[1];
                long a = vals[2];
                if (s < 1 || s > 31 || nv < 3 + s) { fclose(f); return -1; }
                uint32_t v[O_NCOLS + 1];
                memset(v, 0, sizeof v);
                for (int i = 1; i <= s; ++i) v[i] = (uint32_t)vals[2 + i] << (32 - i);
                for (int i = s + 1; i <= O_NCOLS; ++i) {
                        v[i] = v[i - s] ^ (v[i - s] >> s);
                        for (int j = 1; j < s; ++j)
                                if ((a >> (s - 1 - j)) & 1) v[i] ^= v[i - j];
                }
                for (int i = 0; i < O_NCOLS; ++i) out[dim * O_NCOLS + i] = v[i + 1];
                ++dim;
        }
        fclose(f);
        return dim == O_NDIMS ? 0 : -1;
}

/* ZSobolSampler.swift:6-31 */
float orc_sobol_sample(int64_t index, int dim, uint32_t seed, const uint32_t *m) {
        uint32_t acc = 0;
        int64_t cur = index;
        const uint32_t *col = m + (int64_t)dim * O_NCOLS;
        for (int b = 0; b < O_NCOLS; ++b) {
                if (cur & 1) acc ^= col[b];
                cur >>= 1;
                if (cur == 0) break;
        }
        uint32_t r = orc_owen(acc, seed);
        float f = (float)r * 2.32830643653869628906e-10f;
        const float oneMinusEps = 1.0f - 1e-6f;
        return f < oneMinusEps ? f : oneMinusEps;
}

typedef struct {
        int64_t seed;
        int32_t log2spp;
        int32_t nBase4Digits;
        uint64_t morton;
        int32_t dim;
        int32_t pad;
} o_zsobol;

/* ZSobolSampler.swift:114-122 */
void orc_zsobol_init(o_zsobol *z, int64_t spp, int64_t resX, int64_t resY, int64_t seed) {
        memset(z, 0, sizeof *z);
        int tz = 0;
        if (spp != 0) { uint64_t s = (uint64_t)spp; while (!(s & 1)) { s >>= 1; ++tz; } }
        else tz = 64;
        z->log2spp = tz;
        z->seed = seed;
        int64_t res = orc_next_pow2(resX > resY ? resX : resY);
        int rtz = 0;
        if (res != 0) { uint64_t r = (uint64_t)res; while (!(r & 1)) { r >>= 1; ++rtz; } }
        else rtz = 64;
        z->nBase4Digits = rtz + (z->log2spp + 1) / 2;
}

/* ZSobolSampler.swift:130-136 */
void orc_zsobol_start(o_zsobol *z, int32_t px, int32_t py, int64_t index, int32_t dim) {
        z->dim = dim;
        uint64_t pm = orc_encode_morton2((uint32_t)px, (uint32_t)py);
        z->morton = (z->log2spp >= 64 ? 0 : pm << z->log2spp) | (uint64_t)index;
}

static const uint8_t o_perm[24][4] = {
        {0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 2, 1}, {0, 3, 1, 2},
        {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0}, {1, 3, 2, 0}, {1, 3, 0, 2},
        {2, 1, 0, 3}, {2, 1, 3, 0}, {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 3, 0, 1}, {2, 3, 1, 0},
        {3, 1, 2, 0}, {3, 1, 0, 2}, {3, 2, 1, 0}, {3, 2, 0, 1}, {3, 0, 2, 1}, {3, 0, 1, 2},
};

/* ZSobolSampler.swift:138-166 */
uint64_t orc_zsobol_sample_index(const o_zsobol *z) {
        uint64_t out = 0;
        int odd = (z->log2spp & 1) == 1;
        int last = odd ? 1 : 0;
        uint64_t dimMix = 0x55555555ull * (uint64_t)(int64_t)z->dim;
        for (int i = z->nBase4Digits - 1; i >= last; --i) {
                int shift = 2 * i - (odd ? 1 : 0);
                int digit = (int)((shift >= 64 ? 0 : z->morton >> shift) & 3);
                uint64_t higher = shift + 2 >= 64 ? 0 : z->morton >> (shift + 2);
                uint32_t h = orc_mix_bits(higher ^ dimMix);
                int p = (int)((h >> 24) % 24);
                digit = o_perm[p][digit];
                if (shift < 64) out |= (uint64_t)digit << shift;
        }
        if (odd) {
                int digit = (int)(z->morton & 1);
                uint32_t h = orc_mix_bits((z->morton >> 1) ^ dimMix);
                digit ^= (int)(h & 1);
                out |= (uint64_t)digit;
        }
        return out;
}

/* ZSobolSampler.swift:188-199: (mixBits(dim + off ^ seed), 0) */
static uint32_t o_hash(const o_zsobol *z, int off) {
        return orc_mix_bits((uint64_t)(int64_t)(z->dim + off) ^ (uint64_t)z->seed);
}

/* ZSobolSampler.swift:203-219 */
float orc_zsobol_get1d(o_zsobol *z, const uint32_t *m) {
        int64_t idx = (int64_t)orc_zsobol_sample_index(z);
        uint32_t h = o_hash(z, 0);
        if (z->dim >= O_NDIMS) z->dim = 2;
        z->dim += 1;
        return orc_sobol_sample(idx, z->dim - 1, h, m);
}

/* ZSobolSampler.swift:221-244 */
void orc_zsobol_get2d(o_zsobol *z, const uint32_t *m, float *out) {
        int64_t idx = (int64_t)orc_zsobol_sample_index(z);
        uint32_t h0 = o_hash(z, 0), h1 = 0;
        if (z->dim + 1 >= O_NDIMS) z->dim = 2;
        z->dim += 2;
        out[0] = orc_sobol_sample(idx, z->dim - 2, h0, m);
        out[1] = orc_sobol_sample(idx, z->dim - 1, h1, m);
}

/* ZSobolSampler.swift:249-280 */
void orc_zsobol_get3d(o_zsobol *z, const uint32_t *m, float *out) {
        int64_t idx = (int64_t)orc_zsobol_sample_index(z);
        uint32_t h0 = o_hash(z, 0), h1 = 0, h2 = o_hash(z, 2);
        if (z->dim + 2 >= O_NDIMS) z->dim = 2;
        z->dim += 3;
        out[0] = orc_sobol_sample(idx, z->dim - 3, h0, m);
        out[1] = orc_sobol_sample(idx, z->dim - 2, h1, m);
        out[2] = orc_sobol_sample(idx, z->dim - 1, h2, m);
}

/* Batch: for each (px[i], py[i], sampleIndex[i]) start a pixel sample at dimension startDim and
 * run the calls in `pattern` (each 1, 2 or 3 = get1D/get2D/get3D), writing their values in order
 * to out[i * stride ...], stride = sum(pattern). */
void orc_zsobol_batch(int64_t spp, int64_t resX, int64_t resY, int64_t seed, const uint32_t *m,
                      const int32_t *px, const int32_t *py, const int64_t *sampleIndex, int64_t n,
                      int32_t startDim, const int8_t *pattern, int32_t nCalls, float *out) {
        o_zsobol z;
        orc_zsobol_init(&z, spp, resX, resY, seed);
        int stride = 0;
        for (int c = 0; c < nCalls; ++c) stride += pattern[c];
        for (int64_t i = 0; i < n; ++i) {
                orc_zsobol_start(&z, px[i], py[i], sampleIndex[i], startDim);
                float *o = out + i * stride;
                for (int c = 0; c < nCalls; ++c) {
                        if (pattern[c] == 1) { *o++ = orc_zsobol_get1d(&z, m); }
                        else if (pattern[c] == 2) { orc_zsobol_get2d(&z, m, o); o += 2; }
                        else { orc_zsobol_get3d(&z, m, o); o += 3; }
                }
        }
}
