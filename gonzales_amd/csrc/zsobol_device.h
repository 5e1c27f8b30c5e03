// Z-Sobol sampler, device side (shared by the sampler kernel and the bounce kernels).
//
// Restates Sources/libgonzales/Sampler/ZSobolSampler.swift bit for bit:
//   sobolSample :6-31, reverseBits32 :34-41, FastOwenScrambler :44-67, encodeMorton2 :71-88,
//   mixBits :91-101 (a 32-bit hash: the second seed of each get2D/get3D pair is always 0),
//   startPixelSample :130-136, getSampleIndex :138-166, getSampleHash :188-199,
//   get1D/get2D/get3D :203-280 (index and hash from the dimension before the wrap to 2).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gz {

constexpr int kSobolDims = 1024;
constexpr int kSobolCols = 52;
// The device form of the matrices: per dimension, for each of the 7 bytes of a 52-bit index and
// each value of that byte, the XOR of the columns of its set bits (bits 52..55 of byte 6 have no
// column: 0). 1024 x 7 x 256 words, 7 MiB.
constexpr int kSobolBytes = 7;
constexpr size_t kSobolByteTableWords = (size_t)kSobolDims * kSobolBytes * 256;

struct ZSobolParams {
        int64_t seed;
        int32_t log2spp;
        int32_t nBase4Digits;
};

__device__ __forceinline__ uint32_t zs_mix_bits(uint64_t v) {
        uint32_t x = (uint32_t)v ^ (uint32_t)(v >> 32);
        x ^= x >> 16;
        x *= 0x85ebca77u;
        x ^= x >> 13;
        x *= 0xc2b2ae35u;
        x ^= x >> 16;
        return x;
}

__device__ __forceinline__ uint32_t zs_owen(uint32_t v, uint32_t seed) {
        v = __builtin_bitreverse32(v);
        v ^= v * 0x3d20adeau;
        v += seed;
        v *= (seed >> 16) | 1u;
        v ^= v * 0x05526c56u;
        v ^= v * 0x53a22864u;
        return __builtin_bitreverse32(v);
}

__device__ __forceinline__ uint64_t zs_morton2(uint32_t x, uint32_t y) {
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

// Swift's `>>` on UInt64 yields 0 for shifts of 64 or more.
__device__ __forceinline__ uint64_t zs_shr(uint64_t v, int s) { return s >= 64 ? 0ull : v >> s; }

// The 24 permutations of the base-4 digits (ZSobolSampler.swift:148-156), 2 bits per entry and a
// byte per permutation, packed into three 64-bit words so that a lookup is shifts and masks.
constexpr uint8_t kZsPermTable[24][4] = {
        {0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 2, 1}, {0, 3, 1, 2},
        {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0}, {1, 3, 2, 0}, {1, 3, 0, 2},
        {2, 1, 0, 3}, {2, 1, 3, 0}, {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 3, 0, 1}, {2, 3, 1, 0},
        {3, 1, 2, 0}, {3, 1, 0, 2}, {3, 2, 1, 0}, {3, 2, 0, 1}, {3, 0, 2, 1}, {3, 0, 1, 2},
};

constexpr uint64_t zs_perm_word(int w) {
        uint64_t x = 0;
        for (int p = 0; p < 8; ++p)
                for (int d = 0; d < 4; ++d) x |= (uint64_t)kZsPermTable[8 * w + p][d] << (8 * p + 2 * d);
        return x;
}

__device__ __forceinline__ uint32_t zs_perm(int perm, int digit) {
        constexpr uint64_t w0 = zs_perm_word(0), w1 = zs_perm_word(1), w2 = zs_perm_word(2);
        const uint64_t w = perm < 8 ? w0 : perm < 16 ? w1 : w2;
        return (uint32_t)(w >> (((perm & 7) << 3) + (digit << 1))) & 3u;
}

__device__ __forceinline__ uint64_t zs_sample_index(const ZSobolParams &p, uint64_t morton, int dim) {
        uint64_t out = 0;
        const bool odd = (p.log2spp & 1) == 1;
        const int last = odd ? 1 : 0;
        const uint64_t dimMix = 0x55555555ull * (uint64_t)(int64_t)dim;
        if (morton == 0) {
                // Morton index 0 (the light chooser's sampler never starts a pixel sample): every
                // digit and every higher part is 0, so every level hashes the same value and takes
                // the same permutation of digit 0; the loop below gives the same bits.
                const uint32_t h = zs_mix_bits(dimMix);
                const uint64_t d = zs_perm((int)((h >> 24) % 24u), 0);
                for (int i = p.nBase4Digits - 1; i >= last; --i) {
                        const int shift = 2 * i - (odd ? 1 : 0);
                        out |= (shift < 64 ? d << shift : 0ull);
                }
                if (odd) out |= (uint64_t)(h & 1u);
                return out;
        }
        for (int i = p.nBase4Digits - 1; i >= last; --i) {
                const int shift = 2 * i - (odd ? 1 : 0);
                const int digit = (int)(zs_shr(morton, shift) & 3);
                const uint64_t higher = zs_shr(morton, shift + 2);
                const uint32_t h = zs_mix_bits(higher ^ dimMix);
                const int perm = (int)((h >> 24) % 24u);
                out |= (shift < 64 ? (uint64_t)zs_perm(perm, digit) << shift : 0ull);
        }
        if (odd) {
                int digit = (int)(morton & 1);
                const uint32_t h = zs_mix_bits((morton >> 1) ^ dimMix);
                digit ^= (int)(h & 1);
                out |= (uint64_t)digit;
        }
        return out;
}

// sobolSample: XOR of the columns of the set bits among the 52 lowest, Owen-scrambled, scaled
// by 2^-32 (u32 -> f32 rounds to nearest even) and clamped to 1 - 1e-6. The XOR is taken a byte
// of the index at a time from the byte table (the same XOR regrouped): seven loads that do not
// depend on each other, where a walk over the set bits chains one dependent load per bit.
__device__ __forceinline__ float zs_sobol(uint64_t index, int dim, uint32_t seed, const uint32_t *__restrict__ table) {
        const uint32_t *t = table + (size_t)dim * (kSobolBytes * 256);
        const uint64_t x = index & ((1ull << kSobolCols) - 1);
        uint32_t acc = t[(uint32_t)x & 255u];
#pragma unroll
        for (int k = 1; k < kSobolBytes; ++k) acc ^= t[k * 256 + ((uint32_t)(x >> (8 * k)) & 255u)];
        const float f = (float)zs_owen(acc, seed) * 2.32830643653869628906e-10f;
        const float oneMinusEps = 1.0f - 1e-6f;
        return oneMinusEps < f ? oneMinusEps : f;
}

struct ZSobolState {
        uint64_t morton;
        int dim;
};

__device__ __forceinline__ ZSobolState zs_start(const ZSobolParams &p, int32_t px, int32_t py, int64_t index, int dim) {
        ZSobolState z;
        z.dim = dim;
        const uint64_t m = zs_morton2((uint32_t)px, (uint32_t)py);
        z.morton = (p.log2spp >= 64 ? 0ull : (m << p.log2spp)) | (uint64_t)index;
        return z;
}

__device__ __forceinline__ uint32_t zs_hash(const ZSobolParams &p, int dim) {
        return zs_mix_bits((uint64_t)(int64_t)dim ^ (uint64_t)p.seed);
}

__device__ __forceinline__ float zs_get1d(const ZSobolParams &p, ZSobolState &z, const uint32_t *__restrict__ table) {
        const uint64_t idx = zs_sample_index(p, z.morton, z.dim);
        const uint32_t h = zs_hash(p, z.dim);
        if (z.dim >= kSobolDims) z.dim = 2;
        z.dim += 1;
        return zs_sobol(idx, z.dim - 1, h, table);
}

__device__ __forceinline__ void zs_get2d(const ZSobolParams &p, ZSobolState &z, const uint32_t *__restrict__ table,
                                         float &a, float &b) {
        const uint64_t idx = zs_sample_index(p, z.morton, z.dim);
        const uint32_t h0 = zs_hash(p, z.dim);
        if (z.dim + 1 >= kSobolDims) z.dim = 2;
        z.dim += 2;
        a = zs_sobol(idx, z.dim - 2, h0, table);
        b = zs_sobol(idx, z.dim - 1, 0u, table);
}

__device__ __forceinline__ void zs_get3d(const ZSobolParams &p, ZSobolState &z, const uint32_t *__restrict__ table,
                                         float &a, float &b, float &c) {
        const uint64_t idx = zs_sample_index(p, z.morton, z.dim);
        const uint32_t h0 = zs_hash(p, z.dim);
        const uint32_t h2 = zs_hash(p, z.dim + 2);
        if (z.dim + 2 >= kSobolDims) z.dim = 2;
        z.dim += 3;
        a = zs_sobol(idx, z.dim - 3, h0, table);
        b = zs_sobol(idx, z.dim - 2, 0u, table);
        c = zs_sobol(idx, z.dim - 1, h2, table);
}

}  // namespace gz
