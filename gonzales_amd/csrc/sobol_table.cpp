// Sobol generator matrices for the Z-Sobol sampler: 1024 dimensions x 52 columns of uint32.
//
// Restates the reference's build-time generator (Sources/SobolGenerator/main.swift:11-76),
// which the reference runs as a SwiftPM plugin and embeds as base64 (SobolAccessor.swift:1-20):
//   dimension 0: column i = 1 << (31 - i) for i < 32, 0 above (van der Corput);
//   dimension d >= 1: Joe-Kuo row d - 1 (s, a, m_1..m_s): v_i = m_i << (32 - i) for i <= s,
//   v_i = v_{i-s} ^ (v_{i-s} >> s) ^ XOR_{j=1..s-1, bit (s-1-j) of a} v_{i-j} for s < i <= 52,
//   column i = v_{i+1}.
// The direction numbers are compiled in from gonzales_amd/data/new-joe-kuo-6.1023.
#include <cstdint>
#include <cstring>
#include <mutex>

#include "gonzales_amd.h"
#include "joe_kuo_data.inc"

namespace gz {

static uint32_t g_table[1024 * 52];
static std::once_flag g_once;

static void fill_table() {
        std::memset(g_table, 0, sizeof g_table);
        for (int i = 0; i < 32; ++i) g_table[i] = 1u << (31 - i);
        for (int d = 1; d < 1024; ++d) {
                const int s = kJoeKuo[d - 1].s;
                const int a = kJoeKuo[d - 1].a;
                uint32_t v[53] = {0};
                for (int i = 1; i <= s; ++i) v[i] = kJoeKuo[d - 1].m[i - 1] << (32 - i);
                for (int i = s + 1; i <= 52; ++i) {
                        v[i] = v[i - s] ^ (v[i - s] >> s);
                        for (int j = 1; j < s; ++j)
                                if ((a >> (s - 1 - j)) & 1) v[i] ^= v[i - j];
                }
                for (int i = 0; i < 52; ++i) g_table[d * 52 + i] = v[i + 1];
        }
}

const uint32_t *sobol_table() {
        std::call_once(g_once, fill_table);
        return g_table;
}

}  // namespace gz

extern "C" int gz_sobol_matrices(uint32_t *out) {
        if (!out) return GZ_E_INVALID;
        std::memcpy(out, gz::sobol_table(), sizeof(uint32_t) * 1024 * 52);
        return GZ_OK;
}
