/*
 * Oracle: random-access Owen-scrambled Sobol points (TEST INFRASTRUCTURE).
 *
 * Restates src/qmc_mod.f90 with native 32-bit unsigned arithmetic (the
 * reference emulates uint32 in int64 and masks every step; the low 32 bits of
 * each product are the same).
 *
 * Pinned by: the reference's committed golden vectors tests/qmc/q0_*.txt and
 * q2_*.txt (raw and scrambled, both streams, 1024 points each), compared bit
 * for bit through the integer codes; and by the reference qmc_mod compiled
 * from its own source into oracle/_ref (when present).
 *
 * Also here: the counter-based generator for post-launch draws.  The
 * reference draws those from MT19937 (src/random_mt.f90), which depends on
 * the rank's draw history; src/define.f90:151-165 anticipates a generator
 * keyed by (id, istream) instead, which is what this is.  It has no reference
 * counterpart to pin against; tests check its uniformity and independence.
 */
#include "mc_oracle.h"

/* Joe & Kuo new-joe-kuo-6 table rows for dimensions 2..12: degree s,
 * interior-coefficient bits a, initial direction integers m_1..m_s
 * (src/qmc_mod.f90:80-97; a published table, carried as data). */
static const int jk_s[ORC_QMC_MAXDIM] = {0, 1, 2, 3, 3, 4, 4, 5, 5, 5, 5, 5};
static const int jk_a[ORC_QMC_MAXDIM] = {0, 0, 1, 1, 2, 1, 4, 2, 4, 7, 11, 13};
static const int jk_m[ORC_QMC_MAXDIM][5] = {
    {0}, {1}, {1, 3}, {1, 3, 1}, {1, 1, 1}, {1, 1, 3, 3}, {1, 3, 5, 13},
    {1, 1, 5, 5, 17}, {1, 1, 5, 5, 5}, {1, 1, 7, 11, 19}, {1, 1, 5, 1, 1},
    {1, 1, 1, 3, 11}};

/* lowbias32 avalanche (src/qmc_mod.f90:250-260) */
uint32_t orc_hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

/* qmc_setup (src/qmc_mod.f90:73-139) */
void orc_qmc_setup(orc_qmc *q, int seed)
{
    for (int k = 1; k <= 32; ++k) q->V[k - 1][0] = 1u << (32 - k);
    for (int d = 1; d < ORC_QMC_MAXDIM; ++d) {
        int s = jk_s[d], a = jk_a[d];
        for (int k = 1; k <= s; ++k)
            q->V[k - 1][d] = (uint32_t)((uint64_t)jk_m[d][k - 1] << (32 - k));
        for (int k = s + 1; k <= 32; ++k) {
            uint32_t v = q->V[k - s - 1][d] ^ (q->V[k - s - 1][d] >> s);
            for (int j = 1; j <= s - 1; ++j)
                if ((a >> (s - 1 - j)) & 1) v ^= q->V[k - j - 1][d];
            q->V[k - 1][d] = v;
        }
    }
    for (int d = 1; d <= ORC_QMC_MAXDIM; ++d) {
        uint32_t base = (uint32_t)seed + (uint32_t)d * 0x9E3779B9u;
        q->key[d - 1][0] = orc_hash32(base);
        q->key[d - 1][1] = orc_hash32(base + 0x85EBCA6Bu);
    }
}

/* sobol_code (src/qmc_mod.f90:192-206): index bits beyond 32 contribute
 * nothing */
uint32_t orc_sobol_code(const orc_qmc *q, uint64_t idx, int d)
{
    uint32_t code = 0;
    int k = 0;
    while (idx != 0 && k < 32) {
        if (idx & 1u) code ^= q->V[k][d];
        idx >>= 1;
        ++k;
    }
    return code;
}

static uint32_t reverse32(uint32_t n)
{
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00FF00FFu) << 8) | ((n & 0xFF00FF00u) >> 8);
    n = ((n & 0x0F0F0F0Fu) << 4) | ((n & 0xF0F0F0F0u) >> 4);
    n = ((n & 0x33333333u) << 2) | ((n & 0xCCCCCCCCu) >> 2);
    n = ((n & 0x55555555u) << 1) | ((n & 0xAAAAAAAAu) >> 1);
    return n;
}

/* owen_scramble + laine_karras (src/qmc_mod.f90:212-230) */
uint32_t orc_owen_scramble(uint32_t x, uint32_t seed)
{
    uint32_t y = reverse32(x);
    y += seed;
    y ^= y * 0x6C50B47Cu;
    y ^= y * 0xB82F1E52u;
    y ^= y * 0xC7AFE638u;
    y ^= y * 0x8D22F6E6u;
    return reverse32(y);
}

/* qmc_uniforms_stream (src/qmc_mod.f90:157-168); stream is 1 or 2 */
void orc_qmc_uniforms(const orc_qmc *q, uint64_t idx, int n, int stream, double *u)
{
    for (int d = 0; d < n; ++d) {
        uint32_t code = orc_owen_scramble(orc_sobol_code(q, idx, d), q->key[d][stream - 1]);
        u[d] = ((double)code + 0.5) * (1.0 / 4294967296.0);
    }
}

/* qmc_uniforms_raw (src/qmc_mod.f90:174-184) */
void orc_qmc_raw(const orc_qmc *q, uint64_t idx, int n, double *u)
{
    for (int d = 0; d < n; ++d)
        u[d] = ((double)orc_sobol_code(q, idx, d) + 0.5) * (1.0 / 4294967296.0);
}

/* ---- counter-based generator ------------------------------------------- */
static uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

/* Uniform in (0,1) for draw k of packet (id, stream) under seed:
 * three rounds of the splitmix64 finalizer over the packed key, 53-bit
 * mantissa, centred in its bin so 0 and 1 are never returned. */
double orc_rng_uniform(uint32_t seed, int stream, uint64_t id, uint32_t k)
{
    uint64_t h = mix64((uint64_t)seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
    h = mix64(h ^ id);
    h = mix64(h ^ (((uint64_t)(uint32_t)stream << 32) | (uint64_t)k));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
