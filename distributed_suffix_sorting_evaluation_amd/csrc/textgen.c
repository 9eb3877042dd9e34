/*
 * textgen.c -- seeded synthetic texts for the benchmark workloads and the tests.
 *
 * Generator: splitmix64 seeds xoshiro256** (Blackman & Vigna); bounded integers by Lemire's
 * multiply-shift with rejection.  Texts are produced in independent 1 MiB chunks, each with its own
 * stream seeded from (seed, chunk index), so generation is parallel and the output depends only on
 * (kind, n, seed).
 *
 * kinds
 *   0  uniform i.i.d. over {A,C,G,T}
 *   1  genome-like: order-3 Markov chain over {A,C,G,T} (seeded 64x4 transition table), then 5% of
 *      the length overwritten by copies of earlier segments (lengths log-uniform in [1 KiB, 100 KiB],
 *      source before destination) with 1% point substitutions inside each copy
 *   2  byte alphabet: Zipf(s = 1.1) over 256 symbols (rank -> byte by a seeded permutation), plus one
 *      block of period 17 (17 seeded bytes) of length min(64 MiB, n/4) at a seeded offset
 *   3  {A,C,G,T,N}: uniform ACGT with N at probability 0.001
 *   4  uniform over the bytes 0..sigma-1 (sigma = low byte of the seed's upper half, 0 -> 256)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ 0x5851F42D4C957F2Dull;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* Lemire: uniform in [0, range) */
static inline uint64_t rng_bounded(rng_t *r, uint64_t range) {
    __uint128_t m = (__uint128_t)rng_next(r) * range;
    uint64_t l = (uint64_t)m;
    if (l < range) {
        uint64_t t = (0 - range) % range;
        while (l < t) {
            m = (__uint128_t)rng_next(r) * range;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

#define CHUNK (1ull << 20)
static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};

static void gen_uniform_acgt(uint64_t n, uint64_t seed, uint8_t *out) {
    int64_t chunks = (int64_t)((n + CHUNK - 1) / CHUNK);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)c);
        uint64_t b = (uint64_t)c * CHUNK, e = b + CHUNK < n ? b + CHUNK : n;
        uint64_t w = 0;
        int left = 0;
        for (uint64_t i = b; i < e; i++) {
            if (left == 0) {
                w = rng_next(&r);
                left = 32;
            }
            out[i] = ACGT[w >> 62];
            w <<= 2;
            left--;
        }
    }
}

static void gen_acgtn(uint64_t n, uint64_t seed, uint8_t *out) {
    int64_t chunks = (int64_t)((n + CHUNK - 1) / CHUNK);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)c);
        uint64_t b = (uint64_t)c * CHUNK, e = b + CHUNK < n ? b + CHUNK : n;
        for (uint64_t i = b; i < e; i++) {
            uint64_t x = rng_next(&r);
            /* 1/1000 N using the low 32 bits; base from the top 2 bits */
            out[i] = ((x & 0xFFFFFFFFull) < 4294967ull) ? 'N' : ACGT[x >> 62];
        }
    }
}

static void gen_uniform_sigma(uint64_t n, uint64_t seed, uint32_t sigma, uint8_t *out) {
    int64_t chunks = (int64_t)((n + CHUNK - 1) / CHUNK);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)c);
        uint64_t b = (uint64_t)c * CHUNK, e = b + CHUNK < n ? b + CHUNK : n;
        for (uint64_t i = b; i < e; i++) out[i] = (uint8_t)rng_bounded(&r, sigma);
    }
}

static void gen_genome(uint64_t n, uint64_t seed, uint8_t *out) {
    /* transition table: 64 contexts (last 3 bases) x 4 cumulative 32-bit thresholds */
    uint32_t thr[64][4];
    {
        rng_t r;
        rng_seed(&r, seed, 0xFFFFFFFF00000001ull);
        for (int ctx = 0; ctx < 64; ctx++) {
            uint64_t w[4], s = 0;
            for (int k = 0; k < 4; k++) {
                w[k] = 1 + rng_bounded(&r, 1000);
                s += w[k];
            }
            uint64_t acc = 0;
            for (int k = 0; k < 4; k++) {
                acc += w[k];
                thr[ctx][k] = k == 3 ? 0xFFFFFFFFu : (uint32_t)((acc << 32) / s);
            }
        }
    }
    int64_t chunks = (int64_t)((n + CHUNK - 1) / CHUNK);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)c);
        uint64_t b = (uint64_t)c * CHUNK, e = b + CHUNK < n ? b + CHUNK : n;
        uint32_t ctx = (uint32_t)rng_bounded(&r, 64);
        uint64_t w = 0;
        int half = 0;
        for (uint64_t i = b; i < e; i++) {
            uint32_t u;
            if (half == 0) {
                w = rng_next(&r);
                u = (uint32_t)(w >> 32);
                half = 1;
            } else {
                u = (uint32_t)w;
                half = 0;
            }
            uint32_t k = 0;
            while (k < 3 && u >= thr[ctx][k]) k++;
            out[i] = ACGT[k];
            ctx = ((ctx << 2) | k) & 63;
        }
    }
    /* copies of earlier segments over 5% of the length */
    rng_t r;
    rng_seed(&r, seed, 0xFFFFFFFF00000002ull);
    uint64_t target = n / 20, done = 0;
    const double lmin = log(1024.0), lmax = log(102400.0);
    while (done < target) {
        uint64_t len = (uint64_t)exp(lmin + (lmax - lmin) * rng_unit(&r));
        if (len > n / 4) len = n / 4;
        if (len < 1) break;
        if (len > target - done) len = target - done;
        if (n < 2 * len + 1) break;
        uint64_t dst = len + rng_bounded(&r, n - 2 * len + 1); /* dst in [len, n - len] */
        uint64_t src = rng_bounded(&r, dst - len + 1);         /* src + len <= dst       */
        for (uint64_t k = 0; k < len; k++) {
            uint8_t ch = out[src + k];
            if (rng_bounded(&r, 100) == 0) {
                int idx = ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : 3;
                ch = ACGT[(idx + 1 + (int)rng_bounded(&r, 3)) & 3];
            }
            out[dst + k] = ch;
        }
        done += len;
    }
}

static void gen_zipf256(uint64_t n, uint64_t seed, uint8_t *out) {
    uint64_t thr[256];
    uint8_t perm[256];
    {
        double p[256], s = 0;
        for (int k = 0; k < 256; k++) {
            p[k] = exp(-1.1 * log((double)(k + 1)));
            s += p[k];
        }
        double acc = 0;
        for (int k = 0; k < 256; k++) {
            acc += p[k] / s;
            thr[k] = k == 255 ? ~0ull : (uint64_t)(acc * 18446744073709549568.0);
        }
        rng_t r;
        rng_seed(&r, seed, 0xFFFFFFFF00000003ull);
        for (int k = 0; k < 256; k++) perm[k] = (uint8_t)k;
        for (int k = 255; k > 0; k--) {
            int j = (int)rng_bounded(&r, (uint64_t)k + 1);
            uint8_t t = perm[k];
            perm[k] = perm[j];
            perm[j] = t;
        }
    }
    int64_t chunks = (int64_t)((n + CHUNK - 1) / CHUNK);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        rng_t r;
        rng_seed(&r, seed, (uint64_t)c);
        uint64_t b = (uint64_t)c * CHUNK, e = b + CHUNK < n ? b + CHUNK : n;
        for (uint64_t i = b; i < e; i++) {
            uint64_t u = rng_next(&r);
            int lo = 0, hi = 255;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (u < thr[mid])
                    hi = mid;
                else
                    lo = mid + 1;
            }
            out[i] = perm[lo];
        }
    }
    /* periodic block */
    rng_t r;
    rng_seed(&r, seed, 0xFFFFFFFF00000004ull);
    uint64_t blk = 64ull << 20;
    if (blk > n / 4) blk = n / 4;
    if (blk >= 17) {
        uint8_t per[17];
        for (int k = 0; k < 17; k++) per[k] = (uint8_t)rng_bounded(&r, 256);
        uint64_t off = rng_bounded(&r, n - blk + 1);
        for (uint64_t k = 0; k < blk; k++) out[off + k] = per[k % 17];
    }
}

int dcx_gen_text(int kind, uint64_t n, uint64_t seed, uint8_t *out) {
    if (!out && n) return -6;
    switch (kind) {
        case 0: gen_uniform_acgt(n, seed, out); return 0;
        case 1: gen_genome(n, seed, out); return 0;
        case 2: gen_zipf256(n, seed, out); return 0;
        case 3: gen_acgtn(n, seed, out); return 0;
        case 4: {
            uint32_t sigma = (uint32_t)((seed >> 32) & 0xFF);
            if (sigma == 0) sigma = 256;
            gen_uniform_sigma(n, seed, sigma, out);
            return 0;
        }
        default: return -6;
    }
}

/* lengths log-uniform in [lo, hi] for a batch of `count` texts */
int dcx_gen_lengths(uint32_t count, uint64_t lo, uint64_t hi, uint64_t seed, uint64_t *lens) {
    if (!lens || lo == 0 || hi < lo) return -6;
    rng_t r;
    rng_seed(&r, seed, 0xFFFFFFFF00000005ull);
    double a = log((double)lo), b = log((double)hi + 1.0);
    for (uint32_t i = 0; i < count; i++) {
        uint64_t v = (uint64_t)exp(a + (b - a) * rng_unit(&r));
        lens[i] = v < lo ? lo : v > hi ? hi : v;
    }
    return 0;
}
