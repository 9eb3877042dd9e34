/*
 * dcx_oracle.c -- CPU restatement of lightweight_dcx's PDCX suffix-array construction.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke check and the
 * cpu_baseline leg of bench.py load it.  The product path (the HIP engine) never calls it.
 *
 * It restates, phase by phase and for a single process (p = 1), the algorithm of
 *   lightweight_dcx/src/pdcx/pdcx.hpp            (PDCX::compute_sa and the recursion)
 *   lightweight_dcx/src/pdcx/sample_string.hpp   (sample tuples, padding)
 *   lightweight_dcx/src/pdcx/compute_ranks.hpp   (lexicographic names, unique flags)
 *   lightweight_dcx/src/pdcx/merge_samples.hpp   (merge tuples and the DC comparator)
 *   lightweight_dcx/src/pdcx/difference_cover.hpp (DC sets; tables by the closed form of
 *                                                  B_DC_3_7_13/tools/calc_dcparam.cc:271-301)
 *   lightweight_dcx/src/executables/cli.cpp      (compress_alphabet, packed variant choice)
 *   lightweight_dcx/src/util/string_util.hpp     (slow_suffixarray base case)
 *   lightweight_dcx/src/sa_check.hpp             (SA validity check)
 * Characters are held as uint64_t at every level (level 0: compressed alphabet 1..sigma).
 * Packed comparisons of KPackedInteger words equal lexicographic comparisons of the same
 * number of characters (left-justified packing, char_container.hpp:57-71), so the oracle
 * compares characters directly and exposes the packing closed form separately
 * (oracle_pack_word) for the packing tests.
 *
 * Parity pinning (see tests/test_oracle.py): golden SA fixtures held by the reference
 * (B_DC_3_7_13/src/examples, dPD/tests/test_data, psac mississippi KAT), libsais built from
 * the reference sources (oracle/_ref), brute-force suffix sorting on tiny texts, the DC table
 * digests taken from difference_cover.hpp, and invariants of names/discarding.
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define ORACLE_MAX_LEVELS 64

typedef struct {
    uint64_t n;          /* characters at this level                                    */
    uint64_t samples;    /* |S| incl. dummy (pdcx.hpp:473-510)                          */
    uint64_t names;      /* highest lexicographic name (last_rank, pdcx.hpp:743)         */
    uint64_t reduced;    /* ranks left after discarding (pdcx.hpp:215-228); 0 if unused  */
    uint32_t base_case;  /* solved by slow_suffixarray                                   */
    uint32_t recursed;   /* names not distinct -> recursive call                         */
    uint32_t discarding; /* recursive call used discarding                                */
    uint32_t key_chars;  /* characters per sample key                                    */
    uint32_t merge_chars;/* characters compared before ranks in the merge               */
    uint32_t bits_per_char; /* packing width at level 0 (0 = unpacked)                   */
} oracle_level_stat;

typedef struct {
    uint32_t levels;
    uint32_t sigma;
    oracle_level_stat lv[ORACLE_MAX_LEVELS];
} oracle_stats;

/* ------------------------------------------------------------------------------------------
 * Difference covers (difference_cover.hpp: DC3Param..DC133Param, the DC[D] arrays).
 * ---------------------------------------------------------------------------------------- */
static const uint32_t DC_SETS[][14] = {
    /* X, D, DC... */
    {3, 2, 1, 2},
    {7, 3, 1, 2, 4},
    {13, 4, 1, 2, 4, 10},
    {21, 5, 1, 2, 7, 9, 19},
    {31, 6, 1, 2, 4, 9, 13, 19},
    {39, 7, 1, 2, 17, 21, 23, 28, 31},
    {57, 8, 1, 2, 10, 12, 15, 36, 40, 52},
    {73, 9, 1, 2, 4, 8, 16, 32, 37, 55, 64},
    {91, 10, 1, 2, 8, 17, 28, 57, 61, 69, 71, 74},
    {95, 11, 1, 2, 6, 9, 19, 21, 30, 32, 46, 62, 68},
    {133, 12, 1, 2, 33, 43, 45, 49, 52, 60, 73, 78, 98, 112},
};

typedef struct {
    uint32_t X, D;
    uint32_t dc[16];
    uint8_t in_dc[256];
    uint32_t next_rank[256];
    uint32_t *cmp; /* X*X*3: {depth, r1, r2} */
} dcp_t;

/* DC table: difference_cover.hpp DCxParam::{DC, IN_DC, NEXT_RANK, cmpDepthRanks};
 * cmpDepthRanks by the closed form of B_DC_3_7_13/tools/calc_dcparam.cc:271-301. */
int oracle_dc_params(uint32_t X, uint32_t *D_out, uint32_t *dc_out, uint32_t *next_rank_out,
                     uint32_t *cmp_out) {
    int found = -1;
    for (int k = 0; k < (int)(sizeof(DC_SETS) / sizeof(DC_SETS[0])); k++)
        if (DC_SETS[k][0] == X) found = k;
    if (found < 0) return -1;
    uint32_t D = DC_SETS[found][1];
    uint8_t in_dc[256] = {0};
    for (uint32_t i = 0; i < D; i++) in_dc[DC_SETS[found][2 + i]] = 1;
    if (D_out) *D_out = D;
    if (dc_out)
        for (uint32_t i = 0; i < D; i++) dc_out[i] = DC_SETS[found][2 + i];
    if (next_rank_out) {
        uint32_t c = 0;
        for (uint32_t r = 0; r < X; r++) {
            next_rank_out[r] = c; /* #{d in DC : d < r} */
            c += in_dc[r];
        }
    }
    if (cmp_out) {
        for (uint32_t a = 0; a < X; a++)
            for (uint32_t b = 0; b < X; b++) {
                uint32_t j = 0, r1 = 0, r2 = 0;
                while (j < X && (!in_dc[(a + j) % X] || !in_dc[(b + j) % X])) {
                    r1 += in_dc[(a + j) % X];
                    r2 += in_dc[(b + j) % X];
                    j++;
                }
                uint32_t *e = cmp_out + 3 * (a * X + b);
                e[0] = j;
                e[1] = r1;
                e[2] = r2;
            }
    }
    return 0;
}

static int load_dcp(uint32_t X, dcp_t *p) {
    memset(p, 0, sizeof(*p));
    p->X = X;
    p->cmp = (uint32_t *)malloc(sizeof(uint32_t) * 3 * X * X);
    if (!p->cmp) return -1;
    if (oracle_dc_params(X, &p->D, p->dc, p->next_rank, p->cmp) != 0) {
        free(p->cmp);
        p->cmp = NULL;
        return -1;
    }
    for (uint32_t i = 0; i < p->D; i++) p->in_dc[p->dc[i]] = 1;
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Packing closed form (char_container.hpp:57-71 PackedInteger, :483-545 KPackedInteger):
 * char k of a word sits at bits [64-(k+1)b, 64-kb); chars past `count` are zero.
 * ---------------------------------------------------------------------------------------- */
uint64_t oracle_pack_word(const uint64_t *chars, uint32_t count, uint32_t bits) {
    uint64_t w = 0;
    uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    for (uint32_t k = 0; k < count; k++) {
        uint32_t shift = 64 - (k + 1) * bits;
        w |= (chars[k] & mask) << shift;
    }
    return w;
}

/* cli.cpp:475-539 run_packed_dcx_variant: bits per char from the alphabet size, words per
 * X chars, packing_ratio = PACKED_CHARS / X.  sigma = 256 is outside the reference (it exits,
 * cli.cpp:440-446); this restatement extends it with 9-bit characters. */
int oracle_packing(uint32_t X, uint32_t sigma, uint32_t *bits, uint32_t *chars_per_word,
                   uint32_t *words, double *ratio) {
    uint32_t b = sigma <= 7 ? 3 : sigma <= 31 ? 5 : sigma <= 255 ? 8 : 9;
    uint32_t cpw = 64 / b;
    uint32_t w = (X + cpw - 1) / cpw;
    if (bits) *bits = b;
    if (chars_per_word) *chars_per_word = cpw;
    if (words) *words = w;
    if (ratio) *ratio = (double)(w * cpw) / (double)X;
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * compress_alphabet (cli.cpp:417-463): used bytes -> 1..sigma in byte order, 0 reserved.
 * ---------------------------------------------------------------------------------------- */
int oracle_compress_alphabet(const uint8_t *in, uint64_t n, uint64_t *out, uint32_t *sigma_out) {
    uint64_t counts[256] = {0};
    for (uint64_t i = 0; i < n; i++) counts[in[i]]++;
    uint64_t map[256];
    uint32_t next = 1;
    for (int c = 0; c < 256; c++)
        if (counts[c]) map[c] = next++;
    uint32_t sigma = next - 1;
    if (sigma_out) *sigma_out = sigma;
    for (uint64_t i = 0; i < n; i++) out[i] = map[in[i]];
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Base case: slow_suffixarray (string_util.hpp:87-106) -- std::sort with cmp_substrings.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    const uint64_t *T;
    uint64_t n;
} sub_ctx;

static int cmp_substrings(const void *pa, const void *pb, void *arg) {
    const sub_ctx *c = (const sub_ctx *)arg;
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    uint64_t j = 0;
    while (a + j < c->n && b + j < c->n && c->T[a + j] == c->T[b + j]) j++;
    if (a + j == c->n) return -1; /* a ended first */
    if (b + j == c->n) return 1;
    return c->T[a + j] < c->T[b + j] ? -1 : 1;
}

void oracle_slow_sa(const uint64_t *T, uint64_t n, uint64_t *SA) {
    for (uint64_t i = 0; i < n; i++) SA[i] = i;
    sub_ctx c = {T, n};
    qsort_r(SA, n, sizeof(uint64_t), cmp_substrings, &c);
}

/* ------------------------------------------------------------------------------------------
 * One PDCX level.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    dcp_t dc;
    int pack_samples, pack_merging;
    double discarding_threshold;
    uint64_t base_case_chars;
    uint32_t sigma;
    oracle_stats *st;
} octx;

static inline uint64_t chr(const uint64_t *T, uint64_t n, uint64_t i) { return i < n ? T[i] : 0; }

typedef struct {
    const uint64_t *T;
    uint64_t n;
    uint64_t len;
} key_ctx;

/* DCSampleString::operator< (sample_string.hpp:61): compare the key characters only. */
static int cmp_sample_keys(const void *pa, const void *pb, void *arg) {
    const key_ctx *c = (const key_ctx *)arg;
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    for (uint64_t k = 0; k < c->len; k++) {
        uint64_t x = chr(c->T, c->n, a + k), y = chr(c->T, c->n, b + k);
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

static int keys_equal(const key_ctx *c, uint64_t a, uint64_t b) {
    for (uint64_t k = 0; k < c->len; k++)
        if (chr(c->T, c->n, a + k) != chr(c->T, c->n, b + k)) return 0;
    return 1;
}

typedef struct {
    const uint64_t *T;
    uint64_t n;
    uint64_t len; /* chars compared before ranks */
    const uint64_t *ranks;
    const dcp_t *dc;
} merge_ctx;

/* merge_samples.hpp:181-188 get_global_ranks_pos */
static inline uint64_t rank_pos(const dcp_t *dc, uint64_t i) {
    return (i / dc->X) * dc->D + dc->next_rank[i % dc->X];
}

/* DCMergeSamples::operator< (merge_samples.hpp:87-109). */
static int cmp_merge(const void *pa, const void *pb, void *arg) {
    const merge_ctx *c = (const merge_ctx *)arg;
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    for (uint64_t k = 0; k < c->len; k++) {
        uint64_t x = chr(c->T, c->n, a + k), y = chr(c->T, c->n, b + k);
        if (x != y) return x < y ? -1 : 1;
    }
    const uint32_t *e = c->dc->cmp + 3 * ((a % c->dc->X) * c->dc->X + (b % c->dc->X));
    uint64_t ra = c->ranks[rank_pos(c->dc, a) + e[1]];
    uint64_t rb = c->ranks[rank_pos(c->dc, b) + e[2]];
    if (ra != rb) return ra < rb ? -1 : 1;
    return 0; /* only for a == b */
}

typedef struct {
    uint64_t rank1, rank2, index;
} rri_t;

static int cmp_rri(const void *pa, const void *pb) {
    const rri_t *a = (const rri_t *)pa, *b = (const rri_t *)pb;
    if (a->rank1 != b->rank1) return a->rank1 < b->rank1 ? -1 : 1;
    if (a->rank2 != b->rank2) return a->rank2 < b->rank2 ? -1 : 1;
    return 0;
}

static int level_sa(octx *cx, const uint64_t *T, uint64_t n, int depth, uint64_t *SA);

/* map_back (pdcx.hpp:83-94): recursive index -> text position. */
static inline uint64_t map_back(const dcp_t *dc, const uint64_t *samples_before, uint64_t idx) {
    for (uint32_t i = 0; i < dc->D; i++)
        if (idx < samples_before[i + 1]) return (uint64_t)dc->X * (idx - samples_before[i]) + dc->dc[i];
    return ~0ull;
}

static int level_sa(octx *cx, const uint64_t *T, uint64_t n, int depth, uint64_t *SA) {
    const dcp_t *dc = &cx->dc;
    const uint32_t X = dc->X, D = dc->D;
    oracle_stats *st = cx->st;
    oracle_level_stat *ls = NULL;
    if (st && depth < ORACLE_MAX_LEVELS) {
        ls = &st->lv[depth];
        memset(ls, 0, sizeof(*ls));
        if ((uint32_t)depth + 1 > st->levels) st->levels = depth + 1;
        ls->n = n;
    }

    /* pdcx.hpp:623-633: solve sequentially when n <= max(2pX, 10^4) (p = 1). */
    uint64_t threshold = cx->base_case_chars > 2ull * X ? cx->base_case_chars : 2ull * X;
    if (n <= threshold) {
        if (ls) ls->base_case = 1;
        oracle_slow_sa(T, n, SA);
        return 0;
    }

    /* compute_length_info (pdcx.hpp:473-510): samples are i in [0, n] with i mod X in DC
     * (the dummy sample at n exists iff n mod X in DC). */
    uint64_t samples_before[17];
    uint64_t num_at_mod[256];
    for (uint32_t r = 0; r < X; r++) num_at_mod[r] = (n + X - 1 - r) / X;
    if (dc->in_dc[n % X]) num_at_mod[n % X] += 1;
    samples_before[0] = 0;
    for (uint32_t i = 1; i <= D; i++) samples_before[i] = samples_before[i - 1] + num_at_mod[dc->dc[i - 1]];
    uint64_t m = samples_before[D];

    /* packing configuration (pdcx.hpp:559-560, cli.cpp:475-539, sample_string.hpp:241-249) */
    int packed_samples = cx->pack_samples && depth == 0;
    int packed_merging = cx->pack_merging && depth == 0;
    double ratio = 1.0;
    uint32_t bits = 0;
    if (depth == 0 && (cx->pack_samples || cx->pack_merging)) oracle_packing(X, cx->sigma, &bits, NULL, NULL, &ratio);
    double ratio_s = packed_samples ? ratio : 1.0;
    double ratio_m = packed_merging ? ratio : 1.0;
    /* materialize_sample: begin + i + X * ratio (truncating conversion) */
    uint64_t key_len = (uint64_t)((double)X * ratio_s);
    /* materialize_characters: begin + i + ratio * X - 1 */
    uint64_t merge_len = (uint64_t)((double)X * ratio_m) - 1;
    if (ls) {
        ls->samples = m;
        ls->key_chars = (uint32_t)key_len;
        ls->merge_chars = (uint32_t)merge_len;
        ls->bits_per_char = (packed_samples || packed_merging) ? bits : 0;
    }

    /* Phase 1 (sample_string.hpp:145-164, 225-239): sample keys, sorted by key only. */
    uint64_t *sp = (uint64_t *)malloc(sizeof(uint64_t) * m);
    if (!sp) return -1;
    {
        uint64_t k = 0;
        for (uint64_t q = 0; q * X <= n; q++)
            for (uint32_t t = 0; t < D; t++) {
                uint64_t p = q * X + dc->dc[t];
                if (p <= n) sp[k++] = p;
            }
        /* text order equals rank_pos order */
    }
    key_ctx kc = {T, n, key_len};
    qsort_r(sp, m, sizeof(uint64_t), cmp_sample_keys, &kc);

    /* Phase 2 (compute_ranks.hpp:88-111, 113-134): names and unique flags. */
    uint64_t *name_sorted = (uint64_t *)malloc(sizeof(uint64_t) * m);
    uint8_t *uniq_sorted = (uint8_t *)malloc(m);
    if (!name_sorted || !uniq_sorted) return -1;
    uint64_t rank = 1;
    for (uint64_t j = 0; j < m; j++) {
        if (j > 0 && !keys_equal(&kc, sp[j - 1], sp[j])) rank++;
        name_sorted[j] = rank;
    }
    uint64_t last_rank = m ? name_sorted[m - 1] : 0;
    for (uint64_t j = 0; j < m; j++) {
        int u = 1;
        if (j > 0 && name_sorted[j - 1] == name_sorted[j]) u = 0;
        if (j + 1 < m && name_sorted[j + 1] == name_sorted[j]) u = 0;
        uniq_sorted[j] = (uint8_t)u;
    }
    if (ls) ls->names = last_rank;

    /* text-ordered sample ranks (length m + D zero padding, merge_samples.hpp:162-174) */
    uint64_t *ranks = (uint64_t *)calloc(m + D, sizeof(uint64_t));
    if (!ranks) return -1;

    if (last_rank >= m) {
        /* pdcx.hpp:747-756: names distinct -> sort by index */
        for (uint64_t j = 0; j < m; j++) ranks[rank_pos(dc, sp[j])] = name_sorted[j];
    } else {
        if (ls) ls->recursed = 1;
        /* handle_recursive_call (pdcx.hpp:145-180): order by (mod X, div X). */
        uint64_t *R = (uint64_t *)malloc(sizeof(uint64_t) * m);
        uint8_t *U = (uint8_t *)malloc(m);
        if (!R || !U) return -1;
        for (uint64_t j = 0; j < m; j++) {
            uint64_t p = sp[j];
            uint32_t t = 0;
            while (dc->dc[t] != p % X) t++;
            uint64_t cm = samples_before[t] + p / X;
            R[cm] = name_sorted[j];
            U[cm] = uniq_sorted[j];
        }
        /* num_ranks_after_discarding (pdcx.hpp:215-228) */
        uint64_t discarded = 0;
        for (uint64_t k = 1; k < m; k++) discarded += (U[k] && U[k - 1]);
        uint64_t after = m - discarded;
        double reduction = (double)after / (double)m;
        int use_discarding = reduction <= cx->discarding_threshold;
        if (n <= 80) use_discarding = 0; /* sequential_sa_on_local_ranks path, pdcx.hpp:115-121 */

        if (!use_discarding) {
            /* recursive_call_direct (pdcx.hpp:182-212) or sequential_sa_on_local_ranks
             * (sequential_sa.hpp:34-58): SA of the rank string, rank = 1 + position. */
            uint64_t *SAr = (uint64_t *)malloc(sizeof(uint64_t) * m);
            if (!SAr) return -1;
            if (n <= 80)
                oracle_slow_sa(R, m, SAr);
            else if (level_sa(cx, R, m, depth + 1, SAr))
                return -1;
            for (uint64_t k = 0; k < m; k++) {
                uint64_t p = map_back(dc, samples_before, SAr[k]);
                ranks[rank_pos(dc, p)] = k + 1;
            }
            free(SAr);
        } else {
            if (ls) {
                ls->discarding = 1;
                ls->reduced = after;
            }
            /* recursive_call_with_discarding (pdcx.hpp:231-346) */
            uint64_t *Rr = (uint64_t *)malloc(sizeof(uint64_t) * after);
            uint64_t *red_index = (uint64_t *)malloc(sizeof(uint64_t) * m);
            if (!Rr || !red_index) return -1;
            uint64_t c = 0;
            for (uint64_t k = 0; k < m; k++) {
                int keep = (k == 0) || !(U[k] && U[k - 1]);
                red_index[k] = c;
                if (keep) Rr[c++] = R[k];
            }
            uint64_t *SAr = (uint64_t *)malloc(sizeof(uint64_t) * after);
            uint64_t *ISAr = (uint64_t *)malloc(sizeof(uint64_t) * after);
            if (!SAr || !ISAr) return -1;
            if (level_sa(cx, Rr, after, depth + 1, SAr)) return -1;
            for (uint64_t k = 0; k < after; k++) ISAr[SAr[k]] = k + 1;
            /* RankRankIndex sort by (rank1, rank2), rank2 = 0 for unique (pdcx.hpp:297-323) */
            rri_t *rri = (rri_t *)malloc(sizeof(rri_t) * m);
            if (!rri) return -1;
            for (uint64_t k = 0; k < m; k++) {
                rri[k].rank1 = R[k];
                rri[k].rank2 = U[k] ? 0 : ISAr[red_index[k]];
                rri[k].index = map_back(dc, samples_before, k);
            }
            qsort(rri, m, sizeof(rri_t), cmp_rri);
            for (uint64_t k = 0; k < m; k++) ranks[rank_pos(dc, rri[k].index)] = k + 1;
            free(rri);
            free(SAr);
            free(ISAr);
            free(Rr);
            free(red_index);
        }
        free(R);
        free(U);
    }
    free(sp);
    free(name_sorted);
    free(uniq_sorted);

    /* Phase 4 (merge_samples.hpp:237-260, 335-357): all suffixes by the DC comparator. */
    for (uint64_t i = 0; i < n; i++) SA[i] = i;
    merge_ctx mc = {T, n, merge_len, ranks, dc};
    qsort_r(SA, n, sizeof(uint64_t), cmp_merge, &mc);
    free(ranks);
    return 0;
}

/* PDCX::compute_sa entry (pdcx.hpp:529) behind compress_alphabet (cli.cpp:417-463).
 * Returns 0, or -1 (bad X / allocation), -2 (empty input). */
int oracle_dcx_sa(const uint8_t *text, uint64_t n, uint32_t X, int pack_samples, int pack_merging,
                  double discarding_threshold, uint64_t base_case_chars, uint64_t *SA,
                  oracle_stats *st) {
    octx cx;
    memset(&cx, 0, sizeof(cx));
    if (load_dcp(X, &cx.dc)) return -1;
    if (st) memset(st, 0, sizeof(*st));
    uint64_t *T = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!T) return -1;
    uint32_t sigma = 0;
    oracle_compress_alphabet(text, n, T, &sigma);
    cx.pack_samples = pack_samples;
    cx.pack_merging = pack_merging;
    cx.discarding_threshold = discarding_threshold;
    cx.base_case_chars = base_case_chars;
    cx.sigma = sigma;
    cx.st = st;
    if (st) st->sigma = sigma;
    int rc = level_sa(&cx, T, n, 0, SA);
    free(T);
    free(cx.dc.cmp);
    return rc;
}

/* DCX on an integer string whose characters are already >= 1 (used by tests to drive the
 * recursion levels directly). */
int oracle_dcx_sa_int(const uint64_t *T, uint64_t n, uint32_t X, double discarding_threshold,
                      uint64_t base_case_chars, uint64_t *SA, oracle_stats *st) {
    octx cx;
    memset(&cx, 0, sizeof(cx));
    if (load_dcp(X, &cx.dc)) return -1;
    if (st) memset(st, 0, sizeof(*st));
    cx.discarding_threshold = discarding_threshold;
    cx.base_case_chars = base_case_chars;
    cx.st = st;
    int rc = level_sa(&cx, T, n, 0, SA);
    free(cx.dc.cmp);
    return rc;
}

/* ------------------------------------------------------------------------------------------
 * check_suffixarray (sa_check.hpp:56-184), sequential: permutation test, then the triples
 * (rank[i], rank[i+1], T[i]) ordered by rank[i] must be non-decreasing in (T[i], rank[i+1]).
 * Returns 1 if SA is the suffix array of text.
 * ---------------------------------------------------------------------------------------- */
int oracle_check_sa(const uint8_t *text, uint64_t n, const uint64_t *SA) {
    uint64_t *rank = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    if (!rank) return -1;
    for (uint64_t k = 0; k < n; k++) {
        uint64_t p = SA[k];
        if (p >= n || rank[p] != 0) {
            free(rank);
            return 0; /* no permutation */
        }
        rank[p] = k + 1;
    }
    int ok = 1;
    for (uint64_t k = 0; k + 1 < n && ok; k++) {
        uint64_t p = SA[k], q = SA[k + 1];
        uint64_t r2p = rank[p + 1], r2q = rank[q + 1]; /* rank[n] = 0 */
        if (text[p] != text[q])
            ok = text[p] < text[q];
        else
            ok = r2p <= r2q;
    }
    free(rank);
    return ok;
}

/* io.hpp:80-98 with uint40 (uint_types.hpp:262): 5 bytes little endian per entry. */
void oracle_write_uint40(const uint64_t *SA, uint64_t n, uint8_t *out) {
    for (uint64_t k = 0; k < n; k++)
        for (int b = 0; b < 5; b++) out[5 * k + b] = (uint8_t)(SA[k] >> (8 * b));
}

/* ------------------------------------------------------------------------------------------
 * The C ABI of include/dcx.h restated on the CPU under a dcx_host_ prefix: the same arguments,
 * the same status codes for the same argument errors, and the same output encodings, so a test
 * drives the HIP engine and this restatement with identical calls and compares the bytes.
 * Differences, all because there is no device here: the handle holds no HBM budget (admission
 * never rejects), device-pointer flags are rejected with DCX_E_ARG, and stats carry the level
 * counts, sigma and the host wall time only.
 * ---------------------------------------------------------------------------------------- */
#include <time.h>
#include "dcx.h"

#define HOST_MAX_TEXT (1ull << 32) /* the engine's limit: 32-bit positions 0 .. 2^32 - 1      */
#define HOST_BASE_CASE_MAX (1ull << 31)             /* the engine's base-case limit       */

struct dcx_handle {
    char err[256];
};

static int host_fail(dcx_handle *h, int code, const char *msg) {
    if (h) snprintf(h->err, sizeof(h->err), "%s", msg);
    return code;
}

void dcx_host_default_options(dcx_options *opt) {
    if (!opt) return;
    memset(opt, 0, sizeof(*opt));
    opt->X = 39;
    opt->pack_samples = 1;
    opt->pack_merging = 1;
    opt->discarding_threshold = 0.7;
    opt->base_case_chars = 10000;
}

int dcx_host_create(dcx_handle **out, int device, uint64_t hbm_budget_bytes) {
    (void)device;
    (void)hbm_budget_bytes;
    if (!out) return DCX_E_ARG;
    *out = (dcx_handle *)calloc(1, sizeof(dcx_handle));
    return *out ? DCX_OK : DCX_E_NOMEM;
}

int dcx_host_destroy(dcx_handle *h) {
    if (!h) return DCX_E_ARG;
    free(h);
    return DCX_OK;
}

/* the option checks of the engine, in its order: threshold, X, base case */
static int host_check_options(dcx_handle *h, const dcx_options *o) {
    if (!(o->discarding_threshold >= 0.0 && o->discarding_threshold <= 1.0))
        return host_fail(h, DCX_E_ARG, "discarding_threshold must lie in [0, 1]");
    int known = 0;
    for (size_t i = 0; i < sizeof(DC_SETS) / sizeof(DC_SETS[0]); i++) known |= DC_SETS[i][0] == o->X;
    if (!known) return host_fail(h, DCX_E_INVALID_X, "X is not a supported difference cover");
    if (o->base_case_chars > HOST_BASE_CASE_MAX) return host_fail(h, DCX_E_ARG, "base_case_chars must be <= 2^31");
    return DCX_OK;
}

static void host_store(const uint64_t *SA, uint64_t n, void *out, uint32_t width) {
    uint8_t *o = (uint8_t *)out;
    for (uint64_t k = 0; k < n; k++)
        for (uint32_t b = 0; b < width; b++) o[width * k + b] = (uint8_t)(SA[k] >> (8 * b));
}

int dcx_host_build_sa(dcx_handle *h, const uint8_t *text, uint64_t n, const dcx_options *o, void *sa_out,
                      uint32_t width, dcx_stats *stats) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    if (!h) return DCX_E_ARG;
    if (!o || (!text && n) || (!sa_out && n)) return host_fail(h, DCX_E_ARG, "null argument");
    if (width != 4 && width != 5 && width != 8) return host_fail(h, DCX_E_ARG, "sa_width_bytes must be 4, 5 or 8");
    if (o->text_on_device || o->sa_on_device) return host_fail(h, DCX_E_ARG, "device pointers on the host path");
    int rc0 = host_check_options(h, o);
    if (rc0 != DCX_OK) return rc0;
    if (n > HOST_MAX_TEXT) return host_fail(h, DCX_E_UNSUPPORTED, "text too long for 32-bit positions");
    if (stats) memset(stats, 0, sizeof(*stats));
    if (n == 0) return DCX_OK;
    uint64_t *SA = (uint64_t *)malloc(sizeof(uint64_t) * n);
    oracle_stats *st = (oracle_stats *)calloc(1, sizeof(oracle_stats));
    if (!SA || !st) {
        free(SA);
        free(st);
        return host_fail(h, DCX_E_NOMEM, "host allocation failed");
    }
    int rc = oracle_dcx_sa(text, n, o->X, o->pack_samples != 0, o->pack_merging != 0, o->discarding_threshold,
                           o->base_case_chars, SA, st);
    if (rc == 0) host_store(SA, n, sa_out, width);
    if (rc == 0 && stats) {
        stats->sigma = st->sigma;
        stats->levels = st->levels < DCX_MAX_LEVELS ? st->levels : DCX_MAX_LEVELS;
        for (uint32_t l = 0; l < stats->levels; l++) {
            dcx_level_stats *d = &stats->level[l];
            const oracle_level_stat *s = &st->lv[l];
            d->n = s->n;
            d->samples = s->samples;
            d->names = s->names;
            d->reduced = s->reduced;
            d->base_case = s->base_case;
            d->recursed = s->recursed;
            d->discarding = s->discarding;
            d->key_chars = s->key_chars;
            d->merge_chars = s->merge_chars;
            d->bits_per_char = s->bits_per_char;
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        stats->ms_total = (double)(t1.tv_sec - t0.tv_sec) * 1e3 + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-6;
    }
    free(SA);
    free(st);
    return rc == 0 ? DCX_OK : host_fail(h, DCX_E_NOMEM, "host allocation failed");
}

/* dcx_build_sa_noalloc on the CPU: the engine's argument checks in its order (the text limit first), then
 * dcx_host_build_sa.  Here host memory stands in for the device memory the engine's call takes, so the
 * pointers are host pointers and the pointer flags of the options are ignored, as the engine ignores them. */
int dcx_host_build_sa_noalloc(dcx_handle *h, const uint8_t *text, uint64_t n, const dcx_options *o, void *sa_out,
                              uint32_t width, dcx_stats *stats) {
    if (!h) return DCX_E_ARG;
    if (!o) return host_fail(h, DCX_E_ARG, "null argument");
    if (n > DCX_NOALLOC_MAX_CHARS) return host_fail(h, DCX_E_UNSUPPORTED, "text too long for the call without allocation");
    dcx_options oo = *o;
    oo.text_on_device = 0;
    oo.sa_on_device = 0;
    return dcx_host_build_sa(h, text, n, &oo, sa_out, width, stats);
}

/* dcx_build_sa_batch on the CPU: each text's SA on its own (the joined-string pipeline of the engine
 * must give exactly these), written as 32-bit positions local to the text at sa_out + offsets[k] - offsets[0]. */
int dcx_host_build_sa_batch(dcx_handle *h, const uint8_t *texts, const uint64_t *offsets, uint32_t count,
                            const dcx_options *o, uint32_t *sa_out, dcx_stats *stats) {
    if (!h) return DCX_E_ARG;
    if (!o || !offsets) return host_fail(h, DCX_E_ARG, "null argument");
    if (o->text_on_device || o->sa_on_device) return host_fail(h, DCX_E_ARG, "device pointers on the host path");
    int rc0 = host_check_options(h, o);
    if (rc0 != DCX_OK) return rc0;
    if (count == 0) {
        if (stats) memset(stats, 0, sizeof(*stats));
        return DCX_OK;
    }
    const uint64_t base = offsets[0], total = offsets[count] - offsets[0];
    for (uint32_t k = 0; k < count; k++) {
        if (offsets[k + 1] < offsets[k]) return host_fail(h, DCX_E_ARG, "offsets must be non-decreasing");
        if (offsets[k + 1] - offsets[k] + 1 > HOST_MAX_TEXT) /* with its separator */
            return host_fail(h, DCX_E_UNSUPPORTED, "a text of the batch is too long for 32-bit positions");
    }
    if (total && (!texts || !sa_out)) return host_fail(h, DCX_E_ARG, "null argument");
    for (uint32_t k = 0; k < count; k++) {
        const uint64_t n = offsets[k + 1] - offsets[k];
        int rc = dcx_host_build_sa(h, texts + offsets[k], n, o, sa_out + (offsets[k] - base), 4, NULL);
        if (rc != DCX_OK) return rc;
    }
    if (stats) memset(stats, 0, sizeof(*stats));
    return DCX_OK;
}

int dcx_host_check_sa(dcx_handle *h, const uint8_t *text, uint64_t n, const void *sa, uint32_t width,
                      int on_device, int *ok) {
    if (!h) return DCX_E_ARG;
    if (!ok || (n && (!text || !sa))) return host_fail(h, DCX_E_ARG, "null argument");
    if (width != 4 && width != 5 && width != 8) return host_fail(h, DCX_E_ARG, "sa_width_bytes must be 4, 5 or 8");
    if (on_device) return host_fail(h, DCX_E_ARG, "device pointers on the host path");
    if (n >= (1ull << 32) - 1) return host_fail(h, DCX_E_UNSUPPORTED, "text too long for this build");
    *ok = 0;
    uint64_t *SA = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!SA) return host_fail(h, DCX_E_NOMEM, "host allocation failed");
    const uint8_t *b = (const uint8_t *)sa;
    for (uint64_t k = 0; k < n; k++) {
        uint64_t v = 0;
        for (uint32_t i = 0; i < width; i++) v |= (uint64_t)b[width * k + i] << (8 * i);
        SA[k] = v;
    }
    int r = oracle_check_sa(text, n, SA);
    free(SA);
    if (r < 0) return host_fail(h, DCX_E_NOMEM, "host allocation failed");
    *ok = r;
    return DCX_OK;
}

const char *dcx_host_strerror(int code) {
    switch (code) {
        case DCX_OK: return "ok";
        case DCX_E_INVALID_X: return "unsupported difference cover X";
        case DCX_E_ALPHABET: return "alphabet cannot be represented";
        case DCX_E_NOMEM: return "host memory exhausted";
        case DCX_E_ARG: return "invalid argument";
        case DCX_E_WIDTH: return "SA width too small for the text length";
        case DCX_E_UNSUPPORTED: return "unsupported size";
        default: return "unknown error";
    }
}

const char *dcx_host_last_error(dcx_handle *h) { return h ? h->err : "no handle"; }
