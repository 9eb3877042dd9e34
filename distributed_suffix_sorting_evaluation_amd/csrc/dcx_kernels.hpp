// dcx_kernels.hpp -- DCX-specific device code: difference-cover parameters, packed tuple
// construction, naming, recursion bookkeeping, the DC comparator and the single-workgroup base case.
#pragma once

#include "primitives.hpp"

namespace dcx {

constexpr uint32_t MAX_X = 133;
constexpr uint32_t MAX_D = 12;
constexpr uint32_t BASE_CASE_MAX = 16384;
// the largest base_case_chars: levels of more than BASE_CASE_MAX characters at or below the threshold are
// sorted by the refinement over their whole length, whose prefix lengths are 32-bit
constexpr uint64_t BASE_CASE_CHARS_MAX = 1ull << 31;

// Difference covers of lightweight_dcx/src/pdcx/difference_cover.hpp (DC3Param .. DC133Param, DC[D]).
struct DcSet {
    uint32_t X, D;
    uint32_t dc[MAX_D];
};
static const DcSet kDcSets[] = {
    {3, 2, {1, 2}},
    {7, 3, {1, 2, 4}},
    {13, 4, {1, 2, 4, 10}},
    {21, 5, {1, 2, 7, 9, 19}},
    {31, 6, {1, 2, 4, 9, 13, 19}},
    {39, 7, {1, 2, 17, 21, 23, 28, 31}},
    {57, 8, {1, 2, 10, 12, 15, 36, 40, 52}},
    {73, 9, {1, 2, 4, 8, 16, 32, 37, 55, 64}},
    {91, 10, {1, 2, 8, 17, 28, 57, 61, 69, 71, 74}},
    {95, 11, {1, 2, 6, 9, 19, 21, 30, 32, 46, 62, 68}},
    {133, 12, {1, 2, 33, 43, 45, 49, 52, 60, 73, 78, 98, 112}},
};

inline const DcSet* find_dc(uint32_t X) {
    for (const auto& s : kDcSets)
        if (s.X == X) return &s;
    return nullptr;
}

// cmpDepthRanks[a][b] = {l, r1, r2}: l = least j >= 0 with (a+j) mod X and (b+j) mod X both in DC,
// r1 / r2 = number of DC positions in [a, a+l) / [b, b+l) (closed form of
// B_DC_3_7_13/tools/calc_dcparam.cc:271-301; equals difference_cover.hpp's tables).
// Packed as l | r1 << 8 | r2 << 16.
inline std::vector<uint32_t> make_cmp_table(const DcSet& s) {
    std::vector<uint8_t> in(s.X, 0);
    for (uint32_t i = 0; i < s.D; i++) in[s.dc[i]] = 1;
    std::vector<uint32_t> t(s.X * s.X);
    for (uint32_t a = 0; a < s.X; a++)
        for (uint32_t b = 0; b < s.X; b++) {
            uint32_t j = 0, r1 = 0, r2 = 0;
            while (j < s.X && (!in[(a + j) % s.X] || !in[(b + j) % s.X])) {
                r1 += in[(a + j) % s.X];
                r2 += in[(b + j) % s.X];
                j++;
            }
            t[a * s.X + b] = j | (r1 << 8) | (r2 << 16);
        }
    return t;
}

// Per-level parameters handed to kernels by value.
struct DcDev {
    uint32_t X, D;
    uint32_t dc[MAX_D];
    uint32_t sb[MAX_D + 1];     // samples_before (pdcx.hpp:499-508): class-major offsets
    uint8_t next_rank[MAX_X + 3];  // #{d in DC : d < r}  (difference_cover.hpp NEXT_RANK)
    uint8_t in_dc[MAX_X + 3];      // r in DC (difference_cover.hpp IN_DC)
    uint32_t bnd;                  // first position whose sample-key window reaches past the text end
    uint64_t inv_X, inv_D;         // ceil(2^32 / X), ceil(2^32 / D) for division without a divide
    uint64_t in_mask[3];           // bit r set iff r in DC (X <= 133 < 192)
};

// r in DC, from the mask (no table load)
__device__ __forceinline__ uint32_t dc_in(const DcDev& c, uint32_t r) {
    const uint32_t g = r >> 6;
    const uint64_t w = g == 0 ? c.in_mask[0] : (g == 1 ? c.in_mask[1] : c.in_mask[2]);
    return (uint32_t)(w >> (r & 63)) & 1u;
}
// #{d in DC : d < r} (NEXT_RANK), as population counts of the mask below r
__device__ __forceinline__ uint32_t dc_next_rank(const DcDev& c, uint32_t r) {
    const uint32_t g = r >> 6;
    const uint64_t below = (1ull << (r & 63)) - 1;
    const uint64_t a0 = g == 0 ? (c.in_mask[0] & below) : c.in_mask[0];
    const uint64_t a1 = g == 0 ? 0ull : (g == 1 ? (c.in_mask[1] & below) : c.in_mask[1]);
    const uint64_t a2 = g == 2 ? (c.in_mask[2] & below) : 0ull;
    return (uint32_t)(__popcll(a0) + __popcll(a1) + __popcll(a2));
}

// floor(p / d) for 32-bit p with inv = ceil(2^32 / d): the product estimate is q or q + 1
__device__ __forceinline__ uint32_t fast_div(uint32_t p, uint32_t d, uint64_t inv) {
    uint32_t q = (uint32_t)(((uint64_t)p * inv) >> 32);
    if ((uint64_t)q * d > p) q--;
    return q;
}

// k-th sample in text order (k = rank_pos of the sample, merge_samples.hpp:181-188)
__device__ __forceinline__ uint32_t sample_pos(const DcDev& c, uint32_t k) {
    uint32_t q = fast_div(k, c.D, c.inv_D);
    return q * c.X + c.dc[k - q * c.D];
}
// class-major index of the k-th sample (DCRankIndex::cmp_mod_div order, compute_ranks.hpp:49-56)
__device__ __forceinline__ uint32_t class_major(const DcDev& c, uint32_t k) {
    uint32_t q = fast_div(k, c.D, c.inv_D);
    return c.sb[k - q * c.D] + q;
}
// inverse of class_major (map_back of pdcx.hpp:83-94, expressed as the text-order sample index)
__device__ __forceinline__ uint32_t class_major_inv(const DcDev& c, uint32_t cm) {
    uint32_t t = c.D - 1;
    while (cm < c.sb[t]) t--;
    return (cm - c.sb[t]) * c.D + t;
}
// index of the first sample at or after text position i
__device__ __forceinline__ uint32_t rank_pos(const DcDev& c, uint32_t i) {
    uint32_t q = fast_div(i, c.X, c.inv_X);
    return q * c.D + dc_next_rank(c, i - q * c.X);
}

// ------------------------------------------------------------------------------------------------
// Text views: characters are 1..sigma (level 0, through the alphabet map) or names >= 1 (deeper
// levels); every position at or past n reads as the padding character 0 (sample_string.hpp:112-117).
// ------------------------------------------------------------------------------------------------
struct TextU8 {
    const uint8_t* t;
    uint32_t n;
    const uint16_t* lut;
    __device__ __forceinline__ uint32_t at(uint64_t i) const { return i < n ? (uint32_t)lut[t[i]] : 0u; }
};
struct TextU32 {
    const uint32_t* t;
    uint32_t n;
    __device__ __forceinline__ uint32_t at(uint64_t i) const { return i < n ? t[i] : 0u; }
};

// ------------------------------------------------------------------------------------------------
// byte alphabet (compress_alphabet, cli.cpp:417-463): which byte values occur, as a 256-bit mask.  Only
// presence matters for the alphabet map, so every byte just stores a 1 into its LDS flag (plain stores:
// lanes that hit the same flag all write the same value; no atomics, no per-byte register arithmetic --
// 2.9x faster than or-ing each byte into per-thread register masks).  A workgroup reads 16 KB per step,
// four 16-byte loads per thread in flight.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_byte_mask(const uint8_t* __restrict__ t, uint64_t n,
                                                   unsigned long long* __restrict__ mask) {
    __shared__ uint32_t s_f[256];
    s_f[threadIdx.x] = 0;
    __syncthreads();
    constexpr uint64_t chunk = 256ull * 16 * 4;
    for (uint64_t cb = blockIdx.x * chunk; cb < n; cb += (uint64_t)gridDim.x * chunk) {
        if (cb + chunk <= n && ((reinterpret_cast<uintptr_t>(t + cb) & 15) == 0)) {
            uint4 v[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                v[k] = *reinterpret_cast<const uint4*>(t + cb + (k * 256ull + threadIdx.x) * 16);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t x[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int b = 0; b < 4; b++) s_f[(x[q] >> (8 * b)) & 0xFF] = 1;
            }
        } else {
            for (uint64_t i = cb + threadIdx.x; i < cb + chunk && i < n; i += 256) s_f[t[i]] = 1;
        }
    }
    __syncthreads();
    const uint64_t seen = __ballot(s_f[threadIdx.x] != 0);  // wave w covers byte values [64 w, 64 w + 64)
    if (__lane_id() == 0 && seen) atomicOr(&mask[threadIdx.x >> 6], (unsigned long long)seen);
}

// histogram of the character codes lut[t[q]] at every stride-th position (codes 0..256): the order-0
// entropy of a level-0 text decides how many characters its first sort round takes
__global__ __launch_bounds__(256) void k_code_hist(const uint8_t* __restrict__ t, uint64_t n,
                                                   const uint16_t* __restrict__ lut, uint64_t stride,
                                                   unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_h[257];
    __shared__ uint16_t s_lut[256];
    for (int i = threadIdx.x; i < 257; i += 256) s_h[i] = 0;
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    for (uint64_t q = (blockIdx.x * 256ull + threadIdx.x) * stride; q < n; q += (uint64_t)gridDim.x * 256 * stride)
        atomicAdd(&s_h[s_lut[t[q]]], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < 257; i += 256)
        if (s_h[i]) atomicAdd(&hist[i], (unsigned long long)s_h[i]);
}

// ------------------------------------------------------------------------------------------------
// refinement sort: packed characters, heads, active compaction, later-round keys, write back
// ------------------------------------------------------------------------------------------------
// The reference packs a fixed number of characters per word (PackedInteger, char_container.hpp:57-71);
// here the whole text becomes one bit stream and any window of it is a key.  dense != 0 codes character
// c as c - 1 (the padding also reads 0): one bit less whenever the largest character is a power of two.
// Character codes packed into an MSB-first bit stream, `bits` bits per character: code of character q
// at stream bits [q * bits, (q + 1) * bits), codes lut[c] - dense for q < n and 0 (padding) beyond.
// The first L characters of suffix i are the top L * bits bits of the 64-bit window at bit i * bits
// (StreamKeys): 1/8 to 1/32 of the bytes of one word per position.  A workgroup packs PK_CHARS
// characters (a whole number of words for every width) staged in LDS.
constexpr uint32_t PK_CHARS = 8192;
template <class CharT>
__global__ __launch_bounds__(256) void k_pack_text(const CharT* __restrict__ t, uint64_t n,
                                                   const uint16_t* __restrict__ lut, uint32_t bits, uint32_t dense,
                                                   uint64_t nwords, uint64_t* __restrict__ PT) {
    __shared__ uint32_t s_c[PK_CHARS];
    __shared__ uint16_t s_lut[256];
    const uint32_t tid = threadIdx.x;
    if (sizeof(CharT) == 1) s_lut[tid] = lut[tid];
    __syncthreads();
    const uint32_t wpb = PK_CHARS * bits / 64;  // words per chunk of PK_CHARS characters
    for (uint64_t chunk = blockIdx.x; chunk * wpb < nwords; chunk += gridDim.x) {
        const uint64_t q0 = chunk * PK_CHARS;
        for (uint32_t i = tid; i < PK_CHARS; i += 256) {
            const uint64_t q = q0 + i;
            uint32_t c = 0;
            if (q < n) c = (sizeof(CharT) == 1 ? (uint32_t)s_lut[(uint32_t)t[q]] : (uint32_t)t[q]) - dense;
            s_c[i] = c;
        }
        __syncthreads();
        for (uint32_t wl = tid; wl < wpb; wl += 256) {
            const uint64_t w = chunk * wpb + wl;
            if (w >= nwords) break;
            uint64_t word = 0;
            const uint32_t qa = (64 * wl) / bits, qb = (64 * wl + 63) / bits;  // characters overlapping the word
            for (uint32_t q = qa; q <= qb; q++) {
                // bit position of the character's lowest bit within the word (negative: it continues in the
                // next word)
                const int32_t pos = (int32_t)(64 * (wl + 1)) - (int32_t)((q + 1) * bits);
                const uint64_t c = s_c[q];
                word |= pos >= 0 ? (c << pos) : (c >> (-pos));
            }
            PT[w] = word;
        }
        __syncthreads();
    }
}

// the same for widths that divide 64 (no character straddles two words): a thread packs the 64 / bits
// characters of one word straight from global memory (consecutive loads the compiler merges)
template <class CharT, uint32_t BITS>
__global__ __launch_bounds__(256) void k_pack_aligned(const CharT* __restrict__ t, uint64_t n,
                                                      const uint16_t* __restrict__ lut, uint32_t dense, uint64_t nwords,
                                                      uint64_t* __restrict__ PT) {
    constexpr uint32_t CPW = 64 / BITS;
    __shared__ uint16_t s_lut[256];
    if (sizeof(CharT) == 1) s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    for (uint64_t w = blockIdx.x * 256ull + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * 256) {
        const uint64_t q0 = w * CPW;
        uint32_t c[CPW];
        if (q0 + CPW <= n) {
#pragma unroll
            for (uint32_t k = 0; k < CPW; k++) c[k] = (uint32_t)t[q0 + k];
        } else {
#pragma unroll
            for (uint32_t k = 0; k < CPW; k++) c[k] = q0 + k < n ? (uint32_t)t[q0 + k] : 0xFFFFFFFFu;
        }
        uint64_t word = 0;
#pragma unroll
        for (uint32_t k = 0; k < CPW; k++) {
            const uint32_t code = c[k] == 0xFFFFFFFFu ? 0u : (sizeof(CharT) == 1 ? (uint32_t)s_lut[c[k]] : c[k]) - dense;
            word |= (uint64_t)code << (64 - (k + 1) * BITS);
        }
        PT[w] = word;
    }
}

// words of the stream of `chars` characters, with the slack word StreamKeys reads past a window
__host__ __device__ inline uint64_t stream_words(uint64_t chars, uint32_t bits) { return (chars * bits + 63) / 64 + 2; }

__host__ __device__ inline uint64_t top_mask(uint32_t nbits) { return nbits >= 64 ? ~0ull : (nbits == 0 ? 0ull : ~0ull << (64 - nbits)); }

__device__ __forceinline__ uint64_t top_bits(uint64_t w, uint32_t nbits) {
    return nbits >= 64 ? w : (nbits == 0 ? 0ull : (w & (~0ull << (64 - nbits))));
}

// ------------------------------------------------------------------------------------------------
// Round 0 keyed by an order-preserving prefix code.  A byte text of order-0 entropy H carries H bits per
// 8-bit character, so a key of raw characters spends 8 bits for H bits of ordering.  An alphabetic code
// (codewords in character order, none a prefix of another: CODE_MAX_LEN bits at most) orders the bit
// strings of coded suffixes as the suffixes themselves, because two suffixes first differ inside the
// codewords of their first differing characters.  A round-0 key is the first kbits bits of the codes of
// the suffix's first L characters (zeros after the L-th), so a key of kbits bits resolves about kbits / H
// characters instead of kbits / 8.  Suffixes with equal keys share the same whole characters inside the
// kbits bits (a prefix code decodes one way), k of them, and the group rounds go on from character k,
// which every member finds from its own characters (code_chars).  ctab[c] = (length << 24) | the codeword
// of character code c, left-aligned in the low 24 bits.
// ------------------------------------------------------------------------------------------------
constexpr uint32_t CODE_MAX_LEN = 24;
constexpr uint32_t CODE_KEY_BITS_MAX = CK_BITS;  // round-0 keys are CodedKeys (primitives.hpp)

// whole characters covered by the first kbits code bits of suffix p (at most L)
__device__ __forceinline__ uint32_t code_chars(const uint64_t* __restrict__ s, uint32_t bits, uint64_t p,
                                               const uint32_t* __restrict__ ctab, uint32_t kbits, uint32_t L) {
    // eight characters per step, their lengths loaded together (bits <= 8: a window holds at least eight)
    const StreamKeys<false> win{s, bits, ~0ull};
    const uint32_t cm = (1u << bits) - 1;
    uint32_t k = 0, sum = 0;
    while (k < L) {
        const uint64_t x = win(p + k);
        uint32_t len[8];
#pragma unroll
        for (int i = 0; i < 8; i++) len[i] = ctab[(uint32_t)(x >> (64 - bits * (i + 1))) & cm] >> 24;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if (k == L) return L;
            sum += len[i];
            if (sum > kbits) return k;
            k++;
        }
    }
    return L;
}

// The next characters of a suffix in a group round: `cbits` bits of its characters from C on.  After a
// coded round 0 (ctab != nullptr) a member's characters start at k + C instead, k = code_chars of the
// suffix (the same for every member of a group), and stop at character L.
struct RoundChars {
    const uint64_t* s;
    uint32_t bits, C, cbits;
    const uint32_t* ctab;
    uint32_t kbits, L;
    __device__ __forceinline__ uint64_t operator()(uint64_t p) const {
        return with_k(p, ctab ? code_chars(s, bits, p, ctab, kbits, L) : 0u);
    }
    // the same when the member's k is known (its group's)
    __device__ __forceinline__ uint64_t with_k(uint64_t p, uint32_t k) const {
        uint64_t start = p + C;
        uint32_t nb = cbits;
        if (ctab) {
            const uint32_t c0 = k + C < L ? k + C : L;
            start = p + c0;
            nb = cbits < (L - c0) * bits ? cbits : (L - c0) * bits;
        }
        const StreamKeys<false> w{s, bits, top_mask(nb)};
        return w(start);
    }
    // -1, 0 or 1: suffix p against suffix q (members of one group, so the same k) on the characters after
    // this round's, up to character L
    __device__ __forceinline__ int cmp_after(uint64_t p, uint64_t q, uint32_t k) const {
        const uint32_t cpw = 64 / bits;
        const uint32_t c0 = (ctab ? (k + C < L ? k + C : L) : C) + cbits / bits;
        for (uint32_t t = c0; t < L; t += cpw) {
            const uint32_t nc = L - t < cpw ? L - t : cpw;
            const StreamKeys<false> w{s, bits, top_mask(nc * bits)};
            const uint64_t a = w(p + t), b = w(q + t);
            if (a != b) return a < b ? -1 : 1;
        }
        return 0;
    }
};

// k of every group's first member after a coded round 0 (code_chars), lk[g] for g < G; with kmin, their
// minimum into *kmin (set to L beforehand).  first(g) = order[act_idx[slot of g's first member]] with
// the slot given directly (gfirst: order slots, act_idx == nullptr) or through the large list.
__global__ __launch_bounds__(256) void k_group_k(RoundChars rc, uint32_t G, const uint32_t* __restrict__ gfirst,
                                                 const uint32_t* __restrict__ slots,
                                                 const uint32_t* __restrict__ act_idx,
                                                 const uint32_t* __restrict__ order, uint32_t* __restrict__ lk,
                                                 uint32_t* __restrict__ kmin) {
    uint32_t m = rc.L;
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < G; g += gridDim.x * 256ull) {
        const uint32_t j = act_idx ? act_idx[slots[gfirst[g]]] : gfirst[g];
        const uint32_t k = code_chars(rc.s, rc.bits, order[j], rc.ctab, rc.kbits, rc.L);
        if (lk) lk[g] = k;
        m = k < m ? k : m;
    }
    if (kmin) {
        // the workgroup's minimum, then one atomic
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t y = (uint32_t)__shfl_xor((int)m, o);
            m = y < m ? y : m;
        }
        if (__lane_id() == 0) atomicMin(kmin, m);
    }
}

// group heads of sorted keys, comparing only the top kbits
// A thread per 16 consecutive positions (a grid of N / 16 + 1 threads): eight 16-byte key loads, one
// 16-byte store of flags (keys and H are 256-byte aligned buffers).
__global__ __launch_bounds__(256) void k_heads(const uint64_t* __restrict__ keys, uint64_t N, uint32_t kbits,
                                               uint8_t* __restrict__ H) {
    const uint64_t mask = top_mask(kbits);
    const uint64_t groups = N / 16 + 1;  // covers positions 0..N
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < groups; g += gridDim.x * 256ull) {
        const uint64_t j0 = g * 16;
        if (j0 + 16 <= N) {
            uint64_t w[17];
            w[0] = j0 ? (keys[j0 - 1] & mask) : ~(keys[0] & mask);  // position 0 is always a head
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const uint4 x = reinterpret_cast<const uint4*>(keys + j0)[v];
                w[1 + 2 * v] = ((uint64_t)x.y << 32 | x.x) & mask;
                w[2 + 2 * v] = ((uint64_t)x.w << 32 | x.z) & mask;
            }
            uint32_t h[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 16; k++) h[k >> 2] |= (uint32_t)(w[k + 1] != w[k]) << (8 * (k & 3));
            reinterpret_cast<uint4*>(H + j0)[0] = make_uint4(h[0], h[1], h[2], h[3]);
        } else {
            for (uint64_t j = j0; j < j0 + 16 && j <= N; j++)
                H[j] = (j == 0 || j == N) ? 1 : (top_bits(keys[j], kbits) != top_bits(keys[j - 1], kbits));
        }
    }
}

// H[j] = 1 where a[j] != a[j - 1], and at 0 and N.  Used for the heads from the low words written by the
// last radix pass (sort_pairs_from's low32: equal low words within a run of the last digit; the run starts
// are marked by k_mark_runs) and for the sample groups (group ids in sorted order).  A thread per 16
// consecutive positions: four 16-byte loads, one 16-byte store (a and H are 256-byte aligned buffers).
__global__ __launch_bounds__(256) void k_heads_u32(const uint32_t* __restrict__ low, uint64_t N,
                                                   uint8_t* __restrict__ H) {
    const uint64_t groups = N / 16 + 1;  // covers positions 0..N
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < groups; g += gridDim.x * 256ull) {
        const uint64_t j0 = g * 16;
        if (j0 + 16 <= N) {
            uint32_t w[17];
            w[0] = j0 ? low[j0 - 1] : ~low[0];  // position 0 is always a head
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const uint4 x = reinterpret_cast<const uint4*>(low + j0)[v];
                w[1 + 4 * v] = x.x;
                w[2 + 4 * v] = x.y;
                w[3 + 4 * v] = x.z;
                w[4 + 4 * v] = x.w;
            }
            uint32_t h[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 16; k++) h[k >> 2] |= (uint32_t)(w[k + 1] != w[k]) << (8 * (k & 3));
            reinterpret_cast<uint4*>(H + j0)[0] = make_uint4(h[0], h[1], h[2], h[3]);
        } else {
            for (uint64_t j = j0; j < j0 + 16 && j <= N; j++) H[j] = (j == 0 || j == N) ? 1 : (low[j] != low[j - 1]);
        }
    }
}
// Heads written by the last radix pass (k_onesweep HEADS): an entry 2 is the first element of a tile's run
// of a digit, and its predecessor in the sorted order is the last element of another run; the pass wrote
// the keys of both, so the entry becomes their comparison under hmask (1 at position 0).  A thread per 16
// flags, one 16-byte load each; only entries 2 are rewritten.  H[N] is set by the caller.
__global__ __launch_bounds__(256) void k_heads_fix(const uint64_t* __restrict__ keys, uint64_t N, uint64_t hmask,
                                                   uint8_t* __restrict__ H) {
    const uint64_t groups = (N + 15) / 16;
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < groups; g += gridDim.x * 256ull) {
        const uint64_t j0 = g * 16;
        uint32_t h[4] = {0x02020202u, 0x02020202u, 0x02020202u, 0x02020202u};  // a tail group checks each entry
        if (j0 + 16 <= N) {
            const uint4 x = reinterpret_cast<const uint4*>(H + j0)[0];
            if (((x.x | x.y | x.z | x.w) & 0x02020202u) == 0) continue;
            h[0] = x.x;
            h[1] = x.y;
            h[2] = x.z;
            h[3] = x.w;
        }
        for (uint32_t b = 0; b < 16; b++) {
            const uint64_t j = j0 + b;
            if (j >= N) break;
            if (((h[b >> 2] >> (8 * (b & 3))) & 0xFFu) != 2u) continue;
            if (j0 + 16 > N && H[j] != 2) continue;
            H[j] = j == 0 ? 1 : (((keys[j] ^ keys[j - 1]) & hmask) != 0);
        }
    }
}

__global__ __launch_bounds__(256) void k_mark_runs(const uint64_t* __restrict__ offsets, uint64_t N,
                                                   uint8_t* __restrict__ H) {
    const uint64_t o = offsets[threadIdx.x];  // exclusive digit offsets: every run starts at one
    if (o < N) H[o] = 1;
}

// packed round 0: keys are (characters << pbits) | position; heads compare the characters, order gets the
// positions
// positions, a thread per 16 consecutive ones (a grid of N / 16 + 1 threads) as in k_heads
__global__ __launch_bounds__(256) void k_heads_packed(const uint64_t* __restrict__ keys, uint64_t N, uint32_t kbits,
                                                      uint64_t pmask, uint8_t* __restrict__ H,
                                                      uint32_t* __restrict__ order) {
    const uint64_t mask = top_mask(kbits);
    const uint64_t groups = N / 16 + 1;
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < groups; g += gridDim.x * 256ull) {
        const uint64_t j0 = g * 16;
        if (j0 + 16 <= N) {
            uint64_t w[17], kw[16];
            w[0] = j0 ? (keys[j0 - 1] & mask) : ~(keys[0] & mask);
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const uint4 x = reinterpret_cast<const uint4*>(keys + j0)[v];
                kw[2 * v] = (uint64_t)x.y << 32 | x.x;
                kw[2 * v + 1] = (uint64_t)x.w << 32 | x.z;
                w[1 + 2 * v] = kw[2 * v] & mask;
                w[2 + 2 * v] = kw[2 * v + 1] & mask;
            }
            uint32_t h[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 16; k++) h[k >> 2] |= (uint32_t)(w[k + 1] != w[k]) << (8 * (k & 3));
            reinterpret_cast<uint4*>(H + j0)[0] = make_uint4(h[0], h[1], h[2], h[3]);
#pragma unroll
            for (int v = 0; v < 4; v++)
                reinterpret_cast<uint4*>(order + j0)[v] =
                    make_uint4((uint32_t)(kw[4 * v] & pmask), (uint32_t)(kw[4 * v + 1] & pmask),
                               (uint32_t)(kw[4 * v + 2] & pmask), (uint32_t)(kw[4 * v + 3] & pmask));
        } else {
            for (uint64_t j = j0; j < j0 + 16 && j <= N; j++) {
                if (j < N) order[j] = (uint32_t)(keys[j] & pmask);
                H[j] = (j == 0 || j == N) ? 1 : (top_bits(keys[j], kbits) != top_bits(keys[j - 1], kbits));
            }
        }
    }
}

// active = member of a group of size >= 2; value packs (active, active && head) as two 32-bit lanes.
// Scans either all positions (src == nullptr) or a previous active list (a subsequence of it results).
// members of groups of size >= 2 as (element index, group id among those groups).  GATHER: the heads of
// list entry a are H[src[a]] and H[src[a] + 1] (the heads array of all elements); otherwise H is indexed
// by a itself (the full heads array when src == nullptr, or the compact heads of the list src written by
// k_writeback, whose entry a + 1 stands for element src[a] + 1: both are 1 unless they share a group).
template <bool GATHER>
struct ActiveScanOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only where the low half is 1
    const uint8_t* H;
    const uint32_t* src;
    uint32_t* act_idx;
    uint32_t* gid;
    uint32_t* gstart;  // list index of each group's first member
    uint32_t* gfirst;  // element index (order slot) of each group's first member
    __device__ __forceinline__ uint64_t value(uint64_t a) const {
        const uint64_t j = GATHER ? src[a] : a;
        const uint32_t h = H[j], h1 = H[j + 1];  // both loads unconditional: mergeable
        return (h & h1) ? 0ull : (1ull | ((uint64_t)h << 32));
    }
    __device__ __forceinline__ void write(uint64_t a, uint64_t excl, uint64_t v) const {
        if (v) {
            const uint32_t o = (uint32_t)excl;
            const uint32_t j = src ? src[a] : (uint32_t)a;
            act_idx[o] = j;
            const uint32_t g = (uint32_t)(excl >> 32) + (uint32_t)(v >> 32) - 1u;
            gid[o] = g;
            if (v >> 32) {  // a head: the first member of its group
                gstart[g] = o;
                gfirst[g] = j;
            }
        }
    }
};

// sorted active list back into `order`; with `mark`, new group starts into H and, if Ha is given, the
// compact heads of the list (Ha[a] for entry a, Ha[A] = 1) for the next active set
__global__ __launch_bounds__(256) void k_writeback(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                   const uint32_t* __restrict__ act_idx, uint64_t A,
                                                   uint32_t* __restrict__ order, uint8_t* __restrict__ H, int mark,
                                                   uint8_t* __restrict__ Ha) {
    for (uint64_t a = blockIdx.x * 256ull + threadIdx.x; a < A; a += gridDim.x * 256ull) {
        const uint32_t j = act_idx[a];
        order[j] = vals[a];
        if (mark) {
            // entry 0 starts its group; a differing key starts a new one (the old group id is in the key)
            const bool head = a == 0 || keys[a] != keys[a - 1];
            if (head && a > 0) H[j] = 1;
            if (Ha) Ha[a] = head ? 1 : 0;
        }
    }
    if (Ha && blockIdx.x == 0 && threadIdx.x == 0) Ha[A] = 1;
}

// prefix doubling (deeper levels): group id of every position's current key, in text order
struct GroupScatterOp {
    static constexpr bool kBlocked = true;
    const uint8_t* H;
    const uint32_t* order;
    uint32_t* G;
    __device__ __forceinline__ uint64_t value(uint64_t j) const { return H[j]; }
    __device__ __forceinline__ void write(uint64_t j, uint64_t excl, uint64_t v) const {
        G[order[j]] = (uint32_t)(excl + v);
    }
};

// key of a tied position for a doubling step: (group among tied groups, group of the position `off` ahead)
__global__ __launch_bounds__(256) void k_keys_double(const uint32_t* __restrict__ G, uint64_t n, uint64_t A,
                                                     const uint32_t* __restrict__ act_idx,
                                                     const uint32_t* __restrict__ gid,
                                                     const uint32_t* __restrict__ order, uint32_t off, uint32_t gvb,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint64_t a = blockIdx.x * 256ull + threadIdx.x; a < A; a += gridDim.x * 256ull) {
        uint32_t p = order[act_idx[a]];
        uint64_t q = (uint64_t)p + off;
        uint32_t g2 = q < n ? G[q] : 0u;  // past the end: the empty suffix, smallest
        keys[a] = ((uint64_t)gid[a] << gvb) | g2;
        vals[a] = p;
    }
}

// names = 1 + number of distinct keys before (compute_ranks.hpp:88-111): inclusive count of heads
struct NamesOp {
    static constexpr bool kBlocked = true;
    const uint8_t* H;
    uint32_t* names;
    __device__ __forceinline__ uint64_t value(uint64_t j) const { return H[j]; }
    __device__ __forceinline__ void write(uint64_t j, uint64_t excl, uint64_t v) const {
        names[j] = (uint32_t)(excl + v);
    }
};

// ------------------------------------------------------------------------------------------------
// Phase 1+2 from the sort of all suffixes by their first L_s characters: the samples, in that order,
// with the 1-based id of their key group (sample_string.hpp:145-164 and compute_ranks.hpp:88-111 compute
// the same names from a separate sort of the sample keys).  `shift` leaves slot 0 to the dummy sample.
// ------------------------------------------------------------------------------------------------
struct SampleScanOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only samples (low half 1)
    static constexpr bool kAll = true;      // sees every order entry (the 64-bit output below)
    const uint32_t* order;
    const uint8_t* H;
    DcDev dc;
    uint32_t shift;
    uint32_t* s_order;
    uint32_t* s_gid;
    uint64_t* wide;  // optional: every order entry as a 64-bit SA entry (the level-0 output, 8-byte aligned)
    __device__ __forceinline__ bool wants_all() const { return wide != nullptr; }
    // the tile's entries [tbase, tbase + SC_TILE) from p (LDS): two per lane and store, consecutive lanes on
    // consecutive pairs (1 KB per wave and store where the output is 16-byte aligned)
    __device__ __forceinline__ void all(uint64_t tbase, const uint32_t* p, uint64_t N) const {
        if (tbase + SC_TILE <= N && (reinterpret_cast<uintptr_t>(wide) & 15) == 0) {
#pragma unroll
            for (uint32_t q = 0; q < SC_TILE / 2; q += SC_THREADS) {
                const uint2 pr = reinterpret_cast<const uint2*>(p)[q + threadIdx.x];
                reinterpret_cast<uint4*>(wide + tbase)[q + threadIdx.x] = make_uint4(pr.x, 0u, pr.y, 0u);
            }
        } else {
            for (uint32_t e = threadIdx.x; e < SC_TILE && tbase + e < N; e += SC_THREADS) wide[tbase + e] = p[e];
        }
    }
    __device__ __forceinline__ uint64_t value(uint64_t j) const {
        const uint32_t p = order[j];
        const uint32_t q = fast_div(p, dc.X, dc.inv_X);
        return (uint64_t)dc_in(dc, p - q * dc.X) | ((uint64_t)H[j] << 32);
    }
    __device__ __forceinline__ uint32_t payload(uint64_t j) const { return order[j]; }
    __device__ __forceinline__ void write(uint64_t j, uint64_t excl, uint64_t v) const { write(j, excl, v, order[j]); }
    __device__ __forceinline__ void write(uint64_t, uint64_t excl, uint64_t v, uint32_t p) const {
        if (v & 1u) {
            const uint32_t q = fast_div(p, dc.X, dc.inv_X);
            const uint32_t t = (uint32_t)excl + shift;
            s_order[t] = q * dc.D + dc_next_rank(dc, p - q * dc.X);
            s_gid[t] = (uint32_t)(excl >> 32) + (uint32_t)(v >> 32);
        }
    }
};


// ------------------------------------------------------------------------------------------------
// Dense character codes read the padding like the smallest character, so a suffix whose key window
// reaches past the text end (position >= bnd) can share a key group with suffixes it differs from.  In
// such a group the boundary suffixes are the smaller ones and a later boundary suffix is smaller than an
// earlier one (its window is a proper prefix); every boundary suffix has a key of its own.  These kernels
// move them to the front of their group, in that order, as singleton groups.
// ------------------------------------------------------------------------------------------------

constexpr uint32_t BOUNDARY_CAP = 256;

// The groups that hold boundary suffixes, found from the boundary suffixes themselves (at most L - 1 of
// them) instead of a pass over the tied list: the order is sorted by the L-character keys of the dense
// stream, so each boundary suffix's group is the range of its key there (two binary searches).  Ranges of
// two or more slots go to `ranges` as (first, size).
__device__ __forceinline__ int cmp_key_chars(const uint64_t* __restrict__ PT, uint32_t bits, uint32_t L, uint64_t a,
                                             uint64_t b) {
    const uint32_t cpw = 64 / bits;
    for (uint32_t c = 0; c < L; c += cpw) {
        const uint32_t k = L - c < cpw ? L - c : cpw;
        const StreamKeys<false> w{PT, bits, top_mask(k * bits)};
        const uint64_t x = w(a + c), y = w(b + c);
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}
__global__ __launch_bounds__(64) void k_boundary_find(const uint64_t* __restrict__ PT, uint32_t bits, uint32_t L,
                                                      const uint32_t* __restrict__ order, uint64_t n, uint64_t bnd,
                                                      uint32_t* __restrict__ ranges, uint32_t* __restrict__ count,
                                                      uint32_t cap) {
    const uint64_t p = bnd + blockIdx.x * 64ull + threadIdx.x;
    if (p >= n) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {  // first slot whose key is not below p's
        const uint64_t mid = lo + (hi - lo) / 2;
        if (cmp_key_chars(PT, bits, L, order[mid], p) < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    const uint64_t first = lo;
    hi = n;
    while (lo < hi) {  // first slot whose key is above p's
        const uint64_t mid = lo + (hi - lo) / 2;
        if (cmp_key_chars(PT, bits, L, order[mid], p) <= 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo - first >= 2) {  // as (first slot, size): both below 2^32
        const uint32_t i = atomicAdd(count, 1u);
        if (i < cap) {
            ranges[2 * i] = (uint32_t)first;
            ranges[2 * i + 1] = (uint32_t)(lo - first);
        }
    }
}

// one workgroup per affected group order[j0 .. j0+s)
__global__ __launch_bounds__(1024) void k_boundary_fix(uint32_t* __restrict__ order, uint8_t* __restrict__ H,
                                                       uint64_t j0, uint32_t s, uint64_t bnd,
                                                       uint32_t* __restrict__ tmp) {
    __shared__ uint32_t sB[BOUNDARY_CAP];
    __shared__ uint32_t nB;
    __shared__ uint32_t s_tmp[16];
    if (threadIdx.x == 0) nB = 0;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t c = 0; c < s; c += 1024) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t p = i < s ? order[j0 + i] : 0u;
        const bool isB = i < s && (uint64_t)p >= bnd;
        const bool keep = i < s && !isB;
        if (isB) {
            uint32_t x = atomicAdd(&nB, 1u);
            if (x < BOUNDARY_CAP) sB[x] = p;
        }
        uint32_t total = 0;
        const uint32_t e = block_excl_scan<1024>(keep ? 1u : 0u, s_tmp, &total);
        if (keep) tmp[base + e] = p;
        base += total;
    }
    __syncthreads();
    const uint32_t b = nB < BOUNDARY_CAP ? nB : BOUNDARY_CAP;
    if (threadIdx.x == 0) {
        for (uint32_t t = 1; t < b; t++) {  // descending positions
            uint32_t x = sB[t];
            int u = (int)t - 1;
            while (u >= 0 && sB[u] < x) {
                sB[u + 1] = sB[u];
                u--;
            }
            sB[u + 1] = x;
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < b; t += 1024) {
        order[j0 + t] = sB[t];
        H[j0 + t] = 1;
    }
    if (threadIdx.x == 0) H[j0 + b] = 1;
    for (uint32_t i = threadIdx.x; i < s - b; i += 1024) order[j0 + b + i] = tmp[i];
}

// recursive string in (mod X, div X) order (handle_recursive_call, pdcx.hpp:145-180): the name of the j-th
// sample of the name order goes to class_major(sample index), written by a partitioned scatter
// (scatter_u32).  Names stay below 2^31 (checked on the host).
constexpr uint32_t UNIQ_BIT = 0x80000000u;
__global__ __launch_bounds__(256) void k_rec_pairs(const uint32_t* __restrict__ order, const uint32_t* __restrict__ names,
                                                   uint32_t m, DcDev dc, uint32_t* __restrict__ dest,
                                                   uint32_t* __restrict__ word) {
    for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < m; j += gridDim.x * 256ull) {
        dest[j] = class_major(dc, order[j]);
        word[j] = names[j];
    }
}

// scatters of fewer words than this go straight to their destinations
constexpr uint32_t SCATTER_PARTITION_MIN = 1u << 20;

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ out, uint64_t N) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < N; i += gridDim.x * 256ull) out[i] = (uint32_t)i;
}

// out[dest[i]] = val[i] + add for a permutation dest of [0, N) whose pairs are ordered by dest >> wbits:
// pairs [b << wbits, (b + 1) << wbits) are exactly the destinations of window b, which workgroup b
// assembles in LDS (4 << wbits bytes of dynamic LDS) and writes out whole (coalesced stores instead of one
// cache line per word).  Launched with SCATTER_WIDE_THREADS threads for windows of at least
// 2^SCATTER_WIDE_BITS words, else 256.
constexpr uint32_t SCATTER_WINDOW_BITS = 14;  // at most 16384 words (64 KB of LDS) per window
constexpr uint32_t SCATTER_WIDE_BITS = 12;
constexpr uint32_t SCATTER_WIDE_THREADS = 1024;
__global__ __launch_bounds__(SCATTER_WIDE_THREADS) void k_scatter_window(const uint32_t* __restrict__ dest,
                                                                         const uint32_t* __restrict__ val, uint64_t N,
                                                                         uint32_t add, uint32_t wbits,
                                                                         uint32_t* __restrict__ out) {
    extern __shared__ uint32_t s_w[];
    const uint64_t beg = (uint64_t)blockIdx.x << wbits;
    const uint64_t end = beg + (1ull << wbits) < N ? beg + (1ull << wbits) : N;
    const uint32_t len = (uint32_t)(end - beg);
    for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) {
        const uint32_t o = dest[beg + t] - (uint32_t)beg;
        if (o < len) s_w[o] = val[beg + t] + add;  // always true for a permutation
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) out[beg + t] = s_w[t];
}

// out[dest[i]] = val[i] + add
__global__ __launch_bounds__(256) void k_scatter_pairs(const uint32_t* __restrict__ dest, const uint32_t* __restrict__ val,
                                                       uint64_t N, uint32_t add, uint32_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < N; i += gridDim.x * 256ull) out[dest[i]] = val[i] + add;
}

// Discarding from the non-unique samples alone.  Element c of the recursive string is kept unless c
// and c - 1 are both unique (pdcx.hpp:215-246), so the kept set is {0} plus every non-unique position
// and its successor.  NonUniqueOp counts the non-unique samples (sorted order j, unique iff H[j] &&
// H[j+1]); with `cand` it also writes their class-major positions c and c + 1.
struct NonUniqueOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only where the low half is 1
    const uint8_t* H;
    const uint32_t* order;
    DcDev dc;
    uint32_t m;
    uint32_t* cand;
    __device__ __forceinline__ uint64_t value(uint64_t j) const {
        const uint32_t h = H[j], h1 = H[j + 1];
        return (h & h1) ? 0ull : 1ull;
    }
    __device__ __forceinline__ void write(uint64_t j, uint64_t excl, uint64_t v) const {
        if (v && cand) {
            const uint32_t c = class_major(dc, order[j]);
            cand[2 * excl] = c;
            cand[2 * excl + 1] = c + 1 < m ? c + 1 : c;
        }
    }
};

// distinct values of a sorted array, in order
struct DistinctOp {
    static constexpr bool kBlocked = false;  // value(0) must not read a[-1]
    const uint32_t* a;
    uint32_t* out;
    __device__ __forceinline__ uint64_t value(uint64_t i) const { return (i == 0 || a[i] != a[i - 1]) ? 1u : 0u; }
    __device__ __forceinline__ void write(uint64_t i, uint64_t excl, uint64_t v) const {
        if (v) out[excl] = a[i];
    }
};

// the reduced string: the name of the sample at each kept class-major position K[i] (ranks0: 1 + position
// of each sample in the name order)
__global__ __launch_bounds__(256) void k_reduced_string(const uint32_t* __restrict__ K, uint32_t after, DcDev dc,
                                                        const uint32_t* __restrict__ ranks0,
                                                        const uint32_t* __restrict__ names, uint32_t* __restrict__ Rr) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < after; i += gridDim.x * 256ull)
        Rr[i] = names[ranks0[class_major_inv(dc, K[i])] - 1];
}

// index of c in the sorted array K[0..n) (c is present)
__device__ __forceinline__ uint32_t sorted_index(const uint32_t* __restrict__ K, uint32_t n, uint32_t c) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (K[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// (group, rank2) keys of the non-unique samples; rank2 = ISA of the reduced string at the sample's kept
// index (RankRankIndex, pdcx.hpp:297-323)
__global__ __launch_bounds__(256) void k_keys_rank2_kept(uint64_t A, const uint32_t* __restrict__ act_idx,
                                                         const uint32_t* __restrict__ gid,
                                                         const uint32_t* __restrict__ order, DcDev dc,
                                                         const uint32_t* __restrict__ K, uint32_t after,
                                                         const uint32_t* __restrict__ isar, uint32_t rbits,
                                                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint64_t a = blockIdx.x * 256ull + threadIdx.x; a < A; a += gridDim.x * 256ull) {
        const uint32_t k = order[act_idx[a]];
        const uint32_t r2 = isar[sorted_index(K, after, class_major(dc, k))];
        keys[a] = ((uint64_t)gid[a] << rbits) | r2;
        vals[a] = k;
    }
}

// ranks of the refined non-unique samples: 1 + their new position
__global__ __launch_bounds__(256) void k_patch_ranks(uint64_t A, const uint32_t* __restrict__ act_idx,
                                                     const uint32_t* __restrict__ order, uint32_t* __restrict__ ranks) {
    for (uint64_t a = blockIdx.x * 256ull + threadIdx.x; a < A; a += gridDim.x * 256ull) {
        const uint32_t j = act_idx[a];
        ranks[order[j]] = j + 1;  // sample ranks (m < 2^32)
    }
}

// recursive_call_direct (pdcx.hpp:182-212): the rank of map_back(SA_R[j]) is j + 1; dest[j] = map_back(SA_R[j])
__global__ __launch_bounds__(256) void k_sa_dest(const uint32_t* __restrict__ sar, uint32_t m, DcDev dc,
                                                 uint32_t* __restrict__ dest) {
    for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < m; j += gridDim.x * 256ull) dest[j] = class_major_inv(dc, sar[j]);
}

// ------------------------------------------------------------------------------------------------
// DC comparator (DCMergeSamples::operator<, merge_samples.hpp:87-109) for two suffixes whose first
// compared characters are equal: ranks[rank_pos(a) + r1] < ranks[rank_pos(b) + r2].
// ------------------------------------------------------------------------------------------------
// With dense merge words the padding codes like the smallest character, so two suffixes whose windows
// both reach past the end can tie although they differ; the later (shorter) one is then the smaller:
// its window is a proper prefix of the other's.
__device__ __forceinline__ bool dc_rank_less(uint32_t a, uint32_t b, const DcDev& dc, const uint32_t* __restrict__ cmp,
                                             const uint32_t* __restrict__ ranks) {
    if (a >= dc.bnd && b >= dc.bnd) return a > b;
    uint32_t qa = fast_div(a, dc.X, dc.inv_X), qb = fast_div(b, dc.X, dc.inv_X);
    uint32_t ma = a - qa * dc.X, mb = b - qb * dc.X;
    uint32_t e = cmp[ma * dc.X + mb];
    uint32_t r1 = (e >> 8) & 0xFF, r2 = (e >> 16) & 0xFF;
    uint32_t ra = ranks[qa * dc.D + dc_next_rank(dc, ma) + r1];
    uint32_t rb = ranks[qb * dc.D + dc_next_rank(dc, mb) + r2];
    return ra < rb;
}

// ------------------------------------------------------------------------------------------------
// Refinement rounds by group size: a tied group of at most RF_SMALL suffixes is ordered by one thread
// (insertion sort of the members' next characters), a group of at most RF_WAVE by one wave (a bitonic
// sort in the wave's LDS); the members of larger groups go through the radix sort.  All write the new
// group heads into H and the compact heads Ha of the active list.
// ------------------------------------------------------------------------------------------------
constexpr uint32_t RF_SMALL = 8;
constexpr uint32_t RF_WAVE = 512;

__global__ __launch_bounds__(256) void k_refine_small(const uint32_t* __restrict__ gstart, uint32_t G,
                                                      const uint32_t* __restrict__ gfirst, uint32_t* __restrict__ order,
                                                      uint8_t* __restrict__ H, uint8_t* __restrict__ Ha,
                                                      RoundChars chars, uint8_t* __restrict__ large) {
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < G; g += gridDim.x * 256ull) {
        const uint32_t a0 = gstart[g], s = gstart[g + 1] - a0;
        if (s > RF_SMALL) {
            large[g] = s > RF_WAVE ? 1 : 0;  // the wave path (k_refine_wave) takes the others
            continue;
        }
        large[g] = 0;
        const uint32_t j0 = gfirst[g];  // a group occupies consecutive slots of the order
        if (s == 2) {  // the common case (a copied segment and its source): one comparison
            const uint32_t e0 = order[j0], e1 = order[j0 + 1];
            const uint64_t k0 = chars((uint64_t)e0), k1 = chars((uint64_t)e1);
            if (k1 < k0) {
                order[j0] = e1;
                order[j0 + 1] = e0;
            }
            Ha[a0] = 1;
            Ha[a0 + 1] = k0 != k1 ? 1 : 0;
            if (k0 != k1) H[j0 + 1] = 1;
            continue;
        }
        uint64_t k[RF_SMALL];
        uint32_t e[RF_SMALL];
        for (uint32_t t = 0; t < s; t++) {
            e[t] = order[j0 + t];
            k[t] = chars((uint64_t)e[t]);
        }
        for (uint32_t t = 1; t < s; t++) {  // stable: equal keys keep their order
            const uint64_t kx = k[t];
            const uint32_t ex = e[t];
            int u = (int)t - 1;
            while (u >= 0 && kx < k[u]) {
                k[u + 1] = k[u];
                e[u + 1] = e[u];
                u--;
            }
            k[u + 1] = kx;
            e[u + 1] = ex;
        }
        for (uint32_t t = 0; t < s; t++) {
            order[j0 + t] = e[t];
            const bool head = t == 0 || k[t] != k[t - 1];
            Ha[a0 + t] = head ? 1 : 0;
            if (head && t > 0) H[j0 + t] = 1;
        }
    }
}

// Groups of RF_SMALL + 1 .. RF_WAVE members, by waves: every wave takes 64 consecutive groups at a time
// and sorts those of this size in batches of at most RF_WAVE elements in its LDS.  A batch is the longest
// run (in group order) of such groups whose sizes, padded to powers of two P >= 16, sum to at most
// RF_WAVE; its groups are laid out by decreasing P, so each starts at a multiple of its P.  All members of
// the batch are loaded together (their order slots, then their next characters: a random stream window
// each, as in k_refine_small), so a batch costs two memory round trips instead of two per group.  One
// bitonic network over the batch sorts every group: an element takes part in the stages of block sizes up
// to its group's P (the pairs of those stages stay inside the group), and a group's last merge is
// ascending.  Elements compare by (key, member index), so the order is that of a stable sort; padding
// elements (the largest key, indices from the group size on) sort last.  The LDS accesses of one wave run
// in program order, so the steps only need the compiler barrier between them.
constexpr uint32_t RFW_WAVES = 4;   // waves per workgroup
constexpr uint32_t RFW_MAXG = 32;   // groups in a batch (each pads to at least 16 elements)
__device__ __forceinline__ uint32_t rfw_class_p(uint32_t i, const uint32_t (&cend)[6]) {
    // elements [0, cend[5]) belong to groups of P = 512, then [cend[5], cend[4]) to P = 256, ...
#pragma unroll
    for (int c = 5; c >= 0; c--)
        if (i < cend[c]) return 16u << c;
    return 0u;
}
__global__ __launch_bounds__(64 * RFW_WAVES) void k_refine_wave(const uint32_t* __restrict__ gstart, uint32_t G,
                                                              const uint32_t* __restrict__ gfirst,
                                                              uint32_t* __restrict__ order, uint8_t* __restrict__ H,
                                                              uint8_t* __restrict__ Ha, RoundChars chars) {
    __shared__ uint64_t s_key[RFW_WAVES][RF_WAVE];
    __shared__ uint32_t s_pos[RFW_WAVES][RF_WAVE];
    __shared__ uint16_t s_idx[RFW_WAVES][RF_WAVE];
    __shared__ uint32_t s_boff[RFW_WAVES][RFW_MAXG];  // the batch's groups by offset: offset, size, slot, list index
    __shared__ uint32_t s_bs[RFW_WAVES][RFW_MAXG];
    __shared__ uint32_t s_bj0[RFW_WAVES][RFW_MAXG];
    __shared__ uint32_t s_ba0[RFW_WAVES][RFW_MAXG];
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64;
    uint64_t* key = s_key[w];
    uint32_t* pos = s_pos[w];
    uint16_t* idx = s_idx[w];
    uint32_t* boff = s_boff[w];
    uint32_t* bsz = s_bs[w];
    uint32_t* bj0 = s_bj0[w];
    uint32_t* ba0 = s_ba0[w];
    constexpr int R = RF_WAVE / 64;  // elements per lane
    const uint64_t below = lanemask_lt();
    const uint64_t chunks = ((uint64_t)G + 63) / 64;
    for (uint64_t ch = (uint64_t)blockIdx.x * RFW_WAVES + w; ch < chunks; ch += (uint64_t)gridDim.x * RFW_WAVES) {
        const uint64_t gl = ch * 64 + lane;
        uint32_t sz = 0, a0 = 0, j0 = 0;
        if (gl < G) {
            a0 = gstart[gl];
            sz = gstart[gl + 1] - a0;
        }
        const bool med = sz > RF_SMALL && sz <= RF_WAVE;
        uint32_t lp = 4;  // P = 2^lp >= sz, for the groups of this path only (sizes up to RF_WAVE)
        if (med) {
            j0 = gfirst[gl];
            while ((1u << lp) < sz) lp++;
        }
        uint64_t pending = __ballot(med);
        while (pending) {
            // the batch: pending groups in lane order while their padded sizes sum to at most RF_WAVE
            const bool pend = (pending >> lane) & 1;
            const uint32_t pre = (uint32_t)wave_incl_scan64(pend ? (1ull << lp) : 0ull);
            const bool inb = pend && pre <= RF_WAVE;
            const uint64_t bmask = __ballot(inb);
            // layout by decreasing P; cend[c] = end of the elements of groups with P >= 16 << c
            uint32_t cend[6], base = 0, nb = 0, myoff = 0, myrank = 0;
#pragma unroll
            for (int c = 5; c >= 0; c--) {
                const uint64_t cm = __ballot(inb && lp == (uint32_t)(c + 4));
                if (inb && lp == (uint32_t)(c + 4)) {
                    const uint32_t k = (uint32_t)__popcll(cm & below);
                    myoff = base + (k << (c + 4));
                    myrank = nb + k;
                }
                base += (uint32_t)__popcll(cm) << (c + 4);
                nb += (uint32_t)__popcll(cm);
                cend[c] = base;
            }
            const uint32_t T = base;
            if (inb) {
                boff[myrank] = myoff;
                bsz[myrank] = sz;
                bj0[myrank] = j0;
                ba0[myrank] = a0;
            }
            __builtin_amdgcn_wave_barrier();
            // flattened load: element e = lane + 64 r of the batch -> its group (by offset) and member t
            uint32_t ek[R], et[R], ep[R];
            uint64_t kv[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t e = lane + 64u * r;
                ek[r] = 0;
                et[r] = 0xFFFFu;
                ep[r] = 0;
                if (e < T) {
                    uint32_t lo = 0, hi = nb;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (boff[mid] <= e)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    ek[r] = lo;
                    et[r] = e - boff[lo];
                    if (et[r] < bsz[lo]) ep[r] = order[bj0[lo] + et[r]];
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t e = lane + 64u * r;
                kv[r] = ~0ull;
                if (e < T && et[r] < bsz[ek[r]]) kv[r] = chars((uint64_t)ep[r]);
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t e = lane + 64u * r;
                if (e < T) {
                    key[e] = kv[r];
                    idx[e] = (uint16_t)et[r];
                    pos[e] = ep[r];
                }
            }
            __builtin_amdgcn_wave_barrier();
            // one network for every group of the batch
            const uint32_t pmax = rfw_class_p(0, cend);
            for (uint32_t lk = 1; (1u << lk) <= pmax; lk++) {
                const uint32_t kk = 1u << lk;
                const uint32_t S = cend[lk > 4 ? lk - 4 : 0];  // the elements of groups with P >= kk
                for (int lj = (int)lk - 1; lj >= 0; lj--) {
                    const uint32_t j = 1u << lj;
                    for (uint32_t t = lane; t < S / 2; t += 64) {
                        const uint32_t i = ((t >> lj) << (lj + 1)) | (t & (j - 1)), l = i + j;
                        const uint32_t Pi = rfw_class_p(i, cend);
                        if (kk > Pi) continue;
                        const uint64_t ki = key[i], kl = key[l];
                        const uint32_t ii = idx[i], il = idx[l];
                        const bool gt = ki > kl || (ki == kl && ii > il);
                        const bool asc = kk == Pi || (i & kk) == 0;
                        if (gt == asc) {
                            key[i] = kl;
                            key[l] = ki;
                            idx[i] = (uint16_t)il;
                            idx[l] = (uint16_t)ii;
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // write back: new order, group heads of the order and of the list
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t e = lane + 64u * r;
                if (e < T) {
                    const uint32_t k = ek[r], t = et[r];
                    if (t < bsz[k]) {
                        const uint64_t kk2 = key[e];
                        const uint32_t jj = bj0[k];
                        order[jj + t] = pos[boff[k] + idx[e]];
                        const bool head = t == 0 || kk2 != key[e - 1];
                        Ha[ba0[k] + t] = head ? 1 : 0;
                        if (head && t > 0) H[jj + t] = 1;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            pending &= ~bmask;
        }
    }
}

// members of the large groups of the active list: their list indices, in list order, and the rank of
// their group among the large groups (the high counter counts the first members of large groups; a
// group's members are consecutive in the list)
struct LargeMembersOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only where the low half is 1
    const uint8_t* large;
    const uint32_t* gid;
    uint32_t* slots;
    uint32_t* lrank;
    uint32_t* lfirst;  // per large group: its first member in the list
    __device__ __forceinline__ uint64_t value(uint64_t a) const {
        const uint32_t g = gid[a];
        const bool first = a == 0 || gid[a - 1] != g;
        return large[g] ? (1ull | ((uint64_t)first << 32)) : 0ull;
    }
    __device__ __forceinline__ void write(uint64_t a, uint64_t excl, uint64_t v) const {
        if (v) {
            const uint32_t r = (uint32_t)(excl >> 32) + (uint32_t)(v >> 32) - 1u;
            slots[(uint32_t)excl] = (uint32_t)a;
            lrank[(uint32_t)excl] = r;
            if (v >> 32) lfirst[r] = (uint32_t)excl;
        }
    }
};

// radix keys of the large-group members: (large-group rank << (64 - lbits)) | next cbits bits of characters
// (lbits + cbits <= 64); values are the suffixes.  cls (optional): the member's key against its group's
// first member's, 0 smaller, 1 equal, 2 larger (that key recomputed: one stream window per group, cached)
__global__ __launch_bounds__(256) void k_keys_large(RoundChars chars, uint64_t B, const uint32_t* __restrict__ slots,
                                                    const uint32_t* __restrict__ act_idx,
                                                    const uint32_t* __restrict__ lrank,
                                                    const uint32_t* __restrict__ order, uint32_t lbits,
                                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                    const uint32_t* __restrict__ lfirst, uint8_t* __restrict__ cls,
                                                    const uint32_t* __restrict__ lk, uint32_t deep) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull) {
        const uint32_t p = order[act_idx[slots[b]]];
        const uint32_t k = lk ? lk[lrank[b]] : 0u;  // the group's k after a coded round 0
        const uint64_t ch = chars.with_k((uint64_t)p, k);
        uint64_t key = lbits ? (((uint64_t)lrank[b] << (64 - lbits)) | (ch >> lbits)) : ch;
        vals[b] = p;
        if (cls) {
            const uint32_t fb = lfirst[lrank[b]];
            const uint32_t pf = order[act_idx[slots[fb]]];
            const uint64_t cf = chars.with_k((uint64_t)pf, k);
            const uint64_t x = ch >> lbits, pv = cf >> lbits;  // the same group rank above both
            uint8_t c = x < pv ? 0 : (x == pv ? 1 : 2);
            if (deep) {
                // against the first member up to character L: 1 only for the members equal to it there; the
                // class bit below the characters orders the others of this round's key by that comparison
                if (c == 1) {
                    const int r = chars.cmp_after((uint64_t)p, (uint64_t)pf, k);
                    c = r < 0 ? 0 : (r == 0 ? 1 : 2);
                }
                if (c == 2) key |= 1ull << (63 - lbits - chars.cbits);
            }
            cls[b] = c;
        }
        keys[b] = key;
    }
}

// sorted large-group members back into `order` (a group's members keep their slots); new group starts
// into H and the compact heads Ha
__global__ __launch_bounds__(256) void k_writeback_large(const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals,
                                                         const uint32_t* __restrict__ slots, uint64_t B,
                                                         const uint32_t* __restrict__ act_idx,
                                                         uint32_t* __restrict__ order, uint8_t* __restrict__ H,
                                                         uint8_t* __restrict__ Ha) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull) {
        const uint32_t a = slots[b];
        const uint32_t j = act_idx[a];
        order[j] = vals[b];
        const bool head = b == 0 || keys[b] != keys[b - 1];  // the group rank is in the key
        Ha[a] = head ? 1 : 0;
        if (head) H[j] = 1;
    }
}

// Large groups whose members mostly share one key (a periodic run: all members but the few near the
// run's end have the same next characters) are ordered by a stable three-way partition around the key of
// their first member instead of a sort of all members: the members with another key (the minority) are
// compacted and sorted, the members equal to it keep their list order between the smaller and the larger
// ones.  The result is that of the stable sort.  Members b of the large list are consecutive per group;
// fb, fe are the group's first and end member (lfirst by large-group rank).
struct LargeGroupSpan {
    uint32_t fb, fe;
};
__device__ __forceinline__ LargeGroupSpan large_span(uint32_t b, const uint32_t* __restrict__ lrank,
                                                     const uint32_t* __restrict__ lfirst) {
    const uint32_t r = lrank[b];
    return {lfirst[r], lfirst[r + 1]};  // lfirst holds B after the last group
}

// per member (its class from k_keys_large): minority (another key than the group's first) in the low
// half, smaller in the high half; writes both exclusive prefixes of every member and compacts the
// minority (key, member)
struct PivotScanOp {
    static constexpr bool kBlocked = true;
    const uint8_t* cls;
    const uint64_t* keys;
    uint32_t* minpref;
    uint32_t* c0pref;
    uint64_t* mkeys;
    uint32_t* midx;
    __device__ __forceinline__ uint64_t value(uint64_t b) const {
        const uint32_t c = cls[b];
        return (uint64_t)(c != 1) | ((uint64_t)(c == 0) << 32);
    }
    __device__ __forceinline__ void write(uint64_t b, uint64_t excl, uint64_t v) const {
        minpref[b] = (uint32_t)excl;
        c0pref[b] = (uint32_t)(excl >> 32);
        if (v & 1u) {
            mkeys[(uint32_t)excl] = keys[b];
            midx[(uint32_t)excl] = (uint32_t)b;
        }
    }
};

// member value `pos` to member slot `dest` of the large list, with its head flag
__device__ __forceinline__ void large_put(uint32_t dest, uint32_t pos, bool head, const uint32_t* __restrict__ slots,
                                          const uint32_t* __restrict__ act_idx, uint32_t* __restrict__ order,
                                          uint8_t* __restrict__ H, uint8_t* __restrict__ Ha) {
    const uint32_t ad = slots[dest];
    const uint32_t j = act_idx[ad];
    order[j] = pos;
    Ha[ad] = head ? 1 : 0;
    if (head) H[j] = 1;
}

// the members equal to their group's first key: after the group's smaller minority, in list order
__global__ __launch_bounds__(256) void k_place_equal(uint64_t B, const uint32_t* __restrict__ vals,
                                                     const uint32_t* __restrict__ slots,
                                                     const uint32_t* __restrict__ lrank,
                                                     const uint32_t* __restrict__ lfirst,
                                                     const uint32_t* __restrict__ minpref,
                                                     const uint32_t* __restrict__ c0pref,
                                                     const uint32_t* __restrict__ act_idx, uint32_t* __restrict__ order,
                                                     uint8_t* __restrict__ H, uint8_t* __restrict__ Ha,
                                                     uint32_t park) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull) {
        if (minpref[b + 1] != minpref[b]) continue;  // minority
        const LargeGroupSpan sp = large_span(b, lrank, lfirst);
        const uint32_t n0 = c0pref[sp.fe] - c0pref[sp.fb];
        // rank among the equal members (list indices below 2^32; differences of the prefix arrays modulo 2^32)
        const uint32_t r1 = ((uint32_t)b - sp.fb) - (minpref[b] - minpref[sp.fb]);
        large_put(sp.fb + n0 + r1, vals[b], r1 == 0, slots, act_idx, order, H, Ha);
        // park: members equal through character L leave the rounds (singletons of the next active list),
        // their group stays one group in H
        if (park) Ha[slots[sp.fb + n0 + r1]] = 1;
    }
}

// the sorted minority: the smaller members first in their group, the larger ones after the equal ones
__global__ __launch_bounds__(256) void k_place_minority(uint64_t Bm, const uint64_t* __restrict__ mkeys,
                                                        const uint32_t* __restrict__ midx,
                                                        const uint32_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ slots,
                                                        const uint32_t* __restrict__ lrank,
                                                        const uint32_t* __restrict__ lfirst,
                                                        const uint32_t* __restrict__ minpref,
                                                        const uint32_t* __restrict__ c0pref,
                                                        const uint32_t* __restrict__ act_idx, uint32_t* __restrict__ order,
                                                        uint8_t* __restrict__ H, uint8_t* __restrict__ Ha) {
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < Bm; k += gridDim.x * 256ull) {
        const uint32_t b = midx[k];
        const LargeGroupSpan sp = large_span(b, lrank, lfirst);
        const uint32_t n0 = c0pref[sp.fe] - c0pref[sp.fb];
        const uint32_t nmin = minpref[sp.fe] - minpref[sp.fb];
        const uint32_t local = (uint32_t)k - minpref[sp.fb];  // the group's minority starts there in the sorted list
        const uint32_t dest = local < n0 ? sp.fb + local : sp.fb + (sp.fe - sp.fb - nmin) + local;
        const bool head = local == 0 || local == n0 || mkeys[k] != mkeys[k - 1];
        large_put(dest, vals[b], head, slots, act_idx, order, H, Ha);
    }
}

// groups whose members all share the first member's key: only the compact heads change
// (park: the members are equal through character L and all leave the rounds)
__global__ __launch_bounds__(256) void k_uniform_heads(uint64_t B, const uint32_t* __restrict__ slots,
                                                       const uint32_t* __restrict__ lrank,
                                                       const uint32_t* __restrict__ lfirst, uint8_t* __restrict__ Ha,
                                                       uint32_t park) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull)
        Ha[slots[b]] = park || (uint32_t)b == lfirst[lrank[b]] ? 1 : 0;
}

// The same when rank and characters do not fit one word: an LSD sort by the characters (values: member
// indices b, suffixes kept in P), then a stable sort by the rank.
__global__ __launch_bounds__(256) void k_keys_large_chars(RoundChars chars, uint64_t B,
                                                          const uint32_t* __restrict__ slots,
                                                          const uint32_t* __restrict__ act_idx,
                                                          const uint32_t* __restrict__ order,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ P) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull) {
        const uint32_t p = order[act_idx[slots[b]]];
        keys[b] = chars((uint64_t)p);
        P[b] = p;
    }
}
__global__ __launch_bounds__(256) void k_keys_large_rank(const uint32_t* __restrict__ members,
                                                         const uint32_t* __restrict__ lrank, uint64_t B,
                                                         uint64_t* __restrict__ keys) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < B; i += gridDim.x * 256ull) keys[i] = lrank[members[i]];
}
__global__ __launch_bounds__(256) void k_writeback_large2(const uint64_t* __restrict__ rkeys,
                                                          const uint32_t* __restrict__ members,
                                                          const uint32_t* __restrict__ P, RoundChars chars,
                                                          const uint32_t* __restrict__ slots, uint64_t B,
                                                          const uint32_t* __restrict__ act_idx,
                                                          uint32_t* __restrict__ order, uint8_t* __restrict__ H,
                                                          uint8_t* __restrict__ Ha) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < B; i += gridDim.x * 256ull) {
        const uint32_t a = slots[i];
        const uint32_t j = act_idx[a];
        const uint32_t p = P[members[i]];
        order[j] = p;
        const bool head = i == 0 || rkeys[i] != rkeys[i - 1] ||
                          chars((uint64_t)p) != chars((uint64_t)P[members[i - 1]]);
        Ha[a] = head ? 1 : 0;
        if (head) H[j] = 1;
    }
}

constexpr uint32_t TB_SMALL = 16;

// After an active-set scan whose totals are sc[0] = members A and sc[1] = groups G: gstart[G] = A (the end of
// the last group; 0 when A = 2^32, sizes are differences modulo 2^32), and the group-size counts cleared
__global__ void k_active_finish(uint64_t* sc, uint32_t* __restrict__ gstart) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        gstart[sc[1]] = (uint32_t)sc[0];
        sc[8] = 0;
        sc[9] = 0;
        sc[10] = 0;
        sc[11] = 0;
    }
}

// Counts of the active set's groups by size, for the host to size the next steps without another
// readback: sc[8] = groups with more than RF_WAVE members (the radix path of a group round), sc[9] = their
// members, sc[10] = members of groups with more than TB_SMALL members (the merge-sort path of the
// tie-break), sc[11] = groups of RF_SMALL + 1 .. RF_WAVE members (the wave path of a group round)
__global__ __launch_bounds__(256) void k_group_counts(const uint32_t* __restrict__ gstart, uint64_t* sc) {
    const uint64_t G = sc[1];
    unsigned long long nl = 0, b8 = 0, b16 = 0, nw = 0;
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < G; g += gridDim.x * 256ull) {
        const uint32_t sz = gstart[g + 1] - gstart[g];
        nl += sz > RF_WAVE;
        b8 += sz > RF_WAVE ? sz : 0;
        b16 += sz > TB_SMALL ? sz : 0;
        nw += sz > RF_SMALL && sz <= RF_WAVE;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        nl += __shfl_xor(nl, d, 64);
        b8 += __shfl_xor(b8, d, 64);
        b16 += __shfl_xor(b16, d, 64);
        nw += __shfl_xor(nw, d, 64);
    }
    if (__lane_id() == 0 && (nl | b16 | nw)) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&sc[8]), nl);
        atomicAdd(reinterpret_cast<unsigned long long*>(&sc[9]), b8);
        atomicAdd(reinterpret_cast<unsigned long long*>(&sc[10]), b16);
        atomicAdd(reinterpret_cast<unsigned long long*>(&sc[11]), nw);
    }
}

// groups of size <= TB_SMALL: one thread sorts its group by insertion sort with the DC comparator;
// larger groups are flagged for the merge-sort path
__global__ __launch_bounds__(256) void k_tiebreak_small(const uint32_t* __restrict__ gstart, uint32_t G,
                                                        const uint32_t* __restrict__ gfirst, uint32_t* __restrict__ order,
                                                        DcDev dc, const uint32_t* __restrict__ cmp,
                                                        const uint32_t* __restrict__ ranks, uint8_t* __restrict__ large) {
    for (uint64_t g = blockIdx.x * 256ull + threadIdx.x; g < G; g += gridDim.x * 256ull) {
        uint32_t a0 = gstart[g], s = gstart[g + 1] - a0;
        if (s > TB_SMALL) {
            large[g] = 1;
            continue;
        }
        large[g] = 0;
        uint32_t j0 = gfirst[g];
        uint32_t e[TB_SMALL];
        for (uint32_t t = 0; t < s; t++) e[t] = order[j0 + t];
        for (uint32_t t = 1; t < s; t++) {
            uint32_t x = e[t];
            int u = (int)t - 1;
            while (u >= 0 && dc_rank_less(x, e[u], dc, cmp, ranks)) {
                e[u + 1] = e[u];
                u--;
            }
            e[u + 1] = x;
        }
        for (uint32_t t = 0; t < s; t++) order[j0 + t] = e[t];
    }
}

// elements of the large groups: (group << 32 | suffix) plus the order slot they came from
struct LargeScanOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only where the low half is 1
    const uint8_t* large;
    const uint32_t* gid;
    const uint32_t* act_idx;
    const uint32_t* order;
    uint64_t* items;
    uint32_t* slots;
    __device__ __forceinline__ uint64_t value(uint64_t a) const { return large[gid[a]]; }
    __device__ __forceinline__ void write(uint64_t a, uint64_t excl, uint64_t v) const {
        if (v) {
            uint32_t j = act_idx[a];
            items[excl] = ((uint64_t)gid[a] << 32) | order[j];
            slots[excl] = j;
        }
    }
};

__device__ __forceinline__ bool item_less(uint64_t x, uint64_t y, const DcDev& dc, const uint32_t* __restrict__ cmp,
                                          const uint32_t* __restrict__ ranks) {
    uint32_t gx = (uint32_t)(x >> 32), gy = (uint32_t)(y >> 32);
    if (gx != gy) return gx < gy;
    if (gx == 0xFFFFFFFFu) return false;  // padding
    return dc_rank_less((uint32_t)x, (uint32_t)y, dc, cmp, ranks);
}

constexpr uint32_t TB_TILE = 1024;

// sorts each tile of TB_TILE items in LDS (bitonic network, DC comparator)
__global__ __launch_bounds__(256) void k_tb_tile_sort(uint64_t* __restrict__ items, uint64_t B, DcDev dc,
                                                      const uint32_t* __restrict__ cmp, const uint32_t* __restrict__ ranks) {
    __shared__ uint64_t s[TB_TILE];
    const uint64_t base = (uint64_t)blockIdx.x * TB_TILE;
    for (uint32_t i = threadIdx.x; i < TB_TILE; i += 256) s[i] = base + i < B ? items[base + i] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= TB_TILE; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < TB_TILE / 2; t += 256) {
                uint32_t i = (t / j) * 2 * j + (t % j);
                uint64_t x = s[i], y = s[i + j];
                bool asc = (i & k) == 0;
                bool gt = item_less(y, x, dc, cmp, ranks);
                if (gt == asc) {
                    s[i] = y;
                    s[i + j] = x;
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < TB_TILE; i += 256)
        if (base + i < B) items[base + i] = s[i];
}

// one merge pass: runs of width w are merged pairwise; every item finds its rank in the partner run
__global__ __launch_bounds__(256) void k_tb_merge(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t B,
                                                  uint64_t w, DcDev dc, const uint32_t* __restrict__ cmp,
                                                  const uint32_t* __restrict__ ranks) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < B; i += (uint64_t)gridDim.x * 256) {
        uint64_t s = (i / (2ull * w)) * 2ull * w;
        uint64_t x = in[i];
        bool left = (i - s) < w;
        uint64_t lo, hi;
        if (left) {
            lo = s + w < B ? s + w : B;
            hi = s + 2ull * w < B ? s + 2ull * w : B;
        } else {
            lo = s;
            hi = s + w;
        }
        // count partner items < x (items are distinct)
        uint64_t a = lo, b = hi;
        while (a < b) {
            uint64_t mid = (a + b) >> 1;
            if (item_less(in[mid], x, dc, cmp, ranks))
                a = mid + 1;
            else
                b = mid;
        }
        uint64_t before = a - lo;
        uint64_t own = left ? (i - s) : (i - s - w);
        out[s + own + before] = x;
    }
}

__global__ __launch_bounds__(256) void k_tb_scatter(const uint64_t* __restrict__ items, const uint32_t* __restrict__ slots,
                                                    uint64_t B, uint32_t* __restrict__ order) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull) order[slots[b]] = (uint32_t)items[b];
}

// ------------------------------------------------------------------------------------------------
// Large groups already in order.  The DC comparator is a strict total order, so a group whose consecutive
// members (in their current order) all compare ascending is sorted, and one whose pairs all compare
// descending is sorted once reversed.  The groups of a periodic run are of this kind: two of its suffixes
// first differ where the run ends, and all members meet the same pair of characters there.  Only the
// groups with pairs of both directions go through the merge sort.
// ------------------------------------------------------------------------------------------------

// index of each large group among the large groups (in group order)
struct LargeIndexOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only the large groups
    const uint8_t* large;
    uint32_t* lidx;
    __device__ __forceinline__ uint64_t value(uint64_t g) const { return large[g]; }
    __device__ __forceinline__ void write(uint64_t g, uint64_t excl, uint64_t v) const {
        if (v) lidx[g] = (uint32_t)excl;
    }
};

// mono[large group] |= 1 for a pair of consecutive members in comparator order, 2 for a reversed pair
// (items: group << 32 | suffix, in group order).  A wave takes TB_MONO_CHUNK consecutive members, each lane
// keeps the flags of its current group and ORs them into mono only when its group changes and at the end
// (one OR per group and wave): the groups are few and large, so per-pair atomics would queue on a few words.
constexpr uint32_t TB_MONO_CHUNK = 4096;  // a multiple of 64
__global__ __launch_bounds__(256) void k_tb_monotone(const uint64_t* __restrict__ items, uint64_t B,
                                                     const uint32_t* __restrict__ lidx, DcDev dc,
                                                     const uint32_t* __restrict__ cmp, const uint32_t* __restrict__ ranks,
                                                     uint32_t* __restrict__ mono) {
    const uint32_t lane = __lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * 4;
    for (uint64_t c = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64; c * TB_MONO_CHUNK < B; c += waves) {
        const uint64_t b0 = c * TB_MONO_CHUNK;
        const uint64_t b1 = b0 + TB_MONO_CHUNK < B ? b0 + TB_MONO_CHUNK : B;
        uint32_t gcur = 0, fcur = 0;
        for (uint64_t b = b0 + lane; b < b1; b += 64) {
            if (b + 1 >= B) continue;
            const uint64_t x = items[b], y = items[b + 1];
            if ((x >> 32) != (y >> 32)) continue;
            const uint32_t g = lidx[x >> 32];
            const uint32_t f = dc_rank_less((uint32_t)x, (uint32_t)y, dc, cmp, ranks) ? 1u : 2u;
            if (fcur && g != gcur) {
                atomicOr(&mono[gcur], fcur);
                fcur = 0;
            }
            gcur = g;
            fcur |= f;
        }
        // the lanes still holding flags, one group at a time: one OR per group
        for (uint64_t pend = __ballot(fcur != 0); pend; pend = __ballot(fcur != 0)) {
            const int l0 = __builtin_ctzll(pend);
            const uint32_t g0 = (uint32_t)__shfl((int)gcur, l0, 64);
            const bool mine = fcur != 0 && gcur == g0;
            const uint32_t fw = (__ballot(mine && (fcur & 1u)) ? 1u : 0u) | (__ballot(mine && (fcur & 2u)) ? 2u : 0u);
            if ((int)lane == l0) atomicOr(&mono[g0], fw);
            if (mine) fcur = 0;
        }
    }
}

// members of the groups found in reverse order go to the mirrored slot of their group
__global__ __launch_bounds__(256) void k_tb_reverse(const uint64_t* __restrict__ items, const uint32_t* __restrict__ slots,
                                                    uint64_t B, const uint32_t* __restrict__ lidx,
                                                    const uint32_t* __restrict__ mono, const uint32_t* __restrict__ gstart,
                                                    const uint32_t* __restrict__ gfirst, uint32_t* __restrict__ order) {
    for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < B; b += gridDim.x * 256ull) {
        const uint64_t x = items[b];
        const uint32_t g = (uint32_t)(x >> 32);
        if (mono[lidx[g]] != 2u) continue;
        const uint32_t last = gfirst[g] + (gstart[g + 1] - gstart[g]) - 1;  // the group's last slot
        order[gfirst[g] + (last - slots[b])] = (uint32_t)x;
    }
}

// the members of the groups with pairs of both directions, compacted (group order kept) for the merge sort
struct MixedScanOp {
    static constexpr bool kBlocked = true;
    static constexpr bool kCompact = true;  // writes only members of mixed groups
    const uint64_t* items;
    const uint32_t* slots;
    const uint32_t* lidx;
    const uint32_t* mono;
    uint64_t* items_m;
    uint32_t* slots_m;
    __device__ __forceinline__ uint64_t value(uint64_t b) const { return mono[lidx[items[b] >> 32]] == 3u ? 1ull : 0ull; }
    __device__ __forceinline__ void write(uint64_t b, uint64_t excl, uint64_t v) const {
        if (v) {
            items_m[excl] = items[b];
            slots_m[excl] = slots[b];
        }
    }
};

// ------------------------------------------------------------------------------------------------
// Base case (slow_suffixarray, string_util.hpp:87-106; sequential_sa.hpp:17-28): one workgroup,
// prefix doubling with a bitonic sort of (rank, rank+h, index) keys held in LDS.  n <= 16384.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void lds_bitonic(uint64_t* s, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < P / 2; t += blockDim.x) {
                uint32_t i = (t / j) * 2 * j + (t % j);
                uint64_t x = s[i], y = s[i + j];
                bool asc = (i & k) == 0;
                if ((x > y) == asc) {
                    s[i] = y;
                    s[i + j] = x;
                }
            }
            __syncthreads();
        }
}

// assigns dense ranks (1-based) to the sorted keys; groups compare on key >> 14
__device__ __forceinline__ uint32_t lds_assign_ranks(const uint64_t* s, uint32_t n, uint32_t* rank, uint32_t* s_tmp) {
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t b = threadIdx.x * per;
    uint32_t cnt = 0;
    for (uint32_t j = b; j < b + per && j < n; j++) cnt += (j == 0) || ((s[j] >> 14) != (s[j - 1] >> 14));
    uint32_t total = 0;
    uint32_t r = block_excl_scan<1024>(cnt, s_tmp, &total);
    for (uint32_t j = b; j < b + per && j < n; j++) {
        r += (j == 0) || ((s[j] >> 14) != (s[j - 1] >> 14));
        rank[s[j] & 0x3FFF] = r;
    }
    __syncthreads();
    return total;
}

template <class TV>
__global__ __launch_bounds__(1024) void k_base_case(TV tv, uint32_t n, uint32_t* __restrict__ sa,
                                                    uint32_t* __restrict__ rank) {
    __shared__ uint64_t s[BASE_CASE_MAX];
    __shared__ uint32_t s_tmp[16];
    uint32_t P = 2;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += 1024) s[i] = i < n ? (((uint64_t)tv.at(i) << 14) | i) : ~0ull;
    __syncthreads();
    lds_bitonic(s, P);
    uint32_t maxr = lds_assign_ranks(s, n, rank, s_tmp);
    for (uint32_t h = 1; maxr < n; h <<= 1) {
        for (uint32_t i = threadIdx.x; i < P; i += 1024) {
            if (i < n) {
                uint64_t r1 = rank[i];
                uint64_t r2 = i + h < n ? rank[i + h] : 0;
                s[i] = (r1 << 29) | (r2 << 14) | i;
            } else {
                s[i] = ~0ull;
            }
        }
        __syncthreads();
        lds_bitonic(s, P);
        maxr = lds_assign_ranks(s, n, rank, s_tmp);
    }
    for (uint32_t j = threadIdx.x; j < n; j += 1024) sa[j] = (uint32_t)(s[j] & 0x3FFF);
}

// Empty kernel launched at the phase boundaries of level 0 when phase marks are on (profiling): in a
// per-dispatch counter list, the dispatches between marks P - 1 and P belong to phase P.
template <int P>
__global__ void k_phase_mark() {}

// ------------------------------------------------------------------------------------------------
// output width conversion: uint32 SA -> 4 / 5 (uint40, uint_types.hpp:262) / 8 bytes little endian
// ------------------------------------------------------------------------------------------------
// out[j] = sa[j] as 64 bits for the slots j = slots[i], i < A (the tied slots, whose entries change after
// the sample scan wrote the others)
__global__ __launch_bounds__(256) void k_widen_slots(const uint32_t* __restrict__ sa, const uint32_t* __restrict__ slots,
                                                     uint64_t A, uint64_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < A; i += (uint64_t)gridDim.x * 256) {
        const uint32_t j = slots[i];
        out[j] = sa[j];
    }
}

__global__ __launch_bounds__(256) void k_widen(const uint32_t* __restrict__ sa, uint64_t n, uint32_t width,
                                               uint8_t* __restrict__ out) {
    for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < n; j += (uint64_t)gridDim.x * 256) {
        uint32_t v = sa[j];
        if (width == 8) {
            reinterpret_cast<uint64_t*>(out)[j] = v;
        } else if (width == 4) {
            reinterpret_cast<uint32_t*>(out)[j] = v;
        } else {
            uint8_t* o = out + 5 * j;
            o[0] = (uint8_t)v;
            o[1] = (uint8_t)(v >> 8);
            o[2] = (uint8_t)(v >> 16);
            o[3] = (uint8_t)(v >> 24);
            o[4] = 0;
        }
    }
}

// 40-bit output with 16-byte aligned buffers: a thread turns 16 entries (four 16-byte loads) into 80 bytes
// (five 16-byte stores) instead of five byte stores per entry; four entries make five little-endian words
__device__ __forceinline__ void pack40x4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t* w) {
    w[0] = a;
    w[1] = b << 8;
    w[2] = (b >> 24) | (c << 16);
    w[3] = (c >> 16) | (d << 24);
    w[4] = d >> 8;
}
__global__ __launch_bounds__(256) void k_widen40(const uint32_t* __restrict__ sa, uint64_t n, uint8_t* __restrict__ out) {
    const uint64_t groups = n / 16;
    for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < groups; t += (uint64_t)gridDim.x * 256) {
        const uint4* in = reinterpret_cast<const uint4*>(sa + 16 * t);
        const uint4 q0 = in[0], q1 = in[1], q2 = in[2], q3 = in[3];
        uint32_t w[20];
        pack40x4(q0.x, q0.y, q0.z, q0.w, w);
        pack40x4(q1.x, q1.y, q1.z, q1.w, w + 5);
        pack40x4(q2.x, q2.y, q2.z, q2.w, w + 10);
        pack40x4(q3.x, q3.y, q3.z, q3.w, w + 15);
        uint4* o = reinterpret_cast<uint4*>(out + 80 * t);
#pragma unroll
        for (int k = 0; k < 5; k++) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
    }
    // the last n mod 16 entries
    for (uint64_t j = groups * 16 + blockIdx.x * 256ull + threadIdx.x; j < n; j += (uint64_t)gridDim.x * 256) {
        const uint32_t v = sa[j];
        uint8_t* o = out + 5 * j;
        o[0] = (uint8_t)v;
        o[1] = (uint8_t)(v >> 8);
        o[2] = (uint8_t)(v >> 16);
        o[3] = (uint8_t)(v >> 24);
        o[4] = 0;
    }
}

// ------------------------------------------------------------------------------------------------
// SA validity check (check_suffixarray, sa_check.hpp:56-184)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t load_sa(const uint8_t* sa, uint32_t width, uint64_t k) {
    if (width == 8) return reinterpret_cast<const uint64_t*>(sa)[k];
    if (width == 4) return reinterpret_cast<const uint32_t*>(sa)[k];
    const uint8_t* p = sa + 5 * k;
    return (uint64_t)p[0] | ((uint64_t)p[1] << 8) | ((uint64_t)p[2] << 16) | ((uint64_t)p[3] << 24) |
           ((uint64_t)p[4] << 32);
}

// permutation test: every entry below n and no position twice (one bit per position in `seen`); the
// inverse rank[p] = k is 0-based (n may be 2^32)
__global__ __launch_bounds__(256) void k_check_perm(const uint8_t* __restrict__ sa, uint32_t width, uint64_t n,
                                                    uint32_t* __restrict__ rank, uint32_t* __restrict__ seen,
                                                    uint32_t* __restrict__ bad) {
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256) {
        uint64_t p = load_sa(sa, width, k);
        if (p >= n) {
            atomicOr(bad, 1u);
            continue;
        }
        const uint32_t bit = 1u << (uint32_t)(p & 31);
        if (atomicOr(&seen[p >> 5], bit) & bit) atomicOr(bad, 1u);
        rank[p] = (uint32_t)k;
    }
}

__global__ __launch_bounds__(256) void k_check_order(const uint8_t* __restrict__ sa, uint32_t width, uint64_t n,
                                                     const uint8_t* __restrict__ text, const uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ bad) {
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k + 1 < n; k += (uint64_t)gridDim.x * 256) {
        uint64_t p = load_sa(sa, width, k), q = load_sa(sa, width, k + 1);
        if (p >= n || q >= n) continue;  // reported by the permutation check
        uint8_t cp = text[p], cq = text[q];
        // rank of the next suffix; the empty suffix (p + 1 = n) is the smallest
        const int64_t rp = p + 1 < n ? (int64_t)rank[p + 1] : -1, rq = q + 1 < n ? (int64_t)rank[q + 1] : -1;
        bool ok = cp != cq ? (cp < cq) : (rp <= rq);
        if (!ok) atomicOr(bad, 2u);
    }
}

// ------------------------------------------------------------------------------------------------
// batched mode: texts joined by a separator (value 1) below every character (values >= 2)
// ------------------------------------------------------------------------------------------------
// largest k in [0, count) with starts[k] <= pos
__device__ __forceinline__ uint32_t find_owner(const uint64_t* __restrict__ starts, uint32_t count, uint64_t pos) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Owner lookup table of a text space [base, base + len) cut into chunks of 2^shift positions:
// table[c] = owner of the chunk's first position (c <= len >> shift).  With texts of a kilobyte and more a
// chunk meets one or two texts, so a lookup searches a range of one or two starts instead of all of them.
constexpr uint32_t OWNER_SHIFT = 10;
__global__ __launch_bounds__(256) void k_owner_table(const uint64_t* __restrict__ starts, uint32_t count,
                                                     uint64_t base, uint64_t chunks, uint32_t* __restrict__ table) {
    for (uint64_t c = blockIdx.x * 256ull + threadIdx.x; c <= chunks; c += (uint64_t)gridDim.x * 256)
        table[c] = find_owner(starts, count, base + (c << OWNER_SHIFT));
}
__device__ __forceinline__ uint32_t find_owner_tab(const uint64_t* __restrict__ starts, uint32_t count,
                                                   const uint32_t* __restrict__ table, uint64_t base, uint64_t pos) {
    const uint64_t c = (pos - base) >> OWNER_SHIFT;
    uint32_t lo = table[c], hi = table[c + 1] + 1;  // the owner lies in [owner(chunk), owner(next chunk)]
    if (hi > count) hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <class OutT>
__global__ __launch_bounds__(256) void k_batch_text(const uint8_t* __restrict__ texts, uint64_t begin, uint64_t end,
                                                    const uint64_t* __restrict__ in_off, const uint64_t* __restrict__ co,
                                                    uint32_t count, const uint16_t* __restrict__ lut,
                                                    const uint32_t* __restrict__ table, OutT* __restrict__ out) {
    for (uint64_t i = begin + blockIdx.x * 256ull + threadIdx.x; i < end; i += (uint64_t)gridDim.x * 256) {
        uint32_t k = find_owner_tab(in_off, count, table, begin, i);
        out[co[k] + (i - in_off[k])] = (OutT)(lut[texts[i]] + 1u);
    }
}

template <class OutT>
__global__ __launch_bounds__(256) void k_batch_seps(const uint64_t* __restrict__ co, uint32_t count,
                                                    OutT* __restrict__ out) {
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < count; k += gridDim.x * 256ull) out[co[k + 1] - 1] = (OutT)1;
}

// The character stream of a joined sub-batch straight from the input texts (no joined text in HBM): a
// workgroup packs PK_CHARS joined positions as k_pack_text does.  Joined characters: a text byte b is
// lut[b] + 1, the separator after each text 1, positions from N on are padding 0; the stream code is the
// character minus `dense` (padding stays 0).  Codes go through LDS one byte each, so the largest
// character must be at most 255.  A thread fills 16 positions at a time; when they lie inside one text
// (the common case) their bytes come from two aligned 16-byte loads (an aligned chunk that holds a text
// byte lies inside the input's pages) instead of 16 byte loads.  The texts of the chunk are found in LDS
// (at most BP_TEXTS of them; a chunk that meets more searches the global starts).
constexpr uint32_t BP_TEXTS = 256;
__global__ __launch_bounds__(256) void k_batch_pack(const uint8_t* __restrict__ texts,
                                                    const uint64_t* __restrict__ in_off,
                                                    const uint64_t* __restrict__ co, uint32_t count, uint64_t N,
                                                    const uint16_t* __restrict__ lut, uint32_t bits, uint32_t dense,
                                                    uint64_t nwords, uint64_t* __restrict__ PT) {
    __shared__ __align__(16) uint8_t s_c[PK_CHARS];
    __shared__ uint16_t s_lut[256];
    __shared__ uint64_t s_co[BP_TEXTS + 1];  // joined starts of the chunk's texts
    __shared__ uint64_t s_in[BP_TEXTS];      // their input offsets
    __shared__ uint32_t s_k0, s_nk;
    const uint32_t tid = threadIdx.x;
    s_lut[tid] = lut[tid];
    const uint32_t wpb = PK_CHARS * bits / 64;
    for (uint64_t chunk = blockIdx.x; chunk * wpb < nwords; chunk += gridDim.x) {
        const uint64_t q0 = chunk * PK_CHARS;
        if (tid == 0) {
            uint32_t k0 = 0, nk = 0;
            if (q0 < N) {
                k0 = find_owner(co, count, q0);
                const uint64_t qe = q0 + PK_CHARS < N ? q0 + PK_CHARS : N;
                nk = find_owner(co, count, qe - 1) - k0 + 1;  // > BP_TEXTS: search the global starts
            }
            s_k0 = k0;
            s_nk = nk;
        }
        __syncthreads();
        const uint32_t k0 = s_k0, nk = s_nk;
        const bool local = nk <= BP_TEXTS;
        if (local)
            for (uint32_t i = tid; i <= nk; i += 256) {
                s_co[i] = co[k0 + i];
                if (i < nk) s_in[i] = in_off[k0 + i];
            }
        __syncthreads();
        for (uint32_t g = tid; g < PK_CHARS / 16; g += 256) {
            const uint64_t q = q0 + 16ull * g;
            uint32_t c[16];
            // owner of q: the last text that starts at or before it
            uint32_t k = 0;
            uint64_t cs = 0, ce = 0, io = 0;  // the owner's joined start, separator position, input offset
            if (q < N) {
                if (local) {
                    uint32_t lo = 0, hi = nk;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_co[mid] <= q)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    k = lo;
                    cs = s_co[k];
                    ce = s_co[k + 1] - 1;
                    io = s_in[k];
                } else {
                    k = find_owner(co, count, q);
                    cs = co[k];
                    ce = co[k + 1] - 1;
                    io = in_off[k];
                }
            }
            if (q + 16 <= ce) {
                // 16 text bytes at input address a: two aligned 16-byte chunks, then byte alignment
                // (v_alignbyte) of the 4 words that start at the right one
                const uintptr_t a = reinterpret_cast<uintptr_t>(texts + io + (q - cs));
                const uint4* p = reinterpret_cast<const uint4*>(a & ~(uintptr_t)15);
                const uint32_t sh = (uint32_t)(a & 15), ws = sh >> 2, bs = sh & 3;
                const uint4 x = p[0];
                const uint4 y = sh ? p[1] : make_uint4(0, 0, 0, 0);
                const uint32_t t[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
                uint32_t o[5];
#pragma unroll
                for (int j = 0; j < 5; j++)
                    o[j] = ws == 0 ? t[j] : ws == 1 ? t[j + 1] : ws == 2 ? t[j + 2] : (j + 3 < 8 ? t[j + 3] : 0u);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t wd = __builtin_amdgcn_alignbyte(o[j + 1], o[j], bs);
#pragma unroll
                    for (int b = 0; b < 4; b++) c[4 * j + b] = (uint32_t)s_lut[(wd >> (8 * b)) & 0xFFu] + 1u - dense;
                }
            } else {
                uint32_t kk = k;
                uint64_t ks = cs, ke = ce, ki = io;
                for (int i = 0; i < 16; i++) {
                    const uint64_t pos = q + i;
                    uint32_t ch = 0;
                    if (pos < N) {
                        while (pos > ke) {  // the next text (still one of the chunk's texts when local)
                            kk++;
                            ks = local ? s_co[kk] : co[kk];
                            ke = (local ? s_co[kk + 1] : co[kk + 1]) - 1;
                            ki = local ? s_in[kk] : in_off[kk];
                        }
                        ch = (pos == ke ? 1u : (uint32_t)s_lut[texts[ki + (pos - ks)]] + 1u) - dense;
                    }
                    c[i] = ch;
                }
            }
            uint4 v;
            v.x = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24);
            v.y = c[4] | (c[5] << 8) | (c[6] << 16) | (c[7] << 24);
            v.z = c[8] | (c[9] << 8) | (c[10] << 16) | (c[11] << 24);
            v.w = c[12] | (c[13] << 8) | (c[14] << 16) | (c[15] << 24);
            reinterpret_cast<uint4*>(s_c)[g] = v;
        }
        __syncthreads();
        for (uint32_t wl = tid; wl < wpb; wl += 256) {
            const uint64_t w = chunk * wpb + wl;
            if (w >= nwords) break;
            uint64_t word = 0;
            const uint32_t qa = (64 * wl) / bits, qb = (64 * wl + 63) / bits;  // characters overlapping the word
            for (uint32_t qq = qa; qq <= qb; qq++) {
                const int32_t pos = (int32_t)(64 * (wl + 1)) - (int32_t)((qq + 1) * bits);
                const uint64_t cc = s_c[qq];
                word |= pos >= 0 ? (cc << pos) : (cc >> (-pos));
            }
            PT[w] = word;
        }
        __syncthreads();
    }
}

// Partition keys of the joined SA for at most BK_MAX_TEXTS texts: (text << 32) | position within the
// text, a separator keyed (count << 32) so that it sorts after every text.  Every workgroup holds the
// texts' separator positions in LDS (32-bit: a joined string has at most 2^32 positions) and a table of
// the owner of every 2^cshift-th joined position, so a lookup is a short binary search in LDS instead of
// random reads of global tables.  Four SA entries per thread and step: one 16-byte load, two 16-byte
// stores.
constexpr uint32_t BK_THREADS = 1024;
constexpr uint32_t BK_MAX_TEXTS = 12288;  // u32 separator positions: 48 KB of LDS
constexpr uint32_t BK_TABLE = 8192;       // u16 owners: 16 KB of LDS
__device__ __forceinline__ uint64_t bk_key(const uint32_t* s_end, const uint16_t* s_tab, uint32_t count,
                                           uint32_t cshift, uint32_t tab_n, uint32_t p) {
    const uint32_t c = p >> cshift;
    uint32_t lo = s_tab[c], hi = c + 1 < tab_n ? s_tab[c + 1] : count - 1;  // owner(p) lies in [lo, hi]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_end[mid] >= p)
            hi = mid;
        else
            lo = mid + 1;
    }
    const uint32_t start = lo ? s_end[lo - 1] + 1 : 0u;
    return p == s_end[lo] ? (uint64_t)count << 32 : ((uint64_t)lo << 32) | (p - start);
}
__global__ __launch_bounds__(BK_THREADS) void k_batch_keys_lds(const uint32_t* __restrict__ sa, uint64_t N,
                                                               const uint64_t* __restrict__ co, uint32_t count,
                                                               uint32_t cshift, uint64_t* __restrict__ keys) {
    __shared__ uint32_t s_end[BK_MAX_TEXTS];
    __shared__ uint16_t s_tab[BK_TABLE];
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < count; k += BK_THREADS) s_end[k] = (uint32_t)(co[k + 1] - 1);
    __syncthreads();
    // owner of the joined position c << cshift: the first text whose separator lies at or after it
    const uint32_t tab_n = (uint32_t)((N - 1) >> cshift) + 1;
    for (uint32_t c = tid; c < tab_n; c += BK_THREADS) {
        const uint64_t p = (uint64_t)c << cshift;
        uint32_t lo = 0, hi = count - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint64_t)s_end[mid] >= p)
                hi = mid;
            else
                lo = mid + 1;
        }
        s_tab[c] = (uint16_t)lo;
    }
    __syncthreads();
    const uint64_t N4 = N >> 2;
    const uint64_t stride = (uint64_t)gridDim.x * BK_THREADS;
    for (uint64_t q = blockIdx.x * (uint64_t)BK_THREADS + tid; q < N4; q += stride) {
        const uint4 p = reinterpret_cast<const uint4*>(sa)[q];
        ulonglong2 a, b;
        a.x = bk_key(s_end, s_tab, count, cshift, tab_n, p.x);
        a.y = bk_key(s_end, s_tab, count, cshift, tab_n, p.y);
        b.x = bk_key(s_end, s_tab, count, cshift, tab_n, p.z);
        b.y = bk_key(s_end, s_tab, count, cshift, tab_n, p.w);
        reinterpret_cast<ulonglong2*>(keys)[2 * q] = a;
        reinterpret_cast<ulonglong2*>(keys)[2 * q + 1] = b;
    }
    if (blockIdx.x == 0 && tid < (uint32_t)(N & 3)) {
        const uint64_t j = (N4 << 2) + tid;
        keys[j] = bk_key(s_end, s_tab, count, cshift, tab_n, sa[j]);
    }
}

// the same for any number of texts, through the global owner table (keys as above)
__global__ __launch_bounds__(256) void k_batch_keys64(const uint32_t* __restrict__ sa, uint64_t N,
                                                      const uint64_t* __restrict__ co, uint32_t count,
                                                      const uint32_t* __restrict__ table, uint64_t* __restrict__ keys) {
    for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < N; j += gridDim.x * 256ull) {
        const uint32_t p = sa[j];
        const uint32_t k = find_owner_tab(co, count, table, 0, p);
        const bool sep = (uint64_t)p == co[k + 1] - 1;
        keys[j] = sep ? (uint64_t)count << 32 : ((uint64_t)k << 32) | (uint32_t)(p - co[k]);
    }
}

}  // namespace dcx
