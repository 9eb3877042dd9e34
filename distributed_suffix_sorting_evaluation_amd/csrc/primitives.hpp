// primitives.hpp -- device-wide building blocks of the DCX engine, written for gfx950 (wave64):
//
//  * LSD radix sort of (u64 key, u32 value) pairs (and u32 keys for permutation scatters), one kernel
//    launch per 8-bit digit.  Each pass is a "onesweep" pass: a workgroup ranks a 7168-key tile with
//    wavefront ballots (no per-key atomics), reorders it by digit in LDS, obtains its per-digit output
//    offsets with a decoupled lookback over the preceding tiles of its segment (one segment, and one
//    lookback chain, per XCD) and writes each digit run contiguously.  One upfront kernel builds the
//    joint (segment, digit) histograms of every pass from a single read of the first-pass keys, which
//    may come from a key source (e.g. windows of a packed character stream) instead of an array.
//  * device scans over a functor (flags -> positions, names ...): two-pass reduce-then-scan for long
//    inputs, single-pass decoupled lookback for short ones.
//
// These replace the reference's distributed sorts and exclusive prefix sums
// (lightweight_dcx/src/sorters/sample_sort.hpp:37-171, lightweight_dcx/src/mpi/reduce.hpp ex_prefix_sum).
#pragma once

#include <cstring>

#include "common.hpp"

namespace dcx {

// 512 threads x 14 keys: 7168-key tiles at two workgroups per CU.  Fewer, larger tiles mean fewer
// lookbacks; measured on gfx950 for 64-bit keys against 512 x 8/10/12/13/15/16, 384 x 10/12, 256 x 12/16/24,
// 768 x 8 and 1024 x 6/8/10/12/14 (15 and 16 keys per thread fall off sharply; 1024 x 14 at one
// workgroup per CU wins on random 64-bit keys but not on the suffix-sort passes)
#ifndef DCX_RS_ITEMS
#define DCX_RS_ITEMS 14
#endif
#ifndef DCX_RS_THREADS
#define DCX_RS_THREADS 512
#endif
constexpr int RS_THREADS = DCX_RS_THREADS;  // a multiple of 64, at least 256 (one thread per digit)
constexpr int RS_ITEMS = DCX_RS_ITEMS;  // keys per thread of one radix tile
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // keys per workgroup
// 32-bit keys (permutation scatters) take less LDS and fewer registers per key: larger tiles
#ifndef DCX_RS_ITEMS_U32
#define DCX_RS_ITEMS_U32 16
#endif
constexpr int RS_ITEMS_U32 = DCX_RS_ITEMS_U32;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_RADIX = 256;
#ifndef DCX_RS_LOOKBACK
#define DCX_RS_LOOKBACK 1
#endif
// predecessors read per lookback round trip: 1, measured against 2, 4 and 8 and against 1 first then 8
// (the status word of a predecessor a tile does not need is a memory request of 8 useful bytes): the
// 4 GiB Zipf text builds in 209 ms at 1 against 215 at 8, the 1 GiB genome in 47.4 against 50.3
constexpr int RS_LOOKBACK = DCX_RS_LOOKBACK;
static_assert(RS_LOOKBACK >= 1, "lookback width");
// keys per batch of the ranking atomics (7: two batches of a 14-key tile fit 128 VGPRs; 14 does not)
constexpr int RS_RANK_BATCH = 7;

// lookback word: [63:42] tag  [41:40] flag  [39:0] count
constexpr uint64_t LB_VALUE_MASK = (1ull << 40) - 1;
constexpr uint64_t LB_AGG = 1ull << 40;
constexpr uint64_t LB_INC = 2ull << 40;
constexpr uint64_t LB_FLAG_MASK = 3ull << 40;
constexpr uint32_t LB_TAG_LIMIT = (1u << 22) - 2;

__device__ __forceinline__ uint64_t lb_pack(uint32_t tag, uint64_t flag, uint64_t v) {
    return ((uint64_t)tag << 42) | flag | v;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, uint8_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// Key sources of a radix sort's first pass: key(i) for input index i.
// ------------------------------------------------------------------------------------------------
template <class KeyT>
struct ArrayKeys {  // keys[i]
    const KeyT* k;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return (uint64_t)k[i]; }
    uint64_t read_bytes(uint64_t n) const { return n * sizeof(KeyT); }  // bytes a pass reads for n keys
};

// The 64-bit window at bit offset i * bits of an MSB-first bit stream (one element of `bits` bits per
// index), masked to its top bits; PACKPOS: the index in the low bits (below the mask).  The stream
// needs one readable word past the last window's first word.
template <bool PACKPOS>
struct StreamKeys {
    static constexpr bool kPackPos = PACKPOS;
    const uint64_t* s;
    uint32_t bits;
    uint64_t mask;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const {
        const uint64_t o = i * bits;
        const uint64_t w = o >> 6;
        const uint32_t sh = (uint32_t)(o & 63);
        // branch-free funnel shift: (next >> 1) >> (63 - sh) is next >> (64 - sh), and 0 for sh = 0
        const uint64_t x = (s[w] << sh) | ((s[w + 1] >> 1) >> (63 - sh));
        return PACKPOS ? ((x & mask) | i) : (x & mask);
    }
    uint64_t read_bytes(uint64_t n) const { return (n * bits + 63) / 64 * 8; }  // the stream words of n keys
};

// Keys of an order-preserving prefix code of a character stream: key(i) = the first kbits bits
// (top-aligned) of the codewords of characters i, i + 1, ..., i + L - 1 of s (`bits` <= 8 bits per
// character), zeros after the L-th; ctab[c] = (length << 24) | the codeword of character code c,
// left-aligned in the low 24 bits.  Every codeword has at least one bit, so a key reads fewer than
// halo() characters from i on.  With packpos the position i itself fills the bits below the code bits
// (keys-only sorts; i < 2^(64 - kbits)).  The sort's histogram kernel (k_coded_hist_seg) computes the
// keys and writes them to mat (N words), which the first pass then reads.
constexpr uint32_t CK_BITS = 40;  // longest key: codewords are placed at bit CK_BITS - (bits filled) and up
struct CodedKeys {
    const uint64_t* s;
    uint32_t bits;
    const uint32_t* ctab;
    uint32_t kbits, L;
    uint64_t* mat;
    uint32_t packpos;
    uint32_t rtl;  // every window of L characters holds kbits bits (L x the shortest codeword >= kbits)
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return mat[i]; }
    uint64_t read_bytes(uint64_t n) const { return 8 * n; }
    __host__ __device__ uint32_t halo() const { return L < CK_BITS ? L : CK_BITS; }
};

// keys[i] = src(i) for i < N (sorts of a single element)
template <class Src>
__global__ void k_materialize(Src src, uint64_t N, uint64_t* __restrict__ keys) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = src(i);
}

// ------------------------------------------------------------------------------------------------
// Segmented radix passes.  A pass cuts its input into RS_SEGS contiguous segments with one
// decoupled-lookback chain each.  A workgroup takes the next tile of the segment numbered like the XCD
// it runs on (the hardware XCC id), and the other segments' tiles once its own are all taken.  The
// consecutive tiles of a segment then run on one XCD, so the partial cache lines at the ends of their
// digit runs meet in that XCD's L2 and leave it as whole lines: 10% faster passes on gfx950 than one
// chain whose consecutive tiles land on different XCDs.
// Segments of pass 0: RS_SEGS equal ranges of tiles.  Of pass p >= 1, whose input is ordered by the
// digit of pass p - 1: the keys whose pass p - 1 digit lies in [32 s, 32 s + 32).  One joint histogram
// of the first-pass keys therefore gives every pass's (segment, digit) counts.  Within a segment the
// tiles are handed out in order by the segment's own counter, so the chain predecessors of a tile are
// always held by running workgroups.
// ------------------------------------------------------------------------------------------------
constexpr int RS_SEGS = 8;
constexpr uint32_t RS_SEG_SHIFT = 5;             // a segment spans 256 / RS_SEGS digits of the previous pass
constexpr int RS_SEG_BINS = RS_SEGS * RS_RADIX;  // (segment, digit) bins of one pass
constexpr int RS_SEG_INFO = 2 * RS_SEGS + 2;     // per pass: segment starts [0, 8], tile prefixes [9, 17]
constexpr uint32_t RS_HIST_KEYS = 1u << 13;      // histogram keys per workgroup (at least)
constexpr uint32_t RS_HIST_WPS = 64;             // at most this many histogram workgroups per segment
constexpr int RS_HIST_THREADS = 1024;            // 512 workgroups of 16 waves fill the machine

// the XCD (accelerator complex die) this wave runs on: hardware register XCC_ID, bits [3:0]
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)); }

// Joint histograms of every pass from one read of the first-pass keys.  Workgroup b covers a
// contiguous share of pass-0 segment b / wps; part[b][p][s][d] counts its keys whose pass p digit is d
// and that lie in segment s of pass p.  Workgroup 0 also clears h2 for k_hist_seg_reduce.  The counts of
// a workgroup fit 32 bits; the joint counts h2 of all keys are 64-bit (a sort may hold 2^32 keys, all of
// them in one bin).
template <class Src>
__global__ __launch_bounds__(RS_HIST_THREADS) void k_radix_hist_seg(Src src, uint64_t N, uint32_t lo_bit, uint32_t np,
                                                        uint32_t seg0_keys, uint32_t wps, uint32_t* __restrict__ part,
                                                        uint64_t* __restrict__ h2) {
    __shared__ uint32_t s_h[8 * RS_SEG_BINS];
    const uint32_t bins = np * RS_SEG_BINS;
    for (uint32_t i = threadIdx.x; i < bins; i += RS_HIST_THREADS) s_h[i] = 0;
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < bins; i += RS_HIST_THREADS) h2[i] = 0;
    __syncthreads();
    const uint32_t sg0 = blockIdx.x / wps, c = blockIdx.x % wps;
    const uint64_t a = (uint64_t)sg0 * seg0_keys < N ? (uint64_t)sg0 * seg0_keys : N;
    const uint64_t e = (uint64_t)(sg0 + 1) * seg0_keys < N ? (uint64_t)(sg0 + 1) * seg0_keys : N;
    const uint64_t len = e - a, chunk = (len + wps - 1) / wps;
    const uint64_t lo = a + (c * chunk < len ? c * chunk : len);
    const uint64_t hi = a + ((c + 1) * chunk < len ? (c + 1) * chunk : len);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += RS_HIST_THREADS) {
        const uint64_t k = src(i);
        uint32_t sg = sg0;
        for (uint32_t p = 0; p < np; p++) {
            const uint32_t dg = (uint32_t)(k >> (lo_bit + 8 * p)) & 0xFFu;
            atomicAdd(&s_h[p * RS_SEG_BINS + sg * RS_RADIX + dg], 1u);
            sg = dg >> RS_SEG_SHIFT;
        }
    }
    __syncthreads();
    uint32_t* out = part + (uint64_t)blockIdx.x * bins;
    for (uint32_t i = threadIdx.x; i < bins; i += RS_HIST_THREADS) out[i] = s_h[i];
}

// The same for prefix-code keys, which it also computes and writes to src.mat.  Each workgroup stages
// its share of the stream, CKH_CHARS positions (plus the characters their keys read beyond) at a time,
// and the code table in LDS.  A thread takes CKH_RUN consecutive positions and slides one bit window along
// them: acc holds the codewords of characters [i, j) from its top bit (filled bits), the key of i is its
// top kbits bits once filled >= kbits or j = i + L, and moving to i + 1 drops character i's codeword and
// appends the next ones: about one table lookup each way per key instead of one per key character.  When
// every window of L characters holds kbits bits (rtl) the keys go right to left instead: one window for the
// position after the run, then each key is its character's codeword followed by the next key.
// The keys leave through a per-wave LDS buffer, 16 lanes' runs (128 consecutive keys) at a time, so each
// store instruction writes 512 contiguous bytes (a store of every lane's r-th key would touch 64 lines).
constexpr uint32_t CKH_RUN = 8;  // a stream word of byte characters
constexpr uint32_t CKH_CHARS = RS_HIST_THREADS * CKH_RUN;
constexpr uint32_t CKH_PASSES = 5;  // CK_BITS / 8
// (8 waves per SIMD: two workgroups per CU, which the LDS allows, need at most 96 scalar registers)
__global__ __launch_bounds__(RS_HIST_THREADS, 8) void k_coded_hist_seg(CodedKeys src, uint64_t N, uint32_t lo_bit,
                                                                    uint32_t np, uint32_t seg0_keys, uint32_t wps,
                                                                    uint32_t* __restrict__ part,
                                                                    uint64_t* __restrict__ h2) {
    __shared__ uint32_t s_h[CKH_PASSES * RS_SEG_BINS];
    __shared__ uint64_t s_w[(CKH_CHARS + CK_BITS) * 8 / 64 + 3];
    static_assert((CKH_CHARS + CK_BITS) * 8 / 64 + 3 <= 2 * RS_HIST_THREADS, "two staged words per thread");
    __shared__ uint32_t s_tab[256];
    __shared__ uint64_t s_k[RS_HIST_THREADS / 64][16 * CKH_RUN];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t bins = np * RS_SEG_BINS;
    for (uint32_t i = threadIdx.x; i < bins; i += RS_HIST_THREADS) s_h[i] = 0;
    for (uint32_t i = threadIdx.x; i < 256; i += RS_HIST_THREADS) s_tab[i] = src.ctab[i];
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < bins; i += RS_HIST_THREADS) h2[i] = 0;
    const uint32_t sg0 = blockIdx.x / wps, c = blockIdx.x % wps;
    const uint64_t a = (uint64_t)sg0 * seg0_keys < N ? (uint64_t)sg0 * seg0_keys : N;
    const uint64_t e = (uint64_t)(sg0 + 1) * seg0_keys < N ? (uint64_t)(sg0 + 1) * seg0_keys : N;
    // shares of whole 8-character runs: segment starts are multiples of the radix tile, so every chunk
    // starts at a multiple of 8
    const uint64_t len = e - a, chunk = ((len + wps - 1) / wps + 7) & ~7ull;
    const uint64_t lo = a + (c * chunk < len ? c * chunk : len);
    const uint64_t hi = a + ((c + 1) * chunk < len ? (c + 1) * chunk : len);
    const uint32_t bits = src.bits, cm = (1u << bits) - 1, kbits = src.kbits, L = src.L;
    const uint64_t km = ~0ull << (64 - kbits);
    // a chunk's stream words: those of characters [base, end - 1 + halo], and the one after (a window's
    // second word), at most 2 * RS_HIST_THREADS of them; each thread loads words threadIdx.x and
    // threadIdx.x + RS_HIST_THREADS of the next chunk into registers while the current one is processed
    auto chunk_words = [&](uint64_t b0, uint64_t& wf, uint32_t& nw) {
        const uint64_t e0 = b0 + CKH_CHARS < hi ? b0 + CKH_CHARS : hi;
        wf = (b0 * bits) >> 6;
        nw = (uint32_t)((((e0 - 1 + src.halo()) * bits) >> 6) + 2 - wf);
    };
    uint64_t pw0 = 0, pw1 = 0;
    if (lo < hi) {
        uint64_t wf;
        uint32_t nw;
        chunk_words(lo, wf, nw);
        if (threadIdx.x < nw) pw0 = src.s[wf + threadIdx.x];
        if (threadIdx.x + RS_HIST_THREADS < nw) pw1 = src.s[wf + threadIdx.x + RS_HIST_THREADS];
    }
    for (uint64_t base = lo; base < hi; base += CKH_CHARS) {
        const uint64_t end = base + CKH_CHARS < hi ? base + CKH_CHARS : hi;
        uint64_t wfirst;
        uint32_t nw;
        chunk_words(base, wfirst, nw);
        __syncthreads();  // the previous chunk's words are read (and the table and counts are set)
        if (threadIdx.x < nw) s_w[threadIdx.x] = pw0;
        if (threadIdx.x + RS_HIST_THREADS < nw) s_w[threadIdx.x + RS_HIST_THREADS] = pw1;
        __syncthreads();
        if (end < hi) {
            uint64_t wf;
            uint32_t nn;
            chunk_words(end, wf, nn);
            if (threadIdx.x < nn) pw0 = src.s[wf + threadIdx.x];
            if (threadIdx.x + RS_HIST_THREADS < nn) pw1 = src.s[wf + threadIdx.x + RS_HIST_THREADS];
        }
        const uint64_t p0 = base + (uint64_t)threadIdx.x * CKH_RUN;
        const uint64_t o0 = p0 * bits - (wfirst << 6);  // bit offset of character p0 in s_w
        // code table entry of character p0 + t
        auto entry = [&](uint32_t t) -> uint32_t {
            const uint64_t q = o0 + (uint64_t)t * bits;
            const uint32_t wi = (uint32_t)(q >> 6), sh = (uint32_t)(q & 63);
            const uint64_t x = (s_w[wi] << sh) | ((s_w[wi + 1] >> 1) >> (63 - sh));
            return s_tab[(uint32_t)(x >> (64 - bits)) & cm];
        };
        const uint32_t cnt = p0 < end ? (uint32_t)(end - p0 < CKH_RUN ? end - p0 : CKH_RUN) : 0u;
        uint64_t acc = 0;
        uint32_t filled = 0, j = 0;  // acc: the codewords of characters [p0 + r, p0 + j)
        uint64_t ks[CKH_RUN];
        if (bits == 8 && (base & 7) == 0 && src.rtl) {
            // byte characters, and no key is cut short by L (every window of L characters holds kbits
            // bits): key(i) is the codeword of character i followed by the top bits of key(i + 1).  One
            // fill gives the key of p0 + 8 (the next run's first position); the run's keys follow from it
            // right to left, a codeword each (no loop whose length varies by lane).  Characters at or past
            // L of a position do not reach its key, so the fill may read past the stream's end and past
            // the staged characters.  Word indices stay below the staged array (clamped).
            constexpr uint32_t SW_LAST = (CKH_CHARS + CK_BITS) * 8 / 64 + 2;
            const uint64_t w = s_w[threadIdx.x];
            uint32_t e8[8];
#pragma unroll
            for (int i = 0; i < 8; i++) e8[i] = s_tab[(uint32_t)(w >> (56 - 8 * i)) & 0xFFu];
            uint64_t k = 0;
            if (cnt) {
                uint64_t w1 = s_w[threadIdx.x + 1], w2 = s_w[threadIdx.x + 2];
                uint32_t wn = threadIdx.x + 3;
                uint32_t en = s_tab[(uint32_t)(w1 >> 56)];  // the next character's entry, one step ahead
                w1 <<= 8;
                uint32_t left = 7, f = 0;
                while (f < kbits) {
                    const uint32_t cur = en;
                    if (left == 0) {
                        w1 = w2;
                        w2 = s_w[wn < SW_LAST ? wn : SW_LAST];
                        wn++;
                        left = 8;
                    }
                    en = s_tab[(uint32_t)(w1 >> 56)];
                    w1 <<= 8;
                    left--;
                    k |= (uint64_t)(cur & 0xFFFFFFu) << (CK_BITS - f);
                    f += cur >> 24;
                }
            }
#pragma unroll
            for (int r = CKH_RUN - 1; r >= 0; r--) {
                k = ((uint64_t)(e8[r] & 0xFFFFFFu) << CK_BITS) | (k >> (e8[r] >> 24));
                ks[r] = (uint32_t)r < cnt ? k & km : 0ull;
            }
        } else if (bits == 8 && (base & 7) == 0) {
            // byte characters: the run's characters are the word s_w[threadIdx.x], and each later word holds
            // the next eight.  The table entry of the next character is read one step ahead (its LDS
            // latency overlaps the current append), the next word one word ahead; the run's own eight
            // entries give the lengths that leave.  Word indices stay below the staged array (clamped).
            constexpr uint32_t SW_LAST = (CKH_CHARS + CK_BITS) * 8 / 64 + 2;
            uint64_t w = s_w[threadIdx.x];
            uint32_t dl[8];
#pragma unroll
            for (int i = 0; i < 8; i++) dl[i] = s_tab[(uint32_t)(w >> (56 - 8 * i)) & 0xFFu] >> 24;
            uint32_t wn = threadIdx.x + 2;
            uint64_t w2 = s_w[threadIdx.x + 1];
            uint32_t en = s_tab[(uint32_t)(w >> 56)];  // entry of character p0 + j
            w <<= 8;
            uint32_t left = 7;  // characters still in w
#pragma unroll
            for (uint32_t r = 0; r < CKH_RUN; r++) {
                ks[r] = 0;
                if (r < cnt) {
                    while (filled < kbits && j < r + L) {
                        const uint32_t cur = en;
                        if (left == 0) {
                            w = w2;
                            w2 = s_w[wn < SW_LAST ? wn : SW_LAST];
                            wn++;
                            left = 8;
                        }
                        en = s_tab[(uint32_t)(w >> 56)];
                        w <<= 8;
                        left--;
                        acc |= (uint64_t)(cur & 0xFFFFFFu) << (CK_BITS - filled);
                        filled += cur >> 24;
                        j++;
                    }
                    ks[r] = acc & km;
                    acc <<= dl[r];
                    filled -= dl[r];
                }
            }
        } else if (src.rtl) {
            // the same right to left for other character widths (characters past L of the last position
            // read stale words of the staged array, which do not reach a key)
            uint32_t e8[CKH_RUN];
#pragma unroll
            for (uint32_t r = 0; r < CKH_RUN; r++) e8[r] = entry(r);
            uint64_t k = 0;
            if (cnt) {
                uint32_t f = 0, t = CKH_RUN;
                while (f < kbits) {
                    const uint32_t en = entry(t++);
                    k |= (uint64_t)(en & 0xFFFFFFu) << (CK_BITS - f);
                    f += en >> 24;
                }
            }
#pragma unroll
            for (int r = CKH_RUN - 1; r >= 0; r--) {
                k = ((uint64_t)(e8[r] & 0xFFFFFFu) << CK_BITS) | (k >> (e8[r] >> 24));
                ks[r] = (uint32_t)r < cnt ? k & km : 0ull;
            }
        } else {
#pragma unroll
            for (uint32_t r = 0; r < CKH_RUN; r++) {
                ks[r] = 0;
                if (r < cnt) {
                    while (filled < kbits && j < r + L) {
                        const uint32_t en = entry(j);
                        acc |= (uint64_t)(en & 0xFFFFFFu) << (CK_BITS - filled);
                        filled += en >> 24;
                        j++;
                    }
                    ks[r] = acc & km;
                    const uint32_t l0 = entry(r) >> 24;  // character p0 + r leaves the window (j > r)
                    acc <<= l0;
                    filled -= l0;
                }
            }
        }
        if (src.packpos) {
#pragma unroll
            for (uint32_t r = 0; r < CKH_RUN; r++) ks[r] |= p0 + r;
        }
#pragma unroll
        for (uint32_t r = 0; r < CKH_RUN; r++) {
            if (r < cnt) {
                uint32_t sg = sg0;
                for (uint32_t p = 0; p < np; p++) {
                    const uint32_t dg = (uint32_t)(ks[r] >> (lo_bit + 8 * p)) & 0xFFu;
                    atomicAdd(&s_h[p * RS_SEG_BINS + sg * RS_RADIX + dg], 1u);
                    sg = dg >> RS_SEG_SHIFT;
                }
            }
        }
        // the wave's 512 positions [wb, wb + 512) out through s_k: lanes 16q .. 16q + 15 hold [wb + 128q,
        // wb + 128q + 128); the LDS accesses of one wave run in order
        const uint64_t wb = base + (uint64_t)wv * 64 * CKH_RUN;
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            if ((lane >> 4) == q) {
#pragma unroll
                for (uint32_t r = 0; r < CKH_RUN; r++) s_k[wv][(lane & 15) * CKH_RUN + r] = ks[r];
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (uint32_t t = 0; t < 2; t++) {
                const uint64_t pos = wb + 128 * q + 64 * t + lane;
                if (pos < end) src.mat[pos] = s_k[wv][64 * t + lane];
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    uint32_t* out = part + (uint64_t)blockIdx.x * bins;
    for (uint32_t i = threadIdx.x; i < bins; i += RS_HIST_THREADS) out[i] = s_h[i];
}

// Character streams whose width divides 8, with every digit on a character boundary.  The digit of
// pass p of key q is the top byte of the stream window at character q + s_p, s_p = (56 - lo - 8 p) /
// bits, and the pass p (segment, digit) pair, for p >= 1, is the top 11 bits of that window.  So one
// histogram of the top 11 bits of the windows at [0, N), kept per pass-0 segment, gives all passes
// after corrections at the ends.  part[b][w] counts the windows at the characters of histogram
// workgroup b (a share of pass-0 segment b / wps) whose top 11 bits are w.
__global__ __launch_bounds__(RS_HIST_THREADS) void k_window_hist_seg(const uint64_t* __restrict__ s, uint32_t bits, uint64_t N,
                                                         uint32_t seg0_keys, uint32_t wps,
                                                         uint32_t* __restrict__ part) {
    __shared__ uint32_t s_h[RS_SEG_BINS];
    for (uint32_t i = threadIdx.x; i < RS_SEG_BINS; i += RS_HIST_THREADS) s_h[i] = 0;
    __syncthreads();
    const uint32_t sg0 = blockIdx.x / wps, c = blockIdx.x % wps;
    const uint64_t a = (uint64_t)sg0 * seg0_keys < N ? (uint64_t)sg0 * seg0_keys : N;
    const uint64_t e = (uint64_t)(sg0 + 1) * seg0_keys < N ? (uint64_t)(sg0 + 1) * seg0_keys : N;
    const uint64_t len = e - a, chunk = (len + wps - 1) / wps;
    const uint64_t lo = a + (c * chunk < len ? c * chunk : len);
    const uint64_t hi = a + ((c + 1) * chunk < len ? (c + 1) * chunk : len);
    // a thread per stream word: its 64 / bits windows from the word and the next one, two loads in all
    // (one load per window pair made the address units the bound)
    const uint32_t ppw = 64 / bits;
    const uint64_t w_lo = lo / ppw, w_hi = (hi + ppw - 1) / ppw;
    for (uint64_t w = w_lo + threadIdx.x; w < w_hi; w += RS_HIST_THREADS) {
        const uint64_t x0 = s[w], x1 = s[w + 1], q0 = w * ppw;
        const uint32_t k0 = q0 < lo ? (uint32_t)(lo - q0) : 0u;
        const uint32_t k1 = q0 + ppw > hi ? (uint32_t)(hi - q0) : ppw;
        for (uint32_t k = k0; k < k1; k++) {
            const uint32_t sh = k * bits;
            const uint64_t x = (x0 << sh) | ((x1 >> 1) >> (63 - sh));
            atomicAdd(&s_h[(uint32_t)(x >> 53)], 1u);
        }
    }
    __syncthreads();
    uint32_t* out = part + (uint64_t)blockIdx.x * RS_SEG_BINS;
    for (uint32_t i = threadIdx.x; i < RS_SEG_BINS; i += RS_HIST_THREADS) out[i] = s_h[i];
}

// h11[s][w] = sum of part[b][w] over the histogram workgroups b of pass-0 segment s (grid 8 x RS_SEGS)
__global__ __launch_bounds__(256) void k_window_seg_reduce(const uint32_t* __restrict__ part, uint32_t wps,
                                                           uint32_t* __restrict__ h11) {
    const uint32_t w = blockIdx.x * 256 + threadIdx.x, sg = blockIdx.y;
    uint32_t c = 0;
    for (uint32_t b = 0; b < wps; b++) c += part[(uint64_t)(sg * wps + b) * RS_SEG_BINS + w];
    h11[sg * RS_SEG_BINS + w] = c;
}

// Workgroup p: the joint counts h2[p][s][d] of pass p from the window histograms.  Pass p >= 1:
// windows at [s_p, N + s_p) = all windows - [0, s_p) + [N, N + s_p).  Pass 0, segment [a, b):
// windows at [a + s_0, b + s_0) = the segment's windows - [a, a + s_0) + [b, b + s_0).
__global__ __launch_bounds__(256) void k_window_seg_joint(const uint64_t* __restrict__ s, uint32_t bits, uint64_t N,
                                                          uint32_t lo_bit, uint32_t seg0_keys,
                                                          const uint32_t* __restrict__ h11,
                                                          uint64_t* __restrict__ h2) {
    __shared__ int s_c[RS_SEG_BINS];
    const uint32_t p = blockIdx.x, d = threadIdx.x;
    const uint32_t sp = (56 - lo_bit - 8 * p) / bits;
    const StreamKeys<false> win{s, bits, ~0ull};
    for (uint32_t i = d; i < RS_SEG_BINS; i += 256) s_c[i] = 0;
    __syncthreads();
    uint64_t* out = h2 + (uint64_t)p * RS_SEG_BINS;
    if (p > 0) {
        for (uint32_t t = d; t < sp; t += 256) {
            atomicSub(&s_c[(uint32_t)(win(t) >> 53)], 1);
            atomicAdd(&s_c[(uint32_t)(win(N + t) >> 53)], 1);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RS_SEGS; j++) {
            const uint32_t w = (d << 3) | j;
            int64_t c = s_c[w];
#pragma unroll
            for (int sg = 0; sg < RS_SEGS; sg++) c += (int64_t)h11[sg * RS_SEG_BINS + w];
            out[j * RS_RADIX + d] = (uint64_t)c;
        }
    } else {
        // s_c[sg][d]: boundary corrections of segment sg
        for (uint32_t t = d; t < RS_SEGS * sp; t += 256) {
            const uint32_t sg = t / sp, k = t % sp;
            const uint64_t a = (uint64_t)sg * seg0_keys < N ? (uint64_t)sg * seg0_keys : N;
            const uint64_t b = (uint64_t)(sg + 1) * seg0_keys < N ? (uint64_t)(sg + 1) * seg0_keys : N;
            atomicSub(&s_c[sg * RS_RADIX + (uint32_t)(win(a + k) >> 56)], 1);
            atomicAdd(&s_c[sg * RS_RADIX + (uint32_t)(win(b + k) >> 56)], 1);
        }
        __syncthreads();
#pragma unroll
        for (int sg = 0; sg < RS_SEGS; sg++) {
            int64_t c = s_c[sg * RS_RADIX + d];
#pragma unroll
            for (int j = 0; j < RS_SEGS; j++) c += (int64_t)h11[sg * RS_SEG_BINS + ((d << 3) | j)];
            out[sg * RS_RADIX + d] = (uint64_t)c;
        }
    }
}

// h2[i] += the sum of part[b][i] over this workgroup's share of the histogram workgroups b
__global__ __launch_bounds__(256) void k_hist_seg_reduce(const uint32_t* __restrict__ part, uint32_t blocks,
                                                         uint32_t bins, uint64_t* __restrict__ h2) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= bins) return;
    uint64_t c = 0;
    for (uint32_t b = blockIdx.y; b < blocks; b += gridDim.y) c += part[(uint64_t)b * bins + i];
    if (c) atomicAdd((unsigned long long*)&h2[i], (unsigned long long)c);
}

// Workgroup p plans pass p: hist[p][d] = global offset of digit d, segoff[p][s][d] = offset of the keys
// of digit d in segment s, info[p] = segment starts and tile prefixes (tiles of `tile` keys); it also
// clears the pass's tile counters.  Counts and offsets are 64-bit: a sort may hold 2^32 keys.
__global__ __launch_bounds__(256) void k_seg_setup(const uint64_t* __restrict__ h2, uint64_t N, uint32_t tile,
                                                   uint32_t seg0_keys, uint64_t* __restrict__ hist,
                                                   uint64_t* __restrict__ segoff, uint64_t* __restrict__ info,
                                                   uint32_t* __restrict__ tick) {
    __shared__ uint64_t s_tmp[4];
    __shared__ uint64_t s_prev[RS_RADIX];
    const uint32_t p = blockIdx.x, d = threadIdx.x;
    const uint64_t* hp = h2 + (uint64_t)p * RS_SEG_BINS;
    uint64_t t = 0;
#pragma unroll
    for (int sg = 0; sg < RS_SEGS; sg++) t += hp[sg * RS_RADIX + d];
    const uint64_t db = block_excl_scan64<256>(t, s_tmp, nullptr);
    hist[p * RS_RADIX + d] = db;
    uint64_t run = db;
#pragma unroll
    for (int sg = 0; sg < RS_SEGS; sg++) {
        segoff[p * RS_SEG_BINS + sg * RS_RADIX + d] = run;
        run += hp[sg * RS_RADIX + d];
    }
    if (d < (uint32_t)RS_SEGS) tick[p * RS_SEGS + d] = 0;
    if (p > 0) {
        // segment starts: the digit offsets of the previous pass at multiples of 32
        __syncthreads();
        const uint64_t* hq = hp - RS_SEG_BINS;
        uint64_t tq = 0;
#pragma unroll
        for (int sg = 0; sg < RS_SEGS; sg++) tq += hq[sg * RS_RADIX + d];
        s_prev[d] = block_excl_scan64<256>(tq, s_tmp, nullptr);
    }
    __syncthreads();
    if (d == 0) {
        uint64_t* in = info + p * RS_SEG_INFO;
        for (int sg = 0; sg <= RS_SEGS; sg++) {
            uint64_t a;
            if (sg == RS_SEGS) a = N;
            else if (p == 0) a = (uint64_t)sg * seg0_keys < N ? (uint64_t)sg * seg0_keys : N;
            else a = s_prev[sg << RS_SEG_SHIFT];
            in[sg] = a;
        }
        uint64_t T = 0;
        for (int sg = 0; sg <= RS_SEGS; sg++) {
            in[RS_SEGS + 1 + sg] = T;
            if (sg < RS_SEGS) T += (in[sg + 1] - in[sg] + tile - 1) / tile;
        }
    }
}

// One LSD pass over digit (key >> shift) & 255.  Stable.  Keys come from `src` (the previous pass's
// output, or a first-pass key source) and go to kout as KeyT: 64-bit sort keys or 32-bit keys (radix
// partitions by destination index).  IDENT: values are the input indices.
// OutT (when not KeyT): the pass writes OutT(key >> out_shift) instead of the key (the last pass of a sort
// whose remaining key bits fit 32 bits below its digit).
// Byte values (VinT or VoutT uint8_t, 64-bit keys whose digits lie at or above bit kpos): an index is
// carried as its low kpos bits in the key's free low bits and its high bits in a byte, 9 bytes per
// element instead of 12.  An IDENT pass splits the index; a pass with VoutT = uint32_t joins it again.
// LIMIT (keys only): output indices at or above out_limit are not written (a partition whose last
// bucket is dropped writes into a buffer that holds only the buckets before it).
// HEADS (the last pass of a sort whose keys are needed only for their group heads): instead of every key
// the pass writes hout[o] = 1 where the key (under hmask) differs from the previous element of the tile's
// run of its digit, 0 where it equals it, and 2 at a run's first element, whose predecessor in the sorted
// order lies in another run; only the runs' first and last keys are written, for k_heads_fix.
template <bool VALS, bool IDENT, class Src, class KeyT = uint64_t, int ITEMS = RS_ITEMS, class OutT = KeyT,
          class VinT = uint32_t, class VoutT = VinT, bool LIMIT = false, bool HEADS = false>
__global__ __launch_bounds__(RS_THREADS) void k_onesweep(Src src, OutT* __restrict__ kout,
                                                         const VinT* __restrict__ vin, VoutT* __restrict__ vout,
                                                         uint32_t shift, const uint64_t* __restrict__ segoff,
                                                         const uint64_t* __restrict__ info, uint32_t* __restrict__ tick,
                                                         uint64_t* __restrict__ status, uint32_t tag,
                                                         uint32_t out_shift = 0, uint32_t kpos = 0,
                                                         uint64_t out_limit = 0, uint8_t* __restrict__ hout = nullptr,
                                                         uint64_t hmask = 0) {
    static_assert(!LIMIT || !VALS, "an output limit applies to key-only passes");
    static_assert(!HEADS || (sizeof(KeyT) == 8 && sizeof(OutT) == 8), "heads come from 64-bit keys");
    constexpr int TILE = RS_THREADS * ITEMS;
    constexpr bool SPLIT = VALS && (sizeof(VinT) == 1 || sizeof(VoutT) == 1);
    static_assert(!SPLIT || sizeof(KeyT) == 8, "split indices need 64-bit keys");
    const uint64_t kmask = (1ull << kpos) - 1;  // kpos <= 24
    __shared__ KeyT s_keys[TILE];
    __shared__ uint8_t s_vb[SPLIT ? TILE : 1];  // split indices: the high bytes, reordered with the keys
    __shared__ uint32_t s_wcnt[RS_WAVES][RS_RADIX];  // per-wave digit counts, then wave offsets
    __shared__ uint32_t s_start[RS_RADIX];
    __shared__ uint32_t s_gofs[RS_RADIX];
    __shared__ uint32_t s_scan[RS_WAVES];
    __shared__ uint32_t s_tile[4];
    __shared__ uint64_t s_base;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t wave = tid / 64;
    if (tid == 0) {
        // the next tile of this XCD's segment, else of the next segment with tiles left (the grid has a
        // workgroup per tile, so every tile is taken)
        // segment starts are 64-bit (a sort may hold 2^32 keys); tile prefixes fit 32 bits
        const uint64_t* tp = info + RS_SEGS + 1;
        const uint32_t x = xcc_id();
        uint32_t sg = 0, i = 0, found = 0;
        for (uint32_t v = 0; v < (uint32_t)RS_SEGS && !found; v++) {
            sg = (x + v) & (RS_SEGS - 1);
            const uint32_t T = (uint32_t)(tp[sg + 1] - tp[sg]);
            if (T) {
                i = atomicAdd(&tick[sg], 1u);
                found = i < T;
            }
        }
        uint32_t n = 0;
        uint64_t b = 0;
        if (found) {
            b = info[sg] + (uint64_t)i * TILE;
            n = info[sg + 1] - b < (uint64_t)TILE ? (uint32_t)(info[sg + 1] - b) : (uint32_t)TILE;
        }
        s_tile[0] = i;                 // tile within the segment
        s_tile[1] = (uint32_t)tp[sg];  // status slot of the segment's first tile
        s_base = b;                    // first key
        s_tile[3] = n | (sg << 29);
    }
    for (int i = tid; i < RS_WAVES * RS_RADIX; i += RS_THREADS) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile_n = s_tile[3] & ((1u << 29) - 1);
    if (tile_n == 0) return;  // all tiles taken
    const uint32_t tile = s_tile[0];
    const uint32_t slot0 = s_tile[1];
    const uint32_t sg = s_tile[3] >> 29;
    const uint64_t tile_base = s_base;
    const uint64_t tile_end = tile_base + tile_n;
    const uint64_t wbase = tile_base + (uint64_t)wave * (64 * ITEMS);

    KeyT key[ITEMS];  // 32-bit keys stay 32-bit in registers
    uint32_t val[ITEMS];
    uint32_t rk[ITEMS];
    // a character stream of at most 8 bits per character: the wave's stream words (at most 114) go
    // through LDS (the key area is free until the reorder), one global load per lane instead of two per
    // key, whose 64 lanes read the same two or three words (the address units, not the bytes, bound that
    // first pass: 6.4 -> 5.1 ms at 1 GiB)
    constexpr bool STAGE = (std::is_same<Src, StreamKeys<false>>::value || std::is_same<Src, StreamKeys<true>>::value) &&
                           sizeof(KeyT) == 8;
    const uint64_t* sw = nullptr;
    uint64_t wfirst = 0;
    if constexpr (STAGE) {
        if (src.bits <= 8) {
            uint64_t* w = reinterpret_cast<uint64_t*>(s_keys) + wave * 128;
            wfirst = (wbase * src.bits) >> 6;
            const uint64_t wend = wbase + 64 * ITEMS < tile_end ? wbase + 64 * ITEMS : tile_end;
            // words of the positions [wbase, wend): up to the second word of the last one
            const uint32_t nw = wbase < wend ? (uint32_t)((((wend - 1) * src.bits) >> 6) + 2 - wfirst) : 0u;
            for (uint32_t q = lane; q < nw; q += 64) w[q] = src.s[wfirst + q];
            __builtin_amdgcn_wave_barrier();
            sw = w;
        }
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint64_t idx = wbase + k * 64 + lane;
        bool valid = idx < tile_end;
        if constexpr (STAGE) {
            if (sw) {
                const uint64_t o = idx * src.bits - (wfirst << 6);
                const uint32_t wi = (uint32_t)(o >> 6), sh = (uint32_t)(o & 63);
                const uint64_t x = (sw[wi] << sh) | ((sw[wi + 1] >> 1) >> (63 - sh));
                key[k] = valid ? (Src::kPackPos ? ((x & src.mask) | idx) : (x & src.mask)) : 0ull;
            } else {
                key[k] = valid ? src(idx) : 0ull;
            }
        } else {
            key[k] = valid ? src(idx) : 0ull;
        }
        if constexpr (SPLIT && IDENT) {
            const uint32_t i32 = (uint32_t)idx;  // indices of a sort stay below 2^32
            key[k] |= i32 & (uint32_t)kmask;
            val[k] = i32 >> kpos;
        } else if constexpr (SPLIT) {
            val[k] = 0;  // from the staged bytes below
        } else if constexpr (VALS) {
            val[k] = IDENT ? (uint32_t)idx : (valid ? (uint32_t)vin[idx] : 0u);
        }
    }
    if constexpr (SPLIT && !IDENT) {
        // the wave's index bytes by 16-byte loads through LDS (the key area is free until the reorder): a
        // byte load per key would take as many address-unit cycles as a key load for an eighth of the data.
        // The chunks start at a multiple of 16 and end at most 15 bytes past the last byte, inside the
        // buffer (allocations are whole multiples of 256 bytes).
        uint8_t* w = reinterpret_cast<uint8_t*>(s_keys) + wave * 1024;
        const uint64_t wend = wbase + 64 * ITEMS < tile_end ? wbase + 64 * ITEMS : tile_end;
        const uint64_t b16 = wbase & ~15ull;
        const uint32_t nc = wbase < wend ? (uint32_t)((wend - b16 + 15) >> 4) : 0u;  // at most 57
        if (lane < nc)
            reinterpret_cast<uint4*>(w)[lane] = reinterpret_cast<const uint4*>(vin + b16)[lane];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint64_t idx = wbase + k * 64 + lane;
            val[k] = idx < tile_end ? (uint32_t)w[idx - b16] : 0u;
        }
    }
    // rank every key among the keys of its wave with the same digit, in tile order: the lowest lane of
    // each digit adds the digit's lane count to the wave's counter (LDS atomics of one wave apply in
    // order, so the returned count is the number of earlier keys of that digit), the others read it from
    // that lane; RS_RANK_BATCH keys per batch keep the atomics in flight together
#pragma unroll
    for (int k0 = 0; k0 < ITEMS; k0 += RS_RANK_BATCH) {
        uint32_t old[RS_RANK_BATCH];
#pragma unroll
        for (int k = k0; k < k0 + RS_RANK_BATCH && k < ITEMS; k++) {
            uint64_t idx = wbase + k * 64 + lane;
            bool valid = idx < tile_end;
            uint32_t d = (uint32_t)(key[k] >> shift) & 0xFFu;
            // peers: the lanes with this lane's digit, as two 32-bit halves; bit b of the digit as a word of
            // all ones or zeros (m), the wave's ballot of it (bb), and the lanes that agree: ~(bb ^ m)
            const uint64_t vb = __ballot(valid);
            uint32_t plo = (uint32_t)vb, phi = (uint32_t)(vb >> 32);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint32_t m = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
                const uint64_t bb = __ballot(m != 0);
                plo &= ~((uint32_t)bb ^ m);
                phi &= ~((uint32_t)(bb >> 32) ^ m);
            }
            const uint64_t peers = ((uint64_t)phi << 32) | plo;
            // peers in lower lanes (a masked population count in two instructions)
            uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
            uint32_t lead = peers ? (uint32_t)__builtin_ctzll(peers) : lane;
            old[k - k0] = 0;
            if (valid && below == 0) old[k - k0] = atomicAdd(&s_wcnt[wave][d], (uint32_t)__popcll(peers));
            rk[k] = d | (below << 8) | (lead << 16);
        }
#pragma unroll
        for (int k = k0; k < k0 + RS_RANK_BATCH && k < ITEMS; k++) {
            const uint32_t lead = rk[k] >> 16;
            const uint32_t cnt = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(lead << 2), (int)old[k - k0]);
            rk[k] = (rk[k] & 0xFFu) | ((cnt + ((rk[k] >> 8) & 0xFFu)) << 8);
        }
    }
    __syncthreads();
    // per digit (threads 0..255): offsets of the waves and the tile total
    const bool dthr = tid < RS_RADIX;
    const uint32_t d = dthr ? tid : 0;
    uint32_t total = 0;
    if (dthr) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            uint32_t c = s_wcnt[w][d];
            s_wcnt[w][d] = total;
            total += c;
        }
    }
    uint32_t start = block_excl_scan<RS_THREADS>(dthr ? total : 0u, s_scan, nullptr);
    uint64_t* st = status + (uint64_t)(slot0 + tile) * RS_RADIX + d;
    if (dthr) {
        s_start[d] = start;
        // publish this tile's count of the digit: inclusive right away for the segment's first tile
        st_agent(st, lb_pack(tag, tile == 0 ? LB_INC : LB_AGG, total));
    }
    __syncthreads();
    // reorder the tile by digit in LDS (the keys' registers are free for the lookback afterwards)
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint64_t idx = wbase + k * 64 + lane;
        uint32_t dk = rk[k] & 0xFFu;
        if (idx < tile_end) {
            uint32_t lp = s_start[dk] + s_wcnt[wave][dk] + (rk[k] >> 8);
            s_keys[lp] = (KeyT)key[k];
            if constexpr (SPLIT) s_vb[lp] = (uint8_t)val[k];
        }
    }
    if (dthr) {
        // decoupled lookback: exclusive count of this digit over the preceding tiles of the segment.
        // Every read is a round trip to memory; RS_LOOKBACK predecessors are read per round trip.
        uint64_t excl = 0;
        if (tile > 0) {
            int64_t j = (int64_t)tile - 1;  // next predecessor to add; the first tile is always inclusive
            while (true) {
                const int cnt = (int)(j + 1 < RS_LOOKBACK ? j + 1 : RS_LOOKBACK);
                uint64_t wb[RS_LOOKBACK];
#pragma unroll
                for (int b = 0; b < RS_LOOKBACK; b++)
                    wb[b] = b < cnt ? ld_agent(status + (uint64_t)(slot0 + j - b) * RS_RADIX + d) : 0ull;
                bool stop = false, inc = false;
                int used = 0;
#pragma unroll
                for (int b = 0; b < RS_LOOKBACK; b++) {
                    if (stop || b >= cnt) continue;
                    if ((uint32_t)(wb[b] >> 42) != tag || (wb[b] & LB_FLAG_MASK) == 0) {
                        stop = true;  // not published yet: read again from here
                        continue;
                    }
                    excl += wb[b] & LB_VALUE_MASK;
                    used++;
                    if (wb[b] & LB_INC) inc = stop = true;
                }
                if (inc) break;
                if (used == 0) __builtin_amdgcn_s_sleep(1);
                j -= used;
            }
            st_agent(st, lb_pack(tag, LB_INC, excl + total));
        }
        // output index of the tile's position i with this digit: gofs + i (modulo 2^32: every index
        // written is below the key count, at most 2^32)
        s_gofs[d] = (uint32_t)segoff[sg * RS_RADIX + d] + (uint32_t)excl - start;
    }
    __syncthreads();
    // write out: position i of the reordered tile goes to gofs[digit] + i; the digits of a thread's
    // positions stay in registers for the values
    uint32_t dg[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const uint32_t i = tid + r * RS_THREADS;
        dg[r] = 0;
        if (i < tile_n) {
            const KeyT kk = s_keys[i];
            dg[r] = (uint32_t)((uint64_t)kk >> shift) & 0xFFu;
            const uint32_t o = s_gofs[dg[r]] + i;
            if constexpr (HEADS) {
                const bool rs = i == s_start[dg[r]];
                const bool re = i + 1 == tile_n || (((uint64_t)s_keys[i + 1] >> shift) & 0xFFu) != dg[r];
                hout[o] = rs ? (uint8_t)2 : (uint8_t)((((uint64_t)kk ^ (uint64_t)s_keys[i - 1]) & hmask) != 0);
                if (rs || re) kout[o] = (OutT)kk;
            } else if (!LIMIT || (uint64_t)o < out_limit) {
                kout[o] = (OutT)((uint64_t)kk >> out_shift);
            }
            if constexpr (SPLIT) {
                if constexpr (sizeof(VoutT) == 1)
                    vout[o] = s_vb[i];
                else
                    vout[o] = (VoutT)(((uint64_t)s_vb[i] << kpos) | ((uint64_t)kk & kmask));
            }
        }
    }
    if constexpr (VALS && !SPLIT) {
        __syncthreads();
        uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            uint64_t idx = wbase + k * 64 + lane;
            uint32_t dk = rk[k] & 0xFFu;
            if (idx < tile_end) s_vals[s_start[dk] + s_wcnt[wave][dk] + (rk[k] >> 8)] = val[k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const uint32_t i = tid + r * RS_THREADS;
            if (i < tile_n) vout[s_gofs[dg[r]] + i] = s_vals[i];
        }
    }
}

// ------------------------------------------------------------------------------------------------
// single-pass scan (decoupled lookback) over a functor:
//   op.value(i) -> uint64_t     op.write(i, exclusive_prefix, value)
// ------------------------------------------------------------------------------------------------
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

template <class Op>
__global__ __launch_bounds__(SC_THREADS) void k_scan(Op op, uint64_t N, uint64_t* __restrict__ st_lo,
                                                     uint64_t* __restrict__ st_hi, uint32_t tag,
                                                     unsigned long long* __restrict__ ticket,
                                                     unsigned long long ticket_base, uint64_t* __restrict__ total_out) {
    __shared__ uint64_t s_tmp[SC_THREADS / 64];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_v[SC_TILE + SC_TILE / 16];  // padded to spread blocked reads over banks
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_tile = (uint32_t)(atomicAdd(ticket, 1ull) - ticket_base);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tbase = (uint64_t)tile * SC_TILE;
    // striped (coalesced) loads, transposed through LDS to a blocked arrangement for the scan
    uint64_t v[SC_ITEMS];
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t idx = k * SC_THREADS + tid;
        const uint64_t i = tbase + idx;
        v[k] = i < N ? op.value(i) : 0ull;
        s_v[idx + (idx >> 4)] = v[k];
    }
    __syncthreads();
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t idx = tid * SC_ITEMS + k;
        sum += s_v[idx + (idx >> 4)];
    }
    uint64_t tile_total = 0;
    uint64_t texcl = block_excl_scan64<SC_THREADS>(sum, s_tmp, &tile_total);
    if (tid < 64) {
        // wave 0: publish the aggregates of both 32-bit lanes, then look back over 64 predecessors at a
        // time; each lane of the value is an independent lookback (its own nearest inclusive prefix)
        const uint32_t lane = tid;
        const uint64_t tot_lo = tile_total & 0xFFFFFFFFull, tot_hi = tile_total >> 32;
        if (lane == 0) {
            const uint64_t f = tile == 0 ? LB_INC : LB_AGG;
            st_agent(&st_lo[tile], lb_pack(tag, f, tot_lo));
            st_agent(&st_hi[tile], lb_pack(tag, f, tot_hi));
        }
        uint64_t ex_lo = 0, ex_hi = 0;
        if (tile > 0) {
            bool done_lo = false, done_hi = false;
            int64_t base = (int64_t)tile - 1;
            while (!(done_lo && done_hi)) {
                const int64_t j = base - (int64_t)lane;
                uint64_t wl = LB_INC, wh = LB_INC;  // before tile 0: inclusive zeros
                if (j >= 0) {
                    while (true) {
                        wl = ld_agent(&st_lo[j]);
                        wh = ld_agent(&st_hi[j]);
                        if ((uint32_t)(wl >> 42) == tag && (uint32_t)(wh >> 42) == tag && (wl & LB_FLAG_MASK) &&
                            (wh & LB_FLAG_MASK))
                            break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                if (!done_lo) {
                    const uint64_t m = __ballot((wl & LB_INC) != 0);
                    const uint32_t k = m ? (uint32_t)__ffsll((long long)m) - 1 : 63u;
                    uint64_t part = lane <= k ? (wl & LB_VALUE_MASK) : 0ull;
                    ex_lo += __shfl(wave_incl_scan64(part), 63, 64);
                    done_lo = m != 0;
                }
                if (!done_hi) {
                    const uint64_t m = __ballot((wh & LB_INC) != 0);
                    const uint32_t k = m ? (uint32_t)__ffsll((long long)m) - 1 : 63u;
                    uint64_t part = lane <= k ? (wh & LB_VALUE_MASK) : 0ull;
                    ex_hi += __shfl(wave_incl_scan64(part), 63, 64);
                    done_hi = m != 0;
                }
                base -= 64;
            }
        }
        if (lane == 0) {
            if (tile > 0) {
                st_agent(&st_lo[tile], lb_pack(tag, LB_INC, ex_lo + tot_lo));
                st_agent(&st_hi[tile], lb_pack(tag, LB_INC, ex_hi + tot_hi));
            }
            const uint64_t excl = ex_lo | (ex_hi << 32);
            s_excl = excl;
            if (tile == gridDim.x - 1 && total_out) {  // both lane totals in full (a lane may total 2^32)
                total_out[0] = ex_lo + tot_lo;
                total_out[1] = ex_hi + tot_hi;
            }
        }
    }
    __syncthreads();
    uint64_t run = s_excl + texcl;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t idx = tid * SC_ITEMS + k;
        const uint32_t q = idx + (idx >> 4);
        const uint64_t x = s_v[q];
        s_v[q] = run;
        run += x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t idx = k * SC_THREADS + tid;
        const uint64_t i = tbase + idx;
        if (i < N) op.write(i, s_v[idx + (idx >> 4)], v[k]);
    }
}

// ------------------------------------------------------------------------------------------------
// reduce-then-scan for long inputs: block b owns the contiguous tiles [b * tpb, (b + 1) * tpb).
// k_scan_reduce sums each block's values (both 32-bit lanes apart), k_scan_spine turns the block sums
// into exclusive block prefixes, k_scan_down rescans every tile of the block from its prefix and calls
// op.write.  value() is evaluated twice per element, but no tile waits on another one.
//
// Ops with Op::kBlocked read element i (and i + 1) of plain arrays without branches in value(): a thread
// then takes 16 consecutive elements, whose loads the compiler merges into vector loads (one 16-byte
// load instead of 16 byte loads per thread for a flag array).  The exclusive prefixes go through LDS
// so that op.write still runs in striped order (coalesced stores); the value of an element is the
// difference of its neighbouring prefixes.  Other ops load striped and scan through LDS.
// ------------------------------------------------------------------------------------------------
#ifndef DCX_SC_BLOCKS
#define DCX_SC_BLOCKS 2048
#endif
constexpr uint32_t SC_BLOCKS = DCX_SC_BLOCKS;  // most blocks of the two-pass scan
#ifndef DCX_SC_SINGLE_PASS_TILES
#define DCX_SC_SINGLE_PASS_TILES 256
#endif
constexpr uint64_t SC_SINGLE_PASS_TILES = DCX_SC_SINGLE_PASS_TILES;  // up to this many tiles: one pass
static_assert(SC_BLOCKS >= 1 && SC_BLOCKS <= 4 * 1024, "k_scan_spine scans at most 4096 block sums");

// Ops with a payload(i) (a 32-bit word that write() needs, typically loaded by value() anyway) get it
// handed to write(i, excl, v, payload) from LDS instead of loading it a second time.
template <class Op, class = void>
struct ScanPayload {
    static constexpr bool value = false;
};
template <class Op>
struct ScanPayload<Op, decltype((void)&Op::payload)> {
    static constexpr bool value = true;
};

// Ops with kCompact write only for elements whose low value half is 1 (the low half is 0 or 1, the high
// half 0 or 1): the down pass stages just those elements (their offset in the tile and high prefix) in
// LDS, in output order, and hands them to write() striped over the compacted list, so consecutive threads
// write consecutive outputs and LDS traffic follows the element count, not the tile size.
template <class Op, class = void>
struct ScanCompact {
    static constexpr bool value = false;
};
template <class Op>
struct ScanCompact<Op, decltype((void)Op::kCompact)> {
    static constexpr bool value = Op::kCompact;
};

// Ops with kAll see every element's payload in the down pass of a two-pass scan (blocked compacting ops
// with payloads only): when wants_all(), all(tbase, p, N) with the tile's SC_TILE payloads from tbase in
// LDS, called by every thread of the workgroup.
template <class Op, class = void>
struct ScanAll {
    static constexpr bool value = false;
};
template <class Op>
struct ScanAll<Op, decltype((void)Op::kAll)> {
    static constexpr bool value = Op::kAll;
};

template <bool PL = false, class Op>
__device__ __forceinline__ void scan_load(const Op& op, uint64_t N, uint64_t tbase, uint32_t tid,
                                          uint64_t (&v)[SC_ITEMS], uint32_t* pl = nullptr) {
    if (Op::kBlocked) {
        const uint64_t base = tbase + (uint64_t)tid * SC_ITEMS;
        if (tbase + SC_TILE <= N) {
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                v[k] = op.value(base + k);
                if constexpr (PL) pl[k] = op.payload(base + k);
            }
        } else {
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                v[k] = base + k < N ? op.value(base + k) : 0ull;
                if constexpr (PL) pl[k] = base + k < N ? op.payload(base + k) : 0u;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++) {
            const uint64_t i = tbase + (uint64_t)k * SC_THREADS + tid;
            v[k] = i < N ? op.value(i) : 0ull;
        }
    }
}

template <class Op>
__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(Op op, uint64_t N, uint32_t tpb,
                                                            uint64_t* __restrict__ part) {
    __shared__ uint64_t s_red[2][SC_THREADS / 64];
    const uint32_t tid = threadIdx.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * tpb;
    uint64_t lo = 0, hi = 0;
    for (uint64_t t = t0; t < t0 + tpb; t++) {
        const uint64_t tbase = t * SC_TILE;
        if (tbase >= N) break;
        uint64_t v[SC_ITEMS];
        scan_load(op, N, tbase, tid, v);
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++) {
            lo += v[k] & 0xFFFFFFFFull;
            hi += v[k] >> 32;
        }
    }
    lo = __shfl(wave_incl_scan64(lo), 63, 64);
    hi = __shfl(wave_incl_scan64(hi), 63, 64);
    const uint32_t wave = tid / 64;
    if (__lane_id() == 0) {
        s_red[0][wave] = lo;
        s_red[1][wave] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        uint64_t a = 0, b = 0;
#pragma unroll
        for (int w = 0; w < SC_THREADS / 64; w++) {
            a += s_red[0][w];
            b += s_red[1][w];
        }
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

// exclusive prefixes of nb <= 4 * 1024 block sums, packed as lo | hi << 32 (an exclusive prefix of either
// lane is below 2^32); the grand totals of both lanes in full to total_out[0] and total_out[1]
__global__ __launch_bounds__(1024) void k_scan_spine(const uint64_t* __restrict__ part, uint32_t nb,
                                                     uint64_t* __restrict__ prefix, uint64_t* __restrict__ total_out) {
    __shared__ uint64_t s_tmp[1024 / 64];
    const uint32_t tid = threadIdx.x;
    uint64_t lo[4], hi[4], slo = 0, shi = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t b = tid * 4 + k;
        lo[k] = b < nb ? part[2 * b] : 0ull;
        hi[k] = b < nb ? part[2 * b + 1] : 0ull;
        slo += lo[k];
        shi += hi[k];
    }
    uint64_t tlo = 0, thi = 0;
    uint64_t elo = block_excl_scan64<1024>(slo, s_tmp, &tlo);
    uint64_t ehi = block_excl_scan64<1024>(shi, s_tmp, &thi);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t b = tid * 4 + k;
        if (b < nb) prefix[b] = elo | (ehi << 32);
        elo += lo[k];
        ehi += hi[k];
    }
    if (tid == 0 && total_out) {
        total_out[0] = tlo;
        total_out[1] = thi;
    }
}

template <class Op>
__global__ __launch_bounds__(SC_THREADS) void k_scan_down(Op op, uint64_t N, uint32_t tpb,
                                                          const uint64_t* __restrict__ prefix) {
    constexpr bool PL = ScanPayload<Op>::value && Op::kBlocked;
    constexpr bool COMPACT = ScanCompact<Op>::value && Op::kBlocked;
    static_assert(!ScanAll<Op>::value || (PL && COMPACT), "every payload is staged through the compaction buffer");
    constexpr uint32_t PAD = SC_TILE + SC_TILE / 16;  // slot of the tile's end prefix
    __shared__ uint64_t s_tmp[SC_THREADS / 64];
    __shared__ uint64_t s_e[COMPACT ? 1 : PAD + 1];  // padded to spread blocked accesses over banks
    __shared__ uint32_t s_p[PL && !COMPACT ? PAD : 1];
    // COMPACT, 16-bit entries (a tile offset and a count within a tile fit 13 bits; half the LDS of 32-bit
    // entries lets more workgroups share a compute unit): tile offset | high value bit << 15, and the high
    // prefix relative to the tile's
    static_assert(SC_TILE <= (1 << 15), "compaction entries hold tile offsets in 15 bits");
    __shared__ uint16_t s_ci[COMPACT ? SC_TILE : 1];
    __shared__ uint16_t s_chi[COMPACT ? SC_TILE : 1];
    __shared__ __align__(16) uint32_t s_cp[COMPACT && PL ? SC_TILE : 1];  // COMPACT: payloads (16-byte stores)
    const uint32_t tid = threadIdx.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * tpb;
    uint64_t carry = prefix[blockIdx.x];
    for (uint64_t t = t0; t < t0 + tpb; t++) {
        const uint64_t tbase = t * SC_TILE;
        if (tbase >= N) break;
        uint64_t v[SC_ITEMS];
        uint32_t pl[PL ? SC_ITEMS : 1];
        scan_load<PL>(op, N, tbase, tid, v, pl);
        if constexpr (!Op::kBlocked) {
            // striped values -> blocked through LDS
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                const uint32_t idx = k * SC_THREADS + tid;
                s_e[idx + (idx >> 4)] = v[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                const uint32_t idx = tid * SC_ITEMS + k;
                v[k] = s_e[idx + (idx >> 4)];
            }
        }
        uint64_t sum = 0;
#pragma unroll
        for (int k = 0; k < SC_ITEMS; k++) sum += v[k];
        uint64_t tile_total = 0;
        uint64_t run = carry + block_excl_scan64<SC_THREADS>(sum, s_tmp, &tile_total);  // synchronizes
        if constexpr (COMPACT) {
            const uint32_t lo0 = (uint32_t)carry, hi0 = (uint32_t)(carry >> 32);
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                if ((uint32_t)v[k]) {
                    const uint32_t e = (uint32_t)run - lo0;
                    s_ci[e] = (uint16_t)((tid * SC_ITEMS + k) | ((uint32_t)(v[k] >> 32) << 15));
                    s_chi[e] = (uint16_t)((uint32_t)(run >> 32) - hi0);
                    if constexpr (PL) s_cp[e] = pl[k];
                }
                run += v[k];
            }
            __syncthreads();
            const uint32_t cnt = (uint32_t)tile_total;
            for (uint32_t e = tid; e < cnt; e += SC_THREADS) {
                const uint32_t ci = s_ci[e];
                const uint64_t excl = (uint64_t)(lo0 + e) | ((uint64_t)(hi0 + s_chi[e]) << 32);
                const uint64_t vv = 1ull | ((uint64_t)(ci >> 15) << 32);
                if constexpr (PL)
                    op.write(tbase + (ci & 0x7FFFu), excl, vv, s_cp[e]);
                else
                    op.write(tbase + (ci & 0x7FFFu), excl, vv);
            }
            if constexpr (PL && ScanAll<Op>::value) {
                if (op.wants_all()) {
                    // every payload of the tile through s_cp, then written striped (the tile's entries in
                    // order, consecutive lanes on consecutive entries)
                    __syncthreads();  // the compacted payloads are written out
#pragma unroll
                    for (int k = 0; k < SC_ITEMS; k += 4)
                        reinterpret_cast<uint4*>(s_cp)[(tid * SC_ITEMS + k) / 4] =
                            make_uint4(pl[k], pl[k + 1], pl[k + 2], pl[k + 3]);
                    __syncthreads();
                    op.all(tbase, s_cp, N);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                const uint32_t idx = tid * SC_ITEMS + k;
                s_e[idx + (idx >> 4)] = run;
                run += v[k];
                if constexpr (PL) s_p[idx + (idx >> 4)] = pl[k];
            }
            if (tid == SC_THREADS - 1) s_e[PAD] = run;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SC_ITEMS; k++) {
                const uint32_t idx = k * SC_THREADS + tid;
                const uint32_t nx = idx + 1;
                const uint64_t i = tbase + idx;
                const uint64_t e = s_e[idx + (idx >> 4)];
                if constexpr (PL) {
                    if (i < N) op.write(i, e, s_e[nx + (nx >> 4)] - e, s_p[idx + (idx >> 4)]);
                } else {
                    if (i < N) op.write(i, e, s_e[nx + (nx >> 4)] - e);
                }
            }
        }
        carry += tile_total;
        __syncthreads();
    }
}

// Small results for the host (scan totals, counts, a few ranges): one thread copies them into a pinned,
// device-mapped, coherent host buffer with system-scope stores and then stores the call's sequence number
// with release ordering; the host spins on that number.  Neither a copy-engine transfer nor a stream
// synchronization is involved.
constexpr uint32_t MAIL_WORDS = 512;  // payload capacity in 64-bit words (4 KiB)
__global__ void k_publish(const uint64_t* __restrict__ src, uint32_t words, uint64_t* box, uint64_t seq) {
    if (threadIdx.x != 0) return;
    for (uint32_t i = 0; i < words; i++)
        __hip_atomic_store(&box[1 + i], src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&box[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------------------------------------------------------
// Host side: scratch for lookback state and the launch wrappers
// ------------------------------------------------------------------------------------------------
struct Primitives {
    hipStream_t stream = nullptr;
    DeviceCache* cache = nullptr;
    // persistent scratch
    uint64_t* sort_status = nullptr;
    size_t sort_status_words = 0;
    uint32_t sort_tag = 0;
    uint64_t* scan_lo = nullptr;
    uint64_t* scan_hi = nullptr;
    size_t scan_tiles_cap = 0;
    uint32_t scan_tag = 0;
    unsigned long long* ticket = nullptr;  // device counter of scan tile tickets
    unsigned long long ticket_host = 0;
    uint64_t* scalars = nullptr;           // device scalars: scan totals, two per slot (low and high lane)
    uint64_t* scan_part = nullptr;         // two-pass scan: block sums (2 per block) and block prefixes
    uint64_t* scan_prefix = nullptr;
    uint64_t* hist = nullptr;              // 8 passes x 256 global digit offsets of the last sort
    // segmented radix passes (up to 8 passes): per-histogram-workgroup joint counts, their sums, the
    // (segment, digit) offsets, segment starts / tile prefixes and the per-segment tile counters
    uint32_t* seg_part = nullptr;
    uint64_t* seg_h2 = nullptr;
    uint32_t* seg_h11 = nullptr;  // window histograms per pass-0 segment
    uint64_t* seg_off = nullptr;
    uint64_t* seg_info = nullptr;
    uint32_t* seg_tick = nullptr;
    uint64_t* mailbox = nullptr;           // pinned, device-mapped, coherent: [0] sequence, [1..] payload
    uint64_t* dev_mailbox = nullptr;       // the device's address of the same buffer
    uint64_t mail_seq = 0;
    uint32_t cus = 1;                      // compute units of the device
    // set while a call must not allocate: lookback status arrays that would have to grow throw instead
    // (they are sized at handle creation for the largest such call)
    bool no_alloc = false;
    // optional timing of the 64-bit radix passes: one event pair around the passes of each sort (they run
    // back to back; an event per pass would put a gap of a few microseconds after every pass)
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pass_events;
    std::vector<uint32_t> pass_launches;  // passes of each recorded sort
    std::vector<uint64_t> pass_items;     // elements summed over those passes
    std::vector<uint64_t> pass_bytes;     // algorithmic bytes summed over those passes

    // sums the recorded sort times (synchronizes) and releases the events
    void collect_profile(uint64_t* passes, uint64_t* items, uint64_t* bytes, double* ms) {
        *passes = 0;
        *items = 0;
        *bytes = 0;
        *ms = 0;
        for (size_t i = 0; i < pass_events.size(); i++) {
            float a = 0;
            DCX_HIP(hipEventSynchronize(pass_events[i].second));
            DCX_HIP(hipEventElapsedTime(&a, pass_events[i].first, pass_events[i].second));
            *ms += a;
            *passes += pass_launches[i];
            *items += pass_items[i];
            *bytes += pass_bytes[i];
        }
        drop_profile();
    }
    void drop_profile() {
        for (auto& e : pass_events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        pass_events.clear();
        pass_launches.clear();
        pass_items.clear();
        pass_bytes.clear();
    }

    void init(hipStream_t s, DeviceCache* c) {
        stream = s;
        cache = c;
        int dev = 0, ncu = 0;
        DCX_HIP(hipGetDevice(&dev));
        DCX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        cus = ncu > 0 ? (uint32_t)ncu : 1u;
        DCX_HIP(hipMalloc(&ticket, sizeof(unsigned long long)));
        DCX_HIP(hipMemsetAsync(ticket, 0, sizeof(unsigned long long), stream));
        DCX_HIP(hipMalloc(&scalars, 64 * sizeof(uint64_t)));
        DCX_HIP(hipMemsetAsync(scalars, 0, 64 * sizeof(uint64_t), stream));
        DCX_HIP(hipMalloc(&hist, 8 * RS_RADIX * sizeof(uint64_t)));
        DCX_HIP(hipMalloc(&seg_part, (size_t)RS_SEGS * RS_HIST_WPS * 8 * RS_SEG_BINS * sizeof(uint32_t)));
        DCX_HIP(hipMalloc(&seg_h2, 8 * RS_SEG_BINS * sizeof(uint64_t)));
        DCX_HIP(hipMalloc(&seg_h11, RS_SEGS * RS_SEG_BINS * sizeof(uint32_t)));
        DCX_HIP(hipMalloc(&seg_off, 8 * RS_SEG_BINS * sizeof(uint64_t)));
        DCX_HIP(hipMalloc(&seg_info, 8 * RS_SEG_INFO * sizeof(uint64_t)));
        DCX_HIP(hipMalloc(&seg_tick, 8 * RS_SEGS * sizeof(uint32_t)));
        DCX_HIP(hipMalloc(&scan_part, 2 * SC_BLOCKS * sizeof(uint64_t)));
        DCX_HIP(hipMalloc(&scan_prefix, SC_BLOCKS * sizeof(uint64_t)));
        DCX_HIP(hipHostMalloc((void**)&mailbox, (MAIL_WORDS + 1) * sizeof(uint64_t), hipHostMallocMapped | hipHostMallocCoherent));
        memset(mailbox, 0, (MAIL_WORDS + 1) * sizeof(uint64_t));
        DCX_HIP(hipHostGetDevicePointer((void**)&dev_mailbox, mailbox, 0));
        mail_seq = 0;
        ticket_host = 0;
    }
    void destroy() {
        if (sort_status) (void)hipFree(sort_status);
        if (scan_lo) (void)hipFree(scan_lo);
        if (scan_hi) (void)hipFree(scan_hi);
        if (ticket) (void)hipFree(ticket);
        if (scalars) (void)hipFree(scalars);
        if (hist) (void)hipFree(hist);
        for (void* q : {(void*)seg_part, (void*)seg_h2, (void*)seg_h11, (void*)seg_off, (void*)seg_info,
                        (void*)seg_tick})
            if (q) (void)hipFree(q);
        seg_part = seg_h11 = seg_tick = nullptr;
        seg_h2 = seg_off = seg_info = nullptr;
        if (scan_part) (void)hipFree(scan_part);
        if (scan_prefix) (void)hipFree(scan_prefix);
        scan_part = scan_prefix = nullptr;
        if (mailbox) {
            (void)hipStreamSynchronize(stream);
            (void)hipHostFree(mailbox);
        }
        sort_status = nullptr;
        scan_lo = scan_hi = nullptr;
        ticket = nullptr;
        scalars = nullptr;
        hist = nullptr;
        mailbox = dev_mailbox = nullptr;
    }

    static unsigned grid_for(uint64_t n, unsigned threads, unsigned max_blocks = 8192) {
        uint64_t b = (n + threads - 1) / threads;
        if (b == 0) b = 1;
        if (b > max_blocks) b = max_blocks;
        return (unsigned)b;
    }

    void fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
        if (!n) return;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, v, n);
    }
    void fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
        if (!n) return;
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, v, n);
    }
    void fill_u8(uint8_t* p, uint8_t v, uint64_t n) {
        if (!n) return;
        hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, v, n);
    }

    void ensure_sort_status(uint64_t tiles) {
        size_t words = tiles * RS_RADIX;
        if (words <= sort_status_words) return;
        if (no_alloc) throw Error(-3, "radix lookback status would have to grow in a call that must not allocate");
        if (sort_status) DCX_HIP(hipFree(sort_status));
        sort_status = nullptr;
        size_t cap = words + words / 4 + RS_RADIX;
        DCX_HIP(hipMalloc(&sort_status, cap * sizeof(uint64_t)));
        sort_status_words = cap;
        fill_u64(sort_status, 0, cap);
        sort_tag = 0;
    }
    void ensure_scan_status(uint64_t tiles) {
        if (tiles <= scan_tiles_cap) return;
        if (no_alloc) throw Error(-3, "scan lookback status would have to grow in a call that must not allocate");
        if (scan_lo) DCX_HIP(hipFree(scan_lo));
        if (scan_hi) DCX_HIP(hipFree(scan_hi));
        scan_lo = scan_hi = nullptr;
        size_t cap = tiles + tiles / 4 + 16;
        DCX_HIP(hipMalloc(&scan_lo, cap * sizeof(uint64_t)));
        DCX_HIP(hipMalloc(&scan_hi, cap * sizeof(uint64_t)));
        fill_u64(scan_lo, 0, cap);
        fill_u64(scan_hi, 0, cap);
        scan_tiles_cap = cap;
        scan_tag = 0;
    }

    // Key sources whose digits are byte windows of a character stream (histograms from windows)
    template <class Src>
    static bool window_aligned(const Src&, uint64_t, uint32_t, uint32_t) {
        return false;
    }
    static bool window_aligned(const StreamKeys<false>& src, uint64_t N, uint32_t lo_bit, uint32_t np) {
        return (8 % src.bits == 0) && lo_bit % 8 == 0 && lo_bit + 8 * np <= 64 &&
               (src.mask & (~0ull << lo_bit)) == (~0ull << lo_bit) && N > 64;
    }
    template <class Src>
    void hist_windows(const Src&, uint64_t, uint32_t, uint32_t, uint32_t, uint32_t) {}
    void hist_windows(const StreamKeys<false>& src, uint64_t N, uint32_t lo_bit, uint32_t np, uint32_t seg0_keys,
                      uint32_t wps) {
        hipLaunchKernelGGL(k_window_hist_seg, dim3(RS_SEGS * wps), dim3(RS_HIST_THREADS), 0, stream, src.s, src.bits, N, seg0_keys,
                           wps, seg_part);
        hipLaunchKernelGGL(k_window_seg_reduce, dim3(RS_SEG_BINS / 256, RS_SEGS), dim3(256), 0, stream,
                           (const uint32_t*)seg_part, wps, seg_h11);
        hipLaunchKernelGGL(k_window_seg_joint, dim3(np), dim3(256), 0, stream, src.s, src.bits, N, lo_bit, seg0_keys,
                           (const uint32_t*)seg_h11, seg_h2);
    }

    // Plans the segmented passes of a sort of N keys over bits [lo_bit, lo_bit + 8 np) in tiles of
    // `tile` keys from one read of the first-pass keys: hist, seg_off, seg_info, cleared seg_tick.
    // Returns the grid of every pass: one workgroup per tile of all segments (at most N / tile + 8).
    template <class Src>
    uint64_t seg_plan(const Src& first, uint64_t N, uint32_t lo_bit, uint32_t np, uint32_t tile) {
        const uint64_t tiles = ceil_div(N, tile);
        const uint64_t seg0_keys = ceil_div(tiles, (uint64_t)RS_SEGS) * tile;
        const uint64_t grid = tiles + RS_SEGS;
        ensure_sort_status(grid);
        uint64_t wps = ceil_div(seg0_keys, (uint64_t)RS_HIST_KEYS);
        if (wps > RS_HIST_WPS) wps = RS_HIST_WPS;
        const uint32_t bins = np * RS_SEG_BINS;
        if (window_aligned(first, N, lo_bit, np)) {
            hist_windows(first, N, lo_bit, np, (uint32_t)seg0_keys, (uint32_t)wps);
        } else {
            if constexpr (std::is_same<Src, CodedKeys>::value)
                if (np > CKH_PASSES) throw Error(-4, "prefix-code keys have at most 40 bits");
            if constexpr (std::is_same<Src, CodedKeys>::value)
                hipLaunchKernelGGL(k_coded_hist_seg, dim3((unsigned)(RS_SEGS * wps)), dim3(RS_HIST_THREADS), 0, stream,
                                   first, N, lo_bit, np, (uint32_t)seg0_keys, (uint32_t)wps, seg_part, seg_h2);
            else
                hipLaunchKernelGGL(k_radix_hist_seg<Src>, dim3((unsigned)(RS_SEGS * wps)), dim3(RS_HIST_THREADS), 0, stream,
                                   first, N, lo_bit, np, (uint32_t)seg0_keys, (uint32_t)wps, seg_part, seg_h2);
            hipLaunchKernelGGL(k_hist_seg_reduce, dim3(bins / 256, 16), dim3(256), 0, stream,
                               (const uint32_t*)seg_part, (uint32_t)(RS_SEGS * wps), bins, seg_h2);
        }
        hipLaunchKernelGGL(k_seg_setup, dim3(np), dim3(256), 0, stream, (const uint64_t*)seg_h2, N, tile,
                           (uint32_t)seg0_keys, hist, seg_off, seg_info, seg_tick);
        return grid;
    }

    // LSD radix sort of (u64 key, u32 value) pairs by key bits [lo_bit, hi_bit) (rounded up to whole
    // bytes); bits above hi_bit must be equal for all keys.  The first pass reads its keys from `first`
    // (ArrayKeys{keys}: the keys array itself); with ident_vals its values are the input indices (vals
    // is then only an output).  vals == nullptr sorts keys only.  On return keys/vals point at the
    // sorted data (the buffers may have been swapped with the alternates).
    //
    // low32 (optional, with values and at most five passes): the last pass writes the key bits below its
    // digit, (key >> lo_bit) truncated to 32 bits, to low32 instead of the keys (keys then holds the input
    // of that pass).  Two sorted keys are equal iff they lie in the same run of the last digit (the runs
    // start at the last pass's digit offsets, hist + (np - 1) * 256) and their low32 words are equal.
    // Returns the number of passes when low32 was written, else 0.
    //
    // idx_bytes (two arrays of N bytes, optional, with identity values and low32): when an index fits
    // the key bits below the digits plus one byte, the passes carry it split (low bits in the key, high
    // byte in idx_bytes: 9 instead of 12 bytes per element); the last pass joins it into vals.  The
    // passes read idx_bytes in aligned 16-byte chunks: both arrays must be 16-byte aligned and readable
    // up to N rounded up to 16 (DeviceCache blocks are).
    //
    // heads (optional, with values, when neither low32 nor split indices are used and there are at least
    // two passes): the last pass writes the group heads of the top key bits under hmask into heads[0, N)
    // (k_onesweep HEADS; 2 marks entries for k_heads_fix) and only the keys that fix needs; *heads_done
    // tells whether it did.
    template <class Src>
    uint32_t sort_pairs_from(const Src& first, uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals,
                         uint32_t*& vals_alt, uint64_t N, uint32_t lo_bit, uint32_t hi_bit, bool ident_vals,
                         uint32_t* low32 = nullptr, uint8_t* idx_bytes = nullptr, uint8_t* idx_bytes_alt = nullptr,
                         uint8_t* heads = nullptr, uint64_t hmask = 0, bool* heads_done = nullptr) {
        if (heads_done) *heads_done = false;
        if (hi_bit <= lo_bit) hi_bit = lo_bit + 1;
        if (N <= 1) {
            // nothing to order: materialize the inputs in place
            if (N == 1) hipLaunchKernelGGL(k_materialize<Src>, dim3(1), dim3(64), 0, stream, first, N, keys);
            if (N == 1 && ident_vals && vals) fill_u32(vals, 0, 1);
            return 0;
        }
        uint32_t np = (hi_bit - lo_bit + 7) / 8;
        if (np > 8) np = 8;
        if (lo_bit + 8 * np > 64) lo_bit = 64 - 8 * np;
        if (np > 5 || np < 2 || !vals) low32 = nullptr;  // the last pass must not be the first
        const uint32_t kpos = std::min<uint32_t>(lo_bit, 24);
        const bool split = low32 && ident_vals && idx_bytes && idx_bytes_alt && bit_width(N - 1) <= kpos + 8;
        if (low32 || split || np < 2 || !vals) heads = nullptr;
        const uint64_t tiles = seg_plan(first, N, lo_bit, np, RS_TILE);
        uint64_t span_bytes = 0;
        if (profile && !no_alloc) {  // the pair is listed at once, so drop_profile releases it whatever happens next
            hipEvent_t e0 = nullptr, e1 = nullptr;
            DCX_HIP(hipEventCreate(&e0));
            if (hipEventCreate(&e1) != hipSuccess) {
                (void)hipEventDestroy(e0);
                throw Error(-4, "hipEventCreate failed");
            }
            pass_events.emplace_back(e0, e1);
            pass_launches.push_back(0);
            pass_items.push_back(0);
            pass_bytes.push_back(0);
            DCX_HIP(hipEventRecord(e0, stream));
        }
        for (uint32_t p = 0; p < np; p++) {
            if (sort_tag >= LB_TAG_LIMIT) {
                fill_u64(sort_status, 0, sort_status_words);
                sort_tag = 0;
            }
            uint32_t tag = ++sort_tag;
            const uint32_t shift = lo_bit + 8 * p;
            const uint64_t* so = seg_off + p * RS_SEG_BINS;
            const uint64_t* si = seg_info + p * RS_SEG_INFO;
            uint32_t* tk = seg_tick + p * RS_SEGS;
            const dim3 g((unsigned)tiles), b(RS_THREADS);
            if (split) {
                const ArrayKeys<uint64_t> in{keys};
                if (p == 0)
                    hipLaunchKernelGGL((k_onesweep<true, true, Src, uint64_t, RS_ITEMS, uint64_t, uint8_t, uint8_t>), g, b, 0,
                                       stream, first, keys_alt, (const uint8_t*)nullptr, idx_bytes_alt, shift, so, si, tk,
                                       sort_status, tag, 0u, kpos);
                else if (p < np - 1)
                    hipLaunchKernelGGL((k_onesweep<true, false, ArrayKeys<uint64_t>, uint64_t, RS_ITEMS, uint64_t, uint8_t, uint8_t>),
                                       g, b, 0, stream, in, keys_alt, (const uint8_t*)idx_bytes, idx_bytes_alt, shift, so, si,
                                       tk, sort_status, tag, 0u, kpos);
                else
                    hipLaunchKernelGGL((k_onesweep<true, false, ArrayKeys<uint64_t>, uint64_t, RS_ITEMS, uint32_t, uint8_t, uint32_t>),
                                       g, b, 0, stream, in, low32, (const uint8_t*)idx_bytes, vals_alt, shift, so, si, tk,
                                       sort_status, tag, lo_bit, kpos);
            } else if (p == 0) {
                if (!vals)
                    hipLaunchKernelGGL((k_onesweep<false, false, Src>), g, b, 0, stream, first, keys_alt,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, shift, so, si, tk, sort_status, tag);
                else if (ident_vals)
                    hipLaunchKernelGGL((k_onesweep<true, true, Src>), g, b, 0, stream, first, keys_alt,
                                       (const uint32_t*)nullptr, vals_alt, shift, so, si, tk, sort_status, tag);
                else
                    hipLaunchKernelGGL((k_onesweep<true, false, Src>), g, b, 0, stream, first, keys_alt,
                                       (const uint32_t*)vals, vals_alt, shift, so, si, tk, sort_status, tag);
            } else if (low32 && p == np - 1 && vals) {
                const ArrayKeys<uint64_t> in{keys};
                hipLaunchKernelGGL((k_onesweep<true, false, ArrayKeys<uint64_t>, uint64_t, RS_ITEMS, uint32_t>), g, b, 0,
                                   stream, in, low32, (const uint32_t*)vals, vals_alt, shift, so, si, tk, sort_status, tag, lo_bit);
            } else if (heads && p == np - 1) {
                const ArrayKeys<uint64_t> in{keys};
                hipLaunchKernelGGL((k_onesweep<true, false, ArrayKeys<uint64_t>, uint64_t, RS_ITEMS, uint64_t, uint32_t,
                                               uint32_t, false, true>),
                                   g, b, 0, stream, in, keys_alt, (const uint32_t*)vals, vals_alt, shift, so, si, tk,
                                   sort_status, tag, 0u, 0u, 0ull, heads, hmask);
                if (heads_done) *heads_done = true;
            } else {
                const ArrayKeys<uint64_t> in{keys};
                if (!vals)
                    hipLaunchKernelGGL((k_onesweep<false, false, ArrayKeys<uint64_t>>), g, b, 0, stream, in, keys_alt,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, shift, so, si, tk, sort_status, tag, 0u);
                else
                    hipLaunchKernelGGL((k_onesweep<true, false, ArrayKeys<uint64_t>>), g, b, 0, stream, in, keys_alt,
                                       (const uint32_t*)vals, vals_alt, shift, so, si, tk, sort_status, tag, 0u);
            }
            if (profile) {
                // keys in (the first pass: its key source) and out (a low32 last pass: 4 B), values in
                // (not for identity values) and out
                const bool out32 = low32 && p == np - 1;
                const bool outh = heads && p == np - 1;  // a byte of heads instead of the key (the few run-end keys aside)
                uint64_t b = (p == 0 ? first.read_bytes(N) : 8ull * N) + (out32 ? 4ull : outh ? 1ull : 8ull) * N;
                if (split)  // index bytes in (not in the first pass) and out (the last pass: 4-byte indices)
                    b += (p == 0 ? 0ull : 1ull) * N + (p == np - 1 ? 4ull : 1ull) * N;
                else if (vals)
                    b += (p == 0 && ident_vals ? 4ull : 8ull) * N;
                span_bytes += b;
            }
            if (!(low32 && p == np - 1 && vals)) std::swap(keys, keys_alt);
            if (split) {
                if (p < np - 1) std::swap(idx_bytes, idx_bytes_alt);
                else std::swap(vals, vals_alt);
            } else if (vals) {
                std::swap(vals, vals_alt);
            }
        }
        if (profile && !no_alloc) {
            DCX_HIP(hipEventRecord(pass_events.back().second, stream));
            pass_launches.back() = np;
            pass_items.back() = (uint64_t)N * np;
            pass_bytes.back() = span_bytes;
        }
        return low32 ? np : 0;
    }

    // the same with the keys array as the first pass's key source
    void sort_pairs(uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals, uint32_t*& vals_alt, uint64_t N,
                    uint32_t lo_bit, uint32_t hi_bit) {
        sort_pairs_from(ArrayKeys<uint64_t>{keys}, keys, keys_alt, vals, vals_alt, N, lo_bit, hi_bit, false);
    }

    // Stable partition of N >= 2 u64 keys by bits [lo_bit, hi_bit) (lo_bit >= 32, higher bits equal): LSD
    // passes over the keys alone, the last of which writes only each key's low word, to out32[0, out_limit)
    // (the keys whose sorted index is out_limit or more are dropped).  keys and keys_alt are both
    // overwritten.  The bucket id rides in the high word, the payload in the low word: 16 bytes per
    // element in a middle pass, 12 in the last, instead of 24 for (key, value) pairs.
    void sort_keys_low32(uint64_t* keys, uint64_t* keys_alt, uint64_t N, uint32_t lo_bit, uint32_t hi_bit,
                         uint32_t* out32, uint64_t out_limit) {
        if (N < 2 || lo_bit < 32 || hi_bit > 64) throw Error(-6, "sort_keys_low32: bad arguments");
        if (hi_bit <= lo_bit) hi_bit = lo_bit + 1;
        uint32_t np = (hi_bit - lo_bit + 7) / 8;
        if (lo_bit + 8 * np > 64) lo_bit = 64 - 8 * np;
        const uint64_t tiles = seg_plan(ArrayKeys<uint64_t>{keys}, N, lo_bit, np, RS_TILE);
        uint64_t span_bytes = 0;
        if (profile && !no_alloc) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            DCX_HIP(hipEventCreate(&e0));
            if (hipEventCreate(&e1) != hipSuccess) {
                (void)hipEventDestroy(e0);
                throw Error(-4, "hipEventCreate failed");
            }
            pass_events.emplace_back(e0, e1);
            pass_launches.push_back(0);
            pass_items.push_back(0);
            pass_bytes.push_back(0);
            DCX_HIP(hipEventRecord(e0, stream));
        }
        for (uint32_t p = 0; p < np; p++) {
            if (sort_tag >= LB_TAG_LIMIT) {
                fill_u64(sort_status, 0, sort_status_words);
                sort_tag = 0;
            }
            const uint32_t tag = ++sort_tag;
            const uint32_t shift = lo_bit + 8 * p;
            const uint64_t* so = seg_off + p * RS_SEG_BINS;
            const uint64_t* si = seg_info + p * RS_SEG_INFO;
            uint32_t* tk = seg_tick + p * RS_SEGS;
            const dim3 g((unsigned)tiles), b(RS_THREADS);
            const ArrayKeys<uint64_t> in{keys};
            if (p + 1 < np) {
                hipLaunchKernelGGL((k_onesweep<false, false, ArrayKeys<uint64_t>>), g, b, 0, stream, in, keys_alt,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, shift, so, si, tk, sort_status, tag, 0u);
                std::swap(keys, keys_alt);
                span_bytes += 16ull * N;
            } else {
                hipLaunchKernelGGL((k_onesweep<false, false, ArrayKeys<uint64_t>, uint64_t, RS_ITEMS, uint32_t, uint32_t,
                                               uint32_t, true>),
                                   g, b, 0, stream, in, out32, (const uint32_t*)nullptr, (uint32_t*)nullptr, shift, so,
                                   si, tk, sort_status, tag, 0u, 0u, out_limit);
                span_bytes += 8ull * N + 4ull * std::min<uint64_t>(N, out_limit);
            }
        }
        if (profile && !no_alloc) {
            DCX_HIP(hipEventRecord(pass_events.back().second, stream));
            pass_launches.back() = np;
            pass_items.back() = (uint64_t)N * np;
            pass_bytes.back() = span_bytes;
        }
    }

    // Stable LSD radix sort of (u32 key, u32 value) pairs by key bits [lo_bit, lo_bit + 8 * np) (np <= 4).
    // The input arrays are left alone; the result ends in *keys_out / *vals_out, which point to either
    // (k1, v1) or (k2, v2).  vals == nullptr: the values are the input indices.
    void sort_u32(const uint32_t* keys, const uint32_t* vals, uint32_t* k1, uint32_t* v1, uint32_t* k2, uint32_t* v2,
                  uint64_t N, uint32_t lo_bit, uint32_t np, uint32_t** keys_out, uint32_t** vals_out) {
        const uint32_t* kin = keys;
        const uint32_t* vin = vals;
        uint32_t* ko = k1;
        uint32_t* vo = v1;
        *keys_out = k1;
        *vals_out = v1;
        if (N == 0 || np == 0) return;
        const uint64_t tiles = seg_plan(ArrayKeys<uint32_t>{keys}, N, lo_bit, np, RS_THREADS * RS_ITEMS_U32);
        for (uint32_t p = 0; p < np; p++) {
            if (sort_tag >= LB_TAG_LIMIT) {
                fill_u64(sort_status, 0, sort_status_words);
                sort_tag = 0;
            }
            const uint32_t tag = ++sort_tag;
            const ArrayKeys<uint32_t> in{kin};
            if (vin)
                hipLaunchKernelGGL((k_onesweep<true, false, ArrayKeys<uint32_t>, uint32_t, RS_ITEMS_U32>), dim3((unsigned)tiles),
                                   dim3(RS_THREADS), 0, stream, in, ko, vin, vo, lo_bit + 8 * p,
                                   seg_off + p * RS_SEG_BINS, seg_info + p * RS_SEG_INFO, seg_tick + p * RS_SEGS,
                                   sort_status, tag);
            else
                hipLaunchKernelGGL((k_onesweep<true, true, ArrayKeys<uint32_t>, uint32_t, RS_ITEMS_U32>), dim3((unsigned)tiles),
                                   dim3(RS_THREADS), 0, stream, in, ko, (const uint32_t*)nullptr, vo, lo_bit + 8 * p,
                                   seg_off + p * RS_SEG_BINS, seg_info + p * RS_SEG_INFO, seg_tick + p * RS_SEGS,
                                   sort_status, tag);
            *keys_out = ko;
            *vals_out = vo;
            kin = ko;
            vin = vo;
            ko = (ko == k1) ? k2 : k1;
            vo = (vo == v1) ? v2 : v1;
        }
    }

    // Launches a scan; the grand totals of both 32-bit lanes land in device scalar slot `slot` (read with
    // totals(slot)).
    // Blocks of the two-pass scan of Op. Every block scans the same number of tiles, so a round of
    // resident blocks that is only partly filled takes as long as a full one: the count is a whole number
    // of rounds of both kernels (the reduce kernel usually holds more blocks per compute unit than the
    // down kernel; m = d * floor(r / d) blocks per unit is one round of the first, m / d of the second).
    template <class Op>
    uint32_t scan_blocks(uint64_t tiles) const {
        static const uint32_t per_cu = [] {
            int r = 0, d = 0;
            DCX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, k_scan_reduce<Op>, SC_THREADS, 0));
            DCX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&d, k_scan_down<Op>, SC_THREADS, 0));
            r = std::max(r, 1);
            d = std::max(d, 1);
            return (uint32_t)(r >= d ? d * (r / d) : r * (d / r));
        }();
        const uint64_t want = std::min<uint64_t>(SC_BLOCKS, (uint64_t)per_cu * cus);
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, want));
    }

    // whether scan() over N elements takes the two passes (reduce, down) rather than the single pass
    static bool scan_two_pass(uint64_t N) { return ceil_div(N, (uint64_t)SC_TILE) > SC_SINGLE_PASS_TILES; }
    template <class Op>
    void scan(const Op& op, uint64_t N, int slot) {
        if (N == 0) {
            fill_u64(scalars + 2 * slot, 0, 2);
            return;
        }
        uint64_t tiles = ceil_div(N, SC_TILE);
        if (tiles > SC_SINGLE_PASS_TILES) {
            const uint32_t tpb = (uint32_t)ceil_div(tiles, (uint64_t)scan_blocks<Op>(tiles));
            const uint32_t nb = (uint32_t)ceil_div(tiles, (uint64_t)tpb);
            hipLaunchKernelGGL(k_scan_reduce<Op>, dim3(nb), dim3(SC_THREADS), 0, stream, op, N, tpb, scan_part);
            hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(1024), 0, stream, (const uint64_t*)scan_part, nb,
                               scan_prefix, scalars + 2 * slot);
            hipLaunchKernelGGL(k_scan_down<Op>, dim3(nb), dim3(SC_THREADS), 0, stream, op, N, tpb,
                               (const uint64_t*)scan_prefix);
            return;
        }
        ensure_scan_status(tiles);
        if (scan_tag >= LB_TAG_LIMIT) {
            fill_u64(scan_lo, 0, scan_tiles_cap);
            fill_u64(scan_hi, 0, scan_tiles_cap);
            scan_tag = 0;
        }
        uint32_t tag = ++scan_tag;
        hipLaunchKernelGGL(k_scan<Op>, dim3((unsigned)tiles), dim3(SC_THREADS), 0, stream, op, N, scan_lo, scan_hi,
                           tag, ticket, ticket_host, scalars + 2 * slot);
        ticket_host += tiles;
    }

    // The grand total of a scan only (into scalar slot `slot`): long inputs skip the down pass.
    template <class Op>
    void total(const Op& op, uint64_t N, int slot) {
        const uint64_t tiles = ceil_div(N, SC_TILE);
        if (N == 0 || tiles <= SC_SINGLE_PASS_TILES) {
            scan(op, N, slot);
            return;
        }
        const uint32_t tpb = (uint32_t)ceil_div(tiles, (uint64_t)scan_blocks<Op>(tiles));
        const uint32_t nb = (uint32_t)ceil_div(tiles, (uint64_t)tpb);
        hipLaunchKernelGGL(k_scan_reduce<Op>, dim3(nb), dim3(SC_THREADS), 0, stream, op, N, tpb, scan_part);
        hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(1024), 0, stream, (const uint64_t*)scan_part, nb, scan_prefix,
                           scalars + 2 * slot);
    }

    // scan() right after total() over the same N and the same values (ops of one type that differ only in
    // where they write), with no scan in between: the block prefixes total() left in scan_prefix are
    // reused, so only the down pass runs (short inputs take the single pass again)
    template <class Op>
    void scan_after_total(const Op& op, uint64_t N, int slot) {
        const uint64_t tiles = ceil_div(N, SC_TILE);
        if (N == 0 || tiles <= SC_SINGLE_PASS_TILES) {
            scan(op, N, slot);
            return;
        }
        const uint32_t tpb = (uint32_t)ceil_div(tiles, (uint64_t)scan_blocks<Op>(tiles));
        const uint32_t nb = (uint32_t)ceil_div(tiles, (uint64_t)tpb);
        hipLaunchKernelGGL(k_scan_down<Op>, dim3(nb), dim3(SC_THREADS), 0, stream, op, N, tpb,
                           (const uint64_t*)scan_prefix);
    }

    // `bytes` (at most 8 * MAIL_WORDS) of device memory at `src` (8-byte aligned, readable up to the next
    // multiple of 8), ordered after the work on the stream: published by k_publish into the mailbox and
    // awaited by spinning on its sequence number.  The stream is queried now and then, so that a failed
    // stream raises an error instead of a wait that never ends.  The result stays valid until the next call.
    const uint64_t* read_device(const void* src, size_t bytes) {
        const uint32_t words = (uint32_t)((bytes + 7) / 8);
        if (words > MAIL_WORDS) throw Error(-6, "read_device: at most 4 KiB per read");
        const uint64_t seq = ++mail_seq;
        hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, static_cast<const uint64_t*>(src), words,
                           dev_mailbox, seq);
        for (uint64_t spin = 1;; spin++) {
            if (__atomic_load_n(&mailbox[0], __ATOMIC_ACQUIRE) == seq) return mailbox + 1;
            if ((spin & 1023) == 0) {
                const hipError_t e = hipStreamQuery(stream);
                if (e == hipSuccess) {  // drained: the kernel has run and its stores are visible
                    if (__atomic_load_n(&mailbox[0], __ATOMIC_ACQUIRE) == seq) return mailbox + 1;
                    throw Error(-4, "read_device: the stream drained without publishing");
                }
                if (e != hipErrorNotReady) {
                    (void)hipGetLastError();
                    throw Error(-4, std::string("read_device: ") + hipGetErrorString(e));
                }
            }
        }
    }
    // Copies `count` device scalars to the host (ordered after the work on the stream).
    const uint64_t* read_scalars(int count) { return read_device(scalars, count * sizeof(uint64_t)); }
    // The lane totals of the last scan into slot `slot`: lo = sum of the low 32-bit halves of its values,
    // hi = of the high halves (synchronizes the stream).
    struct Totals {
        uint64_t lo, hi;
    };
    Totals totals(int slot) {
        const uint64_t* h = read_scalars(2 * slot + 2);
        return {h[2 * slot], h[2 * slot + 1]};
    }
};

}  // namespace dcx
