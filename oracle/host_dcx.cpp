// host_dcx.cpp -- multi-threaded CPU restatement of lightweight_dcx's PDCX (the "host DCX" baseline).
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the GPU engine, and a second CPU
// path the tests check against libsais.  The product path (the HIP engine) never loads it.
//
// lightweight_dcx runs PDCX as p MPI processes (mpiexec -n p ./build/cli ..., README.md:26-29); it cannot be
// built here (its FetchContent dependencies are absent), so this file restates the same phases for one
// process with T threads, with the data layout of the reference's records:
//   * compress_alphabet (cli.cpp:417-463) and the packed variant (cli.cpp:475-539): characters 1..sigma,
//     b = 3/5/8 bits per packed character (9 for sigma = 256, which the reference refuses);
//   * phase 1 (sample_string.hpp:145-164, 225-239): one record per difference-cover sample, the packed
//     characters of its key (KPackedInteger words, char_container.hpp:483-545, left-justified as
//     PackedInteger, :57-71) and its index, sorted by key with a sample sort (sample_sort.hpp:37-171: a
//     sample of splitters, every thread classifies its share into buckets, the buckets are sorted on their
//     own) instead of the distributed sorter;
//   * phase 2 (compute_ranks.hpp:88-134): names by a scan over key changes, unique flags;
//   * phase 3 (pdcx.hpp:145-346): names distinct -> ranks; otherwise the recursive string in (mod X, div X)
//     order, the discarding decision (num_ranks_after_discarding, :215-228), the recursive call, the ranks
//     by map_back (:83-94) or by the (rank1, rank2) sort of recursive_call_with_discarding (:297-323);
//   * phase 4 (merge_samples.hpp:44-123, 237-260): one record per suffix with its first ratio * X - 1
//     packed characters, sorted by the characters and then by the DC comparator on the sample ranks
//     (operator<, :87-109; cmpDepthRanks from the closed form of B_DC_3_7_13/tools/calc_dcparam.cc).
//     The records carry the suffix index and read the D ranks from the text-ordered rank array
//     (rank_pos, :181-188) rather than copying them into every record.
// The base case (n <= max(2X, base_case_chars), pdcx.hpp:623) sorts the suffixes directly
// (slow_suffixarray, string_util.hpp:87-106).
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace {

const uint32_t kDcSets[][14] = {
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

struct Dc {
    uint32_t X = 0, D = 0;
    uint32_t dc[16] = {};
    uint8_t in_dc[256] = {};
    uint32_t next_rank[256] = {};
    std::vector<uint32_t> cmp;  // X * X entries: r1 | r2 << 16 (cmpDepthRanks without the depth)
    bool load(uint32_t x) {
        int f = -1;
        for (int k = 0; k < (int)(sizeof(kDcSets) / sizeof(kDcSets[0])); k++)
            if (kDcSets[k][0] == x) f = k;
        if (f < 0) return false;
        X = x;
        D = kDcSets[f][1];
        for (uint32_t i = 0; i < D; i++) {
            dc[i] = kDcSets[f][2 + i];
            in_dc[dc[i]] = 1;
        }
        uint32_t c = 0;
        for (uint32_t r = 0; r < X; r++) {
            next_rank[r] = c;
            c += in_dc[r];
        }
        cmp.assign((size_t)X * X, 0);
        for (uint32_t a = 0; a < X; a++)
            for (uint32_t b = 0; b < X; b++) {
                uint32_t j = 0, r1 = 0, r2 = 0;
                while (j < X && (!in_dc[(a + j) % X] || !in_dc[(b + j) % X])) {
                    r1 += in_dc[(a + j) % X];
                    r2 += in_dc[(b + j) % X];
                    j++;
                }
                cmp[a * X + b] = r1 | (r2 << 16);
            }
        return true;
    }
    uint64_t rank_pos(uint64_t i) const { return (i / X) * D + next_rank[i % X]; }
};

// f(lo, hi, t) on T contiguous shares [n t / T, n (t + 1) / T) of [0, n), one thread each
template <class F>
void pfor(uint64_t n, unsigned T, F f) {
    if (T <= 1 || n < (1u << 15)) {
        f(0, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (unsigned t = 0; t < T; t++) th.emplace_back([&, t]() { f(n * t / T, n * (t + 1) / T, t); });
    for (auto& x : th) x.join();
}

// Sample sort with T threads (the structure of sample_sort.hpp:37-171 on one node): splitters from a
// sorted sample, each thread's share classified into buckets and scattered bucket-major, then the buckets
// sorted independently, largest share of work handed out first.
template <class Rec, class Less>
void psort(std::vector<Rec>& a, Less less, unsigned T) {
    const uint64_t n = a.size();
    if (T <= 1 || n < (1u << 16)) {
        std::sort(a.begin(), a.end(), less);
        return;
    }
    const uint32_t B = 16 * T;  // buckets
    const uint32_t over = 32;  // sample per bucket
    std::vector<Rec> sample;
    sample.reserve((size_t)B * over);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (uint32_t s = 0; s < B * over; s++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        sample.push_back(a[x % n]);
    }
    std::sort(sample.begin(), sample.end(), less);
    std::vector<Rec> spl;
    spl.reserve(B - 1);
    for (uint32_t b = 1; b < B; b++) spl.push_back(sample[(size_t)b * over]);
    std::vector<uint16_t> bid(n);
    std::vector<uint64_t> cnt((size_t)T * B, 0);
    pfor(n, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
        uint64_t* c = &cnt[(size_t)t * B];
        for (uint64_t i = lo; i < hi; i++) {
            const uint32_t b = (uint32_t)(std::upper_bound(spl.begin(), spl.end(), a[i], less) - spl.begin());
            bid[i] = (uint16_t)b;
            c[b]++;
        }
    });
    std::vector<uint64_t> off((size_t)T * B), start(B + 1);
    uint64_t s = 0;
    for (uint32_t b = 0; b < B; b++) {
        start[b] = s;
        for (unsigned t = 0; t < T; t++) {
            off[(size_t)t * B + b] = s;
            s += cnt[(size_t)t * B + b];
        }
    }
    start[B] = s;
    std::vector<Rec> tmp(n);
    pfor(n, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
        uint64_t* o = &off[(size_t)t * B];
        for (uint64_t i = lo; i < hi; i++) tmp[o[bid[i]]++] = a[i];
    });
    std::vector<uint32_t> order(B);
    for (uint32_t b = 0; b < B; b++) order[b] = b;
    std::sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) {
        return start[p + 1] - start[p] > start[q + 1] - start[q];
    });
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&]() {
            for (uint32_t k; (k = next.fetch_add(1)) < B;) {
                const uint32_t b = order[k];
                std::sort(tmp.begin() + start[b], tmp.begin() + start[b + 1], less);
            }
        });
    for (auto& y : th) y.join();
    a.swap(tmp);
}

// key words: the first L characters of the suffix at i, b bits each, left-justified per word (characters at
// or past n read as 0, the padding of sample_string.hpp:112-117)
template <class C>
inline void pack_key(const C* T, uint64_t n, uint64_t i, uint32_t L, uint32_t b, uint64_t* w, uint32_t K) {
    const uint32_t cpw = 64 / b;
    uint32_t c = 0;
    for (uint32_t k = 0; k < K; k++) {
        uint64_t x = 0;
        for (uint32_t j = 0; j < cpw && c < L; j++, c++) {
            const uint64_t ch = i + c < n ? (uint64_t)T[i + c] : 0;
            x |= ch << (64 - (j + 1) * b);
        }
        w[k] = x;
    }
}

// the key of suffix i + 1 from the key of suffix i: every word moves up one character, takes the first
// character of the next word into its last slot, and the last word takes character i + L into slot
// (L - 1) mod cpw (the slots after it, and the bits below the last slot, stay zero)
template <class C>
inline void slide_key(const C* T, uint64_t n, uint64_t i, uint32_t L, uint32_t b, uint64_t* w, uint32_t K) {
    const uint32_t cpw = 64 / b;
    const uint32_t last = 64 - cpw * b;  // lowest bit of a word's last slot
    for (uint32_t k = 0; k + 1 < K; k++) w[k] = (w[k] << b) | ((w[k + 1] >> (64 - b)) << last);
    const uint32_t r = L - (K - 1) * cpw;  // characters in the last word
    const uint64_t ch = i + L < n ? (uint64_t)T[i + L] : 0;
    w[K - 1] = (w[K - 1] << b) | (ch << (64 - r * b));
}

template <int K>
struct Rec {
    uint64_t w[K];
    uint64_t idx;
    Rec() {}  // left uninitialized: record arrays are filled right after their allocation
};
template <int K>
struct KeyLess {
    bool operator()(const Rec<K>& a, const Rec<K>& b) const {
        for (int k = 0; k < K; k++)
            if (a.w[k] != b.w[k]) return a.w[k] < b.w[k];
        return false;
    }
};
// DCMergeSamples::operator< (merge_samples.hpp:87-109)
template <int K>
struct MergeLess {
    const Dc* dc;
    const uint64_t* ranks;
    bool operator()(const Rec<K>& a, const Rec<K>& b) const {
        for (int k = 0; k < K; k++)
            if (a.w[k] != b.w[k]) return a.w[k] < b.w[k];
        const uint32_t e = dc->cmp[(a.idx % dc->X) * dc->X + (b.idx % dc->X)];
        return ranks[dc->rank_pos(a.idx) + (e & 0xFFFF)] < ranks[dc->rank_pos(b.idx) + (e >> 16)];
    }
};

struct Ctx {
    Dc dc;
    bool pack_samples = true, pack_merging = true;
    double thr = 0.7;
    uint64_t base_case = 10000;
    uint32_t sigma = 0;
    unsigned T = 1;
};

uint32_t bit_width(uint64_t v) {
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

// slow_suffixarray (string_util.hpp:87-106)
template <class C>
void base_sa(const C* T, uint64_t n, uint64_t* SA) {
    for (uint64_t i = 0; i < n; i++) SA[i] = i;
    std::sort(SA, SA + n, [&](uint64_t a, uint64_t b) {
        while (a < n && b < n && T[a] == T[b]) {
            a++;
            b++;
        }
        if (a == n) return b != n;
        if (b == n) return false;
        return T[a] < T[b];
    });
}

template <class C>
void level(Ctx& cx, const C* T, uint64_t n, uint64_t maxchar, int depth, uint64_t* SA);

// Phase 1 + 2 with K key words: the samples sorted by key, their names (in that order) and unique flags
template <int K, class C>
void sort_samples(Ctx& cx, const C* T, uint64_t n, const std::vector<uint64_t>& sb, uint32_t L, uint32_t b,
                  std::vector<uint64_t>& s_pos, std::vector<uint64_t>& names, std::vector<uint8_t>& uniq,
                  uint64_t& M) {
    const Dc& dc = cx.dc;
    const uint64_t m = sb[dc.D];
    std::vector<Rec<K>> r(m);
    // text order: sample k is (k div D) * X + DC[k mod D]
    pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
        for (uint64_t k = lo; k < hi; k++) {
            const uint64_t p = (k / dc.D) * dc.X + dc.dc[k % dc.D];
            pack_key(T, n, p, L, b, r[k].w, K);
            r[k].idx = p;
        }
    });
    psort(r, KeyLess<K>(), cx.T);
    s_pos.resize(m);
    names.resize(m);
    uniq.resize(m);
    // names = 1 + distinct keys before (an exclusive scan of key changes over the threads' shares)
    std::vector<uint64_t> part(cx.T + 1, 0);
    auto diff = [&](uint64_t j) {
        return j > 0 && KeyLess<K>()(r[j - 1], r[j]);
    };
    pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned t) {
        uint64_t c = 0;
        for (uint64_t j = lo; j < hi; j++) c += diff(j);
        part[t + 1] = c;
    });
    const unsigned parts = m < (1u << 15) ? 1 : cx.T;
    for (unsigned t = 0; t < parts; t++) part[t + 1] += part[t];
    pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned t) {
        uint64_t name = 1 + part[t];
        for (uint64_t j = lo; j < hi; j++) {
            name += diff(j);
            names[j] = name;
            s_pos[j] = r[j].idx;
        }
    });
    M = m ? names[m - 1] : 0;
    pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
        for (uint64_t j = lo; j < hi; j++)
            uniq[j] = !((j > 0 && names[j - 1] == names[j]) || (j + 1 < m && names[j + 1] == names[j]));
    });
}

// Phase 4 with K character words: every suffix ordered by its characters, then the DC comparator
template <int K, class C>
void merge(Ctx& cx, const C* T, uint64_t n, uint32_t L, uint32_t b, const uint64_t* ranks, uint64_t* SA) {
    std::vector<Rec<K>> r(n);
    const uint32_t KL = std::min<uint32_t>(K, (L + 64 / b - 1) / (64 / b));  // words holding the L characters
    pfor(n, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
        if (lo >= hi) return;
        uint64_t w[K];
        pack_key(T, n, lo, L, b, w, K);
        for (uint64_t i = lo; i < hi; i++) {
            if (i > lo) slide_key(T, n, i - 1, L, b, w, KL);
            for (int k = 0; k < K; k++) r[i].w[k] = w[k];
            r[i].idx = i;
        }
    });
    psort(r, MergeLess<K>{&cx.dc, ranks}, cx.T);
    pfor(n, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
        for (uint64_t i = lo; i < hi; i++) SA[i] = r[i].idx;
    });
}

// key word counts the records are instantiated for (a key of K words goes to the next one up; the extra
// words are zero in every record)
#define DCX_HOST_KS(F) F(1) F(2) F(3) F(4) F(5) F(6) F(8) F(10) F(12) F(17) F(20) F(24) F(34) F(48) F(68)

template <class C>
bool sort_samples_k(uint32_t K, Ctx& cx, const C* T, uint64_t n, const std::vector<uint64_t>& sb, uint32_t L,
                    uint32_t b, std::vector<uint64_t>& s_pos, std::vector<uint64_t>& names, std::vector<uint8_t>& uniq,
                    uint64_t& M) {
#define F(k)                                                          \
    if (K <= k) {                                                     \
        sort_samples<k>(cx, T, n, sb, L, b, s_pos, names, uniq, M);   \
        return true;                                                  \
    }
    DCX_HOST_KS(F)
#undef F
    return false;
}
template <class C>
bool merge_k(uint32_t K, Ctx& cx, const C* T, uint64_t n, uint32_t L, uint32_t b, const uint64_t* ranks, uint64_t* SA) {
#define F(k)                                       \
    if (K <= k) {                                  \
        merge<k>(cx, T, n, L, b, ranks, SA);       \
        return true;                               \
    }
    DCX_HOST_KS(F)
#undef F
    return false;
}

struct Rri {
    uint64_t r1, r2, idx;
};

template <class C>
void level(Ctx& cx, const C* T, uint64_t n, uint64_t maxchar, int depth, uint64_t* SA) {
    const Dc& dc = cx.dc;
    const uint32_t X = dc.X, D = dc.D;
    const uint64_t threshold = cx.base_case > 2ull * X ? cx.base_case : 2ull * X;
    if (n <= threshold) {
        base_sa(T, n, SA);
        return;
    }
    // compute_length_info (pdcx.hpp:473-510): samples i in [0, n] with i mod X in DC
    std::vector<uint64_t> sb(D + 1, 0);
    for (uint32_t t = 0; t < D; t++) sb[t + 1] = sb[t] + (dc.dc[t] <= n ? (n - dc.dc[t]) / X + 1 : 0);
    const uint64_t m = sb[D];
    // key lengths and widths (cli.cpp:475-539, materialize_sample / materialize_characters)
    uint32_t b = std::max<uint32_t>(1, bit_width(maxchar));
    double ratio = 1.0;
    uint32_t pb = 0;
    if (depth == 0 && (cx.pack_samples || cx.pack_merging)) {
        pb = cx.sigma <= 7 ? 3 : cx.sigma <= 31 ? 5 : cx.sigma <= 255 ? 8 : 9;
        const uint32_t cpw = 64 / pb, w = (X + cpw - 1) / cpw;
        ratio = (double)(w * cpw) / (double)X;
    }
    const uint32_t Ls = (uint32_t)((double)X * (cx.pack_samples && depth == 0 ? ratio : 1.0));
    const uint32_t Lm = (uint32_t)((double)X * (cx.pack_merging && depth == 0 ? ratio : 1.0)) - 1;
    const uint32_t bs = (cx.pack_samples && depth == 0) ? pb : b, bm = (cx.pack_merging && depth == 0) ? pb : b;
    const uint32_t Ks = (Ls + 64 / bs - 1) / (64 / bs), Km = (Lm + 64 / bm - 1) / (64 / bm);

    std::vector<uint64_t> s_pos, names;
    std::vector<uint8_t> uniq;
    uint64_t M = 0;
    if (!sort_samples_k(Ks, cx, T, n, sb, Ls, bs, s_pos, names, uniq, M)) throw std::bad_alloc();
    std::vector<uint64_t> ranks(m + D, 0);
    auto class_of = [&](uint64_t p) {
        uint32_t t = 0;
        while (dc.dc[t] != p % X) t++;
        return t;
    };
    if (M >= m) {
        pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t j = lo; j < hi; j++) ranks[dc.rank_pos(s_pos[j])] = names[j];
        });
    } else {
        // handle_recursive_call (pdcx.hpp:145-180): names in (mod X, div X) order
        std::vector<uint64_t> R(m);
        std::vector<uint8_t> U(m);
        pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t j = lo; j < hi; j++) {
                const uint64_t p = s_pos[j];
                const uint64_t cm = sb[class_of(p)] + p / X;
                R[cm] = names[j];
                U[cm] = uniq[j];
            }
        });
        std::vector<uint64_t>().swap(s_pos);
        std::vector<uint64_t>().swap(names);
        auto map_back = [&](uint64_t idx) {
            uint32_t t = D - 1;
            while (idx < sb[t]) t--;
            return (uint64_t)X * (idx - sb[t]) + dc.dc[t];
        };
        uint64_t discarded = 0;
        for (uint64_t k = 1; k < m; k++) discarded += (U[k] && U[k - 1]);
        const uint64_t after = m - discarded;
        const bool use_disc = n > 80 && (double)after / (double)m <= cx.thr;
        if (!use_disc) {
            // recursive_call_direct (pdcx.hpp:182-212): rank = 1 + position in the recursive SA
            std::vector<uint32_t> R32(R.begin(), R.end());
            std::vector<uint64_t>().swap(R);
            std::vector<uint64_t> SAr(m);
            if (n <= 80)
                base_sa(R32.data(), m, SAr.data());
            else
                level<uint32_t>(cx, R32.data(), m, M, depth + 1, SAr.data());
            pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t k = lo; k < hi; k++) ranks[dc.rank_pos(map_back(SAr[k]))] = k + 1;
            });
        } else {
            // recursive_call_with_discarding (pdcx.hpp:231-346)
            std::vector<uint32_t> Rr;
            Rr.reserve(after);
            std::vector<uint64_t> red(m);
            for (uint64_t k = 0; k < m; k++) {
                red[k] = Rr.size();
                if (k == 0 || !(U[k] && U[k - 1])) Rr.push_back((uint32_t)R[k]);
            }
            std::vector<uint64_t> SAr(after), ISAr(after);
            level<uint32_t>(cx, Rr.data(), after, M, depth + 1, SAr.data());
            std::vector<uint32_t>().swap(Rr);
            pfor(after, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t k = lo; k < hi; k++) ISAr[SAr[k]] = k + 1;
            });
            std::vector<Rri> rri(m);
            pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t k = lo; k < hi; k++) rri[k] = {R[k], U[k] ? 0 : ISAr[red[k]], map_back(k)};
            });
            psort(rri, [](const Rri& a, const Rri& c) { return a.r1 != c.r1 ? a.r1 < c.r1 : a.r2 < c.r2; }, cx.T);
            pfor(m, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t k = lo; k < hi; k++) ranks[dc.rank_pos(rri[k].idx)] = k + 1;
            });
        }
    }
    if (!merge_k(Km, cx, T, n, Lm, bm, ranks.data(), SA)) throw std::bad_alloc();
}

}  // namespace

extern "C" {

// PDCX::compute_sa behind compress_alphabet with T threads.  Returns 0, -1 for an unsupported X, -3 when
// memory runs out.
int host_dcx_sa(const uint8_t* text, uint64_t n, uint32_t X, int pack_samples, int pack_merging,
                double discarding_threshold, uint64_t base_case_chars, uint32_t threads, uint64_t* SA) {
    Ctx cx;
    if (!cx.dc.load(X)) return -1;
    cx.pack_samples = pack_samples != 0;
    cx.pack_merging = pack_merging != 0;
    cx.thr = discarding_threshold;
    cx.base_case = base_case_chars;
    cx.T = threads ? threads : 1;
    if (n == 0) return 0;
    try {
        // compress_alphabet (cli.cpp:417-463): used bytes -> 1..sigma in byte order
        std::vector<uint64_t> cnt((size_t)cx.T * 256, 0);
        pfor(n, cx.T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            uint64_t* c = &cnt[(size_t)t * 256];
            for (uint64_t i = lo; i < hi; i++) c[text[i]]++;
        });
        uint16_t map[256];
        uint32_t next = 1;
        for (int c = 0; c < 256; c++) {
            uint64_t s = 0;
            for (unsigned t = 0; t < cx.T; t++) s += cnt[(size_t)t * 256 + c];
            map[c] = s ? (uint16_t)next++ : 0;
        }
        cx.sigma = next - 1;
        std::vector<uint16_t> T(n);
        pfor(n, cx.T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t i = lo; i < hi; i++) T[i] = map[text[i]];
        });
        level<uint16_t>(cx, T.data(), n, cx.sigma, 0, SA);
    } catch (const std::bad_alloc&) {
        return -3;
    }
    return 0;
}

}  // extern "C"
