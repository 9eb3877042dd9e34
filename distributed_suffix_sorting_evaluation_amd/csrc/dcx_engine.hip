// dcx_engine.hip -- MI355X DCX suffix-array engine: level driver and the C ABI of include/dcx.h.
//
// One recursion level (PDCX::compute_sa, lightweight_dcx/src/pdcx/pdcx.hpp:529-941) runs as:
//   phase 1+2  character codes packed into a bit stream -> one refinement radix sort of all suffixes by
//              their first L_s characters -> the samples compacted out of that order, names from their
//              group heads
//   phase 3    names distinct: ranks directly; otherwise recursive string in (mod X, div X) order,
//              discarding decision, recursion, ranks from the recursive SA (permutation scatters)
//   phase 4    groups of suffixes still tied after L_s characters ordered by the DC comparator on the
//              sample ranks; the sorted order is the level's SA
// All data stays in HBM; the host only reads back a handful of scalars per step.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "../../include/dcx.h"
#include "alphabetic_code.hpp"
#include "dcx_kernels.hpp"

// Device memory of dcx_build_sa_noalloc: every buffer of such a build is carved from this array, which
// each GPU the library runs on holds from the code object's load until the process ends.  Host code uses
// only the address hipGetSymbolAddress returns for the current device (scratch_address), never the name.
static __device__ __align__(256) uint8_t scratch[DCX_NOALLOC_SCRATCH_BYTES];

namespace dcx {

constexpr int MAX_DEVICES = 64;

// This GPU's copy of `scratch`, looked up once per device with that device current (dcx_create calls it);
// nullptr if the lookup failed.
static void* scratch_address(int device) {
    static std::mutex mu;
    static void* addr[MAX_DEVICES] = {};
    static bool done[MAX_DEVICES] = {};
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;  // tried again by the next handle
        }
        addr[device] = p;
        done[device] = true;
    }
    return addr[device];
}

// dcx_build_sa_noalloc calls on one GPU share its scratch array: they are serialized by this lock, held
// until the call's work on the device has finished
static std::mutex& scratch_call_mutex(int device) {
    static std::mutex mu[MAX_DEVICES];
    return mu[device];
}

struct Options {
    uint32_t X = 39;
    bool pack_samples = true, pack_merging = true;
    double discarding_threshold = 0.7;
    uint64_t base_case_chars = 10000;
};

#ifndef DCX_ROUND0_SLACK
#define DCX_ROUND0_SLACK 5
#endif
constexpr uint32_t ROUND0_SLACK = DCX_ROUND0_SLACK;  // key bits beyond log2(N) in round 0 of large sorts
#ifndef DCX_PACKED_SLACK
#define DCX_PACKED_SLACK 14
#endif
// round 0 packs the position into the key when the characters beside it carry this many bits of
// information beyond log2(N)
constexpr uint32_t PACKED_SLACK = DCX_PACKED_SLACK;
// large-group lists from this many members, in groups of this many on average, try the partition around
// each group's first key (long periodic runs make groups of millions; chance repeats, groups of tens)
constexpr uint32_t LARGE_PARTITION_MIN = 1u << 16;
// after a coded round 0, the exact least k of the active groups is found while at most this many are left
constexpr uint32_t GROUP_KMIN_MAX = 1u << 22;
constexpr uint32_t LARGE_PARTITION_AVG = 64;

// The round-0 key code of a level (ctab == nullptr: raw characters).  ctab: device table of 256 entries
// (length << 24 | codeword left-aligned in 24 bits) indexed by character code; avg_len: the mean code
// length of the sampled characters, entropy: their order-0 entropy.
struct KeyCode {
    const uint32_t* ctab = nullptr;
    uint32_t max_len = 0, min_len = 0;  // longest and shortest codeword
    double avg_len = 0, entropy = 0;
};

// Longest text (and longest joined batch, separators included) of this build: positions and the values of
// the sorts are 32-bit (a text of 2^32 characters has positions 0 .. 2^32 - 1), while counts, offsets and
// the grid-stride loops of the kernels are 64-bit.
constexpr uint64_t MAX_TEXT = 1ull << 32;

class Engine {
  public:
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceCache cache;
    Primitives prim;
    uint64_t budget = 0;
    std::string last_error;
    void* scratch = nullptr;  // this device's dcx_build_sa_noalloc array (scratch_address)
    // DCX_PHASE_MARKS=1 at handle creation: level 0 launches k_phase_mark<P> at its phase boundaries
    // (1 after the alphabet, 2 after the suffix sort, 3 after the names, 4 after the recursion, 5 after the
    // merge, 6 after the output), so that per-dispatch profiles can be summed per phase
    bool phase_marks = false;
    void mark(int P) {
        if (!phase_marks) return;
        switch (P) {
            case 1: k_phase_mark<1><<<1, 64, 0, stream>>>(); break;
            case 2: k_phase_mark<2><<<1, 64, 0, stream>>>(); break;
            case 3: k_phase_mark<3><<<1, 64, 0, stream>>>(); break;
            case 4: k_phase_mark<4><<<1, 64, 0, stream>>>(); break;
            case 5: k_phase_mark<5><<<1, 64, 0, stream>>>(); break;
            default: k_phase_mark<6><<<1, 64, 0, stream>>>(); break;
        }
    }
    hipEvent_t evt[6] = {};   // phase timing events of a call, created with the handle

    // per call
    Options opt;
    const DcSet* dcs = nullptr;
    DBuf<uint32_t> d_cmp;
    uint32_t cmp_X = 0;
    uint32_t sigma = 0;
    dcx_stats* st = nullptr;
    uint64_t alg_bytes = 0;
    uint32_t sa_width = 8;
    // the second stream of batched calls (its own engine: stream, device cache, sort state), made on first
    // use by a batch of several sub-batches
    std::unique_ptr<Engine> twin;

    // pinned staging ring for copies between device memory and pageable host memory (allocated on the
    // first large copy).  The stage size is 64 MiB; the environment variable DCX_STAGE_BYTES (read when
    // the handle is created) sets another one, so that tests can make small copies wrap the ring.  Copies
    // of at least a quarter of a stage go through the ring.
    static constexpr int kStages = 4;
    size_t stage_bytes = 64ull << 20;
    uint8_t* h_stage[kStages] = {};
    hipEvent_t stage_ev[kStages] = {};

    // Joins the copy threads of a staged transfer however the transfer ends.  If it ends while threads
    // still run (an exception while they were being started), `failed` tells them to stop first, so a
    // joinable std::thread is never destroyed (which would terminate the process).
    struct Joiner {
        std::vector<std::thread>& threads;
        std::atomic<int>& failed;
        ~Joiner() {
            bool early = false;
            for (auto& t : threads) early = early || t.joinable();
            if (early) failed = 1;
            for (auto& t : threads)
                if (t.joinable()) t.join();
        }
    };
    static unsigned copy_threads() {
        const unsigned hw = std::thread::hardware_concurrency();
        return std::min(8u, std::max(1u, hw / 2));
    }

    // Synchronous copy of `bytes` from device memory to pageable host memory, ordered after the work on the
    // engine's stream.  The DMA engine writes pinned memory at about 57 GB/s and pageable memory at about
    // 17 GB/s (5 GiB: 94 against 308 ms), so large copies go in 64 MiB chunks through a ring of pinned
    // buffers; host threads copy each chunk out (each its own slice) while the next chunks download.
    void download(void* dst, const void* src, size_t bytes) {
        if (bytes < stage_bytes / 4) {
            DCX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
            DCX_HIP(hipStreamSynchronize(stream));
            return;
        }
        download_staged(src, bytes, 1, [dst](size_t off, const uint8_t* p, size_t len) {
            memcpy(static_cast<uint8_t*>(dst) + off, p, len);
        });
    }

    // n 32-bit SA entries from device memory to pageable host memory as `width`-byte entries (5: the low
    // four bytes and a zero byte, little endian; 8), widened by the host threads as they copy each chunk out
    // of the pinned ring: the link carries 4 bytes per entry instead of `width` (the positions are below
    // 2^32).  For outputs of at least a quarter stage.
    void download_widen(void* dst, const uint32_t* src, uint64_t n, uint32_t width) {
        const bool aligned8 = (reinterpret_cast<uintptr_t>(dst) & 7) == 0;
        download_staged(src, n * 4, 4, [dst, width, aligned8](size_t off, const uint8_t* p, size_t len) {
            const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
            const size_t e0 = off / 4, ne = len / 4;
            uint8_t* d = static_cast<uint8_t*>(dst) + e0 * width;
            if (width == 8 && aligned8) {
                uint64_t* d8 = reinterpret_cast<uint64_t*>(d);
                for (size_t i = 0; i < ne; i++) d8[i] = q[i];
            } else {
                for (size_t i = 0; i < ne; i++) {
                    const uint64_t v = q[i];
                    memcpy(d + i * width, &v, width);  // little endian: the low `width` bytes of v
                }
            }
        });
    }

    // The staged ring of download(): chunk k of the device bytes goes through stage k mod kStages; the host
    // threads split each chunk into slices of whole `align`-byte units and call copy_out(offset in the
    // source, pinned bytes, length) for them.
    template <class CopyOut>
    void download_staged(const void* src, size_t bytes, size_t align, CopyOut&& copy_out) {
        alloc_stages();
        const size_t SB = stage_bytes;
        const size_t nch = (bytes + SB - 1) / SB;
        const unsigned T = copy_threads();
        std::atomic<size_t> posted{0};                // chunks enqueued (their event recorded)
        std::atomic<unsigned> done[kStages];          // workers finished with the stage's current chunk
        std::atomic<int> failed{0};
        std::atomic<int> fail_code{0};
        for (int s = 0; s < kStages; s++) done[s] = T;
        auto chunk_len = [&](size_t k) { return std::min(SB, bytes - k * SB); };
        std::vector<std::thread> workers;
        Joiner joiner{workers, failed};
        workers.reserve(T);
        for (unsigned t = 0; t < T; t++) {
            workers.emplace_back([&, t]() {
                for (size_t k = 0; k < nch; k++) {
                    while (posted.load() <= k && !failed.load()) std::this_thread::yield();
                    if (failed.load()) return;
                    const int s = (int)(k % kStages);
                    const hipError_t e = hipEventSynchronize(stage_ev[s]);
                    if (e != hipSuccess) {
                        fail_code = (int)e;
                        failed = 1;
                        return;
                    }
                    const size_t units = chunk_len(k) / align;
                    const size_t a = units * t / T * align, b = units * (t + 1) / T * align;
                    if (b > a) copy_out(k * SB + a, h_stage[s] + a, b - a);
                    done[s].fetch_add(1);
                }
            });
        }
        for (size_t k = 0; k < nch && !failed.load(); k++) {
            const int s = (int)(k % kStages);
            while (done[s].load() < T && !failed.load()) std::this_thread::yield();
            if (failed.load()) break;
            done[s] = 0;
            hipError_t e = hipMemcpyAsync(h_stage[s], static_cast<const uint8_t*>(src) + k * SB,
                                          chunk_len(k), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipEventRecord(stage_ev[s], stream);
            if (e != hipSuccess) {
                fail_code = (int)e;
                failed = 1;
                break;
            }
            posted = k + 1;
        }
        for (auto& w : workers) w.join();
        if (failed.load())
            throw Error(DCX_E_HIP, std::string("staged download: ") + hipGetErrorString((hipError_t)fail_code.load()));
    }

    // Copy of `bytes` from pageable host memory to device memory, enqueued on the engine's stream (the
    // kernels after it see the data by stream order).  Large copies go through the same pinned ring: host
    // threads fill a stage while the DMA engine uploads the previous ones.  Returns once every stage is
    // free again, so the caller's buffer and the ring can be reused.
    void upload(void* dst, const void* src, size_t bytes) {
        if (bytes < stage_bytes / 4) {
            DCX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
            DCX_HIP(hipStreamSynchronize(stream));
            return;
        }
        alloc_stages();
        const size_t SB = stage_bytes;
        const size_t nch = (bytes + SB - 1) / SB;
        const unsigned T = copy_threads();
        std::atomic<size_t> released{0};             // chunks whose upload has completed
        std::atomic<unsigned> filled[kStages];        // workers done filling the stage's current chunk
        std::atomic<int> failed{0};
        int fail_code = 0;
        for (int s = 0; s < kStages; s++) filled[s] = 0;
        auto chunk_len = [&](size_t k) { return std::min(SB, bytes - k * SB); };
        std::vector<std::thread> workers;
        Joiner joiner{workers, failed};
        workers.reserve(T);
        for (unsigned t = 0; t < T; t++) {
            workers.emplace_back([&, t]() {
                for (size_t k = 0; k < nch; k++) {
                    // the stage's previous chunk (k - kStages) must have left it
                    while (k >= (size_t)kStages && released.load() < k - kStages + 1 && !failed.load())
                        std::this_thread::yield();
                    if (failed.load()) return;
                    const int s = (int)(k % kStages);
                    const size_t len = chunk_len(k);
                    const size_t a = len * t / T, b = len * (t + 1) / T;
                    memcpy(h_stage[s] + a, static_cast<const uint8_t*>(src) + k * SB + a, b - a);
                    filled[s].fetch_add(1);
                }
            });
        }
        size_t next_release = 0;
        for (size_t k = 0; k < nch && !failed.load(); k++) {
            const int s = (int)(k % kStages);
            while (filled[s].load() < T) std::this_thread::yield();
            filled[s] = 0;
            hipError_t e = hipMemcpyAsync(static_cast<uint8_t*>(dst) + k * SB, h_stage[s], chunk_len(k),
                                          hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipEventRecord(stage_ev[s], stream);
            // free the stage the workers need next: chunk k + 1 reuses the stage of chunk k + 1 - kStages
            if (e == hipSuccess && k + 1 >= (size_t)kStages && k + 1 < nch) {
                e = hipEventSynchronize(stage_ev[(k + 1) % kStages]);
                next_release = k + 2 - kStages;
                released = next_release;
            }
            if (e != hipSuccess) {
                fail_code = (int)e;
                failed = 1;
            }
        }
        for (auto& w : workers) w.join();
        if (!failed.load()) {
            const hipError_t e = hipStreamSynchronize(stream);
            if (e != hipSuccess) {
                fail_code = (int)e;
                failed = 1;
            }
        }
        if (failed.load())
            throw Error(DCX_E_HIP, std::string("staged upload: ") + hipGetErrorString((hipError_t)fail_code));
    }

    void alloc_stages() {
        if (h_stage[0]) return;
        for (int s = 0; s < kStages; s++) {
            DCX_HIP(hipHostMalloc((void**)&h_stage[s], stage_bytes, hipHostMallocDefault));
            DCX_HIP(hipEventCreateWithFlags(&stage_ev[s], hipEventDisableTiming));
        }
    }

    explicit Engine(int dev) : device(dev) {
        if (const char* e = getenv("DCX_PHASE_MARKS")) phase_marks = e[0] == '1';
        if (const char* e = getenv("DCX_STAGE_BYTES")) {
            const unsigned long long v = strtoull(e, nullptr, 10);
            if (v >= 4096 && v <= (1ull << 30)) stage_bytes = (size_t)(v & ~4095ull);
        }
        DCX_HIP(hipSetDevice(device));
        DCX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        prim.init(stream, &cache);
        for (auto& e : evt) DCX_HIP(hipEventCreate(&e));
        // lookback status for the largest build without allocation (radix tiles of 64-bit keys, and the
        // single-pass scans), so that such a build never has to grow it
        prim.ensure_sort_status(ceil_div(DCX_NOALLOC_MAX_CHARS, (uint64_t)RS_TILE) + RS_SEGS);
        prim.ensure_scan_status(SC_SINGLE_PASS_TILES);
        scratch = scratch_address(device);
        DCX_HIP(hipStreamSynchronize(stream));
    }
    ~Engine() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (int k = 0; k < kStages; k++) {
            if (h_stage[k]) (void)hipHostFree(h_stage[k]);
            if (stage_ev[k]) (void)hipEventDestroy(stage_ev[k]);
        }
        for (auto& e : evt)
            if (e) (void)hipEventDestroy(e);
        d_cmp.reset();
        prim.destroy();
        cache.release_all();
        if (stream) (void)hipStreamDestroy(stream);
    }

    unsigned grid(uint64_t n, unsigned max_blocks = 8192) const { return Primitives::grid_for(n, 256, max_blocks); }

    void set_dc(uint32_t X) {
        dcs = find_dc(X);
        if (!dcs) throw Error(DCX_E_INVALID_X, "X=" + std::to_string(X) + " is not a supported difference cover");
        if (cmp_X != X) {
            std::vector<uint32_t> t = make_cmp_table(*dcs);
            d_cmp = DBuf<uint32_t>(cache, t.size());
            DCX_HIP(hipMemcpyAsync(d_cmp.get(), t.data(), t.size() * 4, hipMemcpyHostToDevice, stream));
            DCX_HIP(hipStreamSynchronize(stream));
            cmp_X = X;
        }
    }

    DcDev make_dcdev(uint64_t n, uint64_t* m_out) const {
        DcDev d;
        memset(&d, 0, sizeof(d));
        d.X = dcs->X;
        d.D = dcs->D;
        d.inv_X = ((1ull << 32) + d.X - 1) / d.X;
        d.inv_D = ((1ull << 32) + d.D - 1) / d.D;
        std::vector<uint8_t> in(d.X, 0);
        for (uint32_t i = 0; i < d.D; i++) {
            d.dc[i] = dcs->dc[i];
            in[d.dc[i]] = 1;
        }
        uint32_t c = 0;
        for (uint32_t r = 0; r < d.X; r++) {
            d.next_rank[r] = (uint8_t)c;
            d.in_dc[r] = in[r];
            if (in[r]) d.in_mask[r >> 6] |= 1ull << (r & 63);
            c += in[r];
        }
        // samples are positions p in [0, n] with p mod X in DC (compute_length_info, pdcx.hpp:473-510)
        d.sb[0] = 0;
        for (uint32_t t = 0; t < d.D; t++) {
            uint64_t cnt = d.dc[t] <= n ? (n - d.dc[t]) / d.X + 1 : 0;
            d.sb[t + 1] = (uint32_t)(d.sb[t] + cnt);
        }
        *m_out = d.sb[d.D];
        return d;
    }

    // -------------------------------------------------------------------------------------------
    // active sets: members of groups of size >= 2 of the heads array H, as positions and group ids
    // -------------------------------------------------------------------------------------------
    // List indices, group starts and group ids are 32-bit: a list holds at most 2^32 entries (gstart[G] = A
    // is then stored as 0; group sizes are differences of starts modulo 2^32, and no group holds all 2^32
    // positions: round 0 tells the last position's key, which reaches past the text end, from the others')
    // and at most 2^31 groups.
    struct ActiveSet {
        DBuf<uint32_t> act, gid, gstart, gfirst;  // gstart: G + 1 entries, the last one A; gfirst: G entries
        uint64_t A = 0;
        uint32_t G = 0;
        // groups of more than RF_WAVE members and their members, members of groups of more than TB_SMALL,
        // groups of RF_SMALL + 1 .. RF_WAVE members
        uint32_t nlarge = 0;
        uint64_t B8 = 0, B16 = 0;
        uint64_t nwave = 0;
    };

    // restricted to the previous active list `src` (its superset) or to all N positions (src == nullptr).
    // H: the heads of all positions, or with `compact` the heads of the list entries (k_writeback's Ha).
    void active_set(const uint8_t* H, const uint32_t* src, uint64_t Nsrc, ActiveSet& out, bool compact = false) {
        out.act = DBuf<uint32_t>(cache, Nsrc);
        out.gid = DBuf<uint32_t>(cache, Nsrc);
        out.gstart = DBuf<uint32_t>(cache, (size_t)Nsrc / 2 + 1);  // groups have >= 2 members
        out.gfirst = DBuf<uint32_t>(cache, (size_t)Nsrc / 2 + 1);
        if (src && !compact)
            prim.scan(ActiveScanOp<true>{H, src, out.act.get(), out.gid.get(), out.gstart.get(), out.gfirst.get()},
                      Nsrc, 0);
        else
            prim.scan(ActiveScanOp<false>{H, src, out.act.get(), out.gid.get(), out.gstart.get(), out.gfirst.get()},
                      Nsrc, 0);
        // the end sentinel and the group-size counts on the device, then one readback for all of them
        k_active_finish<<<1, 64, 0, stream>>>(prim.scalars, out.gstart.get());
        k_group_counts<<<grid(Nsrc / 2 + 1, 2048), 256, 0, stream>>>(out.gstart.get(), prim.scalars);
        const uint64_t* t = prim.read_scalars(12);
        out.A = t[0];
        out.G = (uint32_t)t[1];
        out.nlarge = (uint32_t)t[8];
        out.B8 = t[9];
        out.B16 = t[10];
        out.nwave = t[11];
    }

    // -------------------------------------------------------------------------------------------
    // refinement sort: orders N elements by their first L characters (`bits` bits each, taken from the
    // text-order packed words W).  Round 0 sorts everything by the first word; every later round sorts
    // only the members of still-tied groups by (group, next characters).  On return `order` holds the
    // element values in sorted order, H[j] = 1 where a group of equal L-character keys starts (H[N] = 1)
    // and `fin` lists the members of the groups that are still tied.
    // -------------------------------------------------------------------------------------------
    // Rounds after the first either append the next packed characters (cheap when many characters fit
    // a word) or, for wide characters, double the covered prefix with the group ids of the positions C
    // ahead (exact: the last step takes only the L - C characters still missing).  Doubling needs the
    // sentinel coding (dense = false): the empty suffix past the end must be the smallest key.
    // `info`: estimated bits of information per character (at most `bits`), which sizes round 0.
    // `code`: an order-preserving prefix code of the character words (at most 8 bits, the padding's code
    // included), which keys round 0 when its bits for log2(N) + ROUND0_SLACK bits of information fit beside
    // the position (keys-only passes), or, in large sorts, when they take fewer radix passes than the raw
    // characters (40 bits at most, beside the split position).
    void refine_sort(const uint64_t* PT, uint64_t N, uint32_t L, uint32_t bits, bool dense, double info,
                     DBuf<uint32_t>& order, DBuf<uint8_t>& H, ActiveSet& fin, const KeyCode& code = KeyCode{}) {
        const uint32_t cpw = 64 / bits;
        uint32_t c0 = L < cpw ? L : cpw;
        bool coded = false;  // round 0 keyed by `code`: kbits code bits per suffix
        uint32_t kbits = 0;
        const uint32_t pbits = bit_width(N - 1);
        const uint32_t cpack = std::min<uint32_t>(c0, (64 - pbits) / bits);
        // code bits for log2(N) + ROUND0_SLACK bits of information (about entropy / avg_len of it per code
        // bit), in whole digits; with the position beside them when they leave it room
        const bool code_ok = code.ctab && bits <= 8 && code.entropy > 0 && N >= (1u << 16) && L > 1;
        const uint32_t ckb =
            code_ok ? 8 * (uint32_t)std::ceil((bit_width(N) + ROUND0_SLACK) * code.avg_len / code.entropy / 8) : 0;
        if (code_ok && ckb <= CODE_KEY_BITS_MAX && ckb <= 64 - pbits) {
            // round 0 keyed by the code with the position in the bits below (keys-only passes of 16 bytes);
            // the histogram kernel writes the keys into ka, the first pass reads them from there
            coded = true;
            kbits = ckb;
            DBuf<uint64_t> ka(cache, N), kb(cache, N);
            uint64_t* k = ka.get();
            uint64_t* k2 = kb.get();
            uint32_t* nv = nullptr;
            uint32_t* nv2 = nullptr;
            prim.sort_pairs_from(CodedKeys{PT, bits, code.ctab, kbits, L, k, 1u, L * code.min_len >= kbits ? 1u : 0u}, k, k2, nv, nv2, N, 64 - kbits, 64,
                                 false);
            H = DBuf<uint8_t>(cache, (size_t)N + 1);
            order = DBuf<uint32_t>(cache, N);
            k_heads_packed<<<grid(N / 16 + 1ull), 256, 0, stream>>>(k, N, kbits, (1ull << pbits) - 1, H.get(), order.get());
        } else if (N >= (1u << 16) && cpack * info >= bit_width(N) + PACKED_SLACK) {
            // round 0 sorts one word per position holding the first characters above the position itself
            // (16 instead of 24 bytes moved per pass) when the characters left beside the position still
            // carry PACKED_SLACK bits of information more than log2(N); the positions they leave tied go
            // through the group rounds
            c0 = cpack;
            const uint32_t kbits = c0 * bits;
            const uint64_t hi = ~0ull << (64 - kbits);
            DBuf<uint64_t> ka(cache, N), kb(cache, N);
            uint64_t* k = ka.get();
            uint64_t* k2 = kb.get();
            uint32_t* nv = nullptr;
            uint32_t* nv2 = nullptr;
            prim.sort_pairs_from(StreamKeys<true>{PT, bits, hi}, k, k2, nv, nv2, N, 64 - kbits, 64, false);
            H = DBuf<uint8_t>(cache, (size_t)N + 1);
            order = DBuf<uint32_t>(cache, N);
            k_heads_packed<<<grid(N / 16 + 1ull), 256, 0, stream>>>(k, N, kbits, (1ull << pbits) - 1, H.get(), order.get());
        } else {
            kbits = c0 * bits;
            if (N >= (1u << 24)) {
                // large sorts: round 0 takes the characters that carry about log2(N) + ROUND0_SLACK bits of
                // information, rounded up to whole 8-bit digits; the positions still tied after it go
                // through the cheaper group rounds
                const uint32_t want = (uint32_t)std::ceil((bit_width(N) + ROUND0_SLACK) / info);
                const uint32_t digits = std::min<uint32_t>(8, (std::max<uint32_t>(1, want) * bits + 7) / 8);
                c0 = std::max<uint32_t>(1, std::min<uint32_t>(c0, 8 * digits / bits));
                kbits = c0 * bits;
                if (code_ok && L > c0) {
                    // the same information from code bits (about entropy / avg_len of it per bit), at
                    // most CODE_KEY_BITS_MAX of them: taken when that saves a radix pass
                    const double need = (bit_width(N) + ROUND0_SLACK) * code.avg_len / code.entropy;
                    const uint32_t cb = std::min<uint32_t>(CODE_KEY_BITS_MAX, 8 * (uint32_t)std::ceil(need / 8));
                    if (cb / 8 < (kbits + 7) / 8) {
                        coded = true;
                        kbits = cb;
                    }
                }
            }
            DBuf<uint64_t> ka(cache, N), kb(cache, N);
            DBuf<uint32_t> va(cache, N), vb(cache, N);
            uint64_t* k = ka.get();
            uint64_t* k2 = kb.get();
            uint32_t* v = va.get();
            uint32_t* v2 = vb.get();
            // every position is an element: the first pass reads the character stream directly; the last
            // pass writes only the 32 key bits below its digit, enough for the heads
            DBuf<uint32_t> low(cache, N);
            DBuf<uint8_t> ia, ib;  // keys of at most 40 bits leave room for the low 24 bits of a position
            if (kbits <= 40) {
                ia = DBuf<uint8_t>(cache, N);
                ib = DBuf<uint8_t>(cache, N);
            }
            // the heads: from the low words (at most five passes), else written by the last pass itself
            H = DBuf<uint8_t>(cache, (size_t)N + 1);
            bool heads_done = false;
            uint32_t np;
            if (coded) {
                // the histogram kernel writes the code keys into k, the first pass reads them from there
                np = prim.sort_pairs_from(CodedKeys{PT, bits, code.ctab, kbits, L, k, 0u, L * code.min_len >= kbits ? 1u : 0u}, k, k2, v, v2, N,
                                          64 - kbits, 64, true, low.get(), ia.get(), ib.get(), H.get(),
                                          top_mask(kbits), &heads_done);
            } else {
                np = prim.sort_pairs_from(StreamKeys<false>{PT, bits, top_mask(kbits)}, k, k2, v, v2, N, 64 - kbits,
                                          64, true, low.get(), ia.get(), ib.get(), H.get(), top_mask(kbits),
                                          &heads_done);
            }
            if (np) {
                k_heads_u32<<<grid(N / 16 + 1ull), 256, 0, stream>>>(low.get(), N, H.get());
                k_mark_runs<<<1, RS_RADIX, 0, stream>>>(prim.hist + (np - 1) * RS_RADIX, N, H.get());
            } else if (heads_done) {
                k_heads_fix<<<grid(N / 16 + 1ull), 256, 0, stream>>>(k, N, top_mask(kbits), H.get());
                prim.fill_u8(H.get() + N, 1, 1);
            } else {
                k_heads<<<grid(N / 16 + 1ull), 256, 0, stream>>>(k, N, kbits, H.get());
            }
            if (v == va.get())
                order = std::move(va);
            else
                order = std::move(vb);
        }
        // After a coded round 0 a group's members share k >= kmin whole characters (k <= L); the rounds
        // then go on from character k + C of each member, and every group is resolved once kmin + C >= L.
        // kmin: a lower bound of the active groups' k, their exact minimum once few groups are left
        uint32_t C = coded ? 0 : c0;
        uint32_t kmin = coded ? std::min<uint32_t>(L, kbits / code.max_len) : 0;
        const RoundChars rk{PT, bits, 0, 0, code.ctab, kbits, L};
        bool parked = false;  // some large groups left the rounds equal through L (refine_round deep)
        DBuf<uint32_t> act0;  // the first round's list (its slots hold every group tied later), A0 entries
        uint64_t A0 = 0;
        ActiveSet cur;
        active_set(H.get(), nullptr, N, cur);
        DBuf<uint32_t> G;
        while (kmin + C < L && cur.A > 0) {
            const uint64_t A = cur.A;
            const uint32_t gbits = bit_width(cur.G - 1);
            uint32_t c = (64 - gbits) / bits;
            if (c > L - kmin - C) c = L - kmin - C;
            if (c > cpw) c = cpw;
            if (!coded && !dense && (L - C) > 3 * std::max<uint32_t>(c, 1)) {
                // doubling step
                if (!G.get()) G = DBuf<uint32_t>(cache, N);
                prim.scan(GroupScatterOp{H.get(), order.get(), G.get()}, N, 1);
                const uint64_t groups = prim.totals(1).lo;
                const uint32_t off = std::min(C, L - C);
                const uint32_t gvb = bit_width(groups);
                DBuf<uint64_t> rka(cache, A), rkb(cache, A);
                DBuf<uint32_t> rva(cache, A), rvb(cache, A);
                k_keys_double<<<grid(A), 256, 0, stream>>>(G.get(), N, A, cur.act.get(), cur.gid.get(), order.get(), off,
                                                           gvb, rka.get(), rva.get());
                uint64_t* x = rka.get();
                uint64_t* x2 = rkb.get();
                uint32_t* y = rva.get();
                uint32_t* y2 = rvb.get();
                prim.sort_pairs(x, x2, y, y2, A, 0, gbits + gvb);
                DBuf<uint8_t> Ha(cache, (size_t)A + 1);
                k_writeback<<<grid(A), 256, 0, stream>>>(x, y, cur.act.get(), A, order.get(), H.get(), 1, Ha.get());
                C += off;
                ActiveSet next;
                active_set(Ha.get(), cur.act.get(), A, next, true);
                cur = std::move(next);
                continue;
            }
            // a character round takes as many characters as one stream window holds, as far as the large
            // groups' keys (rank among the large groups, characters) still fit one word; small groups
            // compare the characters alone
            c = std::min<uint32_t>(cpw, L - kmin - C);
            {
                // (one bit below the characters is kept for the class bit of deep partitions)
                const uint32_t nlarge = cur.nlarge;  // counted with the active set
                const uint32_t lbits = nlarge > 1 ? bit_width(nlarge - 1) : 0;
                if (lbits + c * bits + 1 > 64 && 64 - lbits - 1 >= bits) c = (64 - lbits - 1) / bits;
            }
            DBuf<uint8_t> Ha(cache, (size_t)A + 1);
            // while some group needs another round after this one, the partitions of large groups look ahead to
            // character L and park the members equal to their first one through it
            const bool deep = kmin + C + c < L;
            parked |= refine_round(RoundChars{PT, bits, C, c * bits, coded ? code.ctab : nullptr, kbits, L}, cur, order,
                                   H, Ha, deep);
            C += c;
            ActiveSet next;
            active_set(Ha.get(), cur.act.get(), A, next, true);
            if (!act0.get()) {
                act0 = std::move(cur.act);
                A0 = A;
            }
            cur = std::move(next);
            if (coded && cur.A > 0 && cur.G <= GROUP_KMIN_MAX) {
                DBuf<uint32_t> km(cache, 1);
                prim.fill_u32(km.get(), L, 1);
                k_group_k<<<grid(cur.G), 256, 0, stream>>>(rk, cur.G, cur.gfirst.get(), nullptr, nullptr, order.get(),
                                                           nullptr, km.get());
                kmin = std::max(kmin, (uint32_t)prim.read_device(km.get(), 4)[0]);
            }
        }
        if (parked) {
            // the parked groups are tied through L as well: the tied set is every group of H with two or
            // more members, all of them inside the first round's list
            cur = ActiveSet();
            active_set(H.get(), act0.get(), A0, cur);
        }
        fin = std::move(cur);
    }

    // out[dest[i]] = val[i] + add for i < N, where dest is a permutation of [0, N) (val == nullptr: the values
    // are i).  Word-sized writes to random places cost a cache line each; instead an LSD radix sort by the
    // high bits of dest (all but the low SCATTER_WINDOW_BITS or fewer) orders the pairs by destination
    // window, and each window is assembled in LDS and written out whole.
    void scatter_u32(const uint32_t* dest, const uint32_t* val, uint64_t N, uint32_t add, uint32_t* out) {
        if (N == 0) return;
        const uint32_t bw = bit_width(N - 1);
        if (N < SCATTER_PARTITION_MIN || bw <= SCATTER_WINDOW_BITS) {
            if (val) {
                k_scatter_pairs<<<grid(N), 256, 0, stream>>>(dest, val, N, add, out);
            } else {
                DBuf<uint32_t> idx(cache, N);
                k_iota<<<grid(N), 256, 0, stream>>>(idx.get(), N);
                k_scatter_pairs<<<grid(N), 256, 0, stream>>>(dest, idx.get(), N, add, out);
            }
            return;
        }
        const uint32_t np = (bw - SCATTER_WINDOW_BITS + 7) / 8;
        const uint32_t wbits = bw - 8 * np;
        DBuf<uint32_t> k1(cache, N), v1(cache, N), k2(cache, N), v2(cache, N);
        uint32_t *ks = nullptr, *vs = nullptr;
        prim.sort_u32(dest, val, k1.get(), v1.get(), k2.get(), v2.get(), N, wbits, np, &ks, &vs);
        const unsigned threads = wbits >= SCATTER_WIDE_BITS ? SCATTER_WIDE_THREADS : 256u;
        k_scatter_window<<<(unsigned)ceil_div((uint64_t)N, 1ull << wbits), threads, (size_t)4 << wbits, stream>>>(
            ks, vs, N, add, wbits, out);
    }

    // One character round of the refinement: orders the members of every tied group of `cur` by their
    // next c characters (after the first C, c * bits <= 64), writes the new group starts into H and the
    // compact heads of the list into Ha (A + 1 entries).  Small groups are sorted in place by one thread
    // each, middle ones (up to RF_WAVE members) by one wave each; the members of large groups go through a
    // radix sort keyed by (rank among the large groups, characters), or two stable sorts when those do not
    // fit 64 bits.
    // deep: the large groups' three-way partitions compare their members with the first one up to character
    // L, and the members equal to it there leave the rounds (returns true when it may have parked some).
    bool refine_round(const RoundChars& rc, const ActiveSet& cur, DBuf<uint32_t>& order, DBuf<uint8_t>& H,
                      DBuf<uint8_t>& Ha, bool deep = false) {
        const uint64_t A = cur.A;
        const uint32_t G = cur.G;
        DBuf<uint8_t> large(cache, G);
        k_refine_small<<<grid(G), 256, 0, stream>>>(cur.gstart.get(), G, cur.gfirst.get(), order.get(), H.get(),
                                                    Ha.get(), rc, large.get());
        if (cur.nwave) {
            const uint64_t chunks = ceil_div((uint64_t)G, 64ull);
            k_refine_wave<<<(unsigned)std::min<uint64_t>(ceil_div(chunks, (uint64_t)RFW_WAVES), 8192), 64 * RFW_WAVES, 0,
                            stream>>>(cur.gstart.get(), G, cur.gfirst.get(), order.get(), H.get(), Ha.get(), rc);
        }
        prim.fill_u8(Ha.get() + A, 1, 1);
        // members of the large groups and their count, known from the active set's group counts
        const uint64_t B = cur.B8;
        const uint32_t nlarge = cur.nlarge;
        if (B == 0) return false;
        DBuf<uint32_t> slots(cache, B), lrank(cache, B), lfirst(cache, (size_t)nlarge + 1);
        prim.scan(LargeMembersOp{large.get(), cur.gid.get(), slots.get(), lrank.get(), lfirst.get()}, A, 1);
        prim.fill_u32(lfirst.get() + nlarge, (uint32_t)B, 1);  // B = 2^32 reads as 0: spans are differences
        const uint32_t lbits = bit_width(nlarge - 1);
        const uint32_t cbits = rc.cbits;
        DBuf<uint64_t> rka(cache, B), rkb(cache, B);
        DBuf<uint32_t> rva(cache, B), rvb(cache, B);
        uint64_t* x = rka.get();
        uint64_t* x2 = rkb.get();
        uint32_t* y = rva.get();
        uint32_t* y2 = rvb.get();
        if (lbits + cbits <= 64) {
            const bool part = B >= LARGE_PARTITION_MIN && B >= (uint64_t)LARGE_PARTITION_AVG * nlarge;
            const bool dp = deep && part && lbits + cbits + 1 <= 64;  // a class bit below the characters
            const uint32_t lo = 64 - lbits - cbits - (dp ? 1 : 0);
            DBuf<uint8_t> cls(cache, part ? B : 0);
            DBuf<uint32_t> lk;  // after a coded round 0: k of each large group, decoded once
            if (rc.ctab) {
                lk = DBuf<uint32_t>(cache, nlarge);
                k_group_k<<<grid(nlarge), 256, 0, stream>>>(rc, nlarge, lfirst.get(), slots.get(), cur.act.get(),
                                                            order.get(), lk.get(), nullptr);
            }
            k_keys_large<<<grid(B), 256, 0, stream>>>(rc, B, slots.get(), cur.act.get(), lrank.get(), order.get(),
                                                      lbits, x, y, lfirst.get(), part ? cls.get() : nullptr, lk.get(),
                                                      dp ? 1u : 0u);
            if (part && partition_large(cur, slots.get(), lrank.get(), lfirst.get(), cls.get(), B, lo, x, y, x2, y2,
                                        order, H, Ha, dp))
                return dp;
            prim.sort_pairs(x, x2, y, y2, B, lo, 64);
            k_writeback_large<<<grid(B), 256, 0, stream>>>(x, y, slots.get(), B, cur.act.get(), order.get(), H.get(),
                                                           Ha.get());
            return false;
        }
        // characters first, then a stable sort by the group rank (the members are identified by index)
        DBuf<uint32_t> P(cache, B);
        k_keys_large_chars<<<grid(B), 256, 0, stream>>>(rc, B, slots.get(), cur.act.get(), order.get(), x, P.get());
        prim.sort_pairs_from(ArrayKeys<uint64_t>{x}, x, x2, y, y2, B, 64 - cbits, 64, true);
        k_keys_large_rank<<<grid(B), 256, 0, stream>>>(y, lrank.get(), B, x2);
        std::swap(x, x2);
        prim.sort_pairs(x, x2, y, y2, B, 0, lbits);
        k_writeback_large2<<<grid(B), 256, 0, stream>>>(x, y, P.get(), rc, slots.get(), B, cur.act.get(), order.get(),
                                                        H.get(), Ha.get());
        return false;
    }

    // Large groups by a stable three-way partition around each group's first key (PivotScanOp), when at
    // most half of their members differ from it: only that minority is sorted.  x, y: the members' keys
    // and suffixes; x2, y2 (B entries each) take the compacted minority.  Returns false (nothing written,
    // x and y untouched) when the caller should sort all members instead.
    bool partition_large(const ActiveSet& cur, const uint32_t* slots, const uint32_t* lrank, const uint32_t* lfirst,
                         const uint8_t* cls, uint64_t B, uint32_t lo_bit, const uint64_t* x, const uint32_t* y,
                         uint64_t* x2, uint32_t* y2, DBuf<uint32_t>& order, DBuf<uint8_t>& H, DBuf<uint8_t>& Ha,
                         bool park = false) {
        DBuf<uint32_t> minpref(cache, (size_t)B + 1), c0pref(cache, (size_t)B + 1);
        prim.scan(PivotScanOp{cls, x, minpref.get(), c0pref.get(), x2, y2}, B, 1);
        const auto tot = prim.totals(1);
        const uint64_t Bm = tot.lo;
        if (Bm * 2 > B) return false;
        if (Bm == 0) {
            k_uniform_heads<<<grid(B), 256, 0, stream>>>(B, slots, lrank, lfirst, Ha.get(), park ? 1u : 0u);
            return true;
        }
        prim.fill_u32(minpref.get() + B, (uint32_t)Bm, 1);
        prim.fill_u32(c0pref.get() + B, (uint32_t)tot.hi, 1);
        DBuf<uint64_t> mk2(cache, Bm);
        DBuf<uint32_t> mi2(cache, Bm);
        uint64_t* mk = x2;
        uint64_t* mka = mk2.get();
        uint32_t* mi = y2;
        uint32_t* mia = mi2.get();
        prim.sort_pairs(mk, mka, mi, mia, Bm, lo_bit, 64);
        k_place_equal<<<grid(B), 256, 0, stream>>>(B, y, slots, lrank, lfirst, minpref.get(), c0pref.get(),
                                                   cur.act.get(), order.get(), H.get(), Ha.get(), park ? 1u : 0u);
        k_place_minority<<<grid(Bm), 256, 0, stream>>>(Bm, mk, mi, y, slots, lrank, lfirst, minpref.get(), c0pref.get(),
                                                       cur.act.get(), order.get(), H.get(), Ha.get());
        return true;
    }

    // dense words only: boundary suffixes that share a key group are moved to its front (k_boundary_fix)
    // The key groups (after the refinement over L characters of the dense stream PT) that hold boundary
    // suffixes, found by searching the sorted order for each boundary suffix's key (k_boundary_find).
    void boundary_fixup(const uint64_t* PT, uint32_t bits, uint32_t L, uint64_t n, uint64_t bnd,
                        DBuf<uint32_t>& order, DBuf<uint8_t>& H, ActiveSet& tied) {
        if (bnd >= n) return;
        const uint32_t nb = (uint32_t)(n - bnd);  // boundary suffixes: at most L - 1
        DBuf<uint32_t> found(cache, 2 * (size_t)nb), cnt(cache, 1);
        prim.fill_u32(cnt.get(), 0, 1);
        k_boundary_find<<<(nb + 63) / 64, 64, 0, stream>>>(PT, bits, L, order.get(), n, bnd, found.get(), cnt.get(), nb);
        const uint32_t hc = (uint32_t)prim.read_device(cnt.get(), 4)[0];
        if (hc == 0) return;
        if (hc > nb) throw Error(DCX_E_HIP, "boundary group count exceeds the boundary suffixes");
        std::vector<uint32_t> rg(2 * (size_t)hc);
        memcpy(rg.data(), prim.read_device(found.get(), 8 * (size_t)hc), 8 * (size_t)hc);
        std::vector<std::pair<uint32_t, uint32_t>> g;  // (first slot, size)
        for (uint32_t i = 0; i < hc; i++) g.emplace_back(rg[2 * i], rg[2 * i + 1]);
        std::sort(g.begin(), g.end());
        g.erase(std::unique(g.begin(), g.end()), g.end());
        for (const auto& r : g) {
            const uint64_t j0 = r.first;
            const uint32_t sz = r.second;
            DBuf<uint32_t> tmp(cache, sz);
            k_boundary_fix<<<1, 1024, 0, stream>>>(order.get(), H.get(), j0, sz, bnd, tmp.get());
        }
        ActiveSet next;
        active_set(H.get(), tied.act.get(), tied.A, next);
        tied = std::move(next);
    }

    // groups of equal merge prefixes -> DC comparator on the sample ranks (tie_break_ranks,
    // merge_samples.hpp:265-333)
    void tie_break(const DcDev& dc, DBuf<uint32_t>& order, DBuf<uint8_t>& H, ActiveSet& as, const uint32_t* ranks) {
        const uint64_t A = as.A;
        const uint32_t G = as.G;
        if (A == 0) return;
        DBuf<uint8_t> large(cache, G);
        k_tiebreak_small<<<grid(G), 256, 0, stream>>>(as.gstart.get(), G, as.gfirst.get(), order.get(), dc, d_cmp.get(),
                                                      ranks, large.get());
        // large groups (their member count is known from the active set's group counts): those whose members
        // already follow the comparator, or its reverse, pair by pair are written back at once; the others
        // go through a merge sort with the composite (group, DC comparator) order
        const uint64_t B = as.B16;
        if (B == 0) return;
        DBuf<uint64_t> items(cache, B);
        DBuf<uint32_t> slots(cache, B);
        LargeScanOp lop{large.get(), as.gid.get(), as.act.get(), order.get(), items.get(), slots.get()};
        prim.scan(lop, A, 1);
        DBuf<uint32_t> lidx(cache, G);
        prim.scan(LargeIndexOp{large.get(), lidx.get()}, G, 2);
        const uint64_t GLmax = B / (TB_SMALL + 1) + 1;  // a large group has more than TB_SMALL members
        DBuf<uint32_t> mono(cache, GLmax);
        prim.fill_u32(mono.get(), 0, GLmax);
        const uint64_t chunks = ceil_div(B, (uint64_t)TB_MONO_CHUNK);
        k_tb_monotone<<<(unsigned)std::min<uint64_t>(ceil_div(chunks, 4ull), 65536), 256, 0, stream>>>(
            items.get(), B, lidx.get(), dc, d_cmp.get(), ranks, mono.get());
        k_tb_reverse<<<grid(B), 256, 0, stream>>>(items.get(), slots.get(), B, lidx.get(), mono.get(), as.gstart.get(),
                                                  as.gfirst.get(), order.get());
        DBuf<uint64_t> items_m(cache, B);
        DBuf<uint32_t> slots_m(cache, B);
        prim.scan(MixedScanOp{items.get(), slots.get(), lidx.get(), mono.get(), items_m.get(), slots_m.get()}, B, 1);
        const uint64_t Bm = prim.totals(1).lo;
        if (Bm == 0) return;
        if (Bm > B) throw Error(DCX_E_HIP, "tie-break member count exceeds the large groups");
        k_tb_tile_sort<<<(unsigned)ceil_div(Bm, TB_TILE), 256, 0, stream>>>(items_m.get(), Bm, dc, d_cmp.get(), ranks);
        uint64_t* in = items_m.get();
        uint64_t* out = items.get();
        for (uint64_t w = TB_TILE; w < Bm; w *= 2) {
            k_tb_merge<<<grid(Bm), 256, 0, stream>>>(in, out, Bm, w, dc, d_cmp.get(), ranks);
            std::swap(in, out);
        }
        k_tb_scatter<<<grid(Bm), 256, 0, stream>>>(in, slots_m.get(), Bm, order.get());
    }

    void base_case(const uint8_t* t, const uint16_t* lut, uint32_t n, uint32_t* sa_out) {
        if (n == 0) return;
        DBuf<uint32_t> rank(cache, (size_t)n + 1);
        k_base_case<TextU8><<<1, 1024, 0, stream>>>(TextU8{t, n, lut}, n, sa_out, rank.get());
    }
    void base_case(const uint32_t* t, const uint16_t*, uint32_t n, uint32_t* sa_out) {
        if (n == 0) return;
        DBuf<uint32_t> rank(cache, (size_t)n + 1);
        k_base_case<TextU32><<<1, 1024, 0, stream>>>(TextU32{t, n}, n, sa_out, rank.get());
    }

    // The base case of a level of more than BASE_CASE_MAX characters (a base_case_chars above it): every
    // suffix sorted by the refinement over its whole length in sentinel codes (characters 1..maxchar,
    // the padding 0 below them), whose rounds double the compared prefix once it is long: prefix doubling
    // on the device.  Suffixes of different lengths differ within n characters, so no group stays tied.
    template <class CharT>
    void base_case_large(const CharT* text, const uint16_t* lut, uint64_t n, uint32_t maxchar, DBuf<uint32_t>& sa) {
        const uint32_t bits = bit_width(maxchar);
        DBuf<uint64_t> PT;
        make_stream(text, lut, n, bits, 0, 2 * n + 2, PT);
        DBuf<uint8_t> H;
        ActiveSet fin;
        refine_sort(PT.get(), n, (uint32_t)n, bits, false, info_bits(text, lut, n, bits), sa, H, fin);
        if (fin.A != 0) throw Error(DCX_E_HIP, "base case: suffixes left tied after their whole length");
    }

    dcx_level_stats* level_stat(int depth) {
        if (!st || depth >= DCX_MAX_LEVELS) return nullptr;
        if ((uint32_t)depth + 1 > st->levels) st->levels = depth + 1;
        dcx_level_stats* ls = &st->level[depth];
        memset(ls, 0, sizeof(*ls));
        return ls;
    }

    // phase 3 when the names are not all distinct: recursive string, discarding decision, recursion,
    // sample ranks (handle_recursive_call, pdcx.hpp:145-180).  s_order: the samples in name order (their
    // text-order sample indices); names: their names; Hs: their group heads; ranks: m + D entries out.
    void recursion(const DcDev& dc, uint64_t n, uint32_t m, uint32_t M, int depth, DBuf<uint32_t>& s_order,
                   DBuf<uint32_t>& names, DBuf<uint8_t>& Hs, DBuf<uint32_t>& ranks) {
        dcx_level_stats* ls = st && depth < DCX_MAX_LEVELS ? &st->level[depth] : nullptr;
        const double thr = opt.discarding_threshold;
        // the kept set of discarding, from the non-unique samples: after >= NU and after <= 2 NU + 1
        prim.total(NonUniqueOp{Hs.get(), s_order.get(), dc, m, nullptr}, m, 0);
        const uint32_t nu = (uint32_t)prim.totals(0).lo;
        DBuf<uint32_t> K;
        uint32_t after = m;
        if (n > 80 && (double)nu / (double)m <= thr) {
            const uint32_t nc = 2 * nu + 1;
            DBuf<uint32_t> cand(cache, nc);
            prim.scan_after_total(NonUniqueOp{Hs.get(), s_order.get(), dc, m, cand.get()}, m, 0);
            prim.fill_u32(cand.get() + 2 * nu, 0, 1);  // element 0 is always kept
            DBuf<uint32_t> k1(cache, nc), v1(cache, nc), k2(cache, nc), v2(cache, nc);
            uint32_t *ks = nullptr, *vs = nullptr;
            const uint32_t cb = bit_width(m);
            prim.sort_u32(cand.get(), nullptr, k1.get(), v1.get(), k2.get(), v2.get(), nc, 0, (cb + 7) / 8, &ks, &vs);
            K = DBuf<uint32_t>(cache, nc);
            prim.scan(DistinctOp{ks, K.get()}, nc, 0);
            after = (uint32_t)prim.totals(0).lo;
        }
        const bool use_disc = n > 80 && ((double)after / (double)m) <= thr;
        if (!use_disc) {
            // recursive string of all samples in (mod X, div X) order, recursion, ranks from its SA
            K.reset();
            DBuf<uint32_t> R(cache, m);
            {
                DBuf<uint32_t> dest(cache, m), word(cache, m);
                k_rec_pairs<<<grid(m), 256, 0, stream>>>(s_order.get(), names.get(), m, dc, dest.get(), word.get());
                names.reset();
                scatter_u32(dest.get(), word.get(), m, 0, R.get());
            }
            DBuf<uint32_t> sar;
            if (n <= 80) {
                sar = DBuf<uint32_t>(cache, m);
                base_case(R.get(), nullptr, m, sar.get());  // sequential_sa_on_local_ranks (pdcx.hpp:115-121)
            } else {
                level<uint32_t>(R.get(), nullptr, m, M, depth + 1, nullptr, nullptr, &sar);
            }
            DBuf<uint32_t> dest(cache, m);
            k_sa_dest<<<grid(m), 256, 0, stream>>>(sar.get(), m, dc, dest.get());
            sar.reset();
            scatter_u32(dest.get(), nullptr, m, 1, ranks.get());
            return;
        }
        if (ls) {
            ls->discarding = 1;
            ls->reduced = after;
        }
        // ranks in name order: final for the unique samples, patched below for the others
        scatter_u32(s_order.get(), nullptr, m, 1, ranks.get());
        DBuf<uint32_t> sar, isar(cache, after);
        {
            DBuf<uint32_t> Rr(cache, after);
            k_reduced_string<<<grid(after), 256, 0, stream>>>(K.get(), after, dc, ranks.get(), names.get(), Rr.get());
            names.reset();
            level<uint32_t>(Rr.get(), nullptr, after, M, depth + 1, nullptr, nullptr, &sar);
        }
        scatter_u32(sar.get(), nullptr, after, 1, isar.get());
        sar.reset();
        // order of (name, rank2) = name order refined inside non-unique groups
        ActiveSet stied;
        active_set(Hs.get(), nullptr, m, stied);
        const uint64_t A = stied.A;
        const uint32_t G = stied.G;
        if (A == 0) return;
        const uint32_t gb = bit_width(G - 1), r2b = bit_width(after);
        DBuf<uint64_t> ka(cache, A), kb(cache, A);
        DBuf<uint32_t> va(cache, A), vb(cache, A);
        k_keys_rank2_kept<<<grid(A), 256, 0, stream>>>(A, stied.act.get(), stied.gid.get(), s_order.get(), dc, K.get(),
                                                       after, isar.get(), r2b, ka.get(), va.get());
        uint64_t* x = ka.get();
        uint64_t* x2 = kb.get();
        uint32_t* y = va.get();
        uint32_t* y2 = vb.get();
        prim.sort_pairs(x, x2, y, y2, A, 0, gb + r2b);
        k_writeback<<<grid(A), 256, 0, stream>>>(x, y, stied.act.get(), A, s_order.get(), Hs.get(), 0, (uint8_t*)nullptr);
        k_patch_ranks<<<grid(A), 256, 0, stream>>>(A, stied.act.get(), s_order.get(), ranks.get());
    }

    // SA entries to the output width
    void widen(const uint32_t* sa, uint64_t n, uint32_t width, uint8_t* out) {
        if (width == 5 && (reinterpret_cast<uintptr_t>(out) & 15) == 0 && (reinterpret_cast<uintptr_t>(sa) & 15) == 0)
            k_widen40<<<grid(ceil_div(n, 16)), 256, 0, stream>>>(sa, n, out);
        else
            k_widen<<<grid(n), 256, 0, stream>>>(sa, n, width, out);
    }

    // bits per character of the reference's packed variants (cli.cpp:488-519), extended by 9 bits
    static uint32_t ref_pack_bits(uint32_t sig) { return sig <= 7 ? 3 : sig <= 31 ? 5 : sig <= 255 ? 8 : 9; }

    // -------------------------------------------------------------------------------------------
    // one recursion level: SA of text[0..n) into sa_out (device, n entries).  Level 0 reads bytes
    // through the alphabet map `lut`; deeper levels read names (uint32, >= 1).
    // -------------------------------------------------------------------------------------------
    // order-0 entropy (bits per character, within [1, bits]) of a byte text from a strided sample of its
    // codes; names of deeper levels are taken at their full width
    double info_bits(const uint8_t* text, const uint16_t* lut, uint64_t n, uint32_t bits, double entropy = -1.0) {
        have_counts = false;
        if (n < (1u << 24)) return bits;
        const double e = entropy >= 0 ? entropy : code_entropy(text, lut, n);
        return std::min<double>(bits, std::max(1.0, e));
    }
    double info_bits(const uint32_t*, const uint16_t*, uint64_t, uint32_t bits, double = -1.0) {
        have_counts = false;
        return bits;
    }
    // order-0 entropy of the codes lut[text[i]] over a strided sample of i; the sample's counts stay in
    // code_counts (have_counts) for the round-0 key code
    unsigned long long code_counts[257];
    bool have_counts = false;
    double code_entropy(const uint8_t* text, const uint16_t* lut, uint64_t n) {
        const uint64_t stride = 61;  // odd: no alignment with periodic structure
        DBuf<unsigned long long> hist(cache, 257);
        prim.fill_u64(reinterpret_cast<uint64_t*>(hist.get()), 0, 257);
        k_code_hist<<<grid(ceil_div(n, stride * 256) * 256, 4096), 256, 0, stream>>>(text, n, lut, stride, hist.get());
        unsigned long long* h = code_counts;
        memcpy(h, prim.read_device(hist.get(), sizeof(code_counts)), sizeof(code_counts));
        have_counts = true;
        double total = 0, e = 0;
        for (int i = 0; i < 257; i++) total += (double)h[i];
        for (int i = 0; i < 257; i++)
            if (h[i]) {
                const double p = (double)h[i] / total;
                e -= p * std::log2(p);
            }
        return e;
    }

    // The round-0 key code of a level-0 stream (alphabetic_code).  wj[j]: the sampled count of character j
    // (0: the padding, which counts nothing); the stream codes character j as j - dense (the padding stays
    // 0), so with dense words characters 0 and 1 share code 0.  Codes that the sample missed weigh as one in
    // 2^18 sampled characters, which keeps the codewords short; a code whose longest codeword would exceed
    // CODE_MAX_LEN bits is not used.  The device table (256 entries) goes to d_ctab; its host copy h_ctab
    // must stay until the copy has run (the calls are ordered on the stream, and a level builds one code).
    uint32_t h_ctab[256];
    KeyCode key_code(const std::vector<double>& wj, bool dense, DBuf<uint32_t>& d_ctab) {
        KeyCode kc;
        const uint32_t maxchar = (uint32_t)wj.size() - 1;
        const uint32_t m = maxchar + 1 - (dense ? 1 : 0);  // stream codes 0 .. m - 1
        if (maxchar < 1 || m < 2 || m > 256) return kc;
        std::vector<double> h(m, 0.0);
        for (uint32_t j = 0; j <= maxchar; j++) h[j == 0 ? 0 : j - (dense ? 1 : 0)] += wj[j];
        double total = 0;
        for (double x : h) total += x;
        if (total <= 0) return kc;
        const double floor_w = std::max(1.0, total / (double)(1u << 18));
        std::vector<double> w(m);
        double e = 0;
        for (uint32_t d = 0; d < m; d++) {
            w[d] = std::max(h[d], floor_w);
            if (h[d] > 0) e -= h[d] / total * std::log2(h[d] / total);
        }
        std::vector<uint32_t> len, cw;
        const uint32_t longest = alphabetic_code(w, CODE_MAX_LEN, len, cw);
        if (longest == 0) return kc;
        double avg = 0;
        for (uint32_t d = 0; d < m; d++) avg += h[d] * len[d];
        avg /= total;
        for (uint32_t c = 0; c < 256; c++) {
            const uint32_t d = c < m ? c : m - 1;  // codes past the alphabet do not occur
            h_ctab[c] = (len[d] << 24) | (cw[d] << (24 - len[d]));
        }
        d_ctab = DBuf<uint32_t>(cache, 256);
        DCX_HIP(hipMemcpyAsync(d_ctab.get(), h_ctab, sizeof(h_ctab), hipMemcpyHostToDevice, stream));
        kc.ctab = d_ctab.get();
        kc.max_len = longest;
        kc.min_len = *std::min_element(len.begin(), len.end());
        kc.avg_len = avg;
        kc.entropy = e;
        return kc;
    }
    // the same for a single text: characters 1..maxchar are the alphabet codes lut[b] (code_counts)
    KeyCode key_code_text(uint32_t maxchar, bool dense, DBuf<uint32_t>& d_ctab) {
        if (!have_counts || maxchar > 256) return KeyCode{};
        std::vector<double> wj(maxchar + 1, 0.0);
        for (uint32_t c = 1; c <= maxchar; c++) wj[c] = (double)code_counts[c];
        return key_code(wj, dense, d_ctab);
    }

    // the level's characters as a packed code stream (k_pack_text) of `chars` characters
    template <class CharT>
    void make_stream(const CharT* text, const uint16_t* lut, uint64_t n, uint32_t bits, uint32_t dense, uint64_t chars,
                     DBuf<uint64_t>& PT) {
        const uint64_t nwords = stream_words(chars, bits);
        PT = DBuf<uint64_t>(cache, nwords);
        const unsigned wgrid = (unsigned)std::min<uint64_t>(ceil_div(nwords, 256), 16384);
        switch (bits) {
            case 1: k_pack_aligned<CharT, 1><<<wgrid, 256, 0, stream>>>(text, n, lut, dense, nwords, PT.get()); return;
            case 2: k_pack_aligned<CharT, 2><<<wgrid, 256, 0, stream>>>(text, n, lut, dense, nwords, PT.get()); return;
            case 4: k_pack_aligned<CharT, 4><<<wgrid, 256, 0, stream>>>(text, n, lut, dense, nwords, PT.get()); return;
            case 8: k_pack_aligned<CharT, 8><<<wgrid, 256, 0, stream>>>(text, n, lut, dense, nwords, PT.get()); return;
            default: break;
        }
        const uint64_t chunks = ceil_div(nwords, (uint64_t)PK_CHARS * bits / 64);
        k_pack_text<CharT><<<(unsigned)std::min<uint64_t>(chunks, 16384), 256, 0, stream>>>(text, n, lut, bits, dense,
                                                                                             nwords, PT.get());
    }

    // characters of the level are 1..maxchar (0 is the padding)
    using StreamFn = std::function<void(uint32_t bits, uint32_t dense, uint64_t chars, DBuf<uint64_t>& PT)>;
    // A 64-bit SA in device memory for the build in progress (null outside build_sa): level 0 writes it
    // from the sample scan's second pass when it takes two passes, and sets out64_written
    uint64_t* out64 = nullptr;
    bool out64_written = false;
    struct Out64Scope {  // out64 for one build, cleared however the build ends
        uint64_t*& p;
        Out64Scope(uint64_t*& q, uint64_t* v) : p(q) { p = v; }
        ~Out64Scope() { p = nullptr; }
    };

    template <class CharT>
    // The SA goes to sa_out, or with sa_take != nullptr into *sa_take (the level's own buffer, no copy).
    // A caller that has the level's text only as pieces (a batch) passes `stream_fn`, which packs the
    // level's character stream of the given width, dense flag and length, `entropy` (of the codes) and the
    // round-0 key code of its characters (`batch_code`, or none).
    void level(const CharT* text, const uint16_t* lut, uint64_t n, uint32_t maxchar, int depth, uint32_t* sa_out,
               hipEvent_t* ev, DBuf<uint32_t>* sa_take = nullptr, const StreamFn* stream_fn = nullptr,
               double entropy = -1.0, const KeyCode* batch_code = nullptr) {
        const uint32_t X = dcs->X, D = dcs->D;
        const uint32_t bits = bit_width(maxchar);
        // dense merge words (c -> c - 1) when they save a bit per character
        const uint32_t dbits = maxchar > 1 ? bit_width(maxchar - 1) : 0;
        const bool dense = dbits >= 1 && dbits < bits;
        const uint32_t mbits = dense ? dbits : bits;
        dcx_level_stats* ls = level_stat(depth);
        if (ls) ls->n = n;
        const uint64_t threshold = opt.base_case_chars > 2ull * X ? opt.base_case_chars : 2ull * X;
        if (n <= threshold) {
            if (ls) ls->base_case = 1;
            if (n > BASE_CASE_MAX) {
                DBuf<uint32_t> sa;
                base_case_large(text, lut, n, maxchar, sa);
                if (sa_take)
                    *sa_take = std::move(sa);
                else
                    DCX_HIP(hipMemcpyAsync(sa_out, sa.get(), n * 4, hipMemcpyDeviceToDevice, stream));
                return;
            }
            if (sa_take) {
                *sa_take = DBuf<uint32_t>(cache, n > 0 ? n : 1);
                sa_out = sa_take->get();
            }
            base_case(text, lut, (uint32_t)n, sa_out);
            return;
        }
        uint64_t m64 = 0;
        DcDev dc = make_dcdev(n, &m64);
        const uint32_t m = (uint32_t)m64;

        // key lengths (materialize_sample / materialize_characters with the packing ratio)
        const bool ps = opt.pack_samples && depth == 0;
        const bool pm = opt.pack_merging && depth == 0;
        double ratio = 1.0;
        uint32_t pbits = 0;
        if (depth == 0 && (opt.pack_samples || opt.pack_merging)) {
            pbits = ref_pack_bits(sigma);
            const uint32_t cpw = 64 / pbits;
            const uint32_t w = (X + cpw - 1) / cpw;
            ratio = (double)(w * cpw) / (double)X;
        }
        const uint32_t Ls = (uint32_t)((double)X * (ps ? ratio : 1.0));
        const uint32_t Lm = (uint32_t)((double)X * (pm ? ratio : 1.0)) - 1;
        dc.bnd = (uint32_t)(n - Ls + 1);  // n <= 2^32 and Ls >= 2
        if (ls) {
            ls->samples = m;
            ls->key_chars = Ls;
            ls->merge_chars = Lm;
            ls->bits_per_char = (ps || pm) ? pbits : 0;
        }
        {  // algorithmic bytes of this level
            const uint64_t b = depth == 0 ? (pbits ? pbits : ref_pack_bits(sigma)) : bits;
            const uint64_t c = depth == 0 ? 1 : (bits + 7) / 8;
            const uint64_t K1 = 8 * ceil_div((uint64_t)X * b, 64), K4 = 8 * ceil_div((uint64_t)(X - 1) * b, 64);
            const uint64_t I = 4, S = depth == 0 ? sa_width : I;
            alg_bytes += n * c + 2ull * m * (K1 + I) + 2ull * n * (K4 + I) + n * S;
        }

        // ---------------- phase 1 + 2: all suffixes sorted by their first L_s characters ----------------
        // One radix refinement over every position yields both the sample names (the samples in this order,
        // grouped by equal keys) and the key groups that the merge resolves with the DC comparator.  Words
        // are dense-coded (character c as c - 1) when that saves a bit per character.
        DBuf<uint32_t> order;
        DBuf<uint8_t> H;
        ActiveSet tied;
        uint64_t* wide64 = nullptr;  // level 0: the 64-bit device output this level writes
        if (depth == 0) mark(1);
        {
            DBuf<uint64_t> PT;
            if (stream_fn)
                (*stream_fn)(mbits, dense ? 1 : 0, n + Ls + 2, PT);
            else
                make_stream(text, lut, n, mbits, dense ? 1 : 0, n + Ls + 2, PT);
            const double info = info_bits(text, lut, n, mbits, entropy);
            DBuf<uint32_t> d_ctab;
            const KeyCode kc = depth == 0 && mbits <= 8
                                   ? (batch_code ? *batch_code : !stream_fn ? key_code_text(maxchar, dense, d_ctab) : KeyCode{})
                                   : KeyCode{};
            refine_sort(PT.get(), n, Ls, mbits, dense, info, order, H, tied, kc);
            if (dense && tied.A > 0) boundary_fixup(PT.get(), mbits, Ls, n, dc.bnd, order, H, tied);
        }

        if (depth == 0) mark(2);
        DBuf<uint32_t> ranks(cache, (size_t)m + D);
        prim.fill_u32(ranks.get() + m, 0, D);
        {
            // samples in sorted order; the dummy sample (position n, key of padding only) is the smallest
            const bool dummy = dc.in_dc[n % X] != 0;
            DBuf<uint32_t> s_order(cache, m), s_gid(cache, m);
            if (dummy) {
                prim.fill_u32(s_order.get(), m - 1, 1);
                prim.fill_u32(s_gid.get(), 0, 1);
            }
            // level 0 with a 64-bit device output: the scan's down pass writes it from every order entry
            // (the tied slots change later and are written again after the merge)
            wide64 = depth == 0 && out64 && Primitives::scan_two_pass(n) ? out64 : nullptr;
            prim.scan(SampleScanOp{order.get(), H.get(), dc, dummy ? 1u : 0u, s_order.get(), s_gid.get(), wide64}, n,
                      0);
            DBuf<uint8_t> Hs(cache, (size_t)m + 1);
            k_heads_u32<<<grid(m / 16 + 1ull), 256, 0, stream>>>(s_gid.get(), m, Hs.get());
            s_gid.reset();
            DBuf<uint32_t> names(cache, m);
            prim.scan(NamesOp{Hs.get(), names.get()}, m, 0);
            const uint32_t M = (uint32_t)prim.totals(0).lo;
            if (ls) ls->names = M;
            if (ev && depth == 0) DCX_HIP(hipEventRecord(ev[2], stream));
            if (depth == 0) mark(3);

            // ---------------- phase 3: recursion ----------------
            if (M >= m) {
                scatter_u32(s_order.get(), nullptr, m, 1, ranks.get());
            } else {
                if (ls) ls->recursed = 1;
                if (M >= UNIQ_BIT) throw Error(DCX_E_UNSUPPORTED, "more than 2^31 distinct sample names");
                recursion(dc, n, m, M, depth, s_order, names, Hs, ranks);
            }
        }
        if (ev && depth == 0) DCX_HIP(hipEventRecord(ev[3], stream));
        if (depth == 0) mark(4);

        // ---------------- phase 4: merge ----------------
        // suffixes with different L_s-character keys are already in order; groups of equal keys are ordered
        // by the DC comparator on the sample ranks
        tie_break(dc, order, H, tied, ranks.get());
        if (wide64) {
            if (tied.A > 0) k_widen_slots<<<grid(tied.A), 256, 0, stream>>>(order.get(), tied.act.get(), tied.A, wide64);
            out64_written = true;
        }
        if (depth == 0) mark(5);
        if (sa_take)
            *sa_take = std::move(order);
        else
            DCX_HIP(hipMemcpyAsync(sa_out, order.get(), n * 4, hipMemcpyDeviceToDevice, stream));
    }

    // -------------------------------------------------------------------------------------------
    void set_options(const dcx_options* o) {
        opt.X = o->X;
        opt.pack_samples = o->pack_samples != 0;
        opt.pack_merging = o->pack_merging != 0;
        opt.discarding_threshold = o->discarding_threshold;
        opt.base_case_chars = o->base_case_chars;
        if (!(opt.discarding_threshold >= 0.0 && opt.discarding_threshold <= 1.0))
            throw Error(DCX_E_ARG, "discarding_threshold must lie in [0, 1]");
        set_dc(opt.X);
        if (opt.base_case_chars > BASE_CASE_CHARS_MAX)
            throw Error(DCX_E_ARG, "base_case_chars must be <= " + std::to_string(BASE_CASE_CHARS_MAX));
    }

    // compress_alphabet (cli.cpp:417-463): used bytes -> 1..sigma in byte order; returns sigma and the map
    uint32_t alphabet(const uint8_t* d_text, uint64_t n, DBuf<uint16_t>& dlut) {
        DBuf<unsigned long long> mask(cache, 4);
        prim.fill_u64(reinterpret_cast<uint64_t*>(mask.get()), 0, 4);
        if (n) k_byte_mask<<<grid(ceil_div(n, 16), 2048), 256, 0, stream>>>(d_text, n, mask.get());
        unsigned long long hm[4];
        memcpy(hm, prim.read_device(mask.get(), sizeof(hm)), sizeof(hm));
        uint16_t lut[256];
        uint32_t next = 1;
        for (int c = 0; c < 256; c++) lut[c] = ((hm[c >> 6] >> (c & 63)) & 1) ? (uint16_t)next++ : 0;
        dlut = DBuf<uint16_t>(cache, 256);
        DCX_HIP(hipMemcpyAsync(dlut.get(), lut, sizeof(lut), hipMemcpyHostToDevice, stream));
        return next - 1;
    }

    // -------------------------------------------------------------------------------------------
    // batched mode (dcx_build_sa_batch): sub-batches of texts joined by a separator below every
    // character, one DCX pipeline per sub-batch, then a stable radix partition by text
    // -------------------------------------------------------------------------------------------
    int build_sa_batch(const uint8_t* texts, const uint64_t* offsets, uint32_t count, const dcx_options* o,
                       uint32_t* sa_out, dcx_stats* stats) {
        auto t0 = std::chrono::steady_clock::now();
        DCX_HIP(hipSetDevice(device));
        if (!o || !offsets) throw Error(DCX_E_ARG, "null argument");
        set_options(o);
        st = stats;
        if (st) memset(st, 0, sizeof(*st));
        alg_bytes = 0;
        sa_width = 4;
        cache.reset_peak();
        if (count == 0) return DCX_OK;
        const uint64_t base = offsets[0], total = offsets[count] - offsets[0];
        for (uint32_t k = 0; k < count; k++) {
            if (offsets[k + 1] < offsets[k]) throw Error(DCX_E_ARG, "offsets must be non-decreasing");
            if (offsets[k + 1] - offsets[k] + 1 > MAX_TEXT)  // with its separator
                throw Error(DCX_E_UNSUPPORTED, "a text of the batch is too long for 32-bit positions");
        }
        if (total && (!texts || !sa_out)) throw Error(DCX_E_ARG, "null argument");
        // sub-batches of 2^24 characters by default: a shorter joined string needs fewer key bits in
        // round 0 and its texts fewer partition passes, and with two sub-batches in flight the fixed costs
        // of each overlap with the other's kernels (10,000 ACGTN texts, 5 GB, X = 39: 2^31 381 ms, 2^27
        // 299, 2^26 288, 2^25 294, 2^24 273, 2^23 299, 2^22 335)
        uint64_t limit = o->batch_max_chars ? o->batch_max_chars : (1ull << 24);
        if (limit > MAX_TEXT) limit = MAX_TEXT;  // the joined string, its separators included
        if (budget) {
            const uint64_t by_budget = budget > (64ull << 20) ? (budget - (64ull << 20)) / 64 : 0;
            if (by_budget < limit) limit = by_budget;
        }
        prim.drop_profile();
        prim.profile = o->profile_kernels != 0;
        struct ProfGuard {
            Primitives* p;
            ~ProfGuard() {
                p->profile = false;
                p->drop_profile();
            }
        } profg{&prim};
        hipEvent_t e0 = evt[0], e1 = evt[1];
        // texts into HBM
        DBuf<uint8_t> dtext;
        const uint8_t* d_texts = texts ? texts + base : nullptr;
        if (!o->text_on_device && total) {
            dtext = DBuf<uint8_t>(cache, total + 16);
            upload(dtext.get(), texts + base, total);
            d_texts = dtext.get();
        }
        DCX_HIP(hipEventRecord(e0, stream));
        DBuf<uint16_t> dlut;
        const uint32_t sig = alphabet(d_texts, total, dlut);
        if (st) st->sigma = sig;
        const uint32_t maxchar = sig + 1;  // characters 2..sig+1, separator 1
        DBuf<uint16_t> ident;
        if (maxchar <= 255) {
            uint16_t id[256];
            for (int c = 0; c < 256; c++) id[c] = (uint16_t)c;
            ident = DBuf<uint16_t>(cache, 256);
            DCX_HIP(hipMemcpyAsync(ident.get(), id, sizeof(id), hipMemcpyHostToDevice, stream));
        }
        // entropy of the codes, once for the batch (it sizes round 0 of the sub-batches of 2^24 characters
        // and more; shorter ones key at full width)
        const double entropy = total >= (1ull << 24) ? code_entropy(d_texts, dlut.get(), total) : 8.0;
        // the round-0 key code of the joined characters: text bytes as lut[b] + 1 (the sampled counts),
        // the separator once per text (sampled at the same rate), the padding
        DBuf<uint32_t> d_ctab;
        KeyCode bcode;
        if (total >= (1ull << 24) && have_counts && maxchar <= 255) {
            std::vector<double> wj(maxchar + 1, 0.0);
            double sampled = 0;
            for (uint32_t c = 1; c <= sig; c++) sampled += (double)code_counts[c];
            for (uint32_t c = 1; c <= sig; c++) wj[c + 1] = (double)code_counts[c];
            wj[1] = sampled * (double)count / (double)total;
            const uint32_t dbits = bit_width(maxchar - 1);
            bcode = key_code(wj, dbits >= 1 && dbits < bit_width(maxchar), d_ctab);
            // the second stream reads the table too: it must have arrived before that stream's first kernel
            DCX_HIP(hipStreamSynchronize(stream));
        }
        // sub-batches: greedy, at least one text, at most `limit` joined characters each
        std::vector<uint32_t> cuts{0};
        while (cuts.back() < count) {
            uint32_t k1 = cuts.back() + 1;
            uint64_t joined = offsets[k1] - offsets[k1 - 1] + 1;
            while (k1 < count && joined + (offsets[k1 + 1] - offsets[k1]) + 1 <= limit) {
                joined += offsets[k1 + 1] - offsets[k1] + 1;
                k1++;
            }
            cuts.push_back(k1);
        }
        const size_t nsub = cuts.size() - 1;
        const BatchInput in{d_texts, offsets, base, maxchar, dlut.get(), ident.get(), entropy, o->sa_on_device != 0,
                            sa_out, bcode.ctab ? &bcode : nullptr};
        // Two sub-batches in flight when there are several: a second engine on its own stream (created on
        // first use and kept with the handle) takes every other one from a shared counter, in a host thread
        // of its own, so one sub-batch's kernels run while the other waits for a readback.  Both write
        // disjoint parts of sa_out.  Not with an HBM budget that cannot hold two sub-batches.
        const bool two = nsub >= 2 && (!budget || budget >= 2 * (64 * limit + (64ull << 20)));
        if (!two) {
            for (size_t i = 0; i < nsub; i++) sub_batch(in, cuts[i], cuts[i + 1]);
        } else {
            if (!twin) twin.reset(new Engine(device));
            twin->set_options(o);
            twin->st = nullptr;
            twin->alg_bytes = 0;
            twin->sa_width = 4;
            twin->cache.reset_peak();
            DCX_HIP(hipStreamSynchronize(stream));  // the texts, the map and the identity map are in place
            std::atomic<size_t> next{0};
            std::atomic<bool> failed{false};
            int twin_code = DCX_OK;
            std::string twin_msg;
            const auto take = [&](Engine& e) {
                for (size_t i; !failed.load() && (i = next.fetch_add(1)) < nsub;) e.sub_batch(in, cuts[i], cuts[i + 1]);
            };
            std::thread worker([&] {
                try {
                    DCX_HIP(hipSetDevice(device));
                    take(*twin);
                } catch (const Error& e) {
                    twin_code = e.code;
                    twin_msg = e.what();
                    failed = true;
                } catch (const std::exception& e) {
                    twin_code = DCX_E_HIP;
                    twin_msg = e.what();
                    failed = true;
                }
            });
            struct Join {
                std::thread& t;
                std::atomic<bool>& f;
                ~Join() {
                    if (t.joinable()) {
                        f = true;
                        t.join();
                    }
                }
            } join{worker, failed};
            take(*this);
            worker.join();
            if (twin_code != DCX_OK) throw Error(twin_code, "second sub-batch stream: " + twin_msg);
            alg_bytes += twin->alg_bytes;
        }
        DCX_HIP(hipEventRecord(e1, stream));
        DCX_HIP(hipStreamSynchronize(stream));
        if (st) {
            float a = 0;
            DCX_HIP(hipEventElapsedTime(&a, e0, e1));
            st->ms_device = a;
            st->alg_bytes = alg_bytes;
            if (prim.profile) prim.collect_profile(&st->sort_passes, &st->sort_pass_items, &st->sort_pass_bytes, &st->sort_pass_ms);
            st->peak_device_bytes = cache.peak() + (twin ? twin->cache.peak() : 0);
            st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return DCX_OK;
    }

    // the inputs every sub-batch of one batched call shares (device pointers of the calling engine, read only)
    struct BatchInput {
        const uint8_t* d_texts;   // the batch's bytes, text k at d_texts + offsets[k] - base
        const uint64_t* offsets;  // host
        uint64_t base;
        uint32_t maxchar;         // joined characters: separator 1, text bytes lut[b] + 1 <= maxchar
        const uint16_t* lut;
        const uint16_t* ident;    // identity map (maxchar <= 255)
        double entropy;
        bool sa_on_device;
        uint32_t* sa_out;
        const KeyCode* code;      // round-0 key code of the joined characters, shared by the sub-batches
    };

    // One sub-batch, texts [k0, k1): joined with separators, one DCX pipeline, partition by text into
    // sa_out.  Returns once its part of sa_out is written.
    void sub_batch(const BatchInput& in, uint32_t k0, uint32_t k1) {
        const uint8_t* d_texts = in.d_texts;
        const uint64_t* offsets = in.offsets;
        const uint64_t base = in.base;
        const uint32_t maxchar = in.maxchar;
        const uint64_t base_threshold = opt.base_case_chars > 2ull * opt.X ? opt.base_case_chars : 2ull * opt.X;
        uint64_t joined = 0;
        for (uint32_t k = k0; k < k1; k++) joined += offsets[k + 1] - offsets[k] + 1;
        const uint32_t cnt = k1 - k0;
        std::vector<uint64_t> in_off(cnt + 1), co(cnt + 1);
        for (uint32_t k = 0; k <= cnt; k++) {
            in_off[k] = offsets[k0 + k] - base;
            co[k] = (offsets[k0 + k] - offsets[k0]) + k;
        }
        DBuf<uint64_t> d_in(cache, cnt + 1), d_co(cache, cnt + 1);
        DCX_HIP(hipMemcpyAsync(d_in.get(), in_off.data(), (cnt + 1) * 8, hipMemcpyHostToDevice, stream));
        DCX_HIP(hipMemcpyAsync(d_co.get(), co.data(), (cnt + 1) * 8, hipMemcpyHostToDevice, stream));
        const uint64_t N = joined;
        const uint64_t chars = offsets[k1] - offsets[k0];
        // joined text (short sub-batches and 32-bit characters): owner table of the input positions
        const auto owner_table_in = [&]() {
            const uint64_t in_chunks = (chars >> OWNER_SHIFT) + 1;
            DBuf<uint32_t> t(cache, in_chunks + 1);
            k_owner_table<<<grid(in_chunks + 1), 256, 0, stream>>>(d_in.get(), cnt, in_off[0], in_chunks, t.get());
            return t;
        };
        DBuf<uint32_t> sa;
        sigma = maxchar;
        if (maxchar <= 255 && N > base_threshold) {
            // the level's character stream straight from the texts (no joined text)
            const StreamFn fn = [&](uint32_t bits, uint32_t dense, uint64_t stream_chars, DBuf<uint64_t>& PT) {
                const uint64_t nwords = stream_words(stream_chars, bits);
                PT = DBuf<uint64_t>(cache, nwords);
                const uint64_t chunks = ceil_div(nwords, (uint64_t)PK_CHARS * bits / 64);
                k_batch_pack<<<(unsigned)std::min<uint64_t>(chunks, 16384), 256, 0, stream>>>(
                    d_texts, d_in.get(), d_co.get(), cnt, N, in.lut, bits, dense, nwords, PT.get());
            };
            level<uint8_t>(nullptr, in.ident, N, maxchar, 0, nullptr, nullptr, &sa, &fn, in.entropy, in.code);
        } else if (maxchar <= 255) {
            DBuf<uint32_t> t_in = owner_table_in();
            DBuf<uint8_t> jt(cache, (size_t)N + 16);
            if (chars)
                k_batch_text<uint8_t><<<grid(chars), 256, 0, stream>>>(d_texts, in_off[0], in_off[cnt], d_in.get(),
                                                                        d_co.get(), cnt, in.lut, t_in.get(),
                                                                        jt.get());
            k_batch_seps<uint8_t><<<grid(cnt), 256, 0, stream>>>(d_co.get(), cnt, jt.get());
            level<uint8_t>(jt.get(), in.ident, N, maxchar, 0, nullptr, nullptr, &sa);
        } else {
            DBuf<uint32_t> t_in = owner_table_in();
            DBuf<uint32_t> jt(cache, (size_t)N + 16);
            if (chars)
                k_batch_text<uint32_t><<<grid(chars), 256, 0, stream>>>(d_texts, in_off[0], in_off[cnt],
                                                                         d_in.get(), d_co.get(), cnt, in.lut,
                                                                         t_in.get(), jt.get());
            k_batch_seps<uint32_t><<<grid(cnt), 256, 0, stream>>>(d_co.get(), cnt, jt.get());
            level<uint32_t>(jt.get(), nullptr, N, maxchar, 0, nullptr, nullptr, &sa);
        }
        // stable partition of the joined SA by text: keys (text << 32 | position in the text), the
        // separators keyed past the last text; the last pass keeps the low words of the first `chars`
        // keys, written straight into a device SA
        if (chars) {
            DBuf<uint64_t> ka(cache, N), kb(cache, N);
            if (cnt <= BK_MAX_TEXTS) {
                uint32_t cshift = 0;
                while (((N - 1) >> cshift) + 1 > BK_TABLE) cshift++;
                k_batch_keys_lds<<<2 * prim.cus, BK_THREADS, 0, stream>>>(sa.get(), N, d_co.get(), cnt, cshift,
                                                                          ka.get());
            } else {
                const uint64_t co_chunks = ((uint64_t)N >> OWNER_SHIFT) + 1;
                DBuf<uint32_t> t_co(cache, co_chunks + 1);
                k_owner_table<<<grid(co_chunks + 1), 256, 0, stream>>>(d_co.get(), cnt, 0, co_chunks, t_co.get());
                k_batch_keys64<<<grid(N), 256, 0, stream>>>(sa.get(), N, d_co.get(), cnt, t_co.get(), ka.get());
            }
            sa.reset();
            uint32_t* dst = in.sa_out + (offsets[k0] - base);
            DBuf<uint32_t> staged;
            if (!in.sa_on_device) staged = DBuf<uint32_t>(cache, chars);
            uint32_t* out = in.sa_on_device ? dst : staged.get();
            prim.sort_keys_low32(ka.get(), kb.get(), N, 32, 32 + bit_width(cnt), out, chars);
            if (!in.sa_on_device) download(dst, out, chars * 4);
        }
        DCX_HIP(hipStreamSynchronize(stream));
    }

    // -------------------------------------------------------------------------------------------
    int build_sa(const uint8_t* text, uint64_t n, const dcx_options* o, void* sa_out, uint32_t width, dcx_stats* stats) {
        auto t0 = std::chrono::steady_clock::now();
        DCX_HIP(hipSetDevice(device));
        if (!o || (!text && n) || (!sa_out && n)) throw Error(DCX_E_ARG, "null argument");
        if (width != 4 && width != 5 && width != 8) throw Error(DCX_E_ARG, "sa_width_bytes must be 4, 5 or 8");
        set_options(o);
        if (n > MAX_TEXT)
            throw Error(DCX_E_UNSUPPORTED, "texts longer than 2^32 characters are not supported by this build");
        if (width == 4 && n > (1ull << 32)) throw Error(DCX_E_WIDTH, "32-bit SA cannot index this text");
        if (budget) {
            uint64_t need = estimate_bytes(n);
            if (need > budget)
                throw Error(DCX_E_ADMISSION, "footprint " + std::to_string(need) + " exceeds budget " +
                                                 std::to_string(budget));
        }
        st = stats;
        if (st) memset(st, 0, sizeof(*st));
        alg_bytes = 0;
        sa_width = width;
        cache.reset_peak();

        hipEvent_t* ev = evt;

        DCX_HIP(hipEventRecord(ev[0], stream));
        if (n == 0) {
            if (dcx_level_stats* ls = level_stat(0)) ls->base_case = 1;
            DCX_HIP(hipStreamSynchronize(stream));
            return DCX_OK;
        }
        prim.drop_profile();
        prim.profile = o->profile_kernels != 0;
        struct ProfGuard {
            Primitives* p;
            ~ProfGuard() {
                p->profile = false;
                p->drop_profile();
            }
        } profg{&prim};
        // text into HBM (the engine never modifies the caller's text)
        DBuf<uint8_t> dtext;
        const uint8_t* d_text = text;
        if (!o->text_on_device) {
            dtext = DBuf<uint8_t>(cache, n + 16);
            upload(dtext.get(), text, n);
            d_text = dtext.get();
        }
        DCX_HIP(hipEventRecord(ev[1], stream));  // text resident
        DBuf<uint16_t> dlut;
        sigma = alphabet(d_text, n, dlut);
        if (st) st->sigma = sigma;

        DBuf<uint32_t> sa;
        // phase boundaries default to "empty" (a base-case level never reaches them)
        DCX_HIP(hipEventRecord(ev[2], stream));
        DCX_HIP(hipEventRecord(ev[3], stream));
        // a 64-bit device output is written by level 0 itself when it takes the sample scan's two passes
        out64_written = false;
        {
            Out64Scope scope(out64, o->sa_on_device && width == 8 && (reinterpret_cast<uintptr_t>(sa_out) & 7) == 0
                                        ? static_cast<uint64_t*>(sa_out)
                                        : nullptr);
            level<uint8_t>(d_text, dlut.get(), n, sigma, 0, nullptr, ev, &sa);
        }
        DCX_HIP(hipEventRecord(ev[4], stream));  // SA resident in HBM
        // output in the requested width
        if (o->sa_on_device && out64_written) {
            // written during the level
        } else if (o->sa_on_device) {
            widen(sa.get(), n, width, static_cast<uint8_t*>(sa_out));
        } else if (width == 4) {
            download(sa_out, sa.get(), n * 4);
        } else if (n * 4 >= stage_bytes / 4) {
            download_widen(sa_out, sa.get(), n, width);
        } else {
            DBuf<uint8_t> wide(cache, n * width);
            widen(sa.get(), n, width, wide.get());
            download(sa_out, wide.get(), n * width);
        }
        mark(6);
        DCX_HIP(hipEventRecord(ev[5], stream));
        DCX_HIP(hipStreamSynchronize(stream));
        if (st) {
            float a = 0;
            DCX_HIP(hipEventElapsedTime(&a, ev[1], ev[4]));
            st->ms_device = a;
            DCX_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            float h2d = a;
            DCX_HIP(hipEventElapsedTime(&a, ev[1], ev[2]));
            st->ms_phase[DCX_PHASE_SAMPLES] = a;
            st->ms_phase[DCX_PHASE_PREPARATION] = h2d;
            DCX_HIP(hipEventElapsedTime(&a, ev[2], ev[3]));
            st->ms_phase[DCX_PHASE_RECURSION] = a;
            DCX_HIP(hipEventElapsedTime(&a, ev[3], ev[4]));
            st->ms_phase[DCX_PHASE_MERGE] = a;
            DCX_HIP(hipEventElapsedTime(&a, ev[4], ev[5]));
            st->ms_phase[DCX_PHASE_OUTPUT] = a;
            st->alg_bytes = alg_bytes;
            if (prim.profile) prim.collect_profile(&st->sort_passes, &st->sort_pass_items, &st->sort_pass_bytes, &st->sort_pass_ms);
            st->peak_device_bytes = cache.peak();
            st->ms_total =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return DCX_OK;
    }

    // dcx_build_sa_noalloc: the same build of a device-resident text into a device SA, with every device
    // buffer carved from this GPU's scratch array (the cache in arena mode) and the lookback status sized
    // at handle creation: nothing is allocated.  Calls on one GPU are serialized (they share the array).
    int build_sa_noalloc(const uint8_t* d_text, uint64_t n, const dcx_options* o, void* d_sa, uint32_t width,
                         dcx_stats* stats) {
        if (!o) throw Error(DCX_E_ARG, "null argument");
        if (n > DCX_NOALLOC_MAX_CHARS)
            throw Error(DCX_E_UNSUPPORTED, "dcx_build_sa_noalloc takes texts of at most " +
                                               std::to_string(DCX_NOALLOC_MAX_CHARS) + " characters, not " +
                                               std::to_string(n));
        if (!scratch) throw Error(DCX_E_HIP, "the scratch array of this device could not be looked up");
        dcx_options oo = *o;
        oo.text_on_device = 1;
        oo.sa_on_device = 1;
        oo.profile_kernels = 0;
        std::lock_guard<std::mutex> lock(scratch_call_mutex(device));
        cache.set_arena(scratch, DCX_NOALLOC_SCRATCH_BYTES);
        prim.no_alloc = true;
        struct ArenaGuard {
            Engine* e;
            ~ArenaGuard() {
                // the device is done with the array before the next call may use it; a comparison table
                // carved from it is dropped (the next call builds it again)
                (void)hipStreamSynchronize(e->stream);
                if (e->d_cmp.get() && e->cache.in_arena(e->d_cmp.get())) {
                    e->d_cmp.reset();
                    e->cmp_X = 0;
                }
                e->prim.no_alloc = false;
                e->cache.clear_arena();
            }
        } guard{this};
        return build_sa(d_text, n, &oo, d_sa, width, stats);
    }

    // conservative device footprint of one build (level 0 dominates; deeper levels shrink by D/X)
    uint64_t estimate_bytes(uint64_t n) const {
        return n * 64 + (64ull << 20);
    }

    int check_sa(const uint8_t* text, uint64_t n, const void* sa, uint32_t width, int on_device, int* ok) {
        DCX_HIP(hipSetDevice(device));
        if (!ok || (n && (!text || !sa))) throw Error(DCX_E_ARG, "null argument");
        if (width != 4 && width != 5 && width != 8) throw Error(DCX_E_ARG, "sa_width_bytes must be 4, 5 or 8");
        if (n > MAX_TEXT) throw Error(DCX_E_UNSUPPORTED, "text too long for this build");
        *ok = 1;
        if (n == 0) return DCX_OK;
        DBuf<uint8_t> dt, ds;
        const uint8_t* d_text = text;
        const uint8_t* d_sa = static_cast<const uint8_t*>(sa);
        if (!on_device) {
            dt = DBuf<uint8_t>(cache, n);
            ds = DBuf<uint8_t>(cache, n * width);
            DCX_HIP(hipMemcpyAsync(dt.get(), text, n, hipMemcpyHostToDevice, stream));
            DCX_HIP(hipMemcpyAsync(ds.get(), sa, n * width, hipMemcpyHostToDevice, stream));
            d_text = dt.get();
            d_sa = ds.get();
        }
        // inverse SA (0-based ranks: n may be 2^32) and a bit per position for the permutation test
        DBuf<uint32_t> rank(cache, n);
        DBuf<uint32_t> seen(cache, n / 32 + 1);
        DBuf<uint32_t> bad(cache, 1);
        prim.fill_u32(seen.get(), 0, n / 32 + 1);
        prim.fill_u32(bad.get(), 0, 1);
        k_check_perm<<<grid(n), 256, 0, stream>>>(d_sa, width, n, rank.get(), seen.get(), bad.get());
        k_check_order<<<grid(n), 256, 0, stream>>>(d_sa, width, n, d_text, rank.get(), bad.get());
        const uint32_t hb = (uint32_t)prim.read_device(bad.get(), 4)[0];
        *ok = hb == 0;
        return DCX_OK;
    }
};

}  // namespace dcx

struct dcx_handle {
    dcx::Engine* eng;
};

using dcx::Error;

extern "C" {

void dcx_default_options(dcx_options* opt) {
    if (!opt) return;
    memset(opt, 0, sizeof(*opt));
    opt->X = 39;
    opt->pack_samples = 1;
    opt->pack_merging = 1;
    opt->discarding_threshold = 0.7;
    opt->base_case_chars = 10000;
}

int dcx_create(dcx_handle** out, int device, uint64_t hbm_budget_bytes) {
    if (!out) return DCX_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return DCX_E_NO_GPU;
    }
    if (device < 0 || device >= count) return DCX_E_ARG;
    try {
        dcx_handle* h = new dcx_handle;
        h->eng = new dcx::Engine(device);
        h->eng->budget = hbm_budget_bytes;
        *out = h;
        return DCX_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (...) {
        return DCX_E_HIP;
    }
}

int dcx_destroy(dcx_handle* h) {
    if (!h) return DCX_E_ARG;
    delete h->eng;
    delete h;
    return DCX_OK;
}

int dcx_build_sa(dcx_handle* h, const uint8_t* text, uint64_t n, const dcx_options* opt, void* sa_out,
                 uint32_t sa_width_bytes, dcx_stats* stats) {
    if (!h || !h->eng) return DCX_E_ARG;
    try {
        h->eng->last_error.clear();
        return h->eng->build_sa(text, n, opt, sa_out, sa_width_bytes, stats);
    } catch (const Error& e) {
        h->eng->last_error = e.what();
        (void)hipStreamSynchronize(h->eng->stream);
        (void)hipGetLastError();
        return e.code;
    } catch (const std::exception& e) {
        h->eng->last_error = e.what();
        return DCX_E_HIP;
    }
}

int dcx_build_sa_batch(dcx_handle* h, const uint8_t* texts, const uint64_t* offsets, uint32_t count,
                       const dcx_options* opt, uint32_t* sa_out, dcx_stats* stats) {
    if (!h || !h->eng) return DCX_E_ARG;
    try {
        h->eng->last_error.clear();
        return h->eng->build_sa_batch(texts, offsets, count, opt, sa_out, stats);
    } catch (const Error& e) {
        h->eng->last_error = e.what();
        (void)hipStreamSynchronize(h->eng->stream);
        (void)hipGetLastError();
        return e.code;
    } catch (const std::exception& e) {
        h->eng->last_error = e.what();
        return DCX_E_HIP;
    }
}

int dcx_build_sa_noalloc(dcx_handle* h, const uint8_t* text, uint64_t n, const dcx_options* opt, void* sa_out,
                         uint32_t sa_width_bytes, dcx_stats* stats) {
    if (!h || !h->eng) return DCX_E_ARG;
    try {
        h->eng->last_error.clear();
        return h->eng->build_sa_noalloc(text, n, opt, sa_out, sa_width_bytes, stats);
    } catch (const Error& e) {
        h->eng->last_error = e.what();
        (void)hipGetLastError();
        return e.code;
    } catch (const std::exception& e) {
        h->eng->last_error = e.what();
        return DCX_E_HIP;
    }
}

int dcx_check_sa(dcx_handle* h, const uint8_t* text, uint64_t n, const void* sa, uint32_t sa_width_bytes, int on_device,
                 int* ok) {
    if (!h || !h->eng) return DCX_E_ARG;
    try {
        h->eng->last_error.clear();
        return h->eng->check_sa(text, n, sa, sa_width_bytes, on_device, ok);
    } catch (const Error& e) {
        h->eng->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        h->eng->last_error = e.what();
        return DCX_E_HIP;
    }
}

const char* dcx_strerror(int code) {
    switch (code) {
        case DCX_OK: return "ok";
        case DCX_E_INVALID_X: return "unsupported difference cover X";
        case DCX_E_ALPHABET: return "alphabet cannot be represented";
        case DCX_E_NOMEM: return "device memory exhausted";
        case DCX_E_HIP: return "HIP runtime error";
        case DCX_E_IO: return "file I/O error";
        case DCX_E_ARG: return "invalid argument";
        case DCX_E_WIDTH: return "SA width too small for the text length";
        case DCX_E_ADMISSION: return "request exceeds the HBM budget";
        case DCX_E_UNSUPPORTED: return "unsupported size";
        case DCX_E_CANCELLED: return "request cancelled";
        case DCX_E_NO_GPU: return "no HIP device available";
        default: return "unknown error";
    }
}

const char* dcx_last_error(dcx_handle* h) {
    if (!h || !h->eng) return "";
    return h->eng->last_error.c_str();
}

}  // extern "C"
