// common.hpp -- error handling, device memory cache and wavefront helpers of the DCX engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <iterator>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace dcx {

// Errors never cross the C ABI as exceptions: dcx_build_sa catches these and returns the code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DCX_HIP(x)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw ::dcx::Error(e_ == hipErrorOutOfMemory ? -3 : -4,                                 \
                               std::string(#x) + ": " + hipGetErrorString(e_));                     \
    } while (0)

// Stream-ordered caching allocator.  Every engine kernel runs on the handle's single stream, so a
// block returned to the cache can be handed out again at once: later users are ordered after the
// earlier kernels on that stream.
//
// Arena mode (set_arena): every allocation is carved from a fixed device range given by the caller
// (first fit over the free ranges, frees coalesce with their neighbours) and nothing reaches hipMalloc or
// hipFree; a request that does not fit throws DCX_E_NOMEM.  Blocks handed out before the arena was set go
// back to the cache as usual.
class DeviceCache {
  public:
    ~DeviceCache() { release_all(); }
    void set_arena(void* base, size_t bytes) {
        arena_ = static_cast<char*>(base);
        arena_bytes_ = bytes & ~size_t(255);
        arena_free_.clear();
        arena_used_.clear();
        arena_free_[0] = arena_bytes_;
    }
    // leaves arena mode; returns the number of arena blocks still held (0 unless a buffer outlived its call)
    size_t clear_arena() {
        const size_t held = arena_used_.size();
        in_use_ -= arena_in_use_;
        arena_in_use_ = 0;
        arena_ = nullptr;
        arena_bytes_ = 0;
        arena_free_.clear();
        arena_used_.clear();
        return held;
    }
    bool in_arena(const void* p) const {
        const char* c = static_cast<const char*>(p);
        return arena_ && c >= arena_ && c < arena_ + arena_bytes_;
    }
    void* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        if (arena_) return arena_alloc(bytes);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= bytes * 2 + (1u << 20)) {
            void* p = it->second;
            size_t sz = it->first;
            free_.erase(it);
            used_[p] = sz;
            in_use_ += sz;
            peak_ = in_use_ > peak_ ? in_use_ : peak_;
            return p;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            // drop cached blocks and retry once
            trim();
            e = hipMalloc(&p, bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(-3, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
            }
        }
        used_[p] = bytes;
        in_use_ += bytes;
        total_ += bytes;
        peak_ = in_use_ > peak_ ? in_use_ : peak_;
        return p;
    }
    void free(void* p) {
        if (!p) return;
        if (in_arena(p)) {
            arena_release(p);
            return;
        }
        auto it = used_.find(p);
        if (it == used_.end()) return;
        in_use_ -= it->second;
        free_.emplace(it->second, p);
        used_.erase(it);
    }
    void trim() {
        for (auto& kv : free_) {
            (void)hipFree(kv.second);
            total_ -= kv.first;
        }
        free_.clear();
    }
    void release_all() {
        trim();
        for (auto& kv : used_) (void)hipFree(kv.first);
        used_.clear();
        in_use_ = 0;
        total_ = 0;
    }
    size_t peak() const { return peak_; }
    void reset_peak() { peak_ = in_use_; }
    size_t in_use() const { return in_use_; }

  private:
    void* arena_alloc(size_t bytes) {
        for (auto it = arena_free_.begin(); it != arena_free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            arena_free_.erase(it);
            if (len > bytes) arena_free_[off + bytes] = len - bytes;
            arena_used_[off] = bytes;
            arena_in_use_ += bytes;
            in_use_ += bytes;
            peak_ = in_use_ > peak_ ? in_use_ : peak_;
            return arena_ + off;
        }
        throw Error(-3, "scratch arena of " + std::to_string(arena_bytes_) + " bytes cannot hold another " +
                            std::to_string(bytes) + " bytes");
    }
    void arena_release(void* p) {
        const size_t off = static_cast<size_t>(static_cast<char*>(p) - arena_);
        auto u = arena_used_.find(off);
        if (u == arena_used_.end()) return;
        size_t start = off, len = u->second;
        arena_used_.erase(u);
        arena_in_use_ -= len;
        in_use_ -= len;
        auto next = arena_free_.lower_bound(start);
        if (next != arena_free_.end() && next->first == start + len) {  // merge with the following range
            len += next->second;
            next = arena_free_.erase(next);
        }
        if (next != arena_free_.begin()) {  // and with the preceding one
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                arena_free_.erase(prev);
            }
        }
        arena_free_[start] = len;
    }

    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> used_;
    size_t in_use_ = 0, peak_ = 0, total_ = 0;
    char* arena_ = nullptr;
    size_t arena_bytes_ = 0, arena_in_use_ = 0;
    std::map<size_t, size_t> arena_free_;  // offset -> length of the free ranges
    std::map<size_t, size_t> arena_used_;  // offset -> length of the blocks handed out
};

// RAII device buffer drawn from a DeviceCache.
template <typename T>
struct DBuf {
    DeviceCache* cache = nullptr;
    T* ptr = nullptr;
    size_t count = 0;
    DBuf() = default;
    DBuf(DeviceCache& c, size_t n) : cache(&c), count(n) { ptr = static_cast<T*>(c.alloc(sizeof(T) * (n ? n : 1))); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept { *this = std::move(o); }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            reset();
            cache = o.cache;
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.cache = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (cache && ptr) cache->free(ptr);
        ptr = nullptr;
        count = 0;
    }
    T* get() const { return ptr; }
    operator T*() const { return ptr; }
};

// ------------------------------------------------------------------------------------------------
// Device helpers (wave64)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Exclusive scan across a workgroup of NT threads (NT multiple of 64).  s_tmp holds NT/64 words.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    constexpr int NW = NT / 64;
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        uint32_t x = s_tmp[w];
        if (w < (int)wave) wave_off += x;
        tot += x;
    }
    __syncthreads();
    if (total) *total = tot;
    return wave_off + inc - v;
}

template <int NT>
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* s_tmp, uint64_t* total) {
    constexpr int NW = NT / 64;
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint64_t inc = wave_incl_scan64(v);
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint64_t wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        uint64_t x = s_tmp[w];
        if (w < (int)wave) wave_off += x;
        tot += x;
    }
    __syncthreads();
    if (total) *total = tot;
    return wave_off + inc - v;
}

__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline uint32_t bit_width(uint64_t v) {
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace dcx
