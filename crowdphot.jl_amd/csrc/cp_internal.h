// cp_internal.h -- context, budgeted device memory and error plumbing shared by
// the HIP translation units of libcrowdphot_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/crowdphot_hip.h"

namespace cp {

// Blocks per CU of a launch: the context's tuned value, else the default.
inline int64_t tuned_bpc(int32_t v, int64_t dflt) { return v > 0 ? (int64_t)v : dflt; }

// Thrown inside the library, converted to a return code at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CP_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::cp::Error(CP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CP_ARG(cond, msg)                                    \
    do {                                                     \
        if (!(cond)) throw ::cp::Error(CP_EARG, (msg));      \
    } while (0)

}  // namespace cp

struct cp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;   // 0 = no cap
    uint64_t in_use = 0;   // bytes currently allocated through DevBuf
    uint64_t peak = 0;
    uint64_t peak_resident = 0;   // most in_use + cached
    // released device blocks kept for reuse (size -> pointer); every use is
    // ordered on `stream`, so a block can be handed out again immediately
    std::multimap<size_t, void*> cache;
    uint64_t cached = 0;
    std::string err;
    // pinned host staging for small scalar read-backs
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // pinned double buffer (2 x STAGE_CHUNK) for result read-back
    void* stage = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int n_cu = 0;
    cp_tuning tuning{};   // launch shapes (0 = default), cp_set_tuning
    // per-kernel HIP-event timing (cp_profile)
    bool prof = false;
    struct Pending {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Pending> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, std::pair<double, long>> prof_acc;  // name -> (total ms, launches)
};

namespace cp {

// Allocation granule: small requests round to a power of two (>= 256 B),
// larger ones to 2 MiB, so repeated calls with similar sizes hit the cache.
inline size_t alloc_granule(size_t b) {
    if (b <= (size_t(1) << 21)) {
        size_t g = 256;
        while (g < b) g <<= 1;
        return g;
    }
    const size_t m = size_t(1) << 21;
    return (b + m - 1) / m * m;
}

inline void cache_trim(cp_ctx* c, uint64_t need) {
    // free cached blocks, largest first, until `need` more bytes fit the budget
    while (!c->cache.empty() && (need == UINT64_MAX || c->in_use + c->cached + need > c->budget)) {
        auto it = std::prev(c->cache.end());
        (void)hipFree(it->second);
        c->cached -= it->first;
        c->cache.erase(it);
    }
}

// A cached block of at least g and at most 1.25 g bytes, or null; its size in *got.
inline void* cache_take(cp_ctx* c, size_t g, size_t* got) {
    auto it = c->cache.lower_bound(g);
    if (it == c->cache.end() || it->first > g + g / 4) return nullptr;
    void* p = it->second;
    *got = it->first;
    c->cached -= it->first;
    c->cache.erase(it);
    return p;
}

// Device allocation charged against the context's HBM budget.
class DevBuf {
   public:
    DevBuf() = default;
    DevBuf(cp_ctx* c, size_t bytes) { alloc(c, bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            ctx_ = o.ctx_;
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            gbytes_ = o.gbytes_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
            o.gbytes_ = 0;
        }
        return *this;
    }
    void alloc(cp_ctx* c, size_t bytes) {
        release();
        ctx_ = c;
        if (bytes == 0) return;
        size_t g = alloc_granule(bytes);
        if (c->budget && c->in_use + g > c->budget)
            throw Error(CP_ENOMEM, "HBM budget exceeded: need " + std::to_string(g) +
                                       " more bytes, " + std::to_string(c->budget - c->in_use) +
                                       " left of " + std::to_string(c->budget));
        size_t got = 0;
        void* p = cache_take(c, g, &got);
        if (p && c->budget && c->in_use + got > c->budget) {
            // the cached block is larger than what the budget leaves: put it
            // back and allocate the granule itself (the trim may free it)
            c->cache.emplace(got, p);
            c->cached += got;
            p = nullptr;
        }
        if (p) {
            g = got;   // charge the block's real size
        } else {
            if (c->budget) cache_trim(c, g);
            hipError_t e = hipMalloc(&p, g);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                cache_trim(c, UINT64_MAX);
                e = hipMalloc(&p, g);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(CP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
            }
        }
        ptr_ = p;
        bytes_ = bytes;
        gbytes_ = p ? g : 0;
        c->in_use += g;
        if (c->in_use > c->peak) c->peak = c->in_use;
        if (c->in_use + c->cached > c->peak_resident) c->peak_resident = c->in_use + c->cached;
    }
    void release() {
        if (ptr_) {
            ctx_->in_use -= gbytes_;
            ctx_->cache.emplace(gbytes_, ptr_);
            ctx_->cached += gbytes_;
        }
        ptr_ = nullptr;
        bytes_ = 0;
        gbytes_ = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

   private:
    cp_ctx* ctx_ = nullptr;
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
    size_t gbytes_ = 0;   // size of the underlying block (granule or cached block)
};

// Context of the C-ABI call running on this thread (for launch timing).
extern thread_local cp_ctx* tl_ctx;

// Fold finished launch timings into the context's per-kernel totals.
inline void prof_flush(cp_ctx* c) {
    if (!c || c->prof_pending.empty()) return;
    for (auto& p : c->prof_pending) {
        if (hipEventSynchronize(p.b) == hipSuccess) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                auto& e = c->prof_acc[p.name];
                e.first += ms;
                e.second += 1;
            }
        }
        c->ev_pool.push_back(p.a);
        c->ev_pool.push_back(p.b);
    }
    c->prof_pending.clear();
}

// Run fn, converting exceptions to a return code + message on the context.
template <typename F>
int guarded(cp_ctx* ctx, F&& fn) {
    try {
        tl_ctx = ctx;
        fn();
        if (ctx && ctx->prof) prof_flush(ctx);
        tl_ctx = nullptr;
        if (ctx) ctx->err.clear();
        return CP_OK;
    } catch (const Error& e) {
        tl_ctx = nullptr;
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        tl_ctx = nullptr;
        if (ctx) ctx->err = e.what();
        return CP_EHIP;
    }
}

inline void* pinned(cp_ctx* c, size_t bytes) {
    if (c->pinned_bytes < bytes) {
        if (c->pinned) CP_HIP(hipHostFree(c->pinned));
        c->pinned = nullptr;
        CP_HIP(hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
        c->pinned_bytes = bytes;
    }
    return c->pinned;
}

inline int popcount32(uint32_t v) { return __builtin_popcount(v); }

// Host memcpy split over OpenMP threads (host_copy.cpp).
void par_memcpy(void* dst, const void* src, size_t bytes);

// One device-to-host piece of a result read-back.
struct D2HSeg {
    void* dst;
    const void* src;
    size_t bytes;
};
constexpr size_t STAGE_CHUNK = size_t(4) << 20;
// Copy every segment to its (pageable) host destination through the
// context's pinned double buffer: the DMA of one chunk overlaps the
// multi-threaded host copy of the previous one.  Segments adjacent on the
// device share a chunk.  Returns when every destination is written.
void staged_d2h(cp_ctx* c, hipStream_t st, const std::vector<D2HSeg>& segs);

}  // namespace cp
