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
    std::string err;
    // pinned host staging for small scalar read-backs
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    int n_cu = 0;
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
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void alloc(cp_ctx* c, size_t bytes) {
        release();
        ctx_ = c;
        if (bytes == 0) return;
        if (c->budget && c->in_use + bytes > c->budget)
            throw Error(CP_ENOMEM, "HBM budget exceeded: need " + std::to_string(bytes) +
                                       " more bytes, " + std::to_string(c->budget - c->in_use) +
                                       " left of " + std::to_string(c->budget));
        hipError_t e = hipMalloc(&ptr_, bytes);
        if (e != hipSuccess) {
            ptr_ = nullptr;
            throw Error(CP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        bytes_ = bytes;
        c->in_use += bytes;
        if (c->in_use > c->peak) c->peak = c->in_use;
    }
    void release() {
        if (ptr_) {
            (void)hipFree(ptr_);
            ctx_->in_use -= bytes_;
        }
        ptr_ = nullptr;
        bytes_ = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

   private:
    cp_ctx* ctx_ = nullptr;
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
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

}  // namespace cp
