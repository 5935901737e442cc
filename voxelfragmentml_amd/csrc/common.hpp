// Shared host/device infrastructure for libvfrag: error propagation across the
// C ABI, HIP error checks, per-kernel event timing, and the small device
// helpers (index math, counter hash) every stage uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "vfrag.h"

namespace vf {

// ---------------------------------------------------------------- errors ---
struct Exception {
    int code;
};

void set_error(const char* fmt, ...);
[[noreturn]] void fail(int code, const char* fmt, ...);
[[noreturn]] void fail_hip(hipError_t e, const char* what, const char* file, int line);

#define VF_HIP(call)                                                  \
    do {                                                              \
        hipError_t vf_e_ = (call);                                    \
        if (vf_e_ != hipSuccess) ::vf::fail_hip(vf_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define VF_CHECK_LAUNCH() VF_HIP(hipGetLastError())

// Runs f, converting library exceptions into status codes for the C ABI.
template <class F>
int guard(F&& f)
{
    try {
        f();
        return VF_OK;
    } catch (const Exception& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return VF_ERR_OUT_OF_MEMORY;
    } catch (...) {
        set_error("unexpected C++ exception");
        return VF_ERR_INVALID_ARG;
    }
}

// ---------------------------------------------------------------- timing ---
// When enabled, launches wrapped in TimedLaunch are bracketed by HIP events
// recorded on the launch stream; totals accumulate per kernel name.
bool timing_enabled();
void timing_begin(const char* name, hipStream_t s);
void timing_end(const char* name, hipStream_t s);
bool capturing(hipStream_t s);
// While one lives, the calling thread's launches record no timing events (which would create events).
struct TimingOff {
    TimingOff();
    ~TimingOff();
};

struct TimedLaunch {
    const char* name;
    hipStream_t s;
    bool on;
    TimedLaunch(const char* n, hipStream_t st) : name(n), s(st), on(timing_enabled() && !capturing(st))
    {
        if (on) timing_begin(name, s);
    }
    ~TimedLaunch()
    {
        if (on) timing_end(name, s);
    }
};

// ------------------------------------------------------------- device mem ---
// A grow-only device buffer owned by a grid / batch.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Device blocks by size class, kept for reuse.  A class is the size rounded up
// to an eighth of its power of two (at most 12.5 % more); a released block
// waits in its class's list for the next request of that class, so a steady
// stream of jobs allocates nothing once each class has been seen.  Every use
// is ordered on the engine's one stream, so a block released after queued work
// is reused only by work queued later.
class BlockCache {
public:
    static size_t class_size(size_t n);
    void* get(size_t n, size_t* got);  // a block of class_size(n) bytes; hipMalloc on a miss
    void put(void* p, size_t bytes);   // bytes: the block's class size
    size_t held() const { return held_; }      // in use + cached
    size_t cached() const { return cached_; }
    void trim(size_t keep);                    // hipFree cached blocks, largest first, until at most keep bytes stay cached
    BlockCache() = default;
    BlockCache(const BlockCache&) = delete;
    BlockCache& operator=(const BlockCache&) = delete;
    ~BlockCache() { trim(0); }

private:
    std::map<size_t, std::vector<void*>> free_;
    size_t held_ = 0, cached_ = 0;
};

// A grow-only scratch buffer from the engine's block cache.
struct CacheBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n, BlockCache& c);
    void release(BlockCache& c);
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline uint32_t next_pow2(uint32_t v)
{
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// ------------------------------------------------------------ device math ---
// A generic pointer to device memory as a global-address-space pointer, so the
// access compiles to global_* instead of flat_* instructions.  Pointers read
// from a job descriptor staged in LDS are generic; flat accesses also count
// against the LDS wait counter, which makes every LDS read wait for earlier
// outstanding stores.
#define VF_AS_GLOBAL __attribute__((address_space(1)))
template <typename U, typename T>
__device__ inline VF_AS_GLOBAL U* gcast(T* p)
{
    return (VF_AS_GLOBAL U*)p;
}
// 16- and 8-byte global loads / stores (HIP's vector structs cannot bind global references)
typedef unsigned int vf_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int vf_u32x2 __attribute__((ext_vector_type(2)));
__device__ inline uint4 gload4(const void* p)
{
    const vf_u32x4 v = *(const VF_AS_GLOBAL vf_u32x4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ inline void gstore4(void* p, uint4 q) { *(VF_AS_GLOBAL vf_u32x4*)p = vf_u32x4{ q.x, q.y, q.z, q.w }; }
__device__ inline uint2 gload2(const void* p)
{
    const vf_u32x2 v = *(const VF_AS_GLOBAL vf_u32x2*)p;
    return make_uint2(v.x, v.y);
}
__device__ inline void gstore2(void* p, uint2 q) { *(VF_AS_GLOBAL vf_u32x2*)p = vf_u32x2{ q.x, q.y }; }

__host__ __device__ inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Counter-based hash: murmur3's finaliser applied twice, keyed by 64 bits.
// Drives the flood boundary noise and the seed draws.
__host__ __device__ inline uint32_t hash32(uint64_t key, uint32_t counter)
{
    uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
    return fmix32(fmix32(counter ^ lo) + hi * 0x9E3779B1u + 0x7F4A7C15u);
}

// Uniform integer in [0, amp] from the cell's hash (multiply-shift range map).
__host__ __device__ inline uint32_t noise_delta(uint64_t seed, uint32_t idx, uint32_t amp)
{
    return (uint32_t)(((uint64_t)hash32(seed, idx) * (uint64_t)(amp + 1u)) >> 32);
}

constexpr uint16_t kEmpty = 0;
constexpr uint16_t kFree = 1;
constexpr uint16_t kMask15 = 0x7FFF;

// Pick the per-launch grid for a grid-stride kernel.
inline uint32_t stride_blocks(uint64_t items, uint32_t threads, uint32_t cap = 4096)
{
    uint64_t b = (items + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (uint32_t)b;
}

}  // namespace vf
