// Host side of the shared infrastructure: thread-local error text, HIP error
// mapping, grow-only device buffers and the per-kernel event timer.
#include "common.hpp"

#include <map>
#include <mutex>
#include <vector>

namespace vf {

static thread_local std::string t_error;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_error = buf;
}

void fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_error = buf;
    throw Exception{ code };
}

void fail_hip(hipError_t e, const char* what, const char* file, int line)
{
    (void)hipGetLastError();  // clear sticky launch error state where possible
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    throw Exception{ e == hipErrorOutOfMemory ? VF_ERR_OUT_OF_MEMORY : VF_ERR_DEVICE };
}

size_t BlockCache::class_size(size_t n)
{
    if (n <= 256) return 256;
    size_t p = 1;
    while (p <= n / 2) p <<= 1;  // largest power of two <= n
    const size_t step = std::max<size_t>(p / 8, 256);
    return (n + step - 1) / step * step;
}

void* BlockCache::get(size_t n, size_t* got)
{
    const size_t c = class_size(n);
    *got = c;
    auto it = free_.find(c);
    if (it != free_.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        cached_ -= c;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess && cached_) {  // out of memory with blocks cached: give them back and try once more
        (void)hipGetLastError();
        trim(0);
        e = hipMalloc(&p, c);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(VF_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed: %s", c, hipGetErrorString(e));
    }
    held_ += c;
    return p;
}

void BlockCache::put(void* p, size_t bytes)
{
    if (!p) return;
    free_[bytes].push_back(p);
    cached_ += bytes;
}

void BlockCache::trim(size_t keep)
{
    for (auto it = free_.rbegin(); it != free_.rend() && cached_ > keep; ++it)
        while (!it->second.empty() && cached_ > keep) {
            (void)hipFree(it->second.back());  // synchronises: every queued use has finished
            it->second.pop_back();
            cached_ -= it->first;
            held_ -= it->first;
        }
}

void CacheBuf::ensure(size_t n, BlockCache& c)
{
    if (n <= bytes && p) return;
    release(c);
    p = c.get(n, &bytes);
}

void CacheBuf::release(BlockCache& c)
{
    c.put(p, bytes);
    p = nullptr;
    bytes = 0;
}

void DevBuf::ensure(size_t n)
{
    if (n <= bytes && p) return;
    release();
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        bytes = 0;
        fail(VF_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// ------------------------------------------------------------------ timing
namespace {
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
struct Totals {
    double ms = 0;
    uint64_t n = 0;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<Pending> g_pending;
std::map<std::string, Totals> g_totals;
std::map<std::string, hipEvent_t> g_open;  // name -> start event awaiting its end
std::vector<hipEvent_t> g_free_events;

hipEvent_t take_event()
{
    if (!g_free_events.empty()) {
        hipEvent_t e = g_free_events.back();
        g_free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    VF_HIP(hipEventCreate(&e));
    return e;
}

void drain_locked()
{
    for (auto& p : g_pending) {
        VF_HIP(hipEventSynchronize(p.b));
        float ms = 0;
        VF_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto& t = g_totals[p.name];
        t.ms += ms;
        t.n += 1;
        g_free_events.push_back(p.a);
        g_free_events.push_back(p.b);
    }
    g_pending.clear();
}
}  // namespace

thread_local int t_timing_off = 0;
bool timing_enabled() { return g_timing && !t_timing_off; }
TimingOff::TimingOff() { ++t_timing_off; }
TimingOff::~TimingOff() { --t_timing_off; }

bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
    return st != hipStreamCaptureStatusNone;
}

void timing_begin(const char* name, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_tmu);
    hipEvent_t a = take_event();
    VF_HIP(hipEventRecord(a, s));
    g_open[name] = a;
}

void timing_end(const char* name, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_tmu);
    auto it = g_open.find(name);
    if (it == g_open.end()) return;
    hipEvent_t b = take_event();
    VF_HIP(hipEventRecord(b, s));
    g_pending.push_back({ name, it->second, b });
    g_open.erase(it);
    if (g_pending.size() > 4096) drain_locked();
}

}  // namespace vf

extern "C" {

const char* vf_last_error(void) { return vf::t_error.c_str(); }

int vf_version(void) { return 1; }

int vf_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int vf_timing_enable(int enable)
{
    return vf::guard([&] {
        std::lock_guard<std::mutex> lk(vf::g_tmu);
        vf::g_timing = enable != 0;
    });
}

int vf_timing_reset(void)
{
    return vf::guard([&] {
        std::lock_guard<std::mutex> lk(vf::g_tmu);
        vf::drain_locked();
        vf::g_totals.clear();
    });
}

int vf_timing_read(const char* kernel, double* total_ms, uint64_t* launches)
{
    return vf::guard([&] {
        if (!kernel) vf::fail(VF_ERR_INVALID_ARG, "kernel name is NULL");
        std::lock_guard<std::mutex> lk(vf::g_tmu);
        vf::drain_locked();
        auto it = vf::g_totals.find(kernel);
        if (total_ms) *total_ms = it == vf::g_totals.end() ? 0.0 : it->second.ms;
        if (launches) *launches = it == vf::g_totals.end() ? 0 : it->second.n;
    });
}

}  // extern "C"
