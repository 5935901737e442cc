// Exclusive prefix sum, uint32 in -> uint64 out (n+1 entries, out[n] = total).
// Three launches (block reduce, scan of block sums, block rescan), no host
// synchronisation and no memset, so it can sit inside a captured HIP graph.
#include "jobs.hpp"

namespace vf {

namespace {
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
static_assert(kScanItems % 4 == 0, "a thread's items are loaded as 16-byte vectors");

// a thread's items whole and both arrays 16-byte aligned: 16-byte loads and stores (a quarter and half
// as many accesses as single items)
__device__ inline bool vec_ok(const void* a, const void* b, uint64_t base, uint64_t n)
{
    return base + kScanItems <= n && (((uintptr_t)a | (uintptr_t)b) & 15u) == 0;
}

__device__ inline void load_items(const uint32_t* in, uint64_t base, uint64_t n, bool vec, uint32_t v[kScanItems])
{
    if (vec) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (uint32_t q = 0; q < kScanItems / 4; ++q) {
            const uint4 a = p[q];
            v[4 * q] = a.x; v[4 * q + 1] = a.y; v[4 * q + 2] = a.z; v[4 * q + 3] = a.w;
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
}

__device__ inline unsigned long long block_excl_scan(unsigned long long v, unsigned long long* smem,
                                                     unsigned long long* total)
{
    // wave64 inclusive scan via DPP-free shuffles, then across the 4 waves
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long x = v;
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) smem[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long acc = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
            unsigned long long t = smem[w];
            smem[w] = acc;
            acc += t;
        }
        smem[kScanThreads / 64] = acc;
    }
    __syncthreads();
    unsigned long long r = x - v + smem[wave];
    *total = smem[kScanThreads / 64];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* in, uint64_t n, unsigned long long* sums)
{
    __shared__ unsigned long long smem[kScanThreads / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    load_items(in, base, n, vec_ok(in, in, base, n), v);
    unsigned long long s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) s += v[k];
    unsigned long long total;
    (void)block_excl_scan(s, smem, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(unsigned long long* sums, uint32_t nb)
{
    __shared__ unsigned long long smem[kScanThreads / 64 + 1];
    unsigned long long carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanThreads) {
        uint32_t i = base + threadIdx.x;
        unsigned long long v = i < nb ? sums[i] : 0;
        unsigned long long total;
        unsigned long long e = block_excl_scan(v, smem, &total);
        if (i < nb) sums[i] = e + carry;
        carry += total;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* in, uint64_t n, const unsigned long long* sums,
                                                            unsigned long long* out, uint32_t nb)
{
    __shared__ unsigned long long smem[kScanThreads / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    const bool vec = vec_ok(in, out, base, n);
    uint32_t v[kScanItems];
    load_items(in, base, n, vec, v);
    unsigned long long s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) s += v[k];
    unsigned long long total;
    unsigned long long e = block_excl_scan(s, smem, &total) + sums[blockIdx.x];
    if (vec) {
        uint4* o = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (uint32_t q = 0; q < kScanItems / 2; ++q) {
            const unsigned long long e0 = e, e1 = e + v[2 * q];
            e = e1 + v[2 * q + 1];
            o[q] = make_uint4((uint32_t)e0, (uint32_t)(e0 >> 32), (uint32_t)e1, (uint32_t)(e1 >> 32));
        }
    } else {
        for (uint32_t k = 0; k < kScanItems; ++k) {
            if (base + k < n) out[base + k] = e;
            e += v[k];
        }
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = sums[nb];
}

__global__ void k_scan_empty(unsigned long long* out) { out[0] = 0; }
}  // namespace

size_t scan_tmp_bytes(uint64_t n)
{
    uint64_t nb = (n + kScanTile - 1) / kScanTile;
    return (size_t)(nb + 1) * sizeof(unsigned long long);
}

void scan_exclusive_u32_to_u64(const uint32_t* in, unsigned long long* out, uint64_t n, void* tmp, hipStream_t s)
{
    if (n == 0) {
        TimedLaunch tl("scan", s);
        k_scan_empty<<<1, 1, 0, s>>>(out);
        VF_CHECK_LAUNCH();
        return;
    }
    uint32_t nb = (uint32_t)((n + kScanTile - 1) / kScanTile);
    auto* sums = static_cast<unsigned long long*>(tmp);
    TimedLaunch tl("scan", s);
    k_scan_reduce<<<nb, kScanThreads, 0, s>>>(in, n, sums);
    VF_CHECK_LAUNCH();
    k_scan_sums<<<1, kScanThreads, 0, s>>>(sums, nb);
    VF_CHECK_LAUNCH();
    k_scan_down<<<nb, kScanThreads, 0, s>>>(in, n, sums, out, nb);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
