// Naive fracture of a caller-owned device grid without any device allocation.
//
// vf_fracture_naive works on a grid handle whose occupancy bytes, launch
// descriptor and seed copy live in buffers the library allocated.  This
// variant takes the cells as a plain device pointer and keeps everything else
// in one file-scope device array, `scratch`, of a fixed size: the descriptor,
// the one-job tile prefix, the seeds and the occupancy bytes of at most
// kNoAllocMaxCells cells.  The kernels are the ones vf_fracture_naive runs
// (labels -> occupancy, then the naive labelling), given the array's device
// address, which hipGetSymbolAddress returns once per device.  Every call on a
// device shares that device's array; calls that overlap on one device race.
#include <cstring>
#include <mutex>

#include "jobs.hpp"
#include "tune.hpp"

namespace vf {

namespace {

constexpr uint64_t kNoAllocMaxCells = VF_NOALLOC_MAX_CELLS;
// layout of the array: descriptor, tile prefix and statistics, seeds, occupancy bytes
constexpr size_t kOffJob = 0;
constexpr size_t kOffPrefix = 512;
constexpr size_t kOffStats = 576;
constexpr size_t kOffSeeds = 1024;
constexpr size_t kOffOcc = kOffSeeds + (size_t)kMaxDeviceSeeds * 16;
constexpr size_t kScratchBytes = kOffOcc + kNoAllocMaxCells;
static_assert(sizeof(JobDev) <= kOffPrefix - kOffJob, "descriptor slot");
static_assert(kOffStats + kStatCount * 8 <= kOffSeeds, "statistics slot");
static_assert(kOffOcc % 256 == 0, "occupancy bytes are 256-byte aligned");
static_assert(kScratchBytes == VF_NOALLOC_SCRATCH_BYTES, "the header states the array's size");

// The array itself.  Host code never uses this name as an address: its device address
// comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};

// Device address of this device's array, looked up on the first call on each device.
uint8_t* scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(VF_ERR_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            fail(VF_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        }
        g_ptr[device] = static_cast<uint8_t*>(p);
    }
    return g_ptr[device];
}

// Host staging of the descriptor, prefix, statistics and seeds (thread-local, no heap).
struct Staging {
    uint8_t bytes[kOffOcc];
};
thread_local Staging t_stage;

}  // namespace

}  // namespace vf

using namespace vf;

extern "C" int vf_fracture_naive_noalloc(uint16_t* d_cells, const uint32_t dims[3], const vf_seed* seeds, uint32_t S,
                                         int distance, void* stream)
{
    return guard([&] {
        if (!d_cells || !dims) fail(VF_ERR_INVALID_ARG, "NULL argument");
        if (reinterpret_cast<uintptr_t>(d_cells) % 16 != 0) fail(VF_ERR_INVALID_ARG, "cells must be 16-byte aligned");
        for (int q = 0; q < 3; ++q)
            if (dims[q] == 0 || dims[q] > 2048) fail(VF_ERR_INVALID_ARG, "grid edge %u out of range [1, 2048]", dims[q]);
        const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2];
        if (V > kNoAllocMaxCells)
            fail(VF_ERR_INVALID_ARG, "grid of %llu cells exceeds the %llu cells of the no-allocation scratch",
                 (unsigned long long)V, (unsigned long long)kNoAllocMaxCells);
        if (distance < VF_EUCLIDEAN || distance > VF_CHEBYSHEV) fail(VF_ERR_INVALID_ARG, "unknown distance %d", distance);
        if (S > kMaxDeviceSeeds) fail(VF_ERR_INVALID_ARG, "at most %u seeds", kMaxDeviceSeeds);
        if (S && !seeds) fail(VF_ERR_INVALID_ARG, "seeds is NULL");
        for (uint32_t i = 0; i < S; ++i)
            if (seeds[i].x >= dims[0] || seeds[i].y >= dims[1] || seeds[i].z >= dims[2])
                fail(VF_ERR_INVALID_ARG, "seed %u (%u, %u, %u) lies outside the grid", i, seeds[i].x, seeds[i].y, seeds[i].z);
        if (S == 0) return;  // buildCPU leaves every cell untouched without seeds
        int device = 0;
        VF_HIP(hipGetDevice(&device));
        uint8_t* base = scratch_ptr(device);
        hipStream_t s = static_cast<hipStream_t>(stream);

        // descriptor, one-job tile prefix, zeroed statistics and seeds, staged on the host and copied once
        const NaiveKeyForm kf = naive_key_form(dims, distance, S);
        JobDev J{};
        J.nx = dims[0];
        J.ny = dims[1];
        J.nz = dims[2];
        J.wp = next_pow2(ceil_div(dims[2], 32));
        J.V = V;
        J.labels = d_cells;
        J.occ = base + kOffOcc;
        J.seeds = reinterpret_cast<uint4*>(base + kOffSeeds);
        J.S = S;
        J.S_draw = S;
        J.dist = kf.dist;
        J.key_bits = key_bits_for(S);
        J.naive_key64 = kf.key64;
        J.stats = reinterpret_cast<unsigned long long*>(base + kOffStats);
        J.active = 1;
        J.algo = VF_NAIVE;
        const uint32_t tiles = ceil_div(V, kTileCells);
        uint8_t* st = t_stage.bytes;
        std::memset(st, 0, kOffSeeds);
        std::memcpy(st + kOffJob, &J, sizeof J);
        const uint32_t prefix[2] = { 0, tiles };
        std::memcpy(st + kOffPrefix, prefix, sizeof prefix);
        std::memcpy(st + kOffSeeds, seeds, (size_t)S * sizeof(vf_seed));
        VF_HIP(hipMemcpyAsync(base, st, kOffSeeds + (size_t)S * sizeof(vf_seed), hipMemcpyHostToDevice, s));

        set_active_bucket(tune_bucket(std::max(dims[0], std::max(dims[1], dims[2]))));
        TimingOff no_events;  // per-kernel timing would create HIP events
        Launch L;
        L.jobs = reinterpret_cast<const JobDev*>(base + kOffJob);
        L.njobs = 1;
        L.tile_prefix = reinterpret_cast<const uint32_t*>(base + kOffPrefix);
        L.total_tiles = tiles;
        L.s = s;
        launch_labels_to_occ(L);
        launch_naive(L, kf.key64 != 0);
        VF_HIP(hipStreamSynchronize(s));  // the staging is reused by the thread's next call
    });
}
