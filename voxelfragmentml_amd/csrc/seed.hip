// Device-side uniform seeding: Seeder::uniform (MF/Source/Fracturer/Seeder.cpp:154-208)
// with a counter-based draw so results are reproducible across platforms.
// Candidates are occupied cells (optionally only surface / only interior
// cells by RegularGrid::isBoundary, RegularGrid.cpp:543-558).  The j-th draw
// is rank floor(hash32(key, j) * O / 2^32) among the O candidates in linear
// order; repeated ranks are skipped until S distinct ranks are accepted (at
// most 10^6 draws, Seeder.cpp MAX_TRIES), then seeds are emitted in (x,y,z)
// order with labels 2..S+1.
#include "cellops.hpp"

namespace vf {

namespace {

constexpr uint32_t kSeedThreads = 256;
constexpr uint32_t kMaxDraws = 1000000u;

__device__ inline bool surface_cell(const JobDev* J, uint32_t x, uint32_t y, uint32_t z)
{
    const int nx = (int)J->nx, ny = (int)J->ny, nz = (int)J->nz;
    const int x0 = max((int)x - 1, 0), x1 = min((int)x + 1, nx - 1);
    const int y0 = max((int)y - 1, 0), y1 = min((int)y + 1, ny - 1);
    const int z0 = max((int)z - 1, 0), z1 = min((int)z + 1, nz - 1);
    for (int a = x0; a <= x1; ++a)
        for (int b = y0; b <= y1; ++b)
            for (int c = z0; c <= z1; ++c)
                if (J->occ[((size_t)a * ny + b) * nz + c] == 0) return true;
    return false;
}

__device__ inline bool candidate(const JobDev* J, uint64_t c, uint32_t x, uint32_t y, uint32_t z)
{
    if (J->occ[c] == 0) return false;
    if (J->seed_location == VF_SEED_BOTH) return true;
    const bool s = surface_cell(J, x, y, z);
    return (J->seed_location == VF_SEED_OUTER) == s;
}

__global__ __launch_bounds__(kCellThreads) void k_seed_count(Launch L, uint32_t* tile_cnt)
{
    __shared__ uint32_t wsum[kCellThreads / 64];
    __shared__ JobDev sJ;
    walk_tiles(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t j, uint32_t local, uint32_t t) {
        uint32_t n = 0;
        const uint64_t c0 = (uint64_t)local * kTileCells + threadIdx.x * kCellsPerThread;
        if (J.active && c0 < J.V) {
            const uint64_t o = load8b(J.occ, c0, J.V);
            if (J.seed_location == VF_SEED_BOTH) {
                n = __popcll(o & 0x0101010101010101ull);
            } else if (o) {
                const JobDev* Jg = L.jobs + j;
                uint32_t x, y, z;
                decompose(c0, J, x, y, z);
                for (int k = 0; k < 8 && c0 + k < J.V; ++k) {
                    if ((o >> (8 * k)) & 0xFFu) n += candidate(Jg, c0 + k, x, y, z) ? 1u : 0u;
                    advance(x, y, z, J);
                }
            }
        }
        for (int ofs = 32; ofs > 0; ofs >>= 1) n += __shfl_xor(n, ofs, 64);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < kCellThreads / 64; ++w) tot += wsum[w];
            tile_cnt[t] = tot;
        }
        __syncthreads();
    });
}

// One block per job.
__global__ __launch_bounds__(kSeedThreads) void k_seed_pick(Launch L, const unsigned long long* tile_off)
{
    __shared__ uint32_t chosen[kMaxDeviceSeeds];
    __shared__ uint32_t sorted[kMaxDeviceSeeds];
    __shared__ uint32_t cand[kSeedThreads];
    __shared__ uint32_t wsum[kSeedThreads / 64];
    __shared__ uint32_t s_got;
    const uint32_t j = blockIdx.x;
    const JobDev* J = L.jobs + j;
    if (!J->active) return;
    const uint32_t S = J->S;
    const uint32_t t0 = L.tile_prefix[j], t1 = L.tile_prefix[j + 1];
    const unsigned long long base = tile_off[t0];
    const unsigned long long O = tile_off[t1] - base;
    if (S == 0) return;
    // on failure every seed is parked on cell (0,0,0) so later stages stay in bounds
    auto park = [&]() {
        if (threadIdx.x == 0) atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrSeeding);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) J->seeds[i] = make_uint4(0u, 0u, 0u, 2u + (i & 0xFFu));
    };
    if (O < S || O > 0xFFFFFFFFull || S > kMaxDeviceSeeds) {
        park();
        return;
    }
    if (threadIdx.x == 0) s_got = 0;
    __syncthreads();
    for (uint32_t j0 = 0; j0 < kMaxDraws; j0 += kSeedThreads) {
        const uint32_t got = s_got;
        if (got >= S) break;
        const uint32_t draw = j0 + threadIdx.x;
        const uint32_t r = (uint32_t)(((unsigned long long)hash32(J->seed_key, draw) * O) >> 32);
        cand[threadIdx.x] = r;
        __syncthreads();
        bool dup = draw >= kMaxDraws;
        for (uint32_t k = 0; k < got && !dup; ++k) dup = chosen[k] == r;
        for (uint32_t k = 0; k < threadIdx.x && !dup; ++k) dup = cand[k] == r;
        // block exclusive scan of the accept flags
        uint32_t a = dup ? 0u : 1u, x = a;
        const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t wofs = 0, tot = 0;
        for (uint32_t w = 0; w < kSeedThreads / 64; ++w) {
            if (w < wave) wofs += wsum[w];
            tot += wsum[w];
        }
        const uint32_t pos = got + wofs + x - a;
        if (a && pos < S) chosen[pos] = r;
        __syncthreads();
        if (threadIdx.x == 0) s_got = min(S, got + tot);
        __syncthreads();
    }
    if (s_got < S) {
        park();
        return;
    }
    // sort the distinct ranks: position = number of smaller ranks
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        const uint32_t r = chosen[i];
        uint32_t p = 0;
        for (uint32_t k = 0; k < S; ++k) p += chosen[k] < r ? 1u : 0u;
        sorted[p] = r;
    }
    __syncthreads();
    // rank -> cell: tile by binary search over the tile prefix, then a scan of the tile
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        const unsigned long long want = base + sorted[i];
        uint32_t lo = t0, hi = t1;  // largest t with tile_off[t] <= want
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tile_off[mid] <= want) lo = mid; else hi = mid;
        }
        uint32_t need = (uint32_t)(want - tile_off[lo]);
        uint64_t c = (uint64_t)(lo - t0) * kTileCells;
        uint32_t x, y, z;
        decompose(c, J, x, y, z);
        bool found = false;
        for (; c < J->V; ++c) {
            if (candidate(J, c, x, y, z)) {
                if (need == 0) {
                    found = true;
                    break;
                }
                --need;
            }
            advance(x, y, z, J);
        }
        if (found) {
            J->seeds[i] = make_uint4(x, y, z, 2u + i);
        } else {
            J->seeds[i] = make_uint4(0u, 0u, 0u, 2u + i);
            atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrSeeding);
        }
    }
}

}  // namespace

void launch_seed_count(const Launch& L, uint32_t* tile_cnt)
{
    TimedLaunch tl("seed_count", L.s);
    const uint32_t blocks = std::max(1u, std::min(L.total_tiles, active_shape().cell_blocks));
    k_seed_count<<<blocks, kCellThreads, 0, L.s>>>(L, tile_cnt);
    VF_CHECK_LAUNCH();
}

void launch_seed_pick(const Launch& L, const unsigned long long* tile_off)
{
    TimedLaunch tl("seed_pick", L.s);
    k_seed_pick<<<L.njobs, kSeedThreads, 0, L.s>>>(L, tile_off);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
