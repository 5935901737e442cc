// Flood (region-growing) fracture: FloodFracturer::build
// (MF/Source/Fracturer/FloodFracturer.cpp:98-191, floodFracturer-comp.glsl:24-63,
// undoMask-comp.glsl unmaskRightMost).
//
// The reference grows every seed one BFS step per dispatch and resolves
// simultaneous claims by a race.  Here every cell gets the label of the seed
// minimising (D, seed index), D the node-weighted geodesic distance in the 6-
// or 26-neighbour graph of occupied cells with weight 1 + noise(cell) in
// {1, 2, 3}.  Because weights belong to the target cell, the first level that
// reaches a cell fixes its D, and all claims of that level race only on the
// seed index, which atomicMin on the packed key (D << b | index) settles
// exactly.  So the result is deterministic and equals a brute-force
// per-seed Dijkstra.
//
// Levels are processed level-synchronously with a ring of 4 bucket arrays
// (D, D+1, D+2, D+3 are the only live levels).  The level kernel is
// self-advancing: its last block bumps each job's level, resets the counter
// slot that level D+4 will use and marks drained jobs, so the host (or a
// captured graph) simply replays the same launch.
#include "cellops.hpp"

namespace vf {

namespace {

constexpr uint32_t kFloodThreads = 256;
constexpr uint32_t kEmptyKey = 0u;
constexpr uint32_t kUnreached = 0xFFFFFFFFu;

__constant__ int c_nb26[26][3] = {
    { -1, -1, -1 }, { 1, 1, 1 }, { -1, -1, 0 }, { 1, 1, 0 }, { -1, -1, 1 }, { 1, 1, -1 }, { 0, -1, -1 },
    { 0, 1, 1 },    { 0, -1, 0 }, { 0, 1, 0 },  { 0, -1, 1 }, { 0, 1, -1 }, { 1, -1, -1 }, { -1, 1, 1 },
    { 1, -1, 0 },   { -1, 1, 0 }, { 1, -1, 1 }, { -1, 1, -1 }, { -1, 0, -1 }, { 1, 0, 1 }, { -1, 0, 0 },
    { 1, 0, 0 },    { -1, 0, 1 }, { 1, 0, -1 }, { 0, 0, -1 }, { 0, 0, 1 }
};
__constant__ int c_nb6[6][3] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };

// Keys live in 4x4x4 bricks (64 cells, 256 bytes): a cell's 26 neighbours fall
// in one or two bricks instead of nine z-rows.  Dims are padded to multiples
// of 4; padded cells are never visited (neighbour bounds use the real dims).
__device__ inline uint32_t kidx(const JobDev* J, uint32_t x, uint32_t y, uint32_t z)
{
    const uint32_t BY = (J->ny + 3) >> 2, BZ = (J->nz + 3) >> 2;
    return ((((x >> 2) * BY + (y >> 2)) * BZ + (z >> 2)) << 6) | ((x & 3u) << 4) | ((y & 3u) << 2) | (z & 3u);
}

__device__ inline void kdecode(const JobDev* J, uint32_t e, uint32_t& x, uint32_t& y, uint32_t& z)
{
    const uint32_t BY = (J->ny + 3) >> 2, BZ = (J->nz + 3) >> 2;
    const uint32_t brick = e >> 6, off = e & 63u;
    const uint32_t bz = brick % BZ, r = brick / BZ;
    const uint32_t by = r % BY, bx = r / BY;
    x = (bx << 2) | (off >> 4);
    y = (by << 2) | ((off >> 2) & 3u);
    z = (bz << 2) | (off & 3u);
}

__global__ __launch_bounds__(kCellThreads) void k_flood_init(Launch L)
{
    for (uint32_t t = blockIdx.x; t < L.total_tiles; t += gridDim.x) {
        const uint32_t j = find_job(L.tile_prefix, L.njobs, t);
        const JobDev* J = L.jobs + j;
        if (!J->active || J->algo != VF_FLOOD) continue;
        const uint64_t c0 = (uint64_t)(t - L.tile_prefix[j]) * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J->V) continue;
        const uint64_t o = load8b(J->occ, c0, J->V);
        uint32_t x, y, z;
        decompose(c0, J, x, y, z);
        for (int k = 0; k < 8; ++k) {
            if (c0 + k >= J->V) break;
            J->keys[kidx(J, x, y, z)] = ((o >> (8 * k)) & 0xFFu) ? kUnreached : kEmptyKey;
            advance(x, y, z, J);
        }
    }
}

// One block per job: stamp the seeds (a later seed on the same cell wins, as
// FloodFracturer.cpp:99-105 writes them in list order), mark seed cells
// occupied, and push level 1 (level 0 is never used so that 0 can mean EMPTY).
__global__ __launch_bounds__(kFloodThreads) void k_flood_seed(Launch L)
{
    const JobDev* J = L.jobs + blockIdx.x;
    if (!J->active || J->algo != VF_FLOOD) return;
    const uint32_t S = J->S, b = J->key_bits;
    for (uint32_t i = threadIdx.x; i < kFloodCtlWords; i += blockDim.x) J->fctl[i] = 0;
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        J->keys[kidx(J, q.x, q.y, q.z)] = 0u;
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        atomicMax(J->keys + kidx(J, q.x, q.y, q.z), (1u << b) | s);
    }
    __syncthreads();
    __shared__ uint32_t npush;
    if (threadIdx.x == 0) npush = 0;
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        const uint32_t kc = kidx(J, q.x, q.y, q.z);
        if (J->keys[kc] == ((1u << b) | s)) {
            const uint32_t slot = atomicAdd(&npush, 1u);
            J->buckets[(size_t)(1 % kLevelRing) * J->bucket_cap + slot] = kc;
            J->occ[((size_t)q.x * J->ny + q.y) * J->nz + q.z] = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        J->fctl[0] = 1;
        J->fctl[1] = npush ? 0u : 1u;
        J->stats[kStatFloodDone] = npush ? 0ull : 1ull;
        J->fctl[2 + 1 % 8] = npush;
    }
}

// Wave-aggregated append of cell u to bucket (job j, level du).  All lanes of
// the wave call it; `want` marks lanes with something to push.
__device__ inline void push_aggregated(bool want, const JobDev* jobs, uint32_t j, uint32_t du, uint32_t u)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint64_t pending = __ballot(want);
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((long long)pending) - 1;
        const uint32_t lj = __shfl(j, leader, 64), ld = __shfl(du, leader, 64);
        const bool mine = want && j == lj && du == ld;
        const uint64_t grp = __ballot(mine);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(jobs[lj].fctl + 2 + (ld & 7u), (uint32_t)__popcll(grp));
        base = __shfl(base, leader, 64);
        if (mine) {
            const JobDev* J = jobs + j;
            const uint32_t slot = base + (uint32_t)__popcll(grp & lt);
            if (slot < J->bucket_cap) J->buckets[(size_t)(du % kLevelRing) * J->bucket_cap + slot] = u;
            else atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrBucket);
            want = false;
        }
        pending &= ~grp;
    }
}

// One BFS level for every job: one thread per (frontier cell, neighbour)
// pair, so a level costs one chain of dependent memory accesses (bucket entry,
// keys, atomicMin) instead of one per neighbour.  Each block owns a contiguous
// range of pairs (so it touches one or two jobs) and stages the cells it
// discovers in LDS; a flush reserves bucket space with one global atomic per
// (job, level) the block touched, instead of one per wave, which keeps the
// per-job bucket counters free of contention.
constexpr uint32_t kStage = 2048;     // staged pushes per block before a flush
constexpr uint32_t kStageKeys = 16;   // distinct (job, level) groups per flush

__global__ __launch_bounds__(kFloodThreads) void k_flood_level(Launch L)
{
    __shared__ uint32_t pre[kFloodMaxJobs + 1];  // prefix of (entry, neighbour) pairs per job
    __shared__ uint32_t lev[kFloodMaxJobs];
    __shared__ uint32_t wsum[kFloodThreads / 64];
    __shared__ uint32_t st_u[kStage], st_key[kStage], st_off[kStage];
    __shared__ uint8_t st_grp[kStage];
    __shared__ uint32_t g_key[kStageKeys], g_cnt[kStageKeys], g_base[kStageKeys];
    __shared__ uint32_t st_n;
    const uint32_t nj = L.njobs;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nj; base += kFloodThreads) {
        const uint32_t j = base + threadIdx.x;
        uint32_t n = 0;
        if (j < nj) {
            const JobDev* J = L.jobs + j;
            const uint32_t d = J->fctl[0];
            lev[j] = d;
            uint32_t ne = (J->active && J->algo == VF_FLOOD && J->fctl[1] == 0) ? J->fctl[2 + (d & 7u)] : 0u;
            if (ne > J->bucket_cap) ne = J->bucket_cap;
            n = ne * (uint32_t)J->neighbourhood;
        }
        uint32_t x = n;
        const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t wofs = 0, tot = 0;
        for (uint32_t w = 0; w < kFloodThreads / 64; ++w) {
            if (w < wave) wofs += wsum[w];
            tot += wsum[w];
        }
        if (j < nj) pre[j] = carry + wofs + x - n;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pre[nj] = carry;
        st_n = 0;
    }
    __syncthreads();
    const uint32_t total = pre[nj];
    // contiguous, 256-aligned pair range per block
    const uint32_t per = (((total + gridDim.x - 1) / gridDim.x) + 255u) & ~255u;
    const uint32_t pb = min(total, blockIdx.x * per), pe = min(total, pb + per);
    if (pb >= pe) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;

    auto flush = [&]() {
        __syncthreads();
        const uint32_t n = st_n;
        if (threadIdx.x < kStageKeys) {
            g_key[threadIdx.x] = 0xFFFFFFFFu;
            g_cnt[threadIdx.x] = 0;
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
            const uint32_t k = st_key[e];
            uint32_t slot = kStageKeys;
            for (uint32_t q = 0; q < kStageKeys; ++q) {
                uint32_t cur = g_key[q];
                if (cur == 0xFFFFFFFFu) cur = atomicCAS(&g_key[q], 0xFFFFFFFFu, k) == 0xFFFFFFFFu ? k : g_key[q];
                if (cur == k) {
                    slot = q;
                    break;
                }
            }
            st_grp[e] = (uint8_t)slot;
            if (slot < kStageKeys) st_off[e] = atomicAdd(&g_cnt[slot], 1u);
        }
        __syncthreads();
        if (threadIdx.x < kStageKeys && g_cnt[threadIdx.x]) {
            const uint32_t k = g_key[threadIdx.x];
            const uint32_t j = k >> 2, du = lev[j] + 1u + (k & 3u);
            g_base[threadIdx.x] = atomicAdd(L.jobs[j].fctl + 2 + (du & 7u), g_cnt[threadIdx.x]);
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
            const uint32_t k = st_key[e];
            const uint32_t j = k >> 2, du = lev[j] + 1u + (k & 3u);
            const JobDev* J = L.jobs + j;
            uint32_t slot;
            if (st_grp[e] < kStageKeys) slot = g_base[st_grp[e]] + st_off[e];
            else slot = atomicAdd(J->fctl + 2 + (du & 7u), 1u);  // more than kStageKeys groups: rare
            if (slot < J->bucket_cap) J->buckets[(size_t)(du % kLevelRing) * J->bucket_cap + slot] = st_u[e];
            else atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrBucket);
        }
        __syncthreads();
        if (threadIdx.x == 0) st_n = 0;
        __syncthreads();
    };

    for (uint32_t p0 = pb; p0 < pe; p0 += blockDim.x) {
        const uint32_t p = p0 + threadIdx.x;
        bool want = false;
        uint32_t j = 0, du = 0, u = 0;
        if (p < pe) {
            uint32_t lo = 0, hi = nj;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (pre[mid] <= p) lo = mid; else hi = mid;
            }
            j = lo;
            const JobDev* J = L.jobs + j;
            const uint32_t d = lev[j];
            const uint32_t local = p - pre[j];
            const int nb = J->neighbourhood;
            const uint32_t e = nb == 6 ? local / 6u : local / 26u;
            const int k = (int)(local - e * (uint32_t)nb);
            const uint32_t v = J->buckets[(size_t)(d % kLevelRing) * J->bucket_cap + e];
            const int* o = nb == 6 ? c_nb6[k] : c_nb26[k];
            uint32_t x, y, z;
            kdecode(J, v, x, y, z);
            const int xx = (int)x + o[0], yy = (int)y + o[1], zz = (int)z + o[2];
            if (xx >= 0 && yy >= 0 && zz >= 0 && xx < (int)J->nx && yy < (int)J->ny && zz < (int)J->nz) {
                u = kidx(J, (uint32_t)xx, (uint32_t)yy, (uint32_t)zz);
                const uint32_t b = J->key_bits;
                const uint32_t ku = J->keys[u];
                // quick reject: EMPTY (0) or already reached at a level <= d
                if (ku != kEmptyKey && (ku >> b) > d) {
                    const uint32_t idx = J->keys[v] & ((1u << b) - 1u);
                    const uint32_t lin = ((uint32_t)xx * J->ny + (uint32_t)yy) * J->nz + (uint32_t)zz;
                    du = d + 1u + noise_delta(J->noise_seed, lin, J->noise_amp);
                    if (du >= (1u << (32 - b)) - 1u) {
                        atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrKeyOverflow);
                    } else {
                        const uint32_t cand = (du << b) | idx;
                        if (cand < ku) {
                            const uint32_t old = atomicMin(J->keys + u, cand);
                            want = cand < old && (old >> b) != du;
                        }
                    }
                }
            }
        }
        // stage this chunk's discoveries in LDS (one LDS atomic per wave)
        const uint64_t m = __ballot(want);
        if (m) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&st_n, (uint32_t)__popcll(m));
            base = __shfl(base, 0, 64);
            if (want) {
                const uint32_t at = base + (uint32_t)__popcll(m & lt);
                st_u[at] = u;
                st_key[at] = (j << 2) | (du - lev[j] - 1u);
            }
        }
        __syncthreads();
        if (st_n > kStage - blockDim.x) flush();
    }
    flush();
}

// Advance every flood job to its next level: recycle the counter slot that
// level d+4 will use (last used by level d-4) and mark jobs whose next three
// levels are empty as drained.
__global__ void k_flood_advance(Launch L)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < L.njobs; j += gridDim.x * blockDim.x) {
        const JobDev* J = L.jobs + j;
        if (!J->active || J->algo != VF_FLOOD || J->fctl[1] != 0) continue;
        uint32_t* c = J->fctl + 2;
        const uint32_t d = J->fctl[0];
        c[(d + 4) & 7u] = 0u;
        J->fctl[0] = d + 1;
        if (c[(d + 1) & 7u] == 0u && c[(d + 2) & 7u] == 0u && c[(d + 3) & 7u] == 0u) {
            J->fctl[1] = d + 1;
            J->stats[kStatFloodDone] = d + 1;
        }
    }
}

__global__ __launch_bounds__(kCellThreads) void k_flood_finish(Launch L)
{
    for (uint32_t t = blockIdx.x; t < L.total_tiles; t += gridDim.x) {
        const uint32_t j = find_job(L.tile_prefix, L.njobs, t);
        const JobDev* J = L.jobs + j;
        if (!J->active || J->algo != VF_FLOOD) continue;
        const uint64_t c0 = (uint64_t)(t - L.tile_prefix[j]) * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J->V) continue;
        const uint32_t mask = (1u << J->key_bits) - 1u;
        uint16_t out[8];
        uint32_t x, y, z;
        decompose(c0, J, x, y, z);
        for (int k = 0; k < 8; ++k) {
            const uint32_t key = c0 + k < J->V ? J->keys[kidx(J, x, y, z)] : 0u;
            advance(x, y, z, J);
            if (key == kEmptyKey) out[k] = kEmpty;
            else if (key == kUnreached) out[k] = kFree;
            else out[k] = (uint16_t)(J->seeds[key & mask].w & 0xFFu);  // unmaskRightMost(value, 8)
        }
        store8(J->labels, c0, J->V, out);
    }
}

}  // namespace

void launch_flood_init(const Launch& L)
{
    TimedLaunch tl("flood_init", L.s);
    uint32_t blocks = L.total_tiles < 2048 ? (L.total_tiles ? L.total_tiles : 1) : 2048;
    k_flood_init<<<blocks, kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_flood_seed(const Launch& L, uint32_t /*max_seeds*/)
{
    TimedLaunch tl("flood_seed", L.s);
    k_flood_seed<<<L.njobs, kFloodThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_flood_level(const Launch& L, uint32_t* /*blocks_done*/, uint32_t blocks)
{
    if (L.njobs > kFloodMaxJobs) fail(VF_ERR_INVALID_ARG, "flood launch with %u jobs (max %u)", L.njobs, kFloodMaxJobs);
    {
        TimedLaunch tl("flood_level", L.s);
        k_flood_level<<<blocks, kFloodThreads, 0, L.s>>>(L);
        VF_CHECK_LAUNCH();
    }
    TimedLaunch tl("flood_advance", L.s);
    k_flood_advance<<<ceil_div(L.njobs, 256), 256, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_flood_finish(const Launch& L)
{
    TimedLaunch tl("flood_finish", L.s);
    uint32_t blocks = L.total_tiles < 2048 ? (L.total_tiles ? L.total_tiles : 1) : 2048;
    k_flood_finish<<<blocks, kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
