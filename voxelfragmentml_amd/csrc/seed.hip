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

// skip_both: jobs that seed anywhere were counted by the voxelizer's fill
__global__ __launch_bounds__(kCellThreads) void k_seed_count(Launch L, uint32_t* tile_cnt, bool skip_both)
{
    constexpr uint32_t kRun = 4;  // tiles per step: loads first, one block reduction for all of them
    __shared__ uint32_t wsum[kRun][kCellThreads / 64];
    __shared__ JobDev sJ;
    walk_tile_runs<kRun>(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t j, uint32_t local, uint32_t nrun) {
        if (skip_both && J.seed_location == VF_SEED_BOTH) return;  // block-uniform
        uint32_t n[kRun];
        uint64_t o[kRun];
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint64_t c0 = (uint64_t)(local + r) * kTileCells + threadIdx.x * kCellsPerThread;
            o[r] = J.active && r < nrun && c0 < J.V ? load8b(J.occ, c0, J.V) : 0ull;
        }
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            n[r] = 0;
            if (J.seed_location == VF_SEED_BOTH) {
                n[r] = __popcll(o[r] & 0x0101010101010101ull);
            } else if (o[r]) {
                const JobDev* Jg = L.jobs + j;
                const uint64_t c0 = (uint64_t)(local + r) * kTileCells + threadIdx.x * kCellsPerThread;
                uint32_t x, y, z;
                decompose(c0, J, x, y, z);
                for (int k = 0; k < 8 && c0 + k < J.V; ++k) {
                    if ((o[r] >> (8 * k)) & 0xFFu) n[r] += candidate(Jg, c0 + k, x, y, z) ? 1u : 0u;
                    advance(x, y, z, J);
                }
            }
            for (int ofs = 32; ofs > 0; ofs >>= 1) n[r] += __shfl_xor(n[r], ofs, 64);
        }
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (uint32_t r = 0; r < kRun; ++r) wsum[r][threadIdx.x >> 6] = n[r];
        }
        __syncthreads();
        if (threadIdx.x < nrun) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < kCellThreads / 64; ++w) tot += wsum[threadIdx.x][w];
            tile_cnt[L.tile_prefix[j] + local + threadIdx.x] = tot;
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
    const uint32_t S = J->S_draw;
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
    // rank -> tile: one binary search over the tile prefix per seed (the ranks are sorted, so
    // chosen[] is free to hold the tile index)
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        const unsigned long long want = base + sorted[i];
        uint32_t lo = t0, hi = t1;  // largest t with tile_off[t] <= want
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tile_off[mid] <= want) lo = mid; else hi = mid;
        }
        chosen[i] = lo;
    }
    __syncthreads();
    // tile -> cell: one wave per seed; lane l counts the candidates among the tile's cells
    // [32 l, 32 l + 32), a wave scan finds the lane holding the wanted one, and that lane walks
    // its 32 cells.  Cells outside the grid are no candidates.
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t i = wave; i < S; i += kSeedThreads / 64u) {
        const uint32_t t = chosen[i];
        const uint32_t need = (uint32_t)(base + sorted[i] - tile_off[t]);
        const uint64_t c0 = (uint64_t)(t - t0) * kTileCells + (uint64_t)lane * 32u;
        uint32_t cnt = 0;
        if (c0 < J->V) {
            if (J->seed_location == VF_SEED_BOTH) {
                if (c0 + 32u <= J->V) {
                    const uint4 a = gload4(J->occ + c0);
                    const uint4 b = gload4(J->occ + c0 + 16u);
                    const uint64_t m = 0x0101010101010101ull;
                    cnt = __popcll((((uint64_t)a.y << 32) | a.x) & m) + __popcll((((uint64_t)a.w << 32) | a.z) & m) +
                          __popcll((((uint64_t)b.y << 32) | b.x) & m) + __popcll((((uint64_t)b.w << 32) | b.z) & m);
                } else {
                    for (uint64_t c = c0; c < J->V; ++c) cnt += J->occ[c] != 0;
                }
            } else {
                uint32_t x, y, z;
                decompose(c0, J, x, y, z);
                for (uint64_t c = c0; c < c0 + 32u && c < J->V; ++c) {
                    cnt += candidate(J, c, x, y, z) ? 1u : 0u;
                    advance(x, y, z, J);
                }
            }
        }
        uint32_t inc = cnt;  // wave inclusive scan
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += v;
        }
        const bool holder = inc > need && inc - cnt <= need;
        if (holder) {
            uint32_t left = need - (inc - cnt);
            uint32_t x, y, z;
            decompose(c0, J, x, y, z);
            for (uint64_t c = c0; c < c0 + 32u && c < J->V; ++c) {
                if (candidate(J, c, x, y, z)) {
                    if (left == 0) {
                        J->seeds[i] = make_uint4(x, y, z, 2u + i);
                        break;
                    }
                    --left;
                }
                advance(x, y, z, J);
            }
        }
        if (__ballot(holder) == 0ull && lane == 0) {  // the counts disagree with the tile prefix
            J->seeds[i] = make_uint4(0u, 0u, 0u, 2u + i);
            atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrSeeding);
        }
    }
}

// Seeder::mergeSeeds (Seeder.cpp:115-152) in CADScene::fractureModel's order (CADScene.cpp:647-655): the
// merged list is the fragment seeds followed by the extra draw (at seeds[xoff..]); each entry takes the label of its nearest
// fragment seed (the reference's float distances on the coordinates, strict '<': the first fragment wins
// ties) with that fragment's running sub-seed id (1, 2, ... in list order) in bits 8..15.  The fragments
// stay first, the merged list follows them.  One block per job; entries staged in LDS.
constexpr uint32_t kMaxMerge = 2048;  // fragments + extra seeds of one job
__global__ __launch_bounds__(kSeedThreads) void k_seed_merge(Launch L)
{
    __shared__ uint32_t px[kMaxMerge], py[kMaxMerge], pz[kMaxMerge];
    __shared__ uint16_t near[kMaxMerge];
    const JobDev* J = L.jobs + blockIdx.x;
    if (!J->active || J->extra == 0) return;
    const uint32_t F = J->S0, M = J->S0 + J->extra;
    uint4* seeds = J->seeds;
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        const uint4 q = i < F ? seeds[i] : seeds[J->xoff + (i - F)];
        px[i] = q.x;
        py[i] = q.y;
        pz[i] = q.z;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        uint32_t best = 0xFFFFFFFFu, arg = 0;
        for (uint32_t f = 0; f < F; ++f) {
            const uint32_t dx = px[i] > px[f] ? px[i] - px[f] : px[f] - px[i];
            const uint32_t dy = py[i] > py[f] ? py[i] - py[f] : py[f] - py[i];
            const uint32_t dz = pz[i] > pz[f] ? pz[i] - pz[f] : pz[f] - pz[i];
            // integer forms of the float distances (coordinates below 2^11): Manhattan and Chebyshev
            // exact, Euclidean as the correctly rounded float root of the squared distance
            const uint32_t d = J->merge_dist == VF_MANHATTAN ? dx + dy + dz
                               : J->merge_dist == VF_CHEBYSHEV ? max(dx, max(dy, dz))
                                                               : sqrt_rn_bits(dx * dx + dy * dy + dz * dz);
            if (d < best) {
                best = d;
                arg = f;
            }
        }
        near[i] = (uint16_t)arg;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        const uint32_t f = near[i];
        uint32_t id = 1;
        for (uint32_t k = 0; k < i; ++k) id += near[k] == f ? 1u : 0u;
        if (id > 255u) atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrMerge);
        seeds[F + i] = make_uint4(px[i], py[i], pz[i], (seeds[f].w & 0xFFu) | ((id & 0xFFu) << 8));
    }
}

// floor(mn + (mx - mn) * h / 2^32): static_cast<int>(min + (max - min) * u) of
// RandomUtilities::getUniformRandomInt (RandomUtilities.h:141-144) for
// u = h / 2^32 wherever that is not negative.
__device__ inline int64_t uniform_int(int64_t mn, int64_t mx, uint32_t h)
{
    const int64_t t = (mx - mn) * (int64_t)h;
    return mn + (t >= 0 ? (t >> 32) : -((-t + 0xFFFFFFFFll) >> 32));
}

// Seeder::nearSeeds (Seeder.cpp:49-113) for one grid, one block.  Impact i
// draws with key k_i = key + (i + 1) * 0x9E3779B97F4A7C15: the fragment seed
// uniform_int(0, F - 1, h(k_i, 0)), the seed count min(uniform_int(1,
// pending, h(k_i, 1)), pending), and candidate t's axis a as
// n_a / 2 - sum_{j < spreading} h(k_i, 2 + (3t + a) spreading + j) mod
// (n_a / spreading) (getBiasedRandomInt) wrapped around the fragment seed.  A
// candidate is accepted, in draw order, when it lies within min(n) / 2 of the
// fragment seed, is occupied, on the surface (isBoundary) and new.  Batches of
// kSeedThreads candidates are tested in parallel; one lane accepts them in
// order.  out receives the fragments (unless out == frags) and then the new
// seeds; returns the number of new seeds, err = 1 if an impact hit the draw cap.
__device__ __forceinline__ uint32_t near_draw(const JobDev* J, const uint4* frags, uint32_t F, uint32_t impacts,
                                              uint32_t num_seeds, uint32_t spreading, uint64_t key, uint4* out,
                                              uint32_t& err)
{
    __shared__ uint32_t chosen[kMaxDeviceSeeds];
    __shared__ uint32_t sorted[kMaxDeviceSeeds];
    __shared__ uint32_t cand[kSeedThreads];
    __shared__ uint32_t s_n, s_got, s_err;
    const uint32_t n[3] = { J->nx, J->ny, J->nz };
    const uint32_t m[3] = { n[0] / spreading, n[1] / spreading, n[2] / spreading };
    const uint32_t mind = min(n[0], min(n[1], n[2])) / 2;
    if (threadIdx.x == 0) {
        s_n = 0;
        s_err = 0;
    }
    __syncthreads();
    uint32_t pending = num_seeds;
    for (uint32_t i = 0; i < impacts && !s_err; ++i) {
        const uint64_t k = key + (uint64_t)(i + 1u) * 0x9E3779B97F4A7C15ull;
        const uint4 fr = frags[uniform_int(0, (int64_t)F - 1, hash32(k, 0))];
        const int64_t w = uniform_int(1, pending, hash32(k, 1));
        const uint32_t want = (uint32_t)max<int64_t>(0, min<int64_t>(w, pending));
        if (threadIdx.x == 0) s_got = 0;
        __syncthreads();
        for (uint32_t t0 = 0; s_got < want; t0 += kSeedThreads) {
            if (t0 >= kMaxDraws) {
                if (threadIdx.x == 0) s_err = 1;
                break;
            }
            const uint32_t t = t0 + threadIdx.x;
            const uint32_t f[3] = { fr.x, fr.y, fr.z };
            uint32_t c[3];
            int64_t d2 = 0;
            for (int a = 0; a < 3; ++a) {
                uint32_t b = 0;
                for (uint32_t j = 0; j < spreading; ++j) b += hash32(k, 2u + (3u * t + (uint32_t)a) * spreading + j) % m[a];
                const int64_t o = (int64_t)(n[a] / 2) - (int64_t)b;
                c[a] = (uint32_t)(((int64_t)f[a] + o + n[a]) % n[a]);
                d2 += ((int64_t)c[a] - f[a]) * ((int64_t)c[a] - f[a]);
            }
            uint32_t id = 0xFFFFFFFFu;
            if (t < kMaxDraws && d2 <= (int64_t)mind * mind) {
                const size_t cell = ((size_t)c[0] * n[1] + c[1]) * n[2] + c[2];
                if (J->occ[cell] && surface_cell(J, c[0], c[1], c[2])) {
                    id = (uint32_t)cell;
                    for (uint32_t q = 0; q < s_n && id != 0xFFFFFFFFu; ++q)
                        if (chosen[q] == id) id = 0xFFFFFFFFu;
                }
            }
            cand[threadIdx.x] = id;
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint32_t first = s_n;
                for (uint32_t q = 0; q < kSeedThreads && s_got < want; ++q) {
                    const uint32_t v = cand[q];
                    if (v == 0xFFFFFFFFu) continue;
                    bool dup = false;
                    for (uint32_t r = first; r < s_n && !dup; ++r) dup = chosen[r] == v;
                    if (dup) continue;
                    chosen[s_n++] = v;
                    ++s_got;
                }
            }
            __syncthreads();
        }
        pending -= want;
        __syncthreads();
    }
    const uint32_t N = s_n;
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
        const uint32_t v = chosen[i];
        uint32_t p = 0;
        for (uint32_t q = 0; q < N; ++q) p += chosen[q] < v ? 1u : 0u;
        sorted[p] = v;
    }
    __syncthreads();
    if (out != frags)
        for (uint32_t i = threadIdx.x; i < F; i += blockDim.x) out[i] = frags[i];
    const uint32_t label0 = frags[F - 1].w;
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
        const uint32_t v = sorted[i];
        out[F + i] = make_uint4(v / (n[1] * n[2]), (v / n[2]) % n[1], v % n[2], label0 + 1u + i);
    }
    err = s_err;
    return N;
}

// One grid's nearSeeds.  result: [0] seeds written (fragments first), [1] 1 if an impact hit the draw cap.
__global__ __launch_bounds__(kSeedThreads) void k_seed_near(const JobDev* J, const uint4* frags, uint32_t F,
                                                             uint32_t impacts, uint32_t num_seeds, uint32_t spreading,
                                                             uint64_t key, uint4* out, uint32_t* result)
{
    uint32_t err = 0;
    const uint32_t N = near_draw(J, frags, F, impacts, num_seeds, spreading, key, out, err);
    if (threadIdx.x == 0) {
        result[0] = F + N;
        result[1] = err;
    }
}

// nearSeeds of every job with impacts (CADScene::fractureModel, CADScene.cpp:636-646), one block per job, in
// place: the new seeds follow the F_in fragment seeds, and the job's fragment count S0 and seed count S
// (fragments, then their merged copies and the extra draw when mergeSeeds follows) are set in the table.
__global__ __launch_bounds__(kSeedThreads) void k_seed_near_jobs(Launch L)
{
    JobDev* J = const_cast<JobDev*>(L.jobs + blockIdx.x);
    if (!J->active || J->impacts == 0) return;
    const uint32_t F = J->F_in;
    uint32_t err = 0;
    const uint32_t N = near_draw(J, J->seeds, F, J->impacts, J->near_seeds, J->spreading, J->near_key, J->seeds, err);
    if (threadIdx.x == 0) {
        const uint32_t S0 = F + N;
        J->S0 = S0;
        J->S = S0 + (J->extra ? S0 + J->extra : 0u);
        J->fctl[kCtlSeedCounts] = J->S;  // for tables made later
        J->fctl[kCtlSeedCounts + 1] = S0;
        if (err) atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrSeeding);
    }
}

__global__ void k_seed_counts(Launch L)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= L.njobs) return;
    JobDev* J = const_cast<JobDev*>(L.jobs + j);
    if (!J->impacts) return;
    J->S = J->fctl[kCtlSeedCounts];
    J->S0 = J->fctl[kCtlSeedCounts + 1];
}

}  // namespace

void launch_seed_merge(const Launch& L)
{
    TimedLaunch tl("seed_merge", L.s);
    k_seed_merge<<<L.njobs, kSeedThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_seed_near(const Launch& L, const uint4* frags, uint32_t F, uint32_t impacts, uint32_t num_seeds,
                      uint32_t spreading, uint64_t key, uint4* out, uint32_t* result)
{
    if (!F || !spreading || spreading > 64 || num_seeds > kMaxDeviceSeeds)
        fail(VF_ERR_INVALID_ARG, "near seeding needs fragment seeds, spreading 1..64 and at most %u seeds",
             kMaxDeviceSeeds);
    TimedLaunch tl("seed_near", L.s);
    k_seed_near<<<1, kSeedThreads, 0, L.s>>>(L.jobs, frags, F, impacts, num_seeds, spreading, key, out, result);
    VF_CHECK_LAUNCH();
}

void launch_seed_near_jobs(const Launch& L)
{
    TimedLaunch tl("seed_near", L.s);
    k_seed_near_jobs<<<L.njobs, kSeedThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_seed_counts(const Launch& L)
{
    TimedLaunch tl("seed_counts", L.s);
    k_seed_counts<<<ceil_div(L.njobs, 64u), 64, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_seed_count(const Launch& L, uint32_t* tile_cnt, bool skip_both)
{
    TimedLaunch tl("seed_count", L.s);
    const uint32_t blocks = std::max(1u, std::min(L.total_tiles, active_shape().cell_blocks));
    k_seed_count<<<blocks, kCellThreads, 0, L.s>>>(L, tile_cnt, skip_both);
    VF_CHECK_LAUNCH();
}

void launch_seed_pick(const Launch& L, const unsigned long long* tile_off)
{
    TimedLaunch tl("seed_pick", L.s);
    k_seed_pick<<<L.njobs, kSeedThreads, 0, L.s>>>(L, tile_off);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
