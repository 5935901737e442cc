// Device helpers for cell-parallel stages: each thread owns 8 consecutive
// cells of one job's linear index range (x*ny*nz + y*nz + z), loaded and
// stored as one 16-byte (uint16 labels) or 8-byte (uint8 occupancy) access
// when the run is complete, element-wise at the tail.
#pragma once

#include <algorithm>

#include "jobs.hpp"
#include "tune.hpp"

namespace vf {

// Bits of the correctly rounded float square root of an integer n < 2^24, exactly: the estimate
// moves to a float neighbour while sqrt(n) lies beyond the midpoint towards it.  Midpoints of
// adjacent floats have at most 25 significant bits, so their sums and squares are exact in double,
// and no midpoint squares to an integer below 2^48, so there are no ties.
__device__ inline uint32_t sqrt_rn_bits(uint32_t n)
{
    uint32_t b = __float_as_uint(__builtin_sqrtf((float)n));
    const double dn = (double)n;
    for (int it = 0; it < 4; ++it) {
        const double f = (double)__uint_as_float(b);
        if (b > 0u) {
            const double lo = 0.5 * ((double)__uint_as_float(b - 1u) + f);
            if (lo * lo > dn) {
                --b;
                continue;
            }
        }
        const double hi = 0.5 * (f + (double)__uint_as_float(b + 1u));
        if (hi * hi < dn) {
            ++b;
            continue;
        }
        break;
    }
    return b;
}


constexpr uint32_t kCellThreads = 256;
constexpr uint32_t kCellsPerThread = 8;
static_assert(kCellThreads * kCellsPerThread == kTileCells, "tile shape");

struct Cell8 {
    uint16_t v[8];
};

__device__ inline void load8(const uint16_t* p, uint64_t c0, uint64_t V, uint16_t v[8])
{
    if (c0 + 8 <= V) {
        const uint4 q = gload4(p + c0);
        v[0] = (uint16_t)q.x; v[1] = (uint16_t)(q.x >> 16);
        v[2] = (uint16_t)q.y; v[3] = (uint16_t)(q.y >> 16);
        v[4] = (uint16_t)q.z; v[5] = (uint16_t)(q.z >> 16);
        v[6] = (uint16_t)q.w; v[7] = (uint16_t)(q.w >> 16);
    } else {
        for (int k = 0; k < 8; ++k) v[k] = c0 + k < V ? gcast<const uint16_t>(p)[c0 + k] : (uint16_t)0;
    }
}

__device__ inline void store8(uint16_t* p, uint64_t c0, uint64_t V, const uint16_t v[8])
{
    if (c0 + 8 <= V) {
        uint4 q;
        q.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
        q.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
        q.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
        q.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
        gstore4(p + c0, q);
    } else {
        for (int k = 0; k < 8; ++k)
            if (c0 + k < V) gcast<uint16_t>(p)[c0 + k] = v[k];
    }
}

// 8 occupancy bytes packed little-endian into a uint64 (0 beyond V)
__device__ inline uint64_t load8b(const uint8_t* p, uint64_t c0, uint64_t V)
{
    if (c0 + 8 <= V) return *gcast<const uint64_t>(p + c0);
    uint64_t r = 0;
    for (int k = 0; k < 8; ++k)
        if (c0 + k < V) r |= (uint64_t)gcast<const uint8_t>(p)[c0 + k] << (8 * k);
    return r;
}

__device__ inline void store8b(uint8_t* p, uint64_t c0, uint64_t V, uint64_t r)
{
    if (c0 + 8 <= V) {
        *gcast<uint64_t>(p + c0) = r;
    } else {
        for (int k = 0; k < 8; ++k)
            if (c0 + k < V) gcast<uint8_t>(p)[c0 + k] = (uint8_t)(r >> (8 * k));
    }
}

// Contiguous tile range of this block (keeps a block on one job for long
// stretches so per-job reductions flush rarely).
__device__ inline void block_tile_range(uint32_t total, uint32_t& b, uint32_t& e)
{
    uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    b = blockIdx.x * per;
    if (b > total) b = total;
    e = b + per;
    if (e > total) e = total;
}

// Copy a job descriptor into LDS (block-wide; every thread must call it).
__device__ inline void stage_job(JobDev* sJ, const JobDev* src)
{
    static_assert(sizeof(JobDev) % 4 == 0, "JobDev is copied as words");
    __syncthreads();
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(sJ);
    for (uint32_t i = threadIdx.x; i < sizeof(JobDev) / 4; i += blockDim.x) d[i] = s[i];
    __syncthreads();
}

// Walk this block's contiguous range of the global tiles (prefix: njobs + 1
// tile prefix over the jobs).  The job is tracked incrementally and its
// descriptor staged in LDS on every change, so a tile costs no lookup and no
// dependent descriptor loads.  f(J, j, local_tile) runs block-uniformly;
// leave(j) runs before the block moves off job j and at the end.
template <class F, class G>
__device__ inline void walk_tiles(const Launch& L, const uint32_t* prefix, uint32_t total, JobDev* sJ, F&& f, G&& leave)
{
    uint32_t tb, te;
    block_tile_range(total, tb, te);
    if (tb >= te) return;
    uint32_t j = find_job(prefix, L.njobs, tb);
    uint32_t jb = prefix[j], je = prefix[j + 1];
    stage_job(sJ, L.jobs + j);
    for (uint32_t t = tb; t < te; ++t) {
        if (t >= je) {
            leave(j);
            do {
                ++j;
                jb = je;
                je = prefix[j + 1];
            } while (t >= je);
            stage_job(sJ, L.jobs + j);
        }
        f(*sJ, j, t - jb, t);
    }
    leave(j);
}

template <class F>
__device__ inline void walk_tiles(const Launch& L, const uint32_t* prefix, uint32_t total, JobDev* sJ, F&& f)
{
    walk_tiles(L, prefix, total, sJ, f, [](uint32_t) {});
}

// The same walk in runs of up to N consecutive tiles of one job:
// f(J, j, first_local_tile, n) handles tiles first .. first + n - 1 (n <= N),
// so a streaming stage can issue all its loads before using any of them.
// leave(j) runs when the block is done with job j (as in walk_tiles).
template <uint32_t N, class F, class G>
__device__ inline void walk_tile_runs(const Launch& L, const uint32_t* prefix, uint32_t total, JobDev* sJ, F&& f, G&& leave)
{
    uint32_t tb, te;
    block_tile_range(total, tb, te);
    if (tb >= te) return;
    uint32_t j = find_job(prefix, L.njobs, tb);
    uint32_t jb = prefix[j], je = prefix[j + 1];
    stage_job(sJ, L.jobs + j);
    for (uint32_t t = tb; t < te;) {
        if (t >= je) {
            leave(j);
            do {
                ++j;
                jb = je;
                je = prefix[j + 1];
            } while (t >= je);
            stage_job(sJ, L.jobs + j);
        }
        const uint32_t n = min(N, min(te, je) - t);
        f(*sJ, j, t - jb, n);
        t += n;
    }
    leave(j);
}

template <uint32_t N, class F>
__device__ inline void walk_tile_runs(const Launch& L, const uint32_t* prefix, uint32_t total, JobDev* sJ, F&& f)
{
    walk_tile_runs<N>(L, prefix, total, sJ, f, [](uint32_t) {});
}

// Job prefixes kept in LDS by the stages that search them per tile (njobs <= 1024).
constexpr uint32_t kWalkMaxJobs = 1024;

// Grid-stride over runs of N consecutive tiles: run r (tiles r*N .. r*N + N - 1, split where a job
// ends) goes to block r % gridDim.x, and f(J, j, first_local_tile, n) handles n <= N tiles of job j,
// so a stage whose tiles differ in cost keeps the round-robin balance and can still issue the
// loads of a whole run before using any of them.
template <uint32_t N, class F>
__device__ inline void walk_tile_runs_strided(const Launch& L, const uint32_t* prefix, uint32_t total, JobDev* sJ,
                                              uint32_t* sp, F&& f)
{
    for (uint32_t i = threadIdx.x; i <= L.njobs && i <= kWalkMaxJobs; i += blockDim.x) sp[i] = prefix[i];
    __syncthreads();
    uint32_t cur = 0xFFFFFFFFu;
    const uint32_t runs = (total + N - 1) / N;
    for (uint32_t r = blockIdx.x; r < runs; r += gridDim.x) {
        uint32_t t = r * N;
        const uint32_t t1 = min(t + N, total);
        while (t < t1) {
            uint32_t lo = 0, hi = L.njobs;  // largest j with sp[j] <= t
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sp[mid] <= t) lo = mid; else hi = mid;
            }
            if (lo != cur) {
                stage_job(sJ, L.jobs + lo);
                cur = lo;
            }
            const uint32_t n = min(t1, sp[lo + 1]) - t;
            f(*sJ, lo, t - sp[lo], n);
            t += n;
        }
    }
}

__device__ inline void decompose(uint64_t c, const JobDev& J, uint32_t& x, uint32_t& y, uint32_t& z)
{
    uint32_t yz = J.ny * J.nz;
    uint32_t cc = (uint32_t)c;
    x = cc / yz;
    uint32_t r = cc - x * yz;
    y = r / J.nz;
    z = r - y * J.nz;
}

__device__ inline void advance(uint32_t& x, uint32_t& y, uint32_t& z, const JobDev& J)
{
    if (++z == J.nz) {
        z = 0;
        if (++y == J.ny) {
            y = 0;
            ++x;
        }
    }
}

__device__ inline void decompose(uint64_t c, const JobDev* J, uint32_t& x, uint32_t& y, uint32_t& z)
{
    uint32_t yz = J->ny * J->nz;
    uint32_t cc = (uint32_t)c;
    x = cc / yz;
    uint32_t r = cc - x * yz;
    y = r / J->nz;
    z = r - y * J->nz;
}

__device__ inline void advance(uint32_t& x, uint32_t& y, uint32_t& z, const JobDev* J)
{
    if (++z == J->nz) {
        z = 0;
        if (++y == J->ny) {
            y = 0;
            ++x;
        }
    }
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace vf
