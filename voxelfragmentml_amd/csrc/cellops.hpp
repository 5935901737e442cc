// Device helpers for cell-parallel stages: each thread owns 8 consecutive
// cells of one job's linear index range (x*ny*nz + y*nz + z), loaded and
// stored as one 16-byte (uint16 labels) or 8-byte (uint8 occupancy) access
// when the run is complete, element-wise at the tail.
#pragma once

#include "jobs.hpp"

namespace vf {

constexpr uint32_t kCellThreads = 256;
constexpr uint32_t kCellsPerThread = 8;
static_assert(kCellThreads * kCellsPerThread == kTileCells, "tile shape");

struct Cell8 {
    uint16_t v[8];
};

__device__ inline void load8(const uint16_t* p, uint64_t c0, uint64_t V, uint16_t v[8])
{
    if (c0 + 8 <= V) {
        uint4 q = *reinterpret_cast<const uint4*>(p + c0);
        v[0] = (uint16_t)q.x; v[1] = (uint16_t)(q.x >> 16);
        v[2] = (uint16_t)q.y; v[3] = (uint16_t)(q.y >> 16);
        v[4] = (uint16_t)q.z; v[5] = (uint16_t)(q.z >> 16);
        v[6] = (uint16_t)q.w; v[7] = (uint16_t)(q.w >> 16);
    } else {
        for (int k = 0; k < 8; ++k) v[k] = c0 + k < V ? p[c0 + k] : (uint16_t)0;
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
        *reinterpret_cast<uint4*>(p + c0) = q;
    } else {
        for (int k = 0; k < 8; ++k)
            if (c0 + k < V) p[c0 + k] = v[k];
    }
}

// 8 occupancy bytes packed little-endian into a uint64 (0 beyond V)
__device__ inline uint64_t load8b(const uint8_t* p, uint64_t c0, uint64_t V)
{
    if (c0 + 8 <= V) return *reinterpret_cast<const uint64_t*>(p + c0);
    uint64_t r = 0;
    for (int k = 0; k < 8; ++k)
        if (c0 + k < V) r |= (uint64_t)p[c0 + k] << (8 * k);
    return r;
}

__device__ inline void store8b(uint8_t* p, uint64_t c0, uint64_t V, uint64_t r)
{
    if (c0 + 8 <= V) {
        *reinterpret_cast<uint64_t*>(p + c0) = r;
    } else {
        for (int k = 0; k < 8; ++k)
            if (c0 + k < V) p[c0 + k] = (uint8_t)(r >> (8 * k));
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
