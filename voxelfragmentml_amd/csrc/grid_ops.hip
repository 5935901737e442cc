// Cell-parallel grid stages: occupancy <-> labels, homogenize, mask removal,
// boundary detection (optionally fused with the per-fragment histogram),
// histogram, occupied count and the order-independent grid checksum.
// All are HBM-bound streaming passes; each thread owns 8 cells.
#include <type_traits>

#include "cellops.hpp"

namespace vf {

namespace {

template <class F>
__device__ inline void for_each_tile(const Launch& L, F&& f)
{
    __shared__ JobDev sJ;
    walk_tiles(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t, uint32_t local, uint32_t) {
        if (!J.active) return;
        const uint64_t c0 = (uint64_t)local * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 < J.V) f(J, c0);
    });
}

__global__ __launch_bounds__(kCellThreads) void k_occ_to_labels(Launch L)
{
    for_each_tile(L, [&](const JobDev& J, uint64_t c0) {
        uint64_t o = load8b(J.occ, c0, J.V);
        uint16_t v[8];
        for (int k = 0; k < 8; ++k) v[k] = (uint16_t)((o >> (8 * k)) & 1u);
        store8(J.labels, c0, J.V, v);
    });
}

__global__ __launch_bounds__(kCellThreads) void k_labels_to_occ(Launch L)
{
    for_each_tile(L, [&](const JobDev& J, uint64_t c0) {
        uint16_t v[8];
        load8(J.labels, c0, J.V, v);
        uint64_t o = 0;
        for (int k = 0; k < 8; ++k) o |= (uint64_t)(v[k] != 0) << (8 * k);
        store8b(J.occ, c0, J.V, o);
    });
}

__global__ __launch_bounds__(kCellThreads) void k_homogenize(Launch L)
{
    for_each_tile(L, [&](const JobDev& J, uint64_t c0) {
        uint16_t v[8];
        load8(J.labels, c0, J.V, v);
        for (int k = 0; k < 8; ++k) v[k] = v[k] != kEmpty ? kFree : kEmpty;
        store8(J.labels, c0, J.V, v);
    });
}

__global__ __launch_bounds__(kCellThreads) void k_undo_mask(Launch L, uint16_t keep)
{
    for_each_tile(L, [&](const JobDev& J, uint64_t c0) {
        uint16_t v[8];
        load8(J.labels, c0, J.V, v);
        for (int k = 0; k < 8; ++k) v[k] = (uint16_t)(v[k] & keep);
        store8(J.labels, c0, J.V, v);
    });
}

__global__ void k_clear_hist(Launch L)
{
    for (uint32_t j = 0; j < L.njobs; ++j) {
        const JobDev* J = L.jobs + j;
        if (!J->active) continue;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 65536; i += gridDim.x * blockDim.x) J->hist[i] = 0;
    }
}

constexpr uint32_t kLdsBins = 4096;
constexpr uint32_t kTiledBins = 1024;  // LDS histogram bins of the tiled boundary kernel (larger labels go to HBM)

// Per-job sum of a per-thread value: wave sums, then one device atomic per block and job.
struct BlockSum {
    unsigned long long part[kCellThreads / 64];
    __device__ void flush(unsigned long long& acc, unsigned long long* dst)
    {
        const unsigned long long s = wave_sum_u64(acc);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long tot = 0;
            for (int w = 0; w < (int)(kCellThreads / 64); ++w) tot += part[w];
            if (tot) atomicAdd(dst, tot);
        }
        __syncthreads();
        acc = 0;
    }
};

// Checksum term of one cell c (linear index) holding v: a = fmix32(c ^ v * 0x9E3779B1),
// term = (a << 32) | fmix32(a ^ c ^ 0x7F4A7C15); EMPTY adds nothing.
__device__ inline unsigned long long checksum_term(uint16_t v, uint32_t c)
{
    if (!v) return 0ull;
    const uint32_t a = fmix32(c ^ ((uint32_t)v * 0x9E3779B1u));
    return ((unsigned long long)a << 32) | fmix32(a ^ c ^ 0x7F4A7C15u);
}


__device__ inline uint32_t ceil_div_d(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Boundary bit of one labelled cell (detectBoundaries-comp.glsl:18-43).
__device__ inline bool is_boundary_cell(const JobDev* J, uint32_t x, uint32_t y, uint32_t z, uint16_t c, int b)
{
    const int nx = (int)J->nx, ny = (int)J->ny, nz = (int)J->nz;
    int x0 = max((int)x - b, 0), x1 = min((int)x + b, nx - 1);
    int y0 = max((int)y - b, 0), y1 = min((int)y + b, ny - 1);
    int z0 = max((int)z - b, 0), z1 = min((int)z + b, nz - 1);
    const uint16_t* g = J->labels;
    for (int xx = x0; xx <= x1; ++xx)
        for (int yy = y0; yy <= y1; ++yy) {
            const uint16_t* row = g + ((size_t)xx * ny + yy) * nz;
            for (int zz = z0; zz <= z1; ++zz) {
                uint16_t v = (uint16_t)(row[zz] & kMask15);
                if (v > kFree && v != c) return true;
            }
        }
    return false;
}

// HIST: accumulate the per-fragment histogram; BOUND: set the boundary bit.
// Generic path (any window size; used for jobs with size > 1).  ONLY_WIDE
// restricts it to those jobs when the tiled kernel handles the rest.
template <bool HIST, bool BOUND, bool ONLY_WIDE = false>
__global__ __launch_bounds__(kCellThreads) void k_cells_post(Launch L, int bsize)
{
    __shared__ uint32_t bins[HIST ? kLdsBins : 1];
    uint32_t tb, te;
    block_tile_range(L.total_tiles, tb, te);
    if (tb >= te) return;
    uint32_t cur_job = find_job(L.tile_prefix, L.njobs, tb);
    if (HIST) {
        for (uint32_t i = threadIdx.x; i < kLdsBins; i += blockDim.x) bins[i] = 0;
        __syncthreads();
    }
    auto flush = [&](uint32_t j) {
        if (!HIST) return;
        __syncthreads();
        uint32_t* h = L.jobs[j].hist;
        for (uint32_t i = threadIdx.x; i < kLdsBins; i += blockDim.x) {
            uint32_t c = bins[i];
            if (c) {
                atomicAdd(h + i, c);
                bins[i] = 0;
            }
        }
        __syncthreads();
    };
    for (uint32_t t = tb; t < te; ++t) {
        uint32_t j = find_job(L.tile_prefix, L.njobs, t);
        if (j != cur_job) {
            flush(cur_job);
            cur_job = j;
        }
        const JobDev* J = L.jobs + j;
        if (!J->active || (ONLY_WIDE && J->bsize <= 1)) continue;
        uint64_t c0 = (uint64_t)(t - L.tile_prefix[j]) * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J->V) continue;
        uint16_t v[8];
        load8(J->labels, c0, J->V, v);
        bool any = false;
        for (int k = 0; k < 8; ++k) any |= v[k] > kFree;
        if (!any) continue;
        bool changed = false;
        if (BOUND) {
            const int bs = bsize >= 0 ? bsize : J->bsize;
            uint32_t x, y, z;
            decompose(c0, J, x, y, z);
            for (int k = 0; k < 8; ++k) {
                if (c0 + k < J->V && v[k] > kFree && is_boundary_cell(J, x, y, z, v[k], bs)) {
                    v[k] = (uint16_t)(v[k] | 0x8000u);
                    changed = true;
                }
                advance(x, y, z, J);
            }
        }
        if (HIST) {
            for (int k = 0; k < 8; ++k) {
                if (c0 + k < J->V && v[k] > kFree) {
                    uint32_t m = v[k] & kMask15;
                    if (m < kLdsBins) atomicAdd(&bins[m], 1u);
                    else atomicAdd(J->hist + m, 1u);
                }
            }
        }
        if (changed) store8(J->labels, c0, J->V, v);
    }
    flush(cur_job);
}

// Tiled detectBoundaries(1) + histogram.  A labelled cell c (2 <= raw < 0x8000)
// is a boundary cell iff its clamped 3x3x3 window holds a label > FREE other than
// its own, i.e. iff the window max of A = (m > 1 ? m : 0) or the window min of
// B = (m > 1 ? m : 0xFFFF) differs from raw(c) (m = value & 0x7FFF; c itself is in
// the window).  Cells whose bit 15 is already set never change.  Clamping the
// window equals replicate padding for max/min, so the halo is loaded with
// clamped coordinates.  8x8x32 bricks: the interior is loaded first with one
// 16-byte load per thread and bricks without a label stop there; live bricks
// load the halo and run separable max/min passes z, y, x in LDS.
constexpr int kBTX = 8, kBTY = 8, kBTZ = 32;
constexpr int kBHX = kBTX + 2, kBHY = kBTY + 2, kBHZ = kBTZ + 2;
// z-pass row stride (words): a multiple of 4, so a thread's 8-value segment is two 16-byte LDS accesses;
// 36 = 9 groups of 4 banks per row, so the 16-byte accesses of 8 consecutive lanes (2 rows x 4 segments)
// fall on distinct bank groups
constexpr int kAZ = kBTZ + 4;

// CSUM: also accumulate the grid checksum of the final cells (jobs with a
// window size <= 1; the others get theirs from k_checksum<true>).
//
// Each thread owns an 8-cell z segment of one of the brick's 64 rows and loads
// it with its two z neighbours (clamped coordinates = replicate padding), so
// the z window is done in registers; 144 of the threads also do one segment
// of the 36 halo rows.  Only the z-window max/min of the 100 rows goes
// through LDS, for the 3x3 x/y window of each labelled cell.
// lane-wise max of two packed u16 pairs
__device__ inline uint32_t pk_max_u16(uint32_t a, uint32_t b)
{
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
// a label as the packed pair (label | 0xFFFF - label)
__device__ inline uint32_t packed_self(uint32_t m) { return m | ((0xFFFFu - m) << 16); }

template <bool HIST, bool CSUM, uint32_t RUN>
__global__ __launch_bounds__(256) void k_boundaries_tiled(Launch L, const uint32_t* bt_prefix, uint32_t total)
{
    // z-window max (low 16 bits) | 0xFFFF - min (high 16) over labelled cells (EMPTY / FREE contribute 0),
    // so one packed u16 max combines both halves; row = hx * kBHY + hy
    __shared__ __attribute__((aligned(16))) uint32_t zmm[kBHX * kBHY * kAZ];
    __shared__ uint32_t bins[HIST ? kTiledBins : 1];
    __shared__ JobDev sJ;
    __shared__ BlockSum csum_part;
    unsigned long long csum = 0;
    if (HIST) {
        for (uint32_t i = threadIdx.x; i < kTiledBins; i += blockDim.x) bins[i] = 0;
    }
    auto flush = [&](uint32_t j) {
        if (CSUM) csum_part.flush(csum, L.jobs[j].stats + kStatChecksum);
        if (!HIST) return;
        __syncthreads();
        uint32_t* h = L.jobs[j].hist;
        for (uint32_t i = threadIdx.x; i < kTiledBins; i += blockDim.x) {
            const uint32_t c = bins[i];
            if (c) {
                atomicAdd(h + i, c);
                bins[i] = 0;
            }
        }
        __syncthreads();
    };
    // this thread's interior segment: row (tx, ty) = tid / 4, z segment (tid % 4) * 8
    const int row = threadIdx.x >> 2, seg = (threadIdx.x & 3) * 8;
    const int tx = row / kBTY, ty = row % kBTY;
    // halo task (threads < 144): ring row r = tid / 4 of the 36 rows with hx in {0, 9} or hy in {0, 9}
    const bool halo_task = threadIdx.x < (2 * kBHY + 2 * kBTX) * 4;
    int hhx = 0, hhy = 0;
    {
        const int r = threadIdx.x >> 2;
        if (r < kBHY) { hhx = 0; hhy = r; }
        else if (r < 2 * kBHY) { hhx = kBHX - 1; hhy = r - kBHY; }
        else if (r < 2 * kBHY + kBTX) { hhx = r - 2 * kBHY + 1; hhy = 0; }
        else { hhx = r - 2 * kBHY - kBTX + 1; hhy = kBHY - 1; }
    }
    // packed z-window (max | 0xFFFF - min) of 8 cells from their 10 raw values w[0..9] (w[k+1] is cell k)
    auto zwin = [](const uint16_t w[10], uint32_t out[8]) {
        uint32_t p[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            const uint32_t m = w[k] & kMask15;
            p[k] = m > kFree ? packed_self(m) : 0u;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = pk_max_u16(pk_max_u16(p[k], p[k + 1]), p[k + 2]);
    };
    constexpr uint32_t kRun = RUN;  // bricks per step: their interior loads are issued together
    walk_tile_runs<kRun>(L, bt_prefix, total, &sJ, [&](const JobDev& J, uint32_t, uint32_t first, uint32_t nrun) {
        if (!J.active || J.bsize > 1) return;
        const int nx = (int)J.nx, ny = (int)J.ny, nz = (int)J.nz;
        const uint32_t tz = ceil_div_d(J.nz, kBTZ), tyn = ceil_div_d(J.ny, kBTY);
        const bool bound = J.bsize == 1;
        const uint16_t* lab = J.labels;
        // ten raw values of row (gx, gy) at z = zs - 1 .. zs + 8, clamped to the grid
        auto load_row = [&](int gx, int gy, int zs, uint16_t w[10]) {
            const int cx = min(max(gx, 0), nx - 1), cy = min(max(gy, 0), ny - 1);
            const uint16_t* r = lab + ((size_t)cx * ny + cy) * nz;
            // zs is a multiple of 8, so with nz a multiple of 4 the 8 cells start 8-byte aligned
            if (zs + 8 <= nz && (nz & 3) == 0) {
                const uint2 a = gload2(r + zs);
                const uint2 q = gload2(r + zs + 4);
                w[1] = (uint16_t)a.x; w[2] = (uint16_t)(a.x >> 16); w[3] = (uint16_t)a.y; w[4] = (uint16_t)(a.y >> 16);
                w[5] = (uint16_t)q.x; w[6] = (uint16_t)(q.x >> 16); w[7] = (uint16_t)q.y; w[8] = (uint16_t)(q.y >> 16);
                w[0] = gcast<const uint16_t>(r)[max(zs - 1, 0)];  // clamped window: z = -1 reads z = 0
                w[9] = gcast<const uint16_t>(r)[min(zs + 8, nz - 1)];
            } else {
#pragma unroll
                for (int k = 0; k < 10; ++k) w[k] = r[min(max(zs - 1 + k, 0), nz - 1)];
            }
        };
        // brick `local`: its origin, this thread's row segment (gx, gy, gz) and how many of its 8 cells lie in the grid
        auto segment = [&](uint32_t local, int& x0, int& y0, int& z0, int& gx, int& gy, int& gz, int& nval) {
            x0 = (int)(local / (tz * tyn)) * kBTX;
            y0 = (int)((local / tz) % tyn) * kBTY;
            z0 = (int)(local % tz) * kBTZ;
            gx = x0 + tx;
            gy = y0 + ty;
            gz = z0 + seg;
            nval = gx < nx && gy < ny ? max(0, min(8, nz - gz)) : 0;
        };
        auto add_checksum = [&](const uint16_t v[8], int gx, int gy, int gz, int nval) {
            if (CSUM)
                if (nval && (v[0] | v[1] | v[2] | v[3] | v[4] | v[5] | v[6] | v[7])) {  // EMPTY cells add nothing
                    const uint32_t c0 = (uint32_t)(((size_t)gx * ny + gy) * nz + gz);
                    for (int k = 0; k < nval; ++k) csum += checksum_term(v[k], c0 + k);
                }
        };
        // the 8 in-grid cells of a segment as packed pairs (cells beyond the grid read 0)
        auto load_seg = [&](int gx, int gy, int gz, int nval, uint32_t p[4]) {
            p[0] = p[1] = p[2] = p[3] = 0u;
            const uint16_t* r = lab + ((size_t)(nval ? gx : 0) * ny + (nval ? gy : 0)) * nz + gz;
            if (nval == 8 && (nz & 3) == 0) {
                const uint2 a = gload2(r), q = gload2(r + 4);
                p[0] = a.x; p[1] = a.y; p[2] = q.x; p[3] = q.y;
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < nval) p[k >> 1] |= (uint32_t)r[k] << (16 * (k & 1));
            }
        };

        // phase 1: the interior segments of the whole run, loaded together; bricks without a label end here
        uint32_t live = 0;
        {
            uint32_t pr[kRun][4];
#pragma unroll
            for (uint32_t r = 0; r < kRun; ++r) {
                int x0, y0, z0, gx, gy, gz, nval;
                segment(first + r, x0, y0, z0, gx, gy, gz, nval);
                load_seg(gx, gy, gz, r < nrun ? nval : 0, pr[r]);
            }
#pragma unroll
            for (uint32_t r = 0; r < kRun; ++r) {
                bool mine = false;
#pragma unroll
                for (int k = 0; k < 4; ++k) mine |= (pr[r][k] & 0xFFFFu) > kFree || (pr[r][k] >> 16) > kFree;
                live |= mine ? 1u << r : 0u;
            }
            live = (uint32_t)__ockl_wgred_or_i32((int)live);  // block-uniform: bit r = brick r holds a label
            if (CSUM) {
#pragma unroll
                for (uint32_t r = 0; r < kRun; ++r) {
                    if (r >= nrun || ((live >> r) & 1u)) continue;
                    int x0, y0, z0, gx, gy, gz, nval;
                    segment(first + r, x0, y0, z0, gx, gy, gz, nval);
                    uint16_t v[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) v[k] = (uint16_t)(pr[r][k >> 1] >> (16 * (k & 1)));
                    add_checksum(v, gx, gy, gz, nval);
                }
            }
        }

        // phase 2: bricks with a label
        for (uint32_t r = 0; r < nrun; ++r) {
            if (!((live >> r) & 1u)) continue;
            int x0, y0, z0, gx, gy, gz, nval;
            segment(first + r, x0, y0, z0, gx, gy, gz, nval);
            uint16_t w[10];
            load_row(gx, gy, gz, w);  // the segment again (a cache hit) with its z neighbours
            uint16_t v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = k < nval ? w[k + 1] : (uint16_t)0;
            if (HIST) {
                // run-length over the 8 cells: one LDS atomic per run of equal labels
                uint32_t cur = 0, run = 0;
                for (int k = 0; k < 8; ++k) {
                    const uint32_t m = v[k] > kFree ? (uint32_t)(v[k] & kMask15) : 0xFFFFFFFFu;
                    if (m != cur && run) {
                        if (cur < kTiledBins) atomicAdd(&bins[cur], run);
                        else atomicAdd(J.hist + cur, run);
                        run = 0;
                    }
                    if (m != 0xFFFFFFFFu) {
                        cur = m;
                        ++run;
                    }
                }
                if (run) {
                    if (cur < kTiledBins) atomicAdd(&bins[cur], run);
                    else atomicAdd(J.hist + cur, run);
                }
            }
            if (!bound) {
                add_checksum(v, gx, gy, gz, nval);
                continue;
            }
            {
                uint32_t zw[8];
                zwin(w, zw);
                uint4* dst = reinterpret_cast<uint4*>(zmm + ((tx + 1) * kBHY + ty + 1) * kAZ + seg);
                dst[0] = make_uint4(zw[0], zw[1], zw[2], zw[3]);
                dst[1] = make_uint4(zw[4], zw[5], zw[6], zw[7]);
            }
            if (halo_task) {
                uint16_t hw[10];
                load_row(x0 + hhx - 1, y0 + hhy - 1, z0 + seg, hw);
                uint32_t zw[8];
                zwin(hw, zw);
                uint4* dst = reinterpret_cast<uint4*>(zmm + (hhx * kBHY + hhy) * kAZ + seg);
                dst[0] = make_uint4(zw[0], zw[1], zw[2], zw[3]);
                dst[1] = make_uint4(zw[4], zw[5], zw[6], zw[7]);
            }
            __syncthreads();
            // x and y windows together: a labelled cell without the bit is interior iff the packed
            // window (max | 0xFFFF - min) equals its own packed value; cells beyond the grid are 0
            uint32_t accs[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    const uint4* row = reinterpret_cast<const uint4*>(zmm + ((tx + dx) * kBHY + ty + dy) * kAZ + seg);
                    const uint4 a = row[0], b = row[1];
                    accs[0] = pk_max_u16(accs[0], a.x); accs[1] = pk_max_u16(accs[1], a.y);
                    accs[2] = pk_max_u16(accs[2], a.z); accs[3] = pk_max_u16(accs[3], a.w);
                    accs[4] = pk_max_u16(accs[4], b.x); accs[5] = pk_max_u16(accs[5], b.y);
                    accs[6] = pk_max_u16(accs[6], b.z); accs[7] = pk_max_u16(accs[7], b.w);
                }
            bool changed = false;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t acc = accs[k];
                const uint32_t c = v[k];
                const bool bnd = c > kFree && (c & 0x8000u) == 0 && acc != packed_self(c);
                v[k] = bnd ? (uint16_t)(c | 0x8000u) : (uint16_t)c;
                changed |= bnd;
            }
            add_checksum(v, gx, gy, gz, nval);
            if (changed) {
                uint16_t* dst = J.labels + ((size_t)gx * ny + gy) * nz + gz;
                if (nval == 8 && (nz & 3) == 0) {
                    gstore2(dst, make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16)));
                    gstore2(dst + 4, make_uint2((uint32_t)v[4] | ((uint32_t)v[5] << 16), (uint32_t)v[6] | ((uint32_t)v[7] << 16)));
                } else {
                    for (int k = 0; k < nval; ++k) dst[k] = v[k];
                }
            }
            __syncthreads();
        }
    }, flush);
}

// Order-independent grid checksum: sum over the non-EMPTY cells of checksum_term.
// ONLY_WIDE: only jobs whose boundary window is wider than 1 (the tiled
// boundary kernel sums the others).
template <bool ONLY_WIDE>
__global__ __launch_bounds__(kCellThreads) void k_checksum(Launch L)
{
    __shared__ BlockSum bs;
    __shared__ JobDev sJ;
    unsigned long long acc = 0;
    walk_tiles(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t, uint32_t local, uint32_t) {
        if (!J.active || (ONLY_WIDE && J.bsize <= 1)) return;
        const uint64_t c0 = (uint64_t)local * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J.V) return;
        uint16_t v[8];
        load8(J.labels, c0, J.V, v);
        for (int k = 0; k < 8; ++k) acc += checksum_term(v[k], (uint32_t)(c0 + k));  // cells beyond V load as 0
    }, [&](uint32_t j) { bs.flush(acc, L.jobs[j].stats + kStatChecksum); });
}

__global__ __launch_bounds__(kCellThreads) void k_count_occupied(Launch L)
{
    __shared__ BlockSum bs;
    __shared__ JobDev sJ;
    unsigned long long acc = 0;
    walk_tiles(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t, uint32_t local, uint32_t) {
        if (!J.active) return;
        const uint64_t c0 = (uint64_t)local * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J.V) return;
        const uint64_t o = load8b(J.occ, c0, J.V);
        acc += __popcll(o & 0x0101010101010101ull);
    }, [&](uint32_t j) { bs.flush(acc, L.jobs[j].stats + kStatOccupied); });
}

// One block per job: number of labels >= 2 with a nonzero count.
// kStatLabels: labels >= 2 with a nonzero count.  kSummaryParts blocks per job, each over 65536 /
// kSummaryParts bins with 16-byte loads issued together, add their counts (the slot is zeroed by job_init).
constexpr uint32_t kSummaryParts = 8;
__global__ __launch_bounds__(256) void k_hist_summary(Launch L)
{
    constexpr uint32_t kBins = 65536 / kSummaryParts, kIters = kBins / 4 / 256;
    __shared__ uint32_t wsum[4];
    const JobDev* J = L.jobs + blockIdx.x / kSummaryParts;
    if (!J->active) return;
    const uint32_t part = blockIdx.x % kSummaryParts;
    const uint4* h = reinterpret_cast<const uint4*>(J->hist + part * kBins);
    uint4 q[kIters];
#pragma unroll
    for (uint32_t i = 0; i < kIters; ++i) q[i] = gload4(h + i * 256 + threadIdx.x);
    uint32_t n = 0;
#pragma unroll
    for (uint32_t i = 0; i < kIters; ++i) {
        const bool first = part == 0 && i == 0 && threadIdx.x == 0;  // bins 0 (EMPTY) and 1 (FREE) are not labels
        n += (q[i].x != 0 && !first) + (q[i].y != 0 && !first) + (q[i].z != 0) + (q[i].w != 0);
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(J->stats + kStatLabels, (unsigned long long)t);
    }
}

// Block per job: zero the per-job statistics, histogram and flood control.
__global__ __launch_bounds__(256) void k_job_init(Launch L)
{
    const JobDev* J = L.jobs + blockIdx.x;
    if (!J->active) return;
    for (uint32_t i = threadIdx.x; i < 65536; i += blockDim.x) J->hist[i] = 0;
    if (threadIdx.x < kStatCount) J->stats[threadIdx.x] = 0;
    if (J->fctl && threadIdx.x < 64) J->fctl[threadIdx.x] = 0;
    if (J->algo == VF_FLOOD && J->bflags) {  // both brick flag parities of the flood rounds
        const uint32_t n = 2u * relax_words_per_parity(J->nbricks);
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) J->bflags[i] = 0ull;
    }
}

// Block per job: stats and flood state into a compact table for one readback.
__global__ void k_gather(Launch L, uint32_t* out)
{
    const JobDev* J = L.jobs + blockIdx.x;
    uint32_t* o = out + 16 * (size_t)blockIdx.x;
    const uint32_t t = threadIdx.x;
    if (t < kStatCount) {
        const unsigned long long v = J->active ? J->stats[t] : 0ull;
        o[2 * t] = (uint32_t)v;
        o[2 * t + 1] = (uint32_t)(v >> 32);
    }
}

inline uint32_t grid_for(const Launch& L) { return std::max(1u, std::min(L.total_tiles, active_shape().cell_blocks)); }

}  // namespace

void launch_occ_to_labels(const Launch& L)
{
    TimedLaunch tl("occ_to_labels", L.s);
    k_occ_to_labels<<<grid_for(L), kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_labels_to_occ(const Launch& L)
{
    TimedLaunch tl("labels_to_occ", L.s);
    k_labels_to_occ<<<grid_for(L), kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_homogenize(const Launch& L)
{
    TimedLaunch tl("homogenize", L.s);
    k_homogenize<<<grid_for(L), kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_undo_mask(const Launch& L, uint16_t keep)
{
    TimedLaunch tl("undo_mask", L.s);
    k_undo_mask<<<grid_for(L), kCellThreads, 0, L.s>>>(L, keep);
    VF_CHECK_LAUNCH();
}

void launch_clear_hist(const Launch& L)
{
    TimedLaunch tl("clear_hist", L.s);
    k_clear_hist<<<64, 256, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_histogram(const Launch& L)
{
    TimedLaunch tl("histogram", L.s);
    k_cells_post<true, false><<<grid_for(L), kCellThreads, 0, L.s>>>(L, 0);
    VF_CHECK_LAUNCH();
}

void launch_boundaries(const Launch& L, int size, int with_hist)
{
    if (with_hist) {
        TimedLaunch tl("boundaries_hist", L.s);
        k_cells_post<true, true><<<grid_for(L), kCellThreads, 0, L.s>>>(L, size);
    } else {
        TimedLaunch tl("boundaries", L.s);
        k_cells_post<false, true><<<grid_for(L), kCellThreads, 0, L.s>>>(L, size);
    }
    VF_CHECK_LAUNCH();
}

void launch_boundaries_tiled(const Launch& L, const uint32_t* bt_prefix, uint32_t total_btiles, bool any_wide,
                             bool with_checksum)
{
    {
        TimedLaunch tl("boundaries_hist", L.s);
        const LaunchShape sh = active_shape();
        const uint32_t blocks = std::max(1u, std::min(total_btiles, sh.brick_blocks));
        auto go = [&](auto run) {
            constexpr uint32_t R = decltype(run)::value;
            if (with_checksum) k_boundaries_tiled<true, true, R><<<blocks, 256, 0, L.s>>>(L, bt_prefix, total_btiles);
            else k_boundaries_tiled<true, false, R><<<blocks, 256, 0, L.s>>>(L, bt_prefix, total_btiles);
        };
        if (sh.brick_run == 2) go(std::integral_constant<uint32_t, 2>{});
        else if (sh.brick_run == 8) go(std::integral_constant<uint32_t, 8>{});
        else go(std::integral_constant<uint32_t, 4>{});
        VF_CHECK_LAUNCH();
    }
    if (any_wide) {
        TimedLaunch tl("boundaries_wide", L.s);
        k_cells_post<true, true, true><<<grid_for(L), kCellThreads, 0, L.s>>>(L, -1);
        VF_CHECK_LAUNCH();
        if (with_checksum) {
            TimedLaunch tc("checksum", L.s);
            k_checksum<true><<<grid_for(L), kCellThreads, 0, L.s>>>(L);
            VF_CHECK_LAUNCH();
        }
    }
}

void launch_checksum(const Launch& L)
{
    TimedLaunch tl("checksum", L.s);
    k_checksum<false><<<grid_for(L), kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_job_init(const Launch& L)
{
    TimedLaunch tl("job_init", L.s);
    k_job_init<<<L.njobs, 256, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_gather(const Launch& L, uint32_t* out)
{
    k_gather<<<L.njobs, 64, 0, L.s>>>(L, out);
    VF_CHECK_LAUNCH();
}

void launch_hist_summary(const Launch& L)
{
    TimedLaunch tl("hist_summary", L.s);
    k_hist_summary<<<L.njobs * kSummaryParts, 256, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_count_occupied(const Launch& L)
{
    TimedLaunch tl("count_occupied", L.s);
    k_count_occupied<<<grid_for(L), kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
