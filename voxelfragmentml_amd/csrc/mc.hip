// Marching cubes of every fragment of one or many grids, with an exact weld
// (RegularGrid::toTriangleMesh RegularGrid.cpp:443-486 ->
//  MarchingCubes::triangulateFieldGPU MarchingCubes.cpp:364-433,
//  marchingCubes-comp.glsl:96-156).
//
// The reference runs one full-grid dispatch per label, appends triangles with
// an atomic counter (nondeterministic order) and welds with 30-bit Morton codes.
// For label L the field is (cell & 0x7FFF) == L on the grid padded by FREE
// cells, isolevel 0.5, so every vertex is the midpoint of a lattice edge whose
// two ends disagree about L: the weld is exact by edge id (3 * padded index +
// axis).  Output is canonical: per grid, labels ascending; per label, vertices
// in edge-id order and triangles in (cube, table slot) order, corners emitted
// (0, 2, 1).
//
// Lattice point p = (px, py, pz) owns its three +axis edges and, when it is a
// valid corner 0, the cube spanning x in [px-1, px], y in [py, py+1],
// z in [pz, pz+1].  Only points next to a label change produce output, a few
// per cent of the lattice, so the work is sparse:
//   k_mc_labels / k_mc_base  each grid's target labels (nonzero histogram
//                 counts above FREE) -> dense index, and one global segment
//                 per (grid, label)
//   k_mc_count   one thread per unit (8 consecutive points of a z column)
//                 loads its columns; units that can have output are queued
//                 in LDS and worked one lane per point: the (segment, point)
//                 pairs with output, and their triangle and vertex counts;
//                 uniform units cost five column loads
//   scan         pair offsets per tile of units
//   (one host synchronisation: totals size the outputs)
//   k_mc_pairs   the pairs, in (grid, point) order, as (segment, point |
//                 counts | cube configuration | vertex axes | boundary flag)
//   radix sort   stable by segment -> (grid, label, point) order
//   scans        triangle and vertex offsets of every pair
//   k_mc_segments per-segment vertex, triangle and pair offsets
//   k_mc_emit    each pair's vertex positions (FP32, aabb_min + cell_size *
//                 (p - 1)) and triangles, with indices local to the label's
//                 vertices (found through the corners' owner pairs)
#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "cellops.hpp"
#include "grid.hpp"
#include "mc_table.hpp"
#include "tune.hpp"

namespace vf {

namespace {

constexpr uint32_t kMcThreads = 256;
constexpr uint32_t kMcZ = 8;                       // lattice points per unit (a z run of one column)
constexpr uint32_t kMcMaxLabels = 4096;            // labels per grid
constexpr uint16_t kMcNoLabel = 0xFFFFu;           // no dense index; also "beyond the lattice" (masked values <= 0x7FFF)
constexpr uint32_t kLindexWords = 65536 + kMcMaxLabels;  // per grid: label -> dense index, then its labels
constexpr uint32_t kMcMaxJobs = 1024;              // grids per run
constexpr uint32_t kMcLdsLabels = 1024;            // label-index entries a block keeps in LDS

__constant__ int c_mc_nb[8][3] = { { 0, 0, 0 }, { 0, 0, 1 }, { -1, 0, 1 }, { -1, 0, 0 },
                                   { 0, 1, 0 }, { 0, 1, 1 }, { -1, 1, 1 }, { -1, 1, 0 } };
__constant__ int c_mc_edge[12][2] = { { 0, 1 }, { 1, 2 }, { 2, 3 }, { 3, 0 }, { 4, 5 }, { 5, 6 },
                                      { 6, 7 }, { 7, 4 }, { 0, 4 }, { 1, 5 }, { 2, 6 }, { 3, 7 } };
__constant__ unsigned long long c_mc_tri[256];

struct McJob {
    const uint16_t* grid;
    const uint32_t* hist;      // countValues histogram (65536 counts)
    uint16_t* lindex;          // kLindexWords: label -> dense index or kMcNoLabel, then the labels ascending
    uint32_t nx, ny, nz;       // grid dims
    uint32_t Nx, Ny, Nz;       // padded lattice dims
    uint32_t cz;               // units per z column
    uint32_t want;             // VF_ALL_LABELS or one label
    uint32_t min_target;       // smallest target label (65536 when none): units below it emit nothing
    uint32_t vec;              // grid rows 16-byte aligned (2), 8-byte aligned (1) or neither (0)
    float inv_cz, inv_Ny;      // 1 / cz and 1 / Ny (small_div)
    float mn[3], cs[3];
};

struct McRun {
    McJob* jobs;
    uint32_t njobs;
    const unsigned long long* unit_prefix;  // njobs + 1
    unsigned long long total_units;
    uint32_t tiles;                         // of kMcThreads units
    uint32_t* label_base;                   // njobs + 1: first segment of each grid
    uint32_t* job_nl;                       // njobs
    uint32_t* job_err;                      // njobs: more than kMcMaxLabels labels
    uint16_t* labels_all;                   // one label per segment
    unsigned long long* job_tot;            // njobs x 4: pairs, triangles, vertices
    uint8_t* unit_pairs;                    // total_units
    uint32_t* tile_pairs;                   // tiles
    const unsigned long long* tile_off;     // tiles + 1
    uint32_t* keys;                         // pairs: segment
    unsigned long long* vals;               // pairs: point | triangles << 32 | vertices << 36 | configuration << 40 | vertex axes << 48 | flag << 51
    const uint32_t* skeys;                  // sorted by segment
    const unsigned long long* svals;
    uint32_t* cnt_t;                        // triangles of each sorted pair
    uint32_t* cnt_v;
    const unsigned long long* tpre;         // npairs + 1
    const unsigned long long* vpre;
    unsigned long long npairs;
    uint32_t segments;
    unsigned long long* voff;               // segments + 1
    unsigned long long* toff;
    uint8_t* tri_flag;
    unsigned long long* pstart;             // segments + 1: first pair of each segment
    const uint32_t* seg_job;                // segments: the grid of each segment
    float* xyz;
    uint32_t* tri_idx;
    unsigned long long nv, nt;
};

__device__ inline uint16_t mask15(uint16_t raw) { return raw == kMcNoLabel ? kMcNoLabel : (uint16_t)(raw & kMask15); }

// The labels with output at one point and their counts.  c: the eight masked cube corners (corner 0 is
// the point itself, corner 1 its +z neighbour, corner 4 its +y neighbour; kMcNoLabel beyond the lattice),
// nbx: its masked +x neighbour; cube: the point is a valid corner 0; ntri: the table's triangle count
// per configuration.  f(dense index, label, triangles, vertices, cube configuration, vertex axes) for every
// label with output; the vertices are the point's own +axis edges with the label on exactly one end; slx holds the first
// kMcLdsLabels entries of J's label index when staged.  Fully unrolled: the candidates stay in registers.
template <class F>
__device__ inline void point_labels(const McJob& J, const uint8_t* ntri, const uint16_t* slx, bool staged, const uint16_t c[8],
                                    uint16_t nbx, bool cube, F&& f)
{
    const uint16_t own = c[0];
    const uint16_t nb[3] = { nbx, c[4], c[1] };
    const uint16_t cand[9] = { c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], nbx };
    // candidates: every corner of a valid cube, else the point and its in-lattice +y / +z neighbours; +x
    uint32_t valid = cube ? 0xFFu : (1u | (c[1] != kMcNoLabel ? 2u : 0u) | (c[4] != kMcNoLabel ? 16u : 0u));
    valid |= nbx != kMcNoLabel ? 256u : 0u;
    // nothing changes around the point: no surface, no vertex
    bool uniform = true;
#pragma unroll
    for (int i = 1; i < 9; ++i) uniform &= !((valid >> i) & 1u) || cand[i] == own;
    if (uniform) return;
    auto one = [&](uint16_t L) {
        if (L < J.min_target) return;
        const uint16_t li = staged && L < kMcLdsLabels ? slx[L] : gcast<const uint16_t>(J.lindex)[L];
        if (li == kMcNoLabel) return;
        uint32_t nt = 0, conf = 0;
        if (cube) {
#pragma unroll
            for (int k = 0; k < 8; ++k) conf |= (uint32_t)(c[k] != L) << k;
            nt = ntri[conf];
        }
        uint32_t nv = 0, vax = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (nb[a] != kMcNoLabel && own != nb[a] && (own == L || nb[a] == L)) {
                ++nv;
                vax |= 1u << a;
            }
        if (nt | nv) f(li, L, nt, nv, conf, vax);
    };
    // the common case: two values meet at the point.  A point's labels may come in any order: each goes to
    // its own segment, and the segments are ordered by the stable sort that follows.
    uint32_t lo = 0xFFFFu, hi = 0u;
#pragma unroll
    for (int i = 0; i < 9; ++i)
        if ((valid >> i) & 1u) {
            lo = min(lo, (uint32_t)cand[i]);
            hi = max(hi, (uint32_t)cand[i]);
        }
    bool two = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) two &= !((valid >> i) & 1u) || cand[i] == lo || cand[i] == hi;
    if (two) {
        one((uint16_t)lo);
        one((uint16_t)hi);
        return;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint16_t L = cand[i];
        if (!((valid >> i) & 1u)) continue;
        bool seen = false;
#pragma unroll
        for (int k = 0; k < i; ++k) seen |= ((valid >> k) & 1u) && cand[k] == L;
        if (!seen) one(L);
    }
}

// Raw values of a unit's five columns: A = (px, py), B = (px - 1, py), C = (px, py + 1), D = (px - 1, py + 1),
// each at lattice z0 .. z0 + 8, and E = (px + 1, py) at z0 .. z0 + 7.
constexpr uint32_t kMcCols = 4 * (kMcZ + 1) + kMcZ;

// A unit's lattice origin (px, py, z0) packed as px | py << 16 and z0: units of a grid number fewer than
// 2^32 (its lattice points do), so 32-bit divisions suffice.
struct UnitOrigin {
    uint32_t pxy, z0;
};
__device__ inline UnitOrigin unit_origin(const McJob& J, uint32_t local)
{
    const uint32_t col = local / J.cz;
    const uint32_t px = col / J.Ny;
    return UnitOrigin{ px | ((col - px * J.Ny) << 16), (local - col * J.cz) * kMcZ };
}

// x / d for x below 2^17: a float estimate off by at most one, corrected.
__device__ inline uint32_t small_div(uint32_t x, uint32_t d, float inv)
{
    uint32_t q = (uint32_t)((float)x * inv);
    const int r = (int)x - (int)(q * d);
    if (r < 0) --q;
    else if (r >= (int)d) ++q;
    return q;
}
// The origin of the unit `lane` units after unit `local0` of grid J, from local0's column (col0 = local0 / cz,
// split as px0 = col0 / Ny, py0 = col0 % Ny) and its remainder rz0 = local0 % cz: the divisions left are of
// numbers below 2^17 (lane < 256, rz0 < cz, py0 < Ny, all below 2^16).
__device__ inline UnitOrigin unit_origin_after(const McJob& J, uint32_t px0, uint32_t py0, uint32_t rz0, uint32_t lane)
{
    const uint32_t t = rz0 + lane, dc = small_div(t, J.cz, J.inv_cz);
    const uint32_t yy = py0 + dc, dx = small_div(yy, J.Ny, J.inv_Ny);
    return UnitOrigin{ (px0 + dx) | ((yy - dx * J.Ny) << 16), (t - dc * J.cz) * kMcZ };
}

// Zeros that lanes off the fast path of load_unit read in place of a grid row (16-byte aligned at +8).
__device__ __align__(16) const uint16_t c_mc_zero[32] = {};

// Loads unit `local`'s columns into v; false when no point of the unit can have output: the unit is uniform
// (outside the vessel, inside one fragment) or holds no target label.  Every lane issues all of its loads
// before waiting on any (interior columns are contiguous grid rows, lattice z0 + k being grid z0 - 1 + k;
// z0 is a multiple of 8, so with 16-byte aligned rows lattice z0 + 1 .. z0 + 8 are one 16-byte load, with
// 8-byte aligned rows two); lanes whose columns touch the padding or the lattice end read zeros and then
// take those columns value by value.
__device__ inline bool load_unit(const McJob& J, UnitOrigin uo, uint16_t v[kMcCols])
{
    const int px = (int)(uo.pxy & 0xFFFFu), py = (int)(uo.pxy >> 16), z0 = (int)uo.z0;
    constexpr int K = kMcZ + 1;
    const int cx[5] = { px, px - 1, px, px - 1, px + 1 }, cy[5] = { py, py, py + 1, py + 1, py };
    const bool zin = z0 >= 1 && z0 + K <= (int)J.Nz - 1;
    bool fast[5];
    const uint16_t* row[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        fast[c] = zin && cx[c] >= 1 && cy[c] >= 1 && cx[c] <= (int)J.Nx - 2 && cy[c] <= (int)J.Ny - 2;
        row[c] = fast[c] ? J.grid + (((size_t)(cx[c] - 1) * J.ny + (size_t)(cy[c] - 1)) * J.nz + (size_t)(z0 - 1))
                         : c_mc_zero + 7;
    }
    auto put = [&](int c, const uint32_t w[4]) {  // lattice z0 + 1 .. from four packed words
        const int n = c < 4 ? K : (int)kMcZ;
#pragma unroll
        for (int k = 1; k < n; ++k) v[c * K + k] = (uint16_t)(w[(k - 1) >> 1] >> (16 * ((k - 1) & 1)));
    };
    if (J.vec == 2) {
        uint4 q[5];
        uint16_t h[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            h[c] = gcast<const uint16_t>(row[c])[0];
            q[c] = gload4(row[c] + 1);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            v[c * K] = h[c];
            const uint32_t w[4] = { q[c].x, q[c].y, q[c].z, q[c].w };
            put(c, w);
        }
    } else if (J.vec == 1) {
        uint2 qa[5], qb[5];
        uint16_t h[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            h[c] = gcast<const uint16_t>(row[c])[0];
            qa[c] = gload2(row[c] + 1);
            qb[c] = gload2(row[c] + 5);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            v[c * K] = h[c];
            const uint32_t w[4] = { qa[c].x, qa[c].y, qb[c].x, qb[c].y };
            put(c, w);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const VF_AS_GLOBAL uint16_t* r = gcast<const uint16_t>(row[c]);
#pragma unroll
            for (int k = 0; k < (c < 4 ? K : (int)kMcZ); ++k) v[c * K + k] = r[k];
        }
    }
    // the units at a column's ends and the columns at or beyond the padding (a few lanes of most tiles): the
    // column's state once, then each value from its lattice z alone
#pragma unroll
    for (int c = 0; c < 5; ++c)
        if (!fast[c]) {
            const bool beyond = cx[c] < 0 || cy[c] < 0 || cx[c] >= (int)J.Nx || cy[c] >= (int)J.Ny;
            const bool pad = cx[c] == 0 || cy[c] == 0 || cx[c] == (int)J.Nx - 1 || cy[c] == (int)J.Ny - 1;
            const VF_AS_GLOBAL uint16_t* r =
                gcast<const uint16_t>(J.grid) + (beyond || pad ? 0 : ((size_t)(cx[c] - 1) * J.ny + (size_t)(cy[c] - 1)) * J.nz);
            for (int k = 0; k < (c < 4 ? K : (int)kMcZ); ++k) {
                const int z = z0 + k;
                uint16_t x;
                if (beyond || z >= (int)J.Nz) x = kMcNoLabel;
                else if (pad || z == 0 || z == (int)J.Nz - 1) x = kFree;
                else x = r[z - 1];
                v[c * K + k] = x;
            }
        }
    uint32_t lo = 0xFFFFu, hi = 0;
#pragma unroll
    for (uint32_t m = 0; m < kMcCols; ++m) {
        const uint16_t x = mask15(v[m]);
        if (x != kMcNoLabel) {
            lo = min(lo, (uint32_t)x);
            hi = max(hi, (uint32_t)x);
        }
    }
    return !(hi < J.min_target || lo == hi);
}

// Point k of unit `local` from the unit's columns v: f(point index, masked corners, masked +x neighbour,
// valid cube, raw value); nothing for a point beyond the lattice.
template <class F>
__device__ inline void unit_point(const McJob& J, UnitOrigin uo, const uint16_t* v, int k, F&& f)
{
    const int px = (int)(uo.pxy & 0xFFFFu), py = (int)(uo.pxy >> 16), z0 = (int)uo.z0;
    const int pz = z0 + k;
    if (pz >= (int)J.Nz) return;
    constexpr int K = kMcZ + 1;
    const uint16_t *A = v, *B = v + K, *C = v + 2 * K, *D = v + 3 * K, *E = v + 4 * K;
    const uint16_t c[8] = { mask15(A[k]), mask15(A[k + 1]), mask15(B[k + 1]), mask15(B[k]),
                            mask15(C[k]), mask15(C[k + 1]), mask15(D[k + 1]), mask15(D[k]) };
    const uint16_t nbx = mask15(E[k]);
    const bool cube = px >= 1 && py <= (int)J.Ny - 2 && pz <= (int)J.Nz - 2;
    const uint32_t p = (uint32_t)px * (J.Ny * J.Nz) + (uint32_t)py * J.Nz + (uint32_t)pz;
    f(p, c, nbx, cube, A[k]);
}

// The 8 lanes of a point group (lanes 8g .. 8g + 7, one per point of a unit): sum, and exclusive prefix.
__device__ inline uint32_t group8_sum(uint32_t x)
{
    x += __shfl_xor(x, 1, 8);
    x += __shfl_xor(x, 2, 8);
    x += __shfl_xor(x, 4, 8);
    return x;
}
__device__ inline uint32_t group8_excl(uint32_t x, uint32_t k)
{
    uint32_t inc = x;
#pragma unroll
    for (uint32_t o = 1; o < 8; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 8);
        if (k >= o) inc += y;
    }
    return inc - x;
}

__device__ inline uint32_t find_job_u64(const unsigned long long* prefix, uint32_t n, unsigned long long u)
{
    uint32_t lo = 0, hi = n;  // largest j with prefix[j] <= u
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= u) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ inline uint32_t find_seg_job(const uint32_t* base, uint32_t n, uint32_t s)
{
    uint32_t lo = 0, hi = n;  // largest j with base[j] <= s
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= s) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ inline uint32_t block_sum_u32(uint32_t v, uint32_t* wsum)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t w = 0; w < kMcThreads / 64; ++w) t += wsum[w];
    __syncthreads();
    return t;
}

// Four block sums in one pass; every thread receives them.
__device__ inline void block_sum4_u32(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d, uint4* wsum4)
{
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
        d += __shfl_xor(d, o, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wsum4[wave] = make_uint4(a, b, c, d);
    __syncthreads();
    a = b = c = d = 0;
    for (uint32_t w = 0; w < kMcThreads / 64; ++w) {
        const uint4 q = wsum4[w];
        a += q.x;
        b += q.y;
        c += q.z;
        d += q.w;
    }
    __syncthreads();
}

__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t* wsum, uint32_t* total)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) wsum[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < kMcThreads / 64; ++w) {
        if (w < wave) off += wsum[w];
        tot += wsum[w];
    }
    *total = tot;
    __syncthreads();
    return off + x - v;
}

// One block per grid: dense indices of its target labels, the values countValues reports (a nonzero
// count; raw values above FREE are counted unmasked, so a flagged EMPTY or FREE cell makes 0 or 1 a
// label too, as in the reference), or the one wanted.
__global__ __launch_bounds__(kMcThreads) void k_mc_labels(McRun R)
{
    __shared__ uint32_t wsum[kMcThreads / 64];
    __shared__ uint32_t first;
    McJob& J = R.jobs[blockIdx.x];
    constexpr uint32_t kPer = 65536 / kMcThreads;
    const uint32_t l0 = threadIdx.x * kPer;
    if (threadIdx.x == 0) first = 65536u;
    auto target = [&](uint32_t l) { return J.hist[l] != 0 && (J.want == VF_ALL_LABELS || J.want == l); };
    uint32_t c = 0;
    for (uint32_t i = 0; i < kPer; ++i) c += target(l0 + i) ? 1u : 0u;
    uint32_t total;
    uint32_t idx = block_excl_scan_u32(c, wsum, &total);
    const bool ok = total <= kMcMaxLabels;
    for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t l = l0 + i;
        uint16_t li = kMcNoLabel;
        if (ok && target(l)) {
            J.lindex[65536 + idx] = (uint16_t)l;
            li = (uint16_t)idx++;
            atomicMin(&first, l);
        }
        J.lindex[l] = li;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        R.job_nl[blockIdx.x] = ok ? total : 0u;
        R.job_err[blockIdx.x] = ok ? 0u : 1u;
        J.min_target = first;
    }
}

// One block: segment bases (prefix of the label counts) and the label of every segment; clears the totals.
__global__ __launch_bounds__(1024) void k_mc_base(McRun R)
{
    __shared__ uint32_t wsum[1024 / 64];
    const uint32_t j = threadIdx.x, lane = j & 63u, wave = j >> 6;
    const uint32_t v = j < R.njobs ? R.job_nl[j] : 0u;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) wsum[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < 1024 / 64; ++w) {
        if (w < wave) off += wsum[w];
        tot += wsum[w];
    }
    if (j < R.njobs) {
        R.label_base[j] = off + x - v;
        for (int k = 0; k < 4; ++k) R.job_tot[4 * j + k] = 0ull;
    }
    if (j == 0) R.label_base[R.njobs] = tot;
    __syncthreads();
    for (uint32_t g = 0; g < R.njobs; ++g) {
        const uint32_t base = R.label_base[g], n = R.job_nl[g];
        for (uint32_t i = j; i < n; i += blockDim.x) R.labels_all[base + i] = R.jobs[g].lindex[65536 + i];
    }
}

// A tile is kMcThreads consecutive units.  Each thread loads one unit's columns; the units that can have
// output are queued in LDS in unit order (a block scan), and the queue is worked 8 lanes per unit, one lane
// per point, so a surface unit costs one pass of its 8 points in parallel and a uniform unit its loads only.
struct McQueue {
    uint16_t v[kMcThreads * kMcCols];  // columns of the queued units
    uint16_t unit[kMcThreads];         // their unit index within the tile
    UnitOrigin org[kMcThreads];        // their lattice origins
    uint32_t off[kMcThreads];          // k_mc_pairs: their first pair's offset within the tile
};

// Pairs with output, their triangle and vertex counts, per unit.  A tile's sums go to its first unit's grid
// through the block's running sums (a few tiles straddle two grids: their other grid's points add their own).
// The grid of a tile's first unit, staged in LDS with the first kMcLdsLabels entries of its label index
// when it changes (block-uniform; the caller's last barrier has freed the previous one).
struct McStage {
    McJob J;
    uint16_t lx[kMcLdsLabels];
};
__device__ inline void stage_grid(McStage& G, const McRun& R, uint32_t j0, uint32_t& staged)
{
    if (j0 == staged) return;
    if (threadIdx.x == 0) G.J = R.jobs[j0];
    const uint16_t* lx = R.jobs[j0].lindex;
    for (uint32_t i = threadIdx.x; i < kMcLdsLabels; i += blockDim.x) G.lx[i] = lx[i];
    __syncthreads();
    staged = j0;
}

// The tiles of the calling block: workgroups are dealt to the 8 XCDs round-robin, so XCD x takes the x-th
// eighth of the tiles and its blocks stride through it together; the columns a tile shares with its
// neighbours in x (the next slice of units) are then read by the same XCD at about the same time, from its
// L2.  Needs gridDim.x >= 8.
constexpr uint32_t kMcXcds = 8;
__device__ inline void xcd_tiles(uint32_t tiles, uint32_t& t0, uint32_t& t1, uint32_t& step)
{
    const uint32_t x = blockIdx.x % kMcXcds;
    t0 = (uint32_t)((unsigned long long)tiles * x / kMcXcds) + blockIdx.x / kMcXcds;
    t1 = (uint32_t)((unsigned long long)tiles * (x + 1) / kMcXcds);
    step = (gridDim.x - x + kMcXcds - 1) / kMcXcds;  // blocks of this XCD
}

// Grid j's descriptor: from LDS when the whole tile lies in the staged grid (one, block-uniform).
__device__ inline McJob job_of(const McStage& G, const McRun& R, bool one, uint32_t j)
{
    if (one) return G.J;
    return R.jobs[j];
}

// 96 registers: 5 waves per SIMD (the compiler's own choice, 103, allows 4; the count is latency-bound)
__global__ __launch_bounds__(kMcThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_mc_count(McRun R)
{
    extern __shared__ unsigned long long sp[];  // njobs + 1 unit prefixes
    __shared__ uint32_t wsum[kMcThreads / 64];
    __shared__ uint4 wsum4[kMcThreads / 64];
    __shared__ uint8_t ntri[256];
    __shared__ McQueue Q;
    __shared__ McStage G;
    for (uint32_t i = threadIdx.x; i <= R.njobs; i += blockDim.x) sp[i] = R.unit_prefix[i];
    ntri[threadIdx.x] = (uint8_t)(c_mc_tri[threadIdx.x] >> 60);
    __syncthreads();
    uint32_t staged = 0xFFFFFFFFu;
    // thread 0: the block's running sums for grid accj, added to the grid's totals when the block moves on
    // to another grid and at its end (same-address atomics serialise: one flush per block and grid)
    uint32_t accj = 0xFFFFFFFFu;
    unsigned long long acc_p = 0, acc_t = 0, acc_v = 0;
    auto flush = [&]() {
        if (accj != 0xFFFFFFFFu && acc_p) {
            atomicAdd(R.job_tot + 4 * accj, acc_p);
            atomicAdd(R.job_tot + 4 * accj + 1, acc_t);
            atomicAdd(R.job_tot + 4 * accj + 2, acc_v);
        }
        acc_p = acc_t = acc_v = 0;
    };
    uint32_t t0, t1, tstep;
    xcd_tiles(R.tiles, t0, t1, tstep);
    for (uint32_t t = t0; t < t1; t += tstep) {
        const unsigned long long ubase = (unsigned long long)t * kMcThreads, u = ubase + threadIdx.x;
        const uint32_t j0 = find_job_u64(sp, R.njobs, ubase);  // block-uniform
        const bool one = sp[j0 + 1] >= ubase + kMcThreads;
        stage_grid(G, R, j0, staged);
        bool live = false;
        uint16_t v[kMcCols];
        UnitOrigin uo{};
        // a tile in one grid: its first unit's column once (scalar), each lane's origin from it
        uint32_t px0 = 0, py0 = 0, rz0 = 0;
        if (one) {
            const uint32_t l0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(ubase - sp[j0]));
            const uint32_t cz = (uint32_t)__builtin_amdgcn_readfirstlane((int)G.J.cz),
                           Ny = (uint32_t)__builtin_amdgcn_readfirstlane((int)G.J.Ny);
            const uint32_t col0 = l0 / cz;
            rz0 = l0 - col0 * cz;
            px0 = col0 / Ny;
            py0 = col0 - px0 * Ny;
        }
        if (u < R.total_units) {
            const uint32_t j = one ? j0 : find_job_u64(sp, R.njobs, u);
            const McJob J = job_of(G, R, one, j);
            uo = one ? unit_origin_after(J, px0, py0, rz0, threadIdx.x) : unit_origin(J, (uint32_t)(u - sp[j]));
            live = load_unit(J, uo, v);
        }
        uint32_t nq = 0;
        const uint32_t pos = block_excl_scan_u32(live ? 1u : 0u, wsum, &nq);
        if (nq == 0) {  // block-uniform: nothing in this tile (k_mc_pairs skips it, unit counts unread)
            if (threadIdx.x == 0) R.tile_pairs[t] = 0;
            continue;
        }
        if (!live && u < R.total_units) R.unit_pairs[u] = 0;
        if (live) {
#pragma unroll
            for (uint32_t m = 0; m < kMcCols; ++m) Q.v[pos * kMcCols + m] = v[m];
            Q.unit[pos] = (uint16_t)threadIdx.x;
            Q.org[pos] = uo;
        }
        __syncthreads();
        uint32_t tp = 0, p0 = 0, t0 = 0, v0 = 0;
        for (uint32_t i = threadIdx.x; i < nq * kMcZ; i += kMcThreads) {  // whole 8-lane groups
            const uint32_t q = i / kMcZ, k = i % kMcZ;
            const unsigned long long uq = ubase + Q.unit[q];
            const uint32_t j = one ? j0 : find_job_u64(sp, R.njobs, uq);
            const bool mine = j == j0;
            const McJob J = job_of(G, R, one, j);
            uint32_t pc = 0, tc = 0, vc = 0;
            unit_point(J, Q.org[q], Q.v + q * kMcCols, (int)k, [&](uint32_t, const uint16_t* c, uint16_t nbx, bool cube, uint16_t) {
                point_labels(J, ntri, G.lx, mine, c, nbx, cube, [&](uint16_t, uint16_t, uint32_t nt, uint32_t nv, uint32_t, uint32_t) {
                    ++pc;
                    tc += nt;
                    vc += nv;
                });
            });
            const uint32_t up = group8_sum(pc);
            if (k == 0) R.unit_pairs[uq] = (uint8_t)up;
            tp += pc;
            if (mine) {
                p0 += pc;
                t0 += tc;
                v0 += vc;
            } else if (pc) {
                atomicAdd(R.job_tot + 4 * j, (unsigned long long)pc);
                atomicAdd(R.job_tot + 4 * j + 1, (unsigned long long)tc);
                atomicAdd(R.job_tot + 4 * j + 2, (unsigned long long)vc);
            }
        }
        block_sum4_u32(tp, p0, t0, v0, wsum4);  // ends with a barrier: the queue is free for the next tile
        if (threadIdx.x == 0) {
            R.tile_pairs[t] = tp;
            if (j0 != accj) {
                flush();
                accj = j0;
            }
            acc_p += p0;
            acc_t += t0;
            acc_v += v0;
        }
    }
    if (threadIdx.x == 0) flush();
}

// The pairs of every unit, in unit order (points in z order, a point's labels in candidate order), from the
// tile offsets and the per-unit counts.
__global__ __launch_bounds__(kMcThreads) void k_mc_pairs(McRun R)
{
    extern __shared__ unsigned long long sp[];  // njobs + 1 unit prefixes
    __shared__ uint32_t wsum[kMcThreads / 64];
    __shared__ uint8_t ntri[256];
    __shared__ McQueue Q;
    __shared__ McStage G;
    for (uint32_t i = threadIdx.x; i <= R.njobs; i += blockDim.x) sp[i] = R.unit_prefix[i];
    ntri[threadIdx.x] = (uint8_t)(c_mc_tri[threadIdx.x] >> 60);
    __syncthreads();
    uint32_t staged = 0xFFFFFFFFu;
    uint32_t t0, t1, tstep;
    xcd_tiles(R.tiles, t0, t1, tstep);
    for (uint32_t t = t0; t < t1; t += tstep) {
        if (R.tile_pairs[t] == 0) continue;  // block-uniform
        const unsigned long long ubase = (unsigned long long)t * kMcThreads, u = ubase + threadIdx.x;
        const uint32_t j0 = find_job_u64(sp, R.njobs, ubase);  // block-uniform
        const bool one = sp[j0 + 1] >= ubase + kMcThreads;
        stage_grid(G, R, j0, staged);
        const uint32_t n = u < R.total_units ? R.unit_pairs[u] : 0u;
        // one scan of both counts: pairs in the low 16 bits (at most 256 x 72), queued units in the high 16
        uint32_t both = 0;
        const uint32_t exb = block_excl_scan_u32(n | (n ? 0x10000u : 0u), wsum, &both);
        const uint32_t ex = exb & 0xFFFFu, pos = exb >> 16, nq = both >> 16;
        if (n) {
            const uint32_t j = one ? j0 : find_job_u64(sp, R.njobs, u);
            const McJob J = job_of(G, R, one, j);
            uint16_t v[kMcCols];
            const UnitOrigin uo = unit_origin(J, (uint32_t)(u - sp[j]));
            load_unit(J, uo, v);
#pragma unroll
            for (uint32_t m = 0; m < kMcCols; ++m) Q.v[pos * kMcCols + m] = v[m];
            Q.unit[pos] = (uint16_t)threadIdx.x;
            Q.org[pos] = uo;
            Q.off[pos] = ex;
        }
        __syncthreads();
        const unsigned long long tile_off = R.tile_off[t];
        for (uint32_t i = threadIdx.x; i < nq * kMcZ; i += kMcThreads) {  // whole 8-lane groups
            const uint32_t q = i / kMcZ, k = i % kMcZ;
            const unsigned long long uq = ubase + Q.unit[q];
            const uint32_t j = one ? j0 : find_job_u64(sp, R.njobs, uq);
            const bool mine = j == j0;
            const McJob J = job_of(G, R, one, j);
            const uint32_t base = R.label_base[j];
            const UnitOrigin local = Q.org[q];
            const uint16_t* v = Q.v + q * kMcCols;
            uint32_t pc = 0;
            unit_point(J, local, v, (int)k, [&](uint32_t, const uint16_t* c, uint16_t nbx, bool cube, uint16_t) {
                point_labels(J, ntri, G.lx, mine, c, nbx, cube,
                             [&](uint16_t, uint16_t, uint32_t, uint32_t, uint32_t, uint32_t) { ++pc; });
            });
            unsigned long long o = tile_off + Q.off[q] + group8_excl(pc, k);
            if (pc == 0) continue;
            unit_point(J, local, v, (int)k, [&](uint32_t p, const uint16_t* c, uint16_t nbx, bool cube, uint16_t raw) {
                const unsigned long long flag = (unsigned long long)((raw >> 15) & 1u) << 51;
                point_labels(J, ntri, G.lx, mine, c, nbx, cube,
                             [&](uint16_t li, uint16_t, uint32_t nt, uint32_t nv, uint32_t conf, uint32_t vax) {
                                 R.keys[o] = base + li;
                                 R.vals[o] = (unsigned long long)p | ((unsigned long long)nt << 32) |
                                             ((unsigned long long)nv << 36) | ((unsigned long long)conf << 40) |
                                             ((unsigned long long)vax << 48) | flag;
                                 ++o;
                             });
            });
        }
        __syncthreads();  // the queue is free for the next tile
    }
}

__global__ void k_mc_split(McRun R)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < R.npairs;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long v = R.svals[i];
        R.cnt_t[i] = (uint32_t)((v >> 32) & 15u);
        R.cnt_v[i] = (uint32_t)((v >> 36) & 15u);
    }
}

// The lower bound of (sg, t) among the sorted pairs (the first pair not below it, npairs if none), given a
// k0 no further than it: a forward gallop, then a binary search.  Lower bounds of increasing targets chain.
__device__ inline unsigned long long gallop_pair(const McRun& R, unsigned long long k0, uint32_t sg, uint32_t t)
{
    auto below = [&](unsigned long long k) {
        const uint32_t sk = R.skeys[k];
        return sk < sg || (sk == sg && (uint32_t)R.svals[k] < t);
    };
    if (k0 >= R.npairs || !below(k0)) return k0;
    unsigned long long lo = k0, hi = k0, step = 1;  // below(lo)
    for (;;) {
        hi = lo + step;
        if (hi >= R.npairs) {
            hi = R.npairs;
            break;
        }
        if (!below(hi)) break;
        lo = hi;
        step <<= 1;
    }
    while (hi - lo > 1) {  // the answer lies in (lo, hi]
        const unsigned long long mid = (lo + hi) >> 1;
        if (below(mid)) lo = mid; else hi = mid;
    }
    return hi;
}

// Each sorted pair writes its vertices (own +axis edges, axis order: positions aabb_min + cell_size * (p - 1),
// the edge's axis at + 0.5) and its triangles (table order (0, 2, 1), boundary flag), from what the pair
// carries: point | triangles << 32 | vertices << 36 | cube configuration << 40 | vertex axes << 48 |
// boundary flag << 51, without reading the grid.  A triangle corner is an edge of the cube whose owner
// (its lower end) is one of the cube's corners; its vertex index within the label is the owner pair's
// vertex offset plus the rank of the edge's axis among the owner's vertex axes.  Pairs are sorted by
// (label, point), so the owners at the pair's own x lie just after it and those at x - 1 just after the
// lower bound of point - Ny * Nz: one binary search per pair, then six short gallops for the cube's corners.
// A block takes chunks of 256 consecutive pairs, whose vertices and triangles are two contiguous output ranges: they
// are built in LDS and copied out with consecutive lanes on consecutive words (scattered 4-byte stores
// straight from each pair's thread reached HBM as partial-line writes, several times the output bytes).
constexpr uint32_t kEmitPairs = 256;
__global__ __launch_bounds__(kEmitPairs) void k_mc_emit(McRun R)
{
    __shared__ float sx[kEmitPairs * 3 * 3];      // at most 3 vertices per pair
    __shared__ uint32_t st[kEmitPairs * 5 * 3];   // at most 5 triangles per pair
    __shared__ uint8_t sf[kEmitPairs * 5];
    // each XCD takes a contiguous eighth of the chunks, so the pairs its searches read (the segment's pairs at
    // x - 1 and the next rows) are mostly in its own L2
    uint32_t c0, c1, cstep;
    xcd_tiles((uint32_t)((R.npairs + kEmitPairs - 1) / kEmitPairs), c0, c1, cstep);
    for (uint32_t c = c0; c < c1; c += cstep) {
        const unsigned long long b0 = (unsigned long long)c * kEmitPairs, b1 = min(b0 + kEmitPairs, R.npairs);
        const unsigned long long vb = R.vpre[b0], tb = R.tpre[b0];
        const unsigned long long i = b0 + threadIdx.x;
        if (i < b1) {
            const uint32_t sg = R.skeys[i];
            const unsigned long long val = R.svals[i];
            const uint32_t p = (uint32_t)val;
            const uint32_t nt = (uint32_t)((val >> 32) & 15u);
            const uint32_t conf = (uint32_t)((val >> 40) & 0xFFu), vax = (uint32_t)((val >> 48) & 7u);
            const uint8_t flag = (uint8_t)((val >> 51) & 1u);
            const uint32_t j = R.seg_job[sg];
            const McJob& J = R.jobs[j];
            const uint32_t yz = J.Ny * J.Nz;
            const int px = (int)(p / yz), py = (int)((p - (uint32_t)px * yz) / J.Nz), pz = (int)(p - (uint32_t)px * yz - (uint32_t)py * J.Nz);
            const int P[3] = { px, py, pz };
            uint32_t vo = (uint32_t)(R.vpre[i] - vb);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if ((vax >> a) & 1u) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const float tq = (float)P[q] - 1.0f + (q == a ? 0.5f : 0.0f);
                        sx[3 * vo + q] = J.mn[q] + J.cs[q] * tq;
                    }
                    ++vo;
                }
            if (nt) {
                const unsigned long long v0 = R.voff[sg];
                // the pairs of the cube's 8 corners as edge owners, o = (x side) * 4 + dy * 2 + dz, x side 1 for
                // the pair's own x and 0 for x - 1: lower bounds, each galloping from the previous one
                unsigned long long own[8];
                own[4] = i;
                own[5] = gallop_pair(R, own[4], sg, p + 1u);
                own[6] = gallop_pair(R, own[5], sg, p + J.Nz);
                own[7] = gallop_pair(R, own[6], sg, p + J.Nz + 1u);
                if (px >= 1) {
                    // the lower bound of point p - yz among the segment's pairs at x - 1
                    const uint32_t t = p - yz;
                    unsigned long long lo = R.pstart[sg], hi = i;
                    while (lo < hi) {
                        const unsigned long long mid = (lo + hi) >> 1;
                        if ((uint32_t)R.svals[mid] < t) lo = mid + 1; else hi = mid;
                    }
                    own[0] = lo;
                    own[1] = gallop_pair(R, own[0], sg, t + 1u);
                    own[2] = gallop_pair(R, own[1], sg, t + J.Nz);
                    own[3] = gallop_pair(R, own[2], sg, t + J.Nz + 1u);
                } else {
                    own[0] = own[1] = own[2] = own[3] = i;  // no cube reaches x - 1 < 0
                }
                // each owner's first vertex within the label and its vertex axes, all loads issued together; a
                // corner picks its owner's by selects on the owner's index bits (own[] indexed by a value
                // would live in scratch memory, several bytes of traffic per pair)
                uint32_t vb[8], ax[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    vb[c] = (uint32_t)(R.vpre[own[c]] - v0);
                    ax[c] = (uint32_t)((R.svals[own[c]] >> 48) & 7u);
                }
                auto pick = [](const uint32_t (&a)[8], int o) {
                    const bool b0 = o & 1, b1 = o & 2, b2 = o & 4;
                    const uint32_t l0 = b0 ? a[1] : a[0], l1 = b0 ? a[3] : a[2], l2 = b0 ? a[5] : a[4], l3 = b0 ? a[7] : a[6];
                    const uint32_t m0 = b1 ? l1 : l0, m1 = b1 ? l3 : l2;
                    return b2 ? m1 : m0;
                };
                auto corner = [&](int ca, int cb) -> uint32_t {
                    const int dxa = c_mc_nb[ca][0], dya = c_mc_nb[ca][1], dza = c_mc_nb[ca][2];
                    const int dxb = c_mc_nb[cb][0], dyb = c_mc_nb[cb][1], dzb = c_mc_nb[cb][2];
                    const int axis = dxa != dxb ? 0 : (dya != dyb ? 1 : 2);
                    const int o = (min(dxa, dxb) + 1) * 4 + min(dya, dyb) * 2 + min(dza, dzb);
                    return pick(vb, o) + (uint32_t)__popc(pick(ax, o) & ((1u << axis) - 1u));
                };
                const unsigned long long w = c_mc_tri[conf];
                uint32_t to = (uint32_t)(R.tpre[i] - tb);
                for (uint32_t q = 0; q < nt; ++q, ++to) {
                    const int e0 = (int)((w >> (12 * q)) & 15), e1 = (int)((w >> (12 * q + 4)) & 15),
                              e2 = (int)((w >> (12 * q + 8)) & 15);
                    st[3 * to + 0] = corner(c_mc_edge[e0][0], c_mc_edge[e0][1]);
                    st[3 * to + 1] = corner(c_mc_edge[e2][0], c_mc_edge[e2][1]);
                    st[3 * to + 2] = corner(c_mc_edge[e1][0], c_mc_edge[e1][1]);
                    sf[to] = flag;
                }
            }
        }
        __syncthreads();
        const uint32_t nxw = (uint32_t)(3 * (R.vpre[b1] - vb)), ntt = (uint32_t)(R.tpre[b1] - tb);
        float* gx = R.xyz + 3 * vb;
        uint32_t* gt = R.tri_idx + 3 * tb;
        uint8_t* gf = R.tri_flag + tb;
        for (uint32_t k = threadIdx.x; k < nxw; k += kEmitPairs) gx[k] = sx[k];
        for (uint32_t k = threadIdx.x; k < 3 * ntt; k += kEmitPairs) gt[k] = st[k];
        for (uint32_t k = threadIdx.x; k < ntt; k += kEmitPairs) gf[k] = sf[k];
        __syncthreads();  // the staging arrays are reused by the next chunk
    }
}

// voff / toff / pstart of every segment: the offsets of its first pair (pairs are sorted by segment).
__global__ void k_mc_segments(McRun R)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s <= R.segments; s += gridDim.x * blockDim.x) {
        unsigned long long lo = 0, hi = R.npairs;  // first pair with key >= s
        while (lo < hi) {
            const unsigned long long mid = (lo + hi) >> 1;
            if (R.skeys[mid] < s) lo = mid + 1; else hi = mid;
        }
        R.voff[s] = R.vpre[lo];
        R.toff[s] = R.tpre[lo];
        R.pstart[s] = lo;
    }
}

__global__ void k_mc_zero_u64(unsigned long long* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

bool g_table_loaded = false;

// ---------------------------------------------------------------- smoothing
// Laplacian smoothing of the fragment meshes (MarchingCubes::smoothSurface,
// MarchingCubes.cpp:491-514; laplacianSmoothing-comp.glsl:24-45,
// finishLaplacianSmoothing-comp.glsl:18-30): a face adds each of its vertices,
// scaled by 10^4 and truncated to int, to its other two vertices; a target
// vertex then moves to mix(p, sum / count / 10^4, weight).  The sums are
// gathered per vertex over a vertex -> triangle adjacency, in integers, so the
// result does not depend on order.  A vertex is a boundary vertex when any of
// its triangles is flagged.
struct SmoothDev {
    uint32_t nv, nt, nl;
    const uint32_t* tri;                 // local indices
    const uint8_t* tri_flag;
    const unsigned long long* voff;      // per label
    const unsigned long long* toff;
    uint32_t* tri_g;                     // 3 * nt global vertex indices
    uint32_t* vflag;                     // nv
    uint32_t* deg;                       // nv incidence counts, then fill cursors
    const unsigned long long* adj_off;   // nv + 1
    uint32_t* adj;                       // 3 * nt triangle ids
    const uint32_t* label_base;          // njobs + 1: first segment of each grid
    uint32_t njobs;
    uint32_t* vjob;                      // nv: grid of each vertex
    const uint4* params;                 // per grid: non-boundary steps, boundary steps, their weights (float bits)
};

__global__ void k_smooth_prep(SmoothDev D)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < D.nt; t += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = D.nl;  // label of triangle t: largest l with toff[l] <= t
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (D.toff[mid] <= t) lo = mid; else hi = mid;
        }
        const uint32_t base = (uint32_t)D.voff[lo];
        const uint32_t job = find_seg_job(D.label_base, D.njobs, lo);
        for (int i = 0; i < 3; ++i) {
            const uint32_t g = base + D.tri[3 * (size_t)t + i];
            D.tri_g[3 * (size_t)t + i] = g;
            D.vjob[g] = job;  // every triangle of a vertex belongs to the same grid
            atomicAdd(D.deg + g, 1u);
            if (D.tri_flag[t]) atomicOr(D.vflag + g, 1u);
        }
    }
}

__global__ void k_smooth_fill(SmoothDev D)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < D.nt; t += gridDim.x * blockDim.x)
        for (int i = 0; i < 3; ++i) {
            const uint32_t g = D.tri_g[3 * (size_t)t + i];
            const uint32_t slot = atomicAdd(D.deg + g, 1u);  // deg was reset to 0 after the scan
            D.adj[D.adj_off[g] + slot] = t;
        }
}

// Step `it` of a pass: a vertex of the pass's class moves when its grid runs more than `it` steps of
// that pass (pass 0: non-boundary vertices over faces of non-boundary vertices, pass 1: boundary
// vertices over all faces), with its grid's weight; every other vertex keeps its position.
__global__ void k_smooth_step(SmoothDev D, const float* src, float* dst, uint32_t target, int check, uint32_t it)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < D.nv; v += gridDim.x * blockDim.x) {
        const float* p = src + 3 * (size_t)v;
        float* o = dst + 3 * (size_t)v;
        float px = p[0], py = p[1], pz = p[2];
        const uint4 prm = D.params[D.vjob[v]];
        const uint32_t steps = target == 0 ? prm.x : prm.y;
        const float weight = __uint_as_float(target == 0 ? prm.z : prm.w);
        if (D.vflag[v] == target && it < steps) {
            int32_t sx = 0, sy = 0, sz = 0, cnt = 0;
            for (unsigned long long e = D.adj_off[v]; e < D.adj_off[v + 1]; ++e) {
                const uint32_t* f = D.tri_g + 3 * (size_t)D.adj[e];
                if (check && !(D.vflag[f[0]] == target && D.vflag[f[1]] == target && D.vflag[f[2]] == target)) continue;
                for (int i = 0; i < 3; ++i) {
                    if (f[i] == v) continue;
                    const float* q = src + 3 * (size_t)f[i];
                    sx += (int32_t)(q[0] * 10000.0f);
                    sy += (int32_t)(q[1] * 10000.0f);
                    sz += (int32_t)(q[2] * 10000.0f);
                    ++cnt;
                }
            }
            if (cnt > 0) {
                const float ax = (float)sx / (float)cnt / 10000.0f, ay = (float)sy / (float)cnt / 10000.0f,
                            az = (float)sz / (float)cnt / 10000.0f;
                px = px * (1.0f - weight) + ax * weight;
                py = py * (1.0f - weight) + ay * weight;
                pz = pz * (1.0f - weight) + az * weight;
            }
        }
        o[0] = px;
        o[1] = py;
        o[2] = pz;
    }
}

}  // namespace

namespace {

// scratch slots
enum : int {
    kBJobs, kBLidx, kBUnitPre, kBBase, kBLabels, kBTot, kBUnit, kBTile, kBTileOff, kBScan, kBKeys0, kBKeys1, kBVals0,
    kBVals1, kBSort, kBCntT, kBCntV, kBTpre, kBVpre, kBSeg, kBFlag, kBXyz, kBTri, kBSegJob,
    // smoothing
    kBSmTriG, kBSmFlag, kBSmAdjOff, kBSmTmp, kBSmAdj, kBSmScan, kBSmPrm, kBCount
};
static_assert(kBCount <= McScratch::kBuffers, "scratch slots");

void load_table()
{
    if (!g_table_loaded) {
        VF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_mc_tri), kMcTriPacked, sizeof(kMcTriPacked)));
        g_table_loaded = true;
    }
}

}  // namespace

void McScratch::ensure(int slot, size_t n) { buf[slot].ensure(n, *cache); }

void McScratch::release_all()
{
    for (auto& b : buf) b.release(*cache);
    if (cache == &own) own.trim(0);
    valid = false;
}

McScratch::~McScratch()
{
    for (auto& b : buf) b.release(*cache);
}

void mc_run_jobs(McScratch& W, hipStream_t s, const McJobSpec* spec, uint32_t n, McJobResult* out,
                 const std::function<bool(size_t)>& reserve)
{
    if (n == 0) return;
    auto refuse = [&] {
        for (uint32_t j = 0; j < n; ++j) out[j] = McJobResult{ 0, 0, 0, 0, 0, 0, VF_ERR_BUDGET };
        W.valid = false;
    };
    if (n > kMcMaxJobs) fail(VF_ERR_INVALID_ARG, "marching cubes of %u grids at once (at most %u)", n, kMcMaxJobs);
    load_table();
    W.valid = false;
    // per grid: lattice, units, label index
    W.ensure(kBLidx, (size_t)n * kLindexWords * 2);
    std::vector<McJob> jobs(n);
    std::vector<unsigned long long> upre(n + 1, 0);
    for (uint32_t j = 0; j < n; ++j) {
        const McJobSpec& q = spec[j];
        McJob& J = jobs[j];
        J.grid = q.grid;
        J.hist = q.hist;
        J.lindex = W.buf[kBLidx].as<uint16_t>() + (size_t)j * kLindexWords;
        J.nx = q.dims[0]; J.ny = q.dims[1]; J.nz = q.dims[2];
        J.Nx = q.dims[0] + 2; J.Ny = q.dims[1] + 2; J.Nz = q.dims[2] + 2;
        const uint64_t P = (uint64_t)J.Nx * J.Ny * J.Nz;
        if (P > 0xFFFFFFFFull) fail(VF_ERR_INVALID_ARG, "grid too large for marching cubes");
        if (J.Nx > 0xFFFFu || J.Ny > 0xFFFFu) fail(VF_ERR_INVALID_ARG, "marching cubes: x and y dims must stay below 65534");
        J.cz = (J.Nz + kMcZ - 1) / kMcZ;
        const uintptr_t a = reinterpret_cast<uintptr_t>(q.grid);
        J.vec = (a % 16 == 0 && J.nz % 8 == 0) ? 2u : (a % 8 == 0 && J.nz % 4 == 0) ? 1u : 0u;
        J.inv_cz = 1.0f / (float)J.cz;
        J.inv_Ny = 1.0f / (float)J.Ny;
        J.want = q.want;
        for (int a = 0; a < 3; ++a) {
            J.mn[a] = q.mn[a];
            J.cs[a] = q.cs[a];
        }
        upre[j + 1] = upre[j] + (unsigned long long)J.Nx * J.Ny * J.cz;
    }
    const unsigned long long units = upre[n];
    const uint64_t tiles64 = (units + kMcThreads - 1) / kMcThreads;
    if (tiles64 > 0xFFFFFFFFull) fail(VF_ERR_INVALID_ARG, "marching-cubes lattice too large");
    // the first phase's buffers: label indices, per-unit pair counts, per-tile offsets and their scan
    const size_t phase1 = (size_t)n * kLindexWords * 2 + n * sizeof(McJob) + (size_t)n * (8 + 12 + 32 + 2 * kMcMaxLabels) +
                          units + tiles64 * 12 + scan_tmp_bytes(tiles64) + 10 * 320;
    if (reserve && !reserve(BlockCache::class_size(phase1) + phase1 / 8)) {
        refuse();
        return;
    }
    McRun R{};
    R.njobs = n;
    R.total_units = units;
    R.tiles = (uint32_t)tiles64;
    W.ensure(kBJobs, n * sizeof(McJob));
    W.ensure(kBUnitPre, (n + 1) * 8);
    W.ensure(kBBase, (3 * (size_t)n + 1) * 4);
    W.ensure(kBTot, (size_t)n * 32);
    W.ensure(kBLabels, (size_t)n * kMcMaxLabels * 2);
    W.ensure(kBUnit, units + 64);
    W.ensure(kBTile, (size_t)R.tiles * 4 + 64);
    W.ensure(kBTileOff, ((size_t)R.tiles + 1) * 8 + 64);
    VF_HIP(hipMemcpyAsync(W.buf[kBJobs].p, jobs.data(), n * sizeof(McJob), hipMemcpyHostToDevice, s));
    VF_HIP(hipMemcpyAsync(W.buf[kBUnitPre].p, upre.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
    R.jobs = W.buf[kBJobs].as<McJob>();
    R.unit_prefix = W.buf[kBUnitPre].as<unsigned long long>();
    R.label_base = W.buf[kBBase].as<uint32_t>();
    R.job_nl = R.label_base + n + 1;
    R.job_err = R.job_nl + n;
    R.labels_all = W.buf[kBLabels].as<uint16_t>();
    R.job_tot = W.buf[kBTot].as<unsigned long long>();
    R.unit_pairs = W.buf[kBUnit].as<uint8_t>();
    R.tile_pairs = W.buf[kBTile].as<uint32_t>();
    R.tile_off = W.buf[kBTileOff].as<unsigned long long>();
    k_mc_labels<<<n, kMcThreads, 0, s>>>(R);
    VF_CHECK_LAUNCH();
    k_mc_base<<<1, 1024, 0, s>>>(R);
    VF_CHECK_LAUNCH();
    {
        TimedLaunch tl("mc_count", s);
        k_mc_count<<<std::max(kMcXcds, std::min(R.tiles, active_shape().mc_blocks)), kMcThreads, (R.njobs + 1) * 8, s>>>(R);
        VF_CHECK_LAUNCH();
    }
    W.ensure(kBScan, scan_tmp_bytes(R.tiles) + 64);
    scan_exclusive_u32_to_u64(R.tile_pairs, W.buf[kBTileOff].as<unsigned long long>(), R.tiles, W.buf[kBScan].p, s);
    // totals: pairs, triangles and vertices of every grid, segments, label-count errors
    std::vector<unsigned long long> tot(4 * (size_t)n);
    std::vector<uint32_t> base(3 * (size_t)n + 1);
    VF_HIP(hipMemcpyAsync(tot.data(), R.job_tot, tot.size() * 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipMemcpyAsync(base.data(), R.label_base, base.size() * 4, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    unsigned long long npairs = 0, NT = 0, NV = 0;
    for (uint32_t j = 0; j < n; ++j) {
        out[j].nl = base[n + 1 + j];
        out[j].seg0 = base[j];
        out[j].status = base[2 * n + 1 + j] ? VF_ERR_INVALID_ARG : VF_OK;
        out[j].nt = (uint32_t)tot[4 * j + 1];
        out[j].nv = (uint32_t)tot[4 * j + 2];
        out[j].t0 = NT;
        out[j].v0 = NV;
        if (tot[4 * j + 1] > 0xFFFFFFFFull || tot[4 * j + 2] > 0xFFFFFFFFull)
            fail(VF_ERR_OVERFLOW, "marching-cubes output of one grid exceeds 2^32 elements");
        npairs += tot[4 * j];
        NT += tot[4 * j + 1];
        NV += tot[4 * j + 2];
    }
    const uint32_t S = base[n];
    W.label_base.assign(base.begin(), base.begin() + n + 1);
    // the grid of every segment (the emit reads it per pair)
    std::vector<uint32_t> sjob((size_t)base[n] + 1, 0);
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t g = base[j]; g < base[j + 1]; ++g) sjob[g] = j;
    if (reserve) {  // the second phase's buffers: pairs (double-buffered, counts, prefixes) and the meshes
        size_t sort_bytes = 0;
        if (npairs)
            VF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                      (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                                      (int)npairs, 0, 32, s));
        const size_t phase2 = (size_t)npairs * (4 + 4 + 8 + 8 + 4 + 4 + 8 + 8) + sort_bytes + scan_tmp_bytes(npairs) +
                              (size_t)NT * (12 + 1) + (size_t)NV * 12 + 3 * ((size_t)S + 1) * 8 + ((size_t)S + 1) * 4 +
                              13 * 320;
        if (!reserve(phase2 + phase2 / 8)) {
            refuse();
            return;
        }
    }
    R.segments = S;
    R.npairs = npairs;
    R.nt = NT;
    R.nv = NV;
    if (npairs > 0xFFFFFFFFull || NT > 0xFFFFFFFFull || NV > 0xFFFFFFFFull)
        fail(VF_ERR_OVERFLOW, "marching-cubes output exceeds 2^32 elements");
    W.ensure(kBSeg, 3 * ((size_t)S + 1) * 8 + 64);
    R.voff = W.buf[kBSeg].as<unsigned long long>();
    R.toff = R.voff + (S + 1);
    R.pstart = R.voff + 2 * ((size_t)S + 1);
    W.ensure(kBSegJob, sjob.size() * 4 + 64);
    VF_HIP(hipMemcpyAsync(W.buf[kBSegJob].p, sjob.data(), sjob.size() * 4, hipMemcpyHostToDevice, s));
    R.seg_job = W.buf[kBSegJob].as<uint32_t>();
    W.ensure(kBXyz, (size_t)NV * 12 + 64);
    W.ensure(kBTri, (size_t)NT * 12 + 64);
    W.ensure(kBFlag, (size_t)NT + 64);
    R.xyz = W.buf[kBXyz].as<float>();
    R.tri_idx = W.buf[kBTri].as<uint32_t>();
    R.tri_flag = W.buf[kBFlag].as<uint8_t>();
    if (npairs == 0) {
        k_mc_zero_u64<<<1, 256, 0, s>>>(R.voff, 3 * (S + 1));
        VF_CHECK_LAUNCH();
    } else {
        const size_t np = (size_t)npairs;
        W.ensure(kBKeys0, np * 4 + 64);
        W.ensure(kBKeys1, np * 4 + 64);
        W.ensure(kBVals0, np * 8 + 64);
        W.ensure(kBVals1, np * 8 + 64);
        R.keys = W.buf[kBKeys0].as<uint32_t>();
        R.vals = W.buf[kBVals0].as<unsigned long long>();
        {
            TimedLaunch tl("mc_pairs", s);
            k_mc_pairs<<<std::max(kMcXcds, std::min(R.tiles, 16384u)), kMcThreads, (R.njobs + 1) * 8, s>>>(R);
            VF_CHECK_LAUNCH();
        }
        // stable sort by segment: (grid, label, point) order
        int bits = 1;
        while (bits < 32 && (1ull << bits) < (unsigned long long)S) ++bits;
        size_t sort_bytes = 0;
        VF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, R.keys, W.buf[kBKeys1].as<uint32_t>(), R.vals,
                                                  W.buf[kBVals1].as<unsigned long long>(), (int)np, 0, bits, s));
        W.ensure(kBSort, sort_bytes + 64);
        {
            TimedLaunch tl("mc_sort", s);
            VF_HIP(hipcub::DeviceRadixSort::SortPairs(W.buf[kBSort].p, sort_bytes, R.keys, W.buf[kBKeys1].as<uint32_t>(),
                                                      R.vals, W.buf[kBVals1].as<unsigned long long>(), (int)np, 0, bits,
                                                      s));
        }
        R.skeys = W.buf[kBKeys1].as<uint32_t>();
        R.svals = W.buf[kBVals1].as<unsigned long long>();
        W.ensure(kBCntT, np * 4 + 64);
        W.ensure(kBCntV, np * 4 + 64);
        W.ensure(kBTpre, (np + 1) * 8 + 64);
        W.ensure(kBVpre, (np + 1) * 8 + 64);
        R.cnt_t = W.buf[kBCntT].as<uint32_t>();
        R.cnt_v = W.buf[kBCntV].as<uint32_t>();
        const uint32_t gp = stride_blocks(np, 256, 16384);
        k_mc_split<<<gp, 256, 0, s>>>(R);
        VF_CHECK_LAUNCH();
        W.ensure(kBScan, scan_tmp_bytes(np) + 64);
        scan_exclusive_u32_to_u64(R.cnt_t, W.buf[kBTpre].as<unsigned long long>(), np, W.buf[kBScan].p, s);
        scan_exclusive_u32_to_u64(R.cnt_v, W.buf[kBVpre].as<unsigned long long>(), np, W.buf[kBScan].p, s);
        R.tpre = W.buf[kBTpre].as<unsigned long long>();
        R.vpre = W.buf[kBVpre].as<unsigned long long>();
        k_mc_segments<<<stride_blocks((uint64_t)S + 1, 256, 4096), 256, 0, s>>>(R);
        VF_CHECK_LAUNCH();
        {
            TimedLaunch tl("mc_emit", s);
            k_mc_emit<<<stride_blocks(np, kEmitPairs, 16384), kEmitPairs, 0, s>>>(R);
            VF_CHECK_LAUNCH();
        }
    }
    VF_HIP(hipStreamSynchronize(s));
    W.nv = (uint32_t)NV;
    W.nt = (uint32_t)NT;
    W.nl = S;
    W.valid = true;
}

void mc_run(McScratch& W, hipStream_t s, const uint16_t* grid, const uint32_t dims[3], const float mn[3], const float cs[3],
            const uint32_t* hist_dev, uint32_t want, vf_mesh_info* info)
{
    McJobSpec q{};
    q.grid = grid;
    q.hist = hist_dev;
    for (int a = 0; a < 3; ++a) {
        q.dims[a] = dims[a];
        q.mn[a] = mn[a];
        q.cs[a] = cs[a];
    }
    q.want = want;
    McJobResult r{};
    mc_run_jobs(W, s, &q, 1, &r, nullptr);
    if (r.status != VF_OK) {
        W.valid = false;
        fail(VF_ERR_INVALID_ARG, "marching cubes supports at most %u labels per grid", kMcMaxLabels);
    }
    info->num_vertices = r.nv;
    info->num_triangles = r.nt;
    info->num_labels = r.nl;
}

void mc_read_range(McScratch& W, hipStream_t s, uint32_t seg0, uint32_t nseg, float* xyz, uint32_t* tri, uint8_t* tri_flag,
                   uint16_t* labels, uint32_t* voff, uint32_t* toff)
{
    auto& B = W.buf;
    if ((uint64_t)seg0 + nseg > W.nl) fail(VF_ERR_INVALID_ARG, "mesh segment range out of bounds");
    std::vector<unsigned long long> vo(nseg + 1), to(nseg + 1);
    const auto* seg = B[kBSeg].as<unsigned long long>();
    VF_HIP(hipMemcpyAsync(vo.data(), seg + seg0, (nseg + 1) * 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipMemcpyAsync(to.data(), seg + (W.nl + 1) + seg0, (nseg + 1) * 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    const unsigned long long v0 = vo[0], nv = vo[nseg] - v0, t0 = to[0], nt = to[nseg] - t0;
    if (xyz && nv) VF_HIP(hipMemcpyAsync(xyz, B[kBXyz].as<float>() + 3 * v0, nv * 12, hipMemcpyDeviceToHost, s));
    if (tri && nt) VF_HIP(hipMemcpyAsync(tri, B[kBTri].as<uint32_t>() + 3 * t0, nt * 12, hipMemcpyDeviceToHost, s));
    if (tri_flag && nt) VF_HIP(hipMemcpyAsync(tri_flag, B[kBFlag].as<uint8_t>() + t0, nt, hipMemcpyDeviceToHost, s));
    if (labels && nseg) VF_HIP(hipMemcpyAsync(labels, B[kBLabels].as<uint16_t>() + seg0, (size_t)nseg * 2, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    for (uint32_t l = 0; l <= nseg; ++l) {
        if (voff) voff[l] = (uint32_t)(vo[l] - v0);
        if (toff) toff[l] = (uint32_t)(to[l] - t0);
    }
}

void mc_read_jobs(McScratch& W, hipStream_t s, const McJobResult* res, vf_mesh_out* const* outs, uint32_t n)
{
    if (!W.valid) return;
    auto& B = W.buf;
    std::vector<unsigned long long> seg(2 * ((size_t)W.nl + 1));
    VF_HIP(hipMemcpyAsync(seg.data(), B[kBSeg].p, seg.size() * 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    const unsigned long long* vo = seg.data();
    const unsigned long long* to = seg.data() + W.nl + 1;
    for (uint32_t j = 0; j < n; ++j) {
        vf_mesh_out* o = outs[j];
        if (!o) continue;
        const McJobResult& r = res[j];
        o->num_vertices = r.nv;
        o->num_triangles = r.nt;
        o->num_labels = r.nl;
        o->copied = 0;
        if (r.status != VF_OK || r.nv > o->cap_vertices || r.nt > o->cap_triangles || r.nl > o->cap_labels) continue;
        if ((r.nv && !o->xyz) || (r.nt && !o->tri)) continue;
        const unsigned long long v0 = vo[r.seg0], t0 = to[r.seg0];
        if (r.nv) VF_HIP(hipMemcpyAsync(o->xyz, B[kBXyz].as<float>() + 3 * v0, (size_t)r.nv * 12, hipMemcpyDeviceToHost, s));
        if (r.nt) VF_HIP(hipMemcpyAsync(o->tri, B[kBTri].as<uint32_t>() + 3 * t0, (size_t)r.nt * 12, hipMemcpyDeviceToHost, s));
        if (r.nt && o->tri_flag) VF_HIP(hipMemcpyAsync(o->tri_flag, B[kBFlag].as<uint8_t>() + t0, r.nt, hipMemcpyDeviceToHost, s));
        if (r.nl && o->labels)
            VF_HIP(hipMemcpyAsync(o->labels, B[kBLabels].as<uint16_t>() + r.seg0, (size_t)r.nl * 2, hipMemcpyDeviceToHost, s));
        for (uint32_t l = 0; l <= r.nl; ++l) {
            if (o->vert_offsets) o->vert_offsets[l] = (uint32_t)(vo[r.seg0 + l] - v0);
            if (o->tri_offsets) o->tri_offsets[l] = (uint32_t)(to[r.seg0 + l] - t0);
        }
        o->copied = 1;
    }
    VF_HIP(hipStreamSynchronize(s));
}

void mc_read(McScratch& W, hipStream_t s, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* voff,
             uint32_t* toff)
{
    mc_read_range(W, s, 0, W.nl, xyz, tri, tri_flag, labels, voff, toff);
}

void mc_smooth_jobs(McScratch& W, hipStream_t s, const McSmooth* per_job, uint32_t njobs)
{
    if (!W.valid) fail(VF_ERR_INVALID_ARG, "no marching-cubes mesh to smooth");
    if (njobs + 1 != W.label_base.size()) fail(VF_ERR_INVALID_ARG, "smoothing parameters for %u grids, mesh of %zu", njobs,
                                                W.label_base.size() - 1);
    uint32_t max_nb = 0, max_b = 0;
    std::vector<uint4> prm(njobs);
    for (uint32_t j = 0; j < njobs; ++j) {
        const McSmooth& q = per_job[j];
        if (!(q.nb_weight >= 0.0f && q.nb_weight <= 1.0f && q.b_weight >= 0.0f && q.b_weight <= 1.0f))
            fail(VF_ERR_INVALID_ARG, "smoothing weights must lie in [0, 1]");
        uint32_t wn, wb;
        std::memcpy(&wn, &q.nb_weight, 4);
        std::memcpy(&wb, &q.b_weight, 4);
        prm[j] = make_uint4(q.nb_iters, q.b_iters, wn, wb);
        max_nb = std::max(max_nb, q.nb_iters);
        max_b = std::max(max_b, q.b_iters);
    }
    if (W.nv == 0 || W.nt == 0 || (max_nb == 0 && max_b == 0)) return;
    auto& B = W.buf;
    SmoothDev D{};
    D.nv = W.nv;
    D.nt = W.nt;
    D.nl = W.nl;
    D.tri = B[kBTri].as<uint32_t>();
    D.tri_flag = B[kBFlag].as<uint8_t>();
    D.voff = B[kBSeg].as<unsigned long long>();
    D.toff = D.voff + (W.nl + 1);
    D.label_base = B[kBBase].as<uint32_t>();
    D.njobs = njobs;
    W.ensure(kBSmTriG, (size_t)W.nt * 12 + 64);
    W.ensure(kBSmFlag, (size_t)W.nv * 12 + 64);  // vflag, deg, vjob
    W.ensure(kBSmScan, scan_tmp_bytes(W.nv) + 64);
    W.ensure(kBSmAdj, (size_t)W.nt * 12 + 64);
    W.ensure(kBSmAdjOff, ((size_t)W.nv + 1) * 8 + 64);
    W.ensure(kBSmTmp, (size_t)W.nv * 12 + 64);  // ping-pong positions
    W.ensure(kBSmPrm, (size_t)njobs * sizeof(uint4) + 64);
    VF_HIP(hipMemcpyAsync(B[kBSmPrm].p, prm.data(), (size_t)njobs * sizeof(uint4), hipMemcpyHostToDevice, s));
    D.params = B[kBSmPrm].as<uint4>();
    D.tri_g = B[kBSmTriG].as<uint32_t>();
    D.vflag = B[kBSmFlag].as<uint32_t>();
    D.deg = D.vflag + W.nv;
    D.vjob = D.deg + W.nv;
    VF_HIP(hipMemsetAsync(D.vflag, 0, (size_t)W.nv * 8, s));
    const uint32_t gt = stride_blocks(W.nt, 256, 8192), gv = stride_blocks(W.nv, 256, 8192);
    k_smooth_prep<<<gt, 256, 0, s>>>(D);
    VF_CHECK_LAUNCH();
    scan_exclusive_u32_to_u64(D.deg, B[kBSmAdjOff].as<unsigned long long>(), W.nv, B[kBSmScan].p, s);
    VF_HIP(hipMemsetAsync(D.deg, 0, (size_t)W.nv * 4, s));
    D.adj_off = B[kBSmAdjOff].as<unsigned long long>();
    D.adj = B[kBSmAdj].as<uint32_t>();
    k_smooth_fill<<<gt, 256, 0, s>>>(D);
    VF_CHECK_LAUNCH();
    float* a = B[kBXyz].as<float>();
    float* b = B[kBSmTmp].as<float>();
    auto pass = [&](uint32_t iters, uint32_t target, int check) {
        for (uint32_t i = 0; i < iters; ++i) {
            TimedLaunch tl("mc_smooth", s);
            k_smooth_step<<<gv, 256, 0, s>>>(D, a, b, target, check, i);
            VF_CHECK_LAUNCH();
            std::swap(a, b);
        }
    };
    pass(max_nb, 0, 1);
    pass(max_b, 1, 0);
    if (a != B[kBXyz].as<float>()) VF_HIP(hipMemcpyAsync(B[kBXyz].p, a, (size_t)W.nv * 12, hipMemcpyDeviceToDevice, s));
    VF_HIP(hipStreamSynchronize(s));
}

void mc_smooth(McScratch& W, hipStream_t s, uint32_t nb_iters, float nb_weight, uint32_t b_iters, float b_weight)
{
    const McSmooth q{ nb_iters, nb_weight, b_iters, b_weight };
    mc_smooth_jobs(W, s, &q, 1);
}

void grid_marching_cubes(vf_grid* g, uint32_t want, vf_mesh_info* info)
{
    Launch L = grid_launch(g);
    launch_clear_hist(L);
    launch_histogram(L);
    g->mc.valid = false;
    mc_run(g->mc, g->stream, g->labels.as<uint16_t>(), g->dims, g->mn, g->cs, g->hist.as<uint32_t>(), want, info);
}

void grid_mesh_read(vf_grid* g, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* voff,
                    uint32_t* toff)
{
    mc_read(g->mc, g->stream, xyz, tri, tri_flag, labels, voff, toff);
}

void grid_mesh_slice(vf_grid* g, uint32_t v0, uint32_t nv, uint32_t t0, uint32_t nt, float* xyz, uint32_t* tri)
{
    auto& B = g->mc.buf;
    if ((uint64_t)v0 + nv > g->mc.nv || (uint64_t)t0 + nt > g->mc.nt) fail(VF_ERR_INVALID_ARG, "mesh slice out of range");
    if (nv) VF_HIP(hipMemcpyAsync(xyz, B[kBXyz].as<float>() + 3ull * v0, (size_t)nv * 12, hipMemcpyDeviceToHost, g->stream));
    if (nt) VF_HIP(hipMemcpyAsync(tri, B[kBTri].as<uint32_t>() + 3ull * t0, (size_t)nt * 12, hipMemcpyDeviceToHost, g->stream));
    VF_HIP(hipStreamSynchronize(g->stream));
}

}  // namespace vf
