// Flood (region-growing) fracture: FloodFracturer::build
// (MF/Source/Fracturer/FloodFracturer.cpp:98-191, floodFracturer-comp.glsl:24-63,
// undoMask-comp.glsl unmaskRightMost).
//
// The reference grows every seed one BFS step per dispatch and resolves
// simultaneous claims by a race.  Here every cell gets the label of the seed
// minimising (D, seed index), D the node-weighted geodesic distance in the 6-
// or 26-neighbour graph of occupied cells with weight 1 + noise(cell) in
// {1, 2, 3}.  With the packed key K = (D << b) | index this is the unique
// fixpoint of
//     K(seed cell) = (1 << b) | index,   K(v) = min_u K(u) + (w(v) << b)
// over the occupied neighbours u of v (positive weights make it unique), so
// any relaxation order that reaches the fixpoint gives the same labels as a
// per-seed Dijkstra.
//
// The relaxation runs on 8x8x8 bricks.  A round processes a deduplicated
// worklist of bricks; one wave loads a brick plus a one-cell halo of keys into
// LDS and iterates min-plus sweeps (column minima from LDS, plus a z sweep in
// registers) until the brick is locally stable, writes the cells that dropped
// and queues the neighbour bricks that see a dropped cell through a face,
// edge or corner.  A round needs no device-scope atomics per cell and the
// rounds needed follow the geodesic depth in bricks, not in cells.
//
// Halo values read while a neighbour brick is being rewritten in the same
// round may be old or new: both are upper bounds of the fixpoint, and a
// brick that lowers a boundary cell queues its neighbours for the next round,
// which reads the value after the kernel boundary.
#include "cellops.hpp"

namespace vf {

namespace {

// Flood key of a cell: 0xFFFFFFFF = EMPTY; otherwise (K << 2) | w with w = 1 + noise in
// 1..3 the cell's node weight and K = (D << b) | seed index, or K = 0x3FFFFFFC
// (key kUnreachedKey | w) while unreached.  Keys compare by K first (the weight bits only break ties
// between equal K, which never changes K), so the minimum over neighbours is
// still the (distance, seed) minimum, and a cell's weight travels with it; EMPTY, all ones, never
// wins a minimum, so the relaxation copies keys into LDS as they are.
constexpr uint32_t kEmptyKey = kFloodEmptyKey;
constexpr uint32_t kUnreachedKey = kFloodUnreachedKey;  // | w
// LDS tile value of EMPTY cells and of cells outside the grid
constexpr uint32_t kTileEmpty = 0xFFFFFFFFu;
static_assert(kEmptyKey == kTileEmpty, "keys are copied into the tile as they are");
// Halo z rows of 10 words (z = -1..8) at an odd stride, so that the row starts
// of a wave's 64 lanes fall on distinct LDS banks.
constexpr uint32_t kRowWords = 11;
// Lanes whose halo-plane position (hy, hz) = (p / 10, p % 10) falls in (y, z) class c = cy * 3 + cz
// (0: below the brick, 1: beside it, 2: above it): the first position p = lane, the second p = lane + 64
// (lanes 0..35; lanes 36..63 repeat their first).
constexpr uint32_t halo_class(uint32_t p)
{
    const uint32_t hy = p / 10u, hz = p % 10u;
    return (hy == 0u ? 0u : (hy == 9u ? 2u : 1u)) * 3u + (hz == 0u ? 0u : (hz == 9u ? 2u : 1u));
}
constexpr uint64_t halo_lanes(uint32_t c, uint32_t off, uint32_t n)
{
    uint64_t m = 0;
    for (uint32_t l = 0; l < n; ++l)
        if (halo_class(l + off) == c) m |= 1ull << l;
    return m;
}
constexpr uint64_t kHaloLanes1[9] = { halo_lanes(0, 0, 64), halo_lanes(1, 0, 64), halo_lanes(2, 0, 64),
                                      halo_lanes(3, 0, 64), halo_lanes(4, 0, 64), halo_lanes(5, 0, 64),
                                      halo_lanes(6, 0, 64), halo_lanes(7, 0, 64), halo_lanes(8, 0, 64) };
constexpr uint64_t kHaloLanes2[9] = { halo_lanes(0, 64, 36), halo_lanes(1, 64, 36), halo_lanes(2, 64, 36),
                                      halo_lanes(3, 64, 36), halo_lanes(4, 64, 36), halo_lanes(5, 64, 36),
                                      halo_lanes(6, 64, 36), halo_lanes(7, 64, 36), halo_lanes(8, 64, 36) };
constexpr uint32_t kTileWords = 10 * 10 * kRowWords;
// 26-neighbour minima, separably: zt = z-window minima of the 100 halo-plane rows (8 per row),
// yt = y-window minima of the 10 x-slabs' 8 rows; rows at an odd stride of 9 words.  yt reuses
// the tile's words (the tile is read only while a visit starts).
constexpr uint32_t kSepStride = 9;
constexpr uint32_t kZtWords = 100 * kSepStride;
constexpr uint32_t kYtWords = 80 * kSepStride;
static_assert(kYtWords <= kTileWords, "yt reuses the tile");
constexpr uint32_t kRelaxLdsWords = kTileWords + kZtWords;
// Consecutive worklist entries a relax workgroup takes at a time (batch step 11.47 ms with one run per
// workgroup, 11.28 / 11.25 / 11.30 / 11.50 / 11.64 ms with runs of 2 / 3 / 4 / 5 / 6, 11.83 with 8)
constexpr uint32_t kRelaxRun = 3;
constexpr uint32_t kMaxSweeps = 2048;              // a brick of 512 cells is stable after at most 513 sweeps
constexpr uint32_t kXcds = kFloodXcds;             // XCDs of the device; workgroups are dealt round-robin

__device__ inline uint32_t bdim(uint32_t n) { return (n + kRelaxBrick - 1) / kRelaxBrick; }

// wave-uniform values in scalar registers
__device__ inline uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
template <typename T>
__device__ inline T* uni_ptr(T* p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint64_t lo = uni((uint32_t)v), hi = uni((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(lo | (hi << 32));
}

// One block per job: stamp the seeds (a later seed on the same cell wins, as
// FloodFracturer.cpp:99-105 writes them in list order), mark seed cells
// occupied and flag, for round 1, every brick that holds a seed cell or one of
// its neighbours.  Both flag arrays were zeroed by an earlier launch.  A seed
// on an EMPTY cell can sit in a brick the key init left without keys: such a
// brick gets EMPTY keys and its liveness bit first.
__global__ __launch_bounds__(256) void k_flood_seed(Launch L)
{
    __shared__ uint32_t todo[256];
    __shared__ uint32_t ntodo;
    const JobDev* J = L.jobs + blockIdx.x;
    if (!J->active || J->algo != VF_FLOOD) return;
    const uint32_t S = J->S, b = J->key_bits;
    const uint32_t nby = bdim(J->ny), nbz = bdim(J->nz);
    unsigned long long* live = J->bflags + 2u * relax_words_per_parity(J->nbricks);
    const uint32_t wpc = live_words_per_column(nbz);
    auto brick_of = [&](const uint4& q) {
        return ((q.x / kRelaxBrick) * nby + q.y / kRelaxBrick) * nbz + q.z / kRelaxBrick;
    };
    // liveness word and bit of a brick id
    auto live_word = [&](uint32_t id) { return (size_t)(id / nbz) * wpc + (id % nbz) / 64u; };
    auto live_bit = [&](uint32_t id) { return 1ull << ((id % nbz) & 63u); };
    for (uint32_t s0 = 0; s0 < S; s0 += blockDim.x) {
        if (threadIdx.x == 0) ntodo = 0;
        __syncthreads();
        if (s0 + threadIdx.x < S) {
            const uint32_t id = brick_of(J->seeds[s0 + threadIdx.x]);
            if (!(live[live_word(id)] & live_bit(id))) todo[atomicAdd(&ntodo, 1u)] = id;
        }
        __syncthreads();
        const uint32_t nt = ntodo;  // a brick listed twice is written twice with the same values
        for (uint32_t t = 0; t < nt; ++t) {
            const uint32_t id = todo[t];
            const uint32_t X0 = id / nbz / nby * kRelaxBrick, Y0 = id / nbz % nby * kRelaxBrick, Z0 = id % nbz * kRelaxBrick;
            for (uint32_t e = threadIdx.x; e < 512u; e += blockDim.x) {
                const uint32_t x = X0 + (e >> 6), y = Y0 + ((e >> 3) & 7u), z = Z0 + (e & 7u);
                if (x < J->nx && y < J->ny && z < J->nz) J->keys[((size_t)x * J->ny + y) * J->nz + z] = kEmptyKey;
            }
        }
        __syncthreads();
        if (threadIdx.x < nt) atomicOr(live + live_word(todo[threadIdx.x]), live_bit(todo[threadIdx.x]));
    }
    if (threadIdx.x < kFloodCtlWords) J->fctl[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        J->keys[((size_t)q.x * J->ny + q.y) * J->nz + q.z] = 0u;
    }
    __syncthreads();
    auto seed_key = [&](uint32_t s, size_t c) {
        return (((1u << b) | s) << 2) | flood_weight(J->noise_seed, (uint32_t)c, J->noise_amp);
    };
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        const size_t c = ((size_t)q.x * J->ny + q.y) * J->nz + q.z;
        atomicMax(J->keys + c, seed_key(s, c));
    }
    __syncthreads();
    unsigned long long* flags1 = J->bflags + relax_words_per_parity(J->nbricks);  // round 1 consumes parity 1
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        const size_t c = ((size_t)q.x * J->ny + q.y) * J->nz + q.z;
        if (J->keys[c] != seed_key(s, c)) continue;
        J->occ[c] = 1;
        // bricks of the 3x3x3 neighbourhood: at most two brick indices per axis
        const uint32_t bx0 = (q.x ? q.x - 1 : 0) / kRelaxBrick, bx1 = min(q.x + 1, J->nx - 1) / kRelaxBrick;
        const uint32_t by0 = (q.y ? q.y - 1 : 0) / kRelaxBrick, by1 = min(q.y + 1, J->ny - 1) / kRelaxBrick;
        const uint32_t bz0 = (q.z ? q.z - 1 : 0) / kRelaxBrick, bz1 = min(q.z + 1, J->nz - 1) / kRelaxBrick;
        for (uint32_t bx = bx0; bx <= bx1; ++bx)
            for (uint32_t by = by0; by <= by1; ++by)
                for (uint32_t bz = bz0; bz <= bz1; ++bz) {
                    const uint32_t id = (bx * nby + by) * nbz + bz;
                    atomicOr(flags1 + (id >> 6), 1ull << (id & 63u));
                }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        J->fctl[0] = 1;
        J->fctl[3] = 1;  // parity-1 flags are non-empty
    }
}

// Per brick: the liveness bits of the brick and its 26 neighbours (bit (dx + 1) * 9 + (dy + 1) * 3 +
// (dz + 1); neighbours outside the grid are not live), read by every relaxation visit.  One thread per
// brick; a job's bricks are dealt over its compaction tasks, 4096 per task.
__global__ __launch_bounds__(256) void k_flood_nbr(Launch L, const uint32_t* task_prefix, uint32_t total_tasks)
{
    const uint64_t slots = (uint64_t)total_tasks * 4096u;
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x; g0 < slots; g0 += (uint64_t)gridDim.x * blockDim.x) {
        // the block's 256 bricks lie in one task (4096 bricks): one job per block
        const uint32_t task = (uint32_t)(g0 >> 12);
        const uint32_t j = uni(find_job(task_prefix, L.njobs, task));
        const JobDev* J = uni_ptr(L.jobs + j);
        if (uni((uint32_t)(J->active && J->algo == VF_FLOOD)) == 0u) continue;
        const uint32_t id = (uint32_t)(g0 - (uint64_t)uni(task_prefix[j]) * 4096u) + threadIdx.x;
        const uint32_t nb = uni(J->nbricks);
        const uint32_t nbx = bdim(uni(J->nx)), nby = bdim(uni(J->ny)), nbz = bdim(uni(J->nz)), wpc = live_words_per_column(nbz);
        unsigned long long* live = uni_ptr(J->bflags) + 2u * relax_words_per_parity(nb);
        const uint32_t idc = min(id, nb - 1u);
        const uint32_t bz = idc % nbz, t1 = idc / nbz, by = t1 % nby, bx = t1 / nby;
        const uint32_t q = bz & 63u, wi = bz >> 6;
        // per neighbouring column: the words holding bz and, where bz sits at a word's end, the next
        // or previous one (only for columns of more than 64 bricks); every index is clamped into range
        uint64_t w0[9], wm[9], wp[9];
        bool in[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const int qx = (int)bx + c / 3 - 1, qy = (int)by + c % 3 - 1;
            in[c] = qx >= 0 && qy >= 0 && (uint32_t)qx < nbx && (uint32_t)qy < nby;
            const size_t base = ((size_t)min(max(qx, 0), (int)nbx - 1) * nby + (uint32_t)min(max(qy, 0), (int)nby - 1)) * wpc;
            w0[c] = live[base + wi];
            wm[c] = wp[c] = 0ull;
        }
        if (wpc > 1u) {
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                const int qx = (int)bx + c / 3 - 1, qy = (int)by + c % 3 - 1;
                const size_t base = ((size_t)min(max(qx, 0), (int)nbx - 1) * nby + (uint32_t)min(max(qy, 0), (int)nby - 1)) * wpc;
                wm[c] = live[base + (wi ? wi - 1u : 0u)];
                wp[c] = live[base + min(wi + 1u, wpc - 1u)];
            }
        }
        uint32_t m = 0;
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            // bits of bz - 1, bz, bz + 1 (bits past the column's last brick are clear in the bitmap)
            const uint32_t lo = q ? (uint32_t)(w0[c] >> (q - 1u)) & 1u : (wi ? (uint32_t)(wm[c] >> 63) : 0u);
            const uint32_t mid = (uint32_t)(w0[c] >> q) & 1u;
            const uint32_t hi = q < 63u ? (uint32_t)(w0[c] >> (q + 1u)) & 1u : (wi + 1u < wpc ? (uint32_t)wp[c] & 1u : 0u);
            if (in[c]) m |= (lo | (mid << 1) | (hi << 2)) << (3 * c);
        }
        if (id < nb) reinterpret_cast<uint32_t*>(live + live_words(nbx, nby, nbz))[id] = m;
    }
}

// Uniform per-job state of a round, held in scalar registers.
struct RelaxJob {
    const JobDev* J;
    uint32_t nx, ny, nz, nby, nbz, nbx, nb, b, amp, r;
    uint64_t noise_seed;
    VF_AS_GLOBAL uint32_t* keys;
    uint32_t nw;                // flag words per parity
    VF_AS_GLOBAL unsigned long long* cur;  // brick flag bits consumed this round
    VF_AS_GLOBAL unsigned long long* nxt;  // brick flag bits for the next round
    const VF_AS_GLOBAL uint32_t* nmask;  // per brick: liveness of the brick and its 26 neighbours
    bool n26;
};


// Relax one brick to local stability against its halo, write the cells that
// dropped and flag the neighbour bricks that see one of them.  Returns the
// sweeps run; *activated is set when a neighbour was flagged.
__device__ inline uint32_t relax_brick(const RelaxJob& R, uint32_t brick, uint32_t* tile, bool& activated, bool& error)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t nx = R.nx, ny = R.ny, nz = R.nz;
    const uint32_t bz = brick % R.nbz, t1 = brick / R.nbz, by = t1 % R.nby, bx = t1 / R.nby;
    const int X0 = (int)(bx * kRelaxBrick), Y0 = (int)(by * kRelaxBrick), Z0 = (int)(bz * kRelaxBrick);

    uint32_t halo_live;  // 27 bits: halo classes holding a cell that is not EMPTY
    // liveness of the brick and its 26 neighbours, bit ix * 9 + iy * 3 + iz (each 0: below, 1: the
    // brick's own range, 2: above), loaded with the halo.  Cells of a brick whose bit is clear (or
    // outside the grid) hold no keys and count as EMPTY.
    const uint32_t nlive = R.nmask[brick];
    // brick + halo -> LDS (cells outside the grid, of bricks without keys and EMPTY cells as
    // kTileEmpty).  Each lane owns one or two (y, z) positions of the 10 x 10 halo plane and loads
    // them in all 10 x slices at once; addresses are clamped into the grid so every load is
    // unconditional, and cells outside the grid or of bricks without keys are masked afterwards.
    {
        // lanes 36..63 have no second position and repeat their first (same cell, same word)
        const uint32_t p1 = lane, p2 = lane < 36u ? lane + 64u : lane;
        const uint32_t hy1 = p1 / 10u, hz1 = p1 - hy1 * 10u, hy2 = p2 / 10u, hz2 = p2 - hy2 * 10u;
        const int gy1 = Y0 + (int)hy1 - 1, gz1 = Z0 + (int)hz1 - 1, gy2 = Y0 + (int)hy2 - 1, gz2 = Z0 + (int)hz2 - 1;
        const bool in1 = (uint32_t)gy1 < ny && (uint32_t)gz1 < nz;
        const bool in2 = (uint32_t)gy2 < ny && (uint32_t)gz2 < nz;
        const uint32_t o1 = (uint32_t)min(max(gy1, 0), (int)ny - 1) * nz + (uint32_t)min(max(gz1, 0), (int)nz - 1);
        const uint32_t o2 = (uint32_t)min(max(gy2, 0), (int)ny - 1) * nz + (uint32_t)min(max(gz2, 0), (int)nz - 1);
        const uint32_t t1 = hy1 * kRowWords + hz1, t2 = hy2 * kRowWords + hz2;  // tile offsets within an x slice
        const uint64_t slice = (uint64_t)ny * nz;
        // each x slice's base address is uniform (64-bit, scalar); the lane offsets within a slice
        // (below 2^22 cells, 2^24 bytes) are 32-bit: a scalar base plus a 32-bit offset per load
        const uint32_t ob1 = o1 * 4u, ob2 = o2 * 4u;
        const VF_AS_GLOBAL uint8_t* kbytes = (const VF_AS_GLOBAL uint8_t*)R.keys;
        uint32_t v1[10], v2[10];
#pragma unroll
        for (uint32_t hx = 0; hx < 10u; ++hx) {
            const uint32_t sx = uni((uint32_t)min(max(X0 + (int)hx - 1, 0), (int)nx - 1));
            const VF_AS_GLOBAL uint8_t* sb = kbytes + sx * slice * 4u;
            v1[hx] = *(const VF_AS_GLOBAL uint32_t*)(sb + ob1);
            v2[hx] = *(const VF_AS_GLOBAL uint32_t*)(sb + ob2);
        }
        // outside the grid or in a brick without keys -> all ones (EMPTY keys are all ones already);
        // per x class ix (slice 0, slices 1..8, slice 9) of the position's (y, z) class
        const uint32_t c1 = halo_class(p1), c2 = halo_class(p2);
        uint32_t bad1[3], bad2[3];
#pragma unroll
        for (uint32_t ix = 0; ix < 3u; ++ix) {
            bad1[ix] = in1 && ((nlive >> (ix * 9u + c1)) & 1u) ? 0u : kTileEmpty;
            bad2[ix] = in2 && ((nlive >> (ix * 9u + c2)) & 1u) ? 0u : kTileEmpty;
        }
        // halo_live bit ix * 9 + iy * 3 + iz: a cell of that halo class (x, y, z each below the brick, beside
        // it, above it) is not EMPTY; a neighbour brick in direction (ix, iy, iz) can only take a drop from
        // this brick through those cells (EMPTY keys never change).  Per position and x class a ballot of
        // "not all EMPTY" (across x slices 1..8 the AND of the keys is all ones iff all are EMPTY), then the
        // lanes of each (y, z) class picked by constant masks.
        uint32_t m1 = ~0u, m2 = ~0u;
        uint64_t b1[3], b2[3];
#pragma unroll
        for (uint32_t hx = 0; hx < 10u; ++hx) {
            const uint32_t xbad = (uint32_t)(X0 + (int)hx - 1) < nx ? 0u : kTileEmpty;
            const uint32_t ix = hx == 0u ? 0u : (hx == 9u ? 2u : 1u);
            const uint32_t a1 = v1[hx] | bad1[ix] | xbad, a2 = v2[hx] | bad2[ix] | xbad;
            tile[hx * 10u * kRowWords + t1] = a1;
            tile[hx * 10u * kRowWords + t2] = a2;
            if (hx == 0u || hx == 9u) {
                b1[hx == 0u ? 0 : 2] = __ballot(a1 != ~0u);
                b2[hx == 0u ? 0 : 2] = __ballot(a2 != ~0u);
            } else {
                m1 &= a1;
                m2 &= a2;
            }
        }
        b1[1] = __ballot(m1 != ~0u);
        b2[1] = __ballot(m2 != ~0u);
        uint32_t live = 0;
#pragma unroll
        for (uint32_t ix = 0; ix < 3u; ++ix)
#pragma unroll
            for (uint32_t c = 0; c < 9u; ++c)
                if ((b1[ix] & kHaloLanes1[c]) | (b2[ix] & kHaloLanes2[c])) live |= 1u << (ix * 9u + c);
        halo_live = live;
    }
    __syncthreads();

    // the lane's own z row of 8 cells
    const uint32_t x = lane >> 3, y = lane & 7u;
    const uint32_t own = ((x + 1u) * 10u + (y + 1u)) * kRowWords;
    const uint32_t wsh = R.b + 2u;
    // During the sweeps keys are held with their weight bits cleared (EMPTY as 0xFFFFFFFC, unreached
    // as kUnreachedKey).  k: the row's keys; wmask: 2 bits per cell, its weight (0 = EMPTY);
    // add: what a neighbour's key gains on entering the cell, w << (b + 2), or all ones for EMPTY
    // cells, so that
    //     relaxed = min(k, sat_add(neighbour, add))
    // leaves EMPTY cells alone and never lets an EMPTY or unreached neighbour win: their sums either
    // saturate or stay at or above every key they could replace.
    uint32_t k[8], add[8];
    uint32_t wmask = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const uint32_t t = tile[own + 1 + z];
        const uint32_t w = t != kTileEmpty ? (t & 3u) : 0u;
        wmask |= w << (2 * z);
        k[z] = t & ~3u;
        add[z] = w ? (w << wsh) : 0xFFFFFFFFu;
    }
    // what the neighbour flags and the write-back need: whether any cell dropped, and the end cells
    const uint32_t k0_first = k[0], k0_last = k[7];
    uint32_t anydiff = 0;
    auto relax = [&](uint32_t cur, uint32_t nb, uint32_t a) { return min(cur, __builtin_elementwise_add_sat(nb, a)); };
    uint32_t* zt = tile + kTileWords;
    uint32_t* yt = tile;
    uint32_t hzlo = 0, hzhi = 0;  // the own row's halo keys at z = -1 and z = 8
    // 26 neighbours: the rows at y = 0 and y = 7 take their outer y neighbour from a halo row of zt (fixed
    // during the visit), the others from the neighbouring lane
    const uint32_t hrow = ((x + 1u) * 10u + (y == 0u ? 0u : 9u)) * kSepStride;
    hzlo = tile[own] & ~3u;
    hzhi = tile[own + 9] & ~3u;
    // 6 neighbours: the halo row beside y = 0 or y = 7 in the tile (not rewritten during the visit)
    const uint32_t hrow6 = ((x + 1u) * 10u + (y == 0u ? 0u : 9u)) * kRowWords + 1u;
    if (R.n26) {
        // the minima that stay fixed during the visit: zt of the 36 halo rows (x-slab 0 or 9, or y-row
        // 0 or 9), then yt of the halo x-slabs 0 and 9
        if (lane < 36u) {
            const uint32_t hx = lane < 10u ? 0u : (lane < 20u ? 9u : (lane < 28u ? lane - 19u : lane - 27u));
            const uint32_t hy = lane < 10u ? lane : (lane < 20u ? lane - 10u : (lane < 28u ? 0u : 9u));
            const uint32_t* row = tile + (hx * 10u + hy) * kRowWords;
            uint32_t r[10];
#pragma unroll
            for (int o = 0; o < 10; ++o) r[o] = row[o];
#pragma unroll
            for (int z = 0; z < 8; ++z) zt[(hx * 10u + hy) * kSepStride + z] = min(min(r[z], r[z + 1]), r[z + 2]) & ~3u;
        }
        __syncthreads();  // the tile is read for the last time above; yt may now overwrite it
        if (lane < 16u) {
            const uint32_t sx = lane < 8u ? 0u : 9u, sy = lane & 7u;
#pragma unroll
            for (int z = 0; z < 8; ++z)
                yt[(sx * 8u + sy) * kSepStride + z] = min(min(zt[(sx * 10u + sy) * kSepStride + z],
                                                              zt[(sx * 10u + sy + 1u) * kSepStride + z]),
                                                          zt[(sx * 10u + sy + 2u) * kSepStride + z]);
        }
    }
    uint32_t sweeps = 0;
    for (;;) {
        uint32_t m[8];
        if (R.n26) {
            // z pass in registers, the y pass across lanes (lane x * 8 + y: the y neighbours are lanes -1 / +1,
            // by DPP row shifts), then the x pass through LDS; every window includes the cell
            uint32_t zo[8], ym[8];
            zo[0] = min(min(hzlo, k[0]), k[1]);
#pragma unroll
            for (int z = 1; z < 7; ++z) zo[z] = min(min(k[z - 1], k[z]), k[z + 1]);
            zo[7] = min(min(k[6], k[7]), hzhi);
#pragma unroll
            for (int z = 0; z < 8; ++z) {
                const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp((int)zo[z], (int)zo[z], 0x111, 0xF, 0xF, false);
                const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp((int)zo[z], (int)zo[z], 0x101, 0xF, 0xF, false);
                const uint32_t hz = zt[hrow + z];
                ym[z] = min(min(y == 0u ? hz : dn, zo[z]), y == 7u ? hz : up);
            }
#pragma unroll
            for (int z = 0; z < 8; ++z) yt[((x + 1u) * 8u + y) * kSepStride + z] = ym[z];
            __syncthreads();
#pragma unroll
            for (int z = 0; z < 8; ++z)
                m[z] = min(min(yt[(x * 8u + y) * kSepStride + z], ym[z]), yt[((x + 2u) * 8u + y) * kSepStride + z]);
        } else {
            // the row's own z neighbours from registers, its y neighbours from the neighbouring lanes (DPP row
            // shifts; the rows at y = 0 and y = 7 take a halo row, fixed during the visit), its x neighbours
            // from the tile, where every row is written back after a sweep that changed it
            uint32_t fm[8];
#pragma unroll
            for (int z = 0; z < 8; ++z) {
                const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp((int)k[z], (int)k[z], 0x111, 0xF, 0xF, false);
                const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp((int)k[z], (int)k[z], 0x101, 0xF, 0xF, false);
                const uint32_t hz = tile[hrow6 + z];
                fm[z] = min(y == 0u ? hz : dn, y == 7u ? hz : up);
            }
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                const uint32_t* row = tile + ((f == 0 ? x : x + 2u) * 10u + y + 1u) * kRowWords + 1u;
#pragma unroll
                for (int o = 0; o < 8; ++o) fm[o] = min(fm[o], row[o]);
            }
#pragma unroll
            for (int z = 0; z < 8; ++z) m[z] = min(min(z == 0 ? hzlo : k[z - 1], z == 7 ? hzhi : k[z + 1]), fm[z]) & ~3u;
        }
        uint32_t prev[8];
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            prev[z] = k[z];
            k[z] = relax(k[z], m[z], add[z]);
        }
        // z sweeps in registers: the row's own neighbours along z
#pragma unroll
        for (int z = 1; z < 8; ++z) k[z] = relax(k[z], k[z - 1], add[z]);
#pragma unroll
        for (int z = 6; z >= 0; --z) k[z] = relax(k[z], k[z + 1], add[z]);
        uint32_t diff = 0;
#pragma unroll
        for (int z = 0; z < 8; ++z) diff |= k[z] ^ prev[z];
        anydiff |= diff;
        if (diff && !R.n26) {  // the 26-neighbour passes read zt / yt, rebuilt every sweep
#pragma unroll
            for (int z = 0; z < 8; ++z) tile[own + 1 + z] = k[z];
        }
        ++sweeps;
        __syncthreads();
        if (__ballot(diff != 0) == 0ull) break;
        if (sweeps >= kMaxSweeps) {
            error = true;
            break;
        }
    }
    // Key overflow: a reached key within one step (at most weight 3) of kUnreachedKey could have a
    // successor that does not fit; flag it (conservatively) where the key was made.
    const bool changed = anydiff != 0;  // a cell of the row dropped in this visit
    const uint32_t zlo = k[0] != k0_first ? 1u : 0u, zhi = k[7] != k0_last ? 1u : 0u;  // which end cells
    const uint32_t headroom = kUnreachedKey - (3u << wsh);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        if (k[z] < kUnreachedKey && k[z] >= headroom) error = true;
        const uint32_t w = (wmask >> (2 * z)) & 3u;
        k[z] = w ? (k[z] | w) : kEmptyKey;  // the stored form: weight bits back
    }

    // write back the rows with a dropped cell, whole: the other cells rewrite the values they hold
    // (a brick is visited at most once per round, so nothing else writes them meanwhile); two
    // 16-byte stores where the row is aligned and complete, else cell by cell within the grid
    if (changed) {
        const uint32_t row = (((uint32_t)X0 + x) * ny + ((uint32_t)Y0 + y)) * nz + (uint32_t)Z0;
        if ((uint32_t)Z0 + 8u <= nz && (nz & 3u) == 0) {
            *(VF_AS_GLOBAL vf_u32x4*)(R.keys + row) = vf_u32x4{ k[0], k[1], k[2], k[3] };
            *(VF_AS_GLOBAL vf_u32x4*)(R.keys + row + 4) = vf_u32x4{ k[4], k[5], k[6], k[7] };
        } else {
#pragma unroll
            for (int z = 0; z < 8; ++z)
                if ((uint32_t)Z0 + (uint32_t)z < nz) R.keys[row + z] = k[z];
        }
    }

    // neighbour bricks that see a dropped cell: bit (dx+1)*9 + (dy+1)*3 + (dz+1).  Lane x*8 + y holds row
    // (x, y): a brick (dx, dy, dz) sees a dropped cell of a row with x = 0 (dx = -1) or 7 (dx = +1) or any
    // x (dx = 0), likewise for y, and a dropped end cell z = 0 (dz = -1) or 7 (dz = +1) or any cell
    // (dz = 0); the 6-neighbour stencil sees faces only.  Built from three ballots with lane masks.
    const uint64_t bzm[3] = { __ballot(changed && zlo), __ballot(changed), __ballot(changed && zhi) };
    constexpr uint64_t kAll = ~0ull, kX0 = 0xFFull, kX7 = 0xFFull << 56, kY0 = 0x0101010101010101ull,
                       kY7 = 0x8080808080808080ull;
    const uint64_t mxs[3] = { kX0, kAll, kX7 }, mys[3] = { kY0, kAll, kY7 };
    uint32_t dirs = 0;
    if (R.n26) {
#pragma unroll
        for (uint32_t ix = 0; ix < 3; ++ix)
#pragma unroll
            for (uint32_t iy = 0; iy < 3; ++iy)
#pragma unroll
                for (uint32_t iz = 0; iz < 3; ++iz)
                    if (ix * 9 + iy * 3 + iz != 13 && (bzm[iz] & mxs[ix] & mys[iy])) dirs |= 1u << (ix * 9 + iy * 3 + iz);
    } else {
        dirs = ((bzm[1] & kX0) ? 1u << 4 : 0u) | ((bzm[1] & kX7) ? 1u << 22 : 0u) | ((bzm[1] & kY0) ? 1u << 10 : 0u) |
               ((bzm[1] & kY7) ? 1u << 16 : 0u) | (bzm[0] ? 1u << 12 : 0u) | (bzm[2] ? 1u << 14 : 0u);
    }
    dirs &= halo_live;  // a neighbour whose cells facing this brick are all EMPTY cannot take a drop from it
    if (lane < 27u && (dirs & (1u << lane))) {
        const int qx = (int)bx + (int)(lane / 9u) - 1, qy = (int)by + (int)((lane / 3u) % 3u) - 1,
                  qz = (int)bz + (int)(lane % 3u) - 1;
        if (qx >= 0 && qy >= 0 && qz >= 0 && (uint32_t)qx < R.nbx && (uint32_t)qy < R.nby && (uint32_t)qz < R.nbz)
        {
            const uint32_t id = ((uint32_t)qx * R.nby + (uint32_t)qy) * R.nbz + (uint32_t)qz;
            __hip_atomic_fetch_or(R.nxt + (id >> 6), 1ull << (id & 63u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    activated |= dirs != 0;
    __syncthreads();  // the tile is reused by the next brick
    return sweeps;
}

// The round-independent part of a flood job's relaxation state, in scalar registers, read in one batch
// of loads; returns the job's flag words (two parities, then the liveness words and neighbour masks).
__device__ inline unsigned long long* load_relax_fields(const JobDev* J, RelaxJob& R)
{
    R.J = uni_ptr(J);
    R.nx = uni(J->nx); R.ny = uni(J->ny); R.nz = uni(J->nz);
    R.nbx = bdim(R.nx); R.nby = bdim(R.ny); R.nbz = bdim(R.nz);
    R.nb = uni(J->nbricks);
    R.b = uni(J->key_bits);
    R.amp = uni(J->noise_amp);
    R.noise_seed = ((uint64_t)uni((uint32_t)(J->noise_seed >> 32)) << 32) | uni((uint32_t)J->noise_seed);
    R.keys = gcast<uint32_t>(uni_ptr(J->keys));
    R.nw = relax_words_per_parity(R.nb);
    unsigned long long* flags = uni_ptr(J->bflags);
    R.nmask = gcast<const uint32_t>(reinterpret_cast<const uint32_t*>(flags + (size_t)R.nw * 2u + live_words(R.nbx, R.nby, R.nbz)));
    R.n26 = uni((uint32_t)J->neighbourhood) != 6u;
    return flags;
}

// Round r's flag words: round r consumes parity r & 1 and flags into the other.
__device__ inline void set_relax_round(RelaxJob& R, unsigned long long* flags, uint32_t r)
{
    R.r = r;
    R.cur = gcast<unsigned long long>(flags + (size_t)R.nw * (r & 1u));
    R.nxt = gcast<unsigned long long>(flags + (size_t)R.nw * ((r + 1u) & 1u));
}

// A flood job is running and not drained: its record's flag (read with its fields, before any branch).
__device__ inline uint32_t flood_job_active(const JobDev* J)
{
    return uni((uint32_t)((J->active != 0) & (J->algo == VF_FLOOD)));
}

// A running job's state for the relaxation step (its round was set by this round's compaction): the job's
// fields, then its control words, two dependent batches of loads.
__device__ inline bool load_relax_job(const Launch& L, uint32_t j, RelaxJob& R)
{
    const JobDev* J = L.jobs + j;
    const uint32_t act = flood_job_active(J);
    const uint32_t* ctl = uni_ptr(J->fctl);
    unsigned long long* flags = load_relax_fields(J, R);
    if (!act) return false;
    const uint32_t done = uni(ctl[1]), r = uni(ctl[0]);
    if (done) return false;
    set_relax_round(R, flags, r);
    return true;
}

// Start job j's next round (in the compaction of chunk round i): the round the previous
// compaction started is read from slot fctl[4 + ((i + 1) & 1)]; when the relaxation of that round
// flagged nothing the job is drained there, else round rp + 1 starts: it goes into slot
// fctl[4 + (i & 1)] and fctl[0], and the flags-non-empty word of round rp is cleared (round
// rp + 1's relaxation writes it for round rp + 2).  Every workgroup of the job computes the same
// values from words no workgroup writes in this launch, and writes them: no ordering is needed.
// The loads come in two dependent batches: the job's fields, then its control words together with
// both parities' flag word w (lane w of the task), of which *m receives the new round's.
__device__ inline bool start_round(const Launch& L, uint32_t j, uint32_t i, uint32_t w, RelaxJob& R, uint64_t& m)
{
    const JobDev* J = L.jobs + j;
    const uint32_t act = flood_job_active(J);
    uint32_t* ctl = uni_ptr(J->fctl);
    unsigned long long* flags = load_relax_fields(J, R);
    if (!act) return false;
    const uint32_t done = uni(ctl[1]), s0 = uni(ctl[4]), s1 = uni(ctl[5]), f0 = uni(ctl[2]), f1 = uni(ctl[3]);
    const bool in = w < R.nw;
    const uint64_t m0 = in ? flags[w] : 0ull, m1 = in ? flags[(size_t)R.nw + w] : 0ull;
    if (done) return false;
    const uint32_t rp = ((i + 1u) & 1u) ? s1 : s0;
    const bool lead = threadIdx.x == 0;
    if ((((rp + 1u) & 1u) ? f1 : f0) == 0u) {
        if (lead) {
            ctl[1] = rp;
            J->stats[kStatFloodDone] = rp;
        }
        return false;
    }
    const uint32_t r = rp + 1u;
    if (lead) {
        ctl[4 + (i & 1u)] = r;
        ctl[0] = r;
        ctl[2 + (rp & 1u)] = 0u;
    }
    set_relax_round(R, flags, r);
    m = (r & 1u) ? m1 : m0;
    return true;
}

// Round step 1: start every running job's next round (or find it drained) and
// turn its flag bitmasks into worklists.
// A task is 64 flag words (4096 bricks, one word per lane) of one job; each
// lane clears its word and appends its flagged bricks as (job << 32 | brick)
// to the worklist of the task's XCD, with one list reservation per wave.
// Tasks are dealt to the XCDs round-robin like workgroups, so the relax step,
// which serves list x from the workgroups of XCD x, finds the halos of
// neighbouring bricks (mostly in the same task) in the same L2.
// prefix: njobs + 1 task prefix over the jobs (0 tasks for non-flood jobs);
// lists: kXcds lists of list_cap entries; counts: kXcds counters, zero on entry;
// stale: the other list set's kXcds counters (the previous round's), cleared here;
// i: the round's index in its chunk.
__global__ __launch_bounds__(64) void k_flood_compact(Launch L, const uint32_t* prefix, uint32_t total_tasks,
                                                      unsigned long long* lists, uint32_t list_cap, uint32_t* counts,
                                                      uint32_t* stale, uint32_t i)
{
    __shared__ uint32_t sp[kFloodMaxJobs + 1];
    const uint32_t lane = threadIdx.x;
    if (blockIdx.x == 0 && lane < kXcds) stale[lane] = 0u;
    for (uint32_t i = lane; i <= L.njobs; i += 64u) sp[i] = prefix[i];
    __syncthreads();
    const uint32_t xcd = blockIdx.x % kXcds, per_xcd = gridDim.x / kXcds, local = blockIdx.x / kXcds;
    if (local >= per_xcd) return;  // gridDim is a multiple of kXcds
    RelaxJob R{};
    uint32_t cur_job = 0xFFFFFFFFu;
    bool running = false;
    for (uint32_t k = local;; k += per_xcd) {
        const uint32_t task = xcd + kXcds * k;
        if (task >= total_tasks) break;
        uint32_t lo = 0, hi = L.njobs;  // job of the task: largest j with sp[j] <= task
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sp[mid] <= task) lo = mid; else hi = mid;
        }
        const uint32_t j = uni(lo);
        const uint32_t w = (task - sp[j]) * 64u + lane;
        uint64_t m = 0ull;
        if (j != cur_job) {
            cur_job = j;
            running = start_round(L, j, i, w, R, m);
        } else if (running) {
            m = w < R.nw ? R.cur[w] : 0ull;
        }
        if (!running) continue;
        const uint32_t n = (uint32_t)__popcll(m);
        uint32_t inc = n;  // wave inclusive scan of the per-lane counts
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += t;
        }
        const uint32_t total = __shfl(inc, 63, 64);
        if (total == 0) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(counts + xcd, total);
        base = __shfl(base, 0, 64) + inc - n;
        if (m) {
            R.cur[w] = 0ull;  // consumed
            uint64_t bits = m;
            unsigned long long* dst = lists + (size_t)xcd * list_cap;
            while (bits) {
                const uint32_t bit = (uint32_t)__ffsll((long long)bits) - 1u;
                bits &= bits - 1ull;
                if (base < list_cap) dst[base] = ((unsigned long long)j << 32) | (w * 64u + bit);
                ++base;
            }
        }
    }
}

// Round step 2: relax the listed bricks; the workgroups of XCD x serve list x.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_flood_relax(Launch L, const unsigned long long* lists, uint32_t list_cap,
                                                    const uint32_t* counts)
{
    __shared__ uint32_t tile[kRelaxLdsWords];  // tile (yt reuses it), then zt
    const uint32_t lane = threadIdx.x;
    const uint32_t xcd = blockIdx.x % kXcds, per_xcd = gridDim.x / kXcds, local = blockIdx.x / kXcds;
    if (local >= per_xcd) return;
    const uint32_t n = min(uni(counts[xcd]), list_cap);
    const unsigned long long* list = lists + (size_t)xcd * list_cap;
    RelaxJob R{};
    uint32_t cur_job = 0xFFFFFFFFu;
    bool activated = false, error = false;
    unsigned long long work = 0;  // (sweeps << 32) + visits for cur_job
    auto flush = [&]() {
        if (cur_job == 0xFFFFFFFFu) return;
        if (lane == 0) {
            if (work) atomicAdd(R.J->stats + kStatFloodWork, work);
            if (activated) R.J->fctl[2 + ((R.r + 1u) & 1u)] = 1u;
            if (error) atomicOr(reinterpret_cast<unsigned int*>(R.J->stats + kStatError), (unsigned int)kDevErrRelax);
        }
        work = 0;
        activated = error = false;
    };
    // the list is cut into runs of kRelaxRun consecutive entries, dealt round-robin to the workgroups: a
    // run is mostly one job's (its state is loaded once per run) and neighbouring bricks of one flag word,
    // so a later brick of a run reads the cells an earlier one just dropped, while the bricks relaxed at
    // one time stay close together in the list and share their halos in L2 (longer runs, up to one run
    // per workgroup, spread them over the whole list)
    for (uint32_t e = local * kRelaxRun; e < n;
         e = (e % kRelaxRun == kRelaxRun - 1u) ? e + 1u + (per_xcd - 1u) * kRelaxRun : e + 1u) {
        const unsigned long long entry = list[e];
        const uint32_t j = uni((uint32_t)(entry >> 32)), brick = uni((uint32_t)entry);
        if (j != cur_job) {
            flush();
            cur_job = j;
            load_relax_job(L, j, R);  // listed jobs are running
        }
        const uint32_t sw = relax_brick(R, brick, tile, activated, error);
        work += ((unsigned long long)sw << 32) | 1ull;
    }
    flush();
}

// End of a chunk (single block): mark drained every running flood job whose
// last round flagged nothing, so the host sees it after the chunk (the next
// chunk's first compaction would find the same).  njobs <= 1024.
__global__ __launch_bounds__(1024) void k_flood_plan(Launch L)
{
    const uint32_t j = threadIdx.x;
    if (j >= L.njobs) return;
    const JobDev* J = L.jobs + j;
    if (!J->active || J->algo != VF_FLOOD || J->fctl[1] != 0u) return;
    const uint32_t r = J->fctl[0];
    if (J->fctl[2 + ((r + 1u) & 1u)] == 0u) {
        J->fctl[1] = r;
        J->stats[kStatFloodDone] = r;
    }
}

// Key init.  A brick with an occupied cell is live: its bit in the job's liveness bitmap is set and
// its keys are written (EMPTY, or unreached with the cell's weight).  The keys of the other bricks
// (most of a vessel's box) are never written and never read: the relaxation takes the cells of
// bricks whose bit is clear as EMPTY, and the other kernels read a key only where the cell's
// occupancy byte is set.
//
// One wave per bitmap word, i.e. per column chunk (bx, by, 64 bricks along z).  First lane l reads
// brick bz = 64 * chunk + l over the column's 64 rows (x, y), each load of the wave 512 contiguous
// bytes of one row, and packs its 512 occupancy bits into 16 words in LDS.  Then the live bricks are
// written one after another, lane x * 8 + y writing row (x, y) of the brick from the packed bits.
// Every word of the bitmap is written whole.  A job's words are dealt over the 64 slots of each of
// its compaction tasks (a task is 64 flag words).
constexpr uint32_t kInitBitStride = 65;  // LDS words between a wave's packed-bit planes (odd: no bank conflicts)

// one bit per byte of v: bit k set iff byte k is not zero
__device__ inline uint32_t nonzero_bytes4(uint32_t v)
{
    const uint32_t h = ((((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u) >> 7;  // 0 or 1 per byte
    return (h * 0x01020408u) >> 24;  // byte k lands on bit 24 + k; the other products fall on distinct lower bits
}

__global__ __launch_bounds__(256) void k_flood_init(Launch L, const uint32_t* task_prefix, uint32_t total_tasks)
{
    constexpr uint32_t kRows = 16;  // rows per step of the packing pass: their loads are issued together
    __shared__ uint32_t packed[4 * 16 * kInitBitStride];
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* pw = packed + (threadIdx.x >> 6) * 16u * kInitBitStride;  // this wave's 16 planes
    const uint32_t slots = total_tasks * 64u;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t g = uni(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); g < slots; g += waves) {
        const uint32_t j = uni(find_job(task_prefix, L.njobs, g >> 6));
        const JobDev* J = uni_ptr(L.jobs + j);
        if (uni((uint32_t)(J->active && J->algo == VF_FLOOD)) == 0u) continue;
        const uint32_t t0 = uni(task_prefix[j]), nslots = (uni(task_prefix[j + 1]) - t0) * 64u;
        const uint32_t nx = uni(J->nx), ny = uni(J->ny), nz = uni(J->nz), amp = uni(J->noise_amp);
        const uint32_t nby = bdim(ny), nbz = bdim(nz), wpc = live_words_per_column(nbz);
        const uint32_t nwords = bdim(nx) * nby * wpc;
        const uint64_t noise_seed = ((uint64_t)uni((uint32_t)(J->noise_seed >> 32)) << 32) | uni((uint32_t)J->noise_seed);
        const VF_AS_GLOBAL uint8_t* occ = gcast<const uint8_t>(uni_ptr(J->occ));
        VF_AS_GLOBAL uint32_t* keys = gcast<uint32_t>(uni_ptr(J->keys));
        VF_AS_GLOBAL unsigned long long* live =
            gcast<unsigned long long>(uni_ptr(J->bflags) + 2u * relax_words_per_parity(uni(J->nbricks)));
        const bool quads = (nz & 3u) == 0u;  // rows of full bricks start on 4-cell boundaries
        for (uint32_t w = g - t0 * 64u; w < nwords; w += nslots) {
            const uint32_t chunk = w % wpc, col = w / wpc, by = col % nby, bx = col / nby;
            // packing: lane l holds brick bz = 64 * chunk + l, cells z0 .. z0 + 7 of each row.  With rows of
            // a multiple of 4 cells every load is unconditional (no wait between rows): two 4-cell loads at
            // addresses clamped into the row, masked where they leave the brick column or the grid.
            const uint32_t bz = chunk * 64u + lane, z0 = bz * kRelaxBrick;
            const bool inz = bz < nbz;
            uint32_t bits[16];
            if (quads) {
                const uint32_t zlo = min(z0, nz - 4u), zhi = min(z0 + 4u, nz - 4u);
                const uint32_t mlo = inz ? ~0u : 0u, mhi = inz && z0 + 4u < nz ? ~0u : 0u;
#pragma unroll
                for (uint32_t r0 = 0; r0 < 64u; r0 += kRows) {
                    uint32_t lo[kRows], hi[kRows];
#pragma unroll
                    for (uint32_t r = 0; r < kRows; ++r) {
                        const uint32_t x = bx * kRelaxBrick + ((r0 + r) >> 3), y = by * kRelaxBrick + ((r0 + r) & 7u);
                        const uint64_t base = ((uint64_t)min(x, nx - 1u) * ny + min(y, ny - 1u)) * nz;
                        lo[r] = *(const VF_AS_GLOBAL uint32_t*)(occ + base + zlo);
                        hi[r] = *(const VF_AS_GLOBAL uint32_t*)(occ + base + zhi);
                    }
                    __builtin_amdgcn_sched_barrier(0);  // every load of the step is issued before the first wait
#pragma unroll
                    for (uint32_t r = 0; r < kRows; ++r) {
                        const uint32_t x = bx * kRelaxBrick + ((r0 + r) >> 3), y = by * kRelaxBrick + ((r0 + r) & 7u);
                        const uint32_t rm = x < nx && y < ny ? ~0u : 0u;
                        lo[r] &= mlo & rm;
                        hi[r] &= mhi & rm;
                    }
#pragma unroll
                    for (uint32_t r = 0; r < kRows; r += 4) {
                        uint32_t word = 0;
#pragma unroll
                        for (uint32_t q = 0; q < 4u; ++q)
                            word |= (nonzero_bytes4(lo[r + q]) | (nonzero_bytes4(hi[r + q]) << 4)) << (8u * q);
                        bits[(r0 + r) >> 2] = word;
                    }
                }
            } else {
                const uint32_t zn = inz ? min(kRelaxBrick, nz - z0) : 0u;
#pragma unroll
                for (uint32_t i = 0; i < 16u; ++i) bits[i] = 0u;
                for (uint32_t r = 0; r < 64u; ++r) {
                    const uint32_t x = bx * kRelaxBrick + (r >> 3), y = by * kRelaxBrick + (r & 7u);
                    if (x >= nx || y >= ny) continue;
                    const uint64_t c = ((uint64_t)x * ny + y) * nz + z0;
                    uint32_t b8 = 0;
                    for (uint32_t k = 0; k < zn; ++k) b8 |= (occ[c + k] ? 1u : 0u) << k;
#pragma unroll
                    for (uint32_t i = 0; i < 16u; ++i)
                        if (i == (r >> 2)) bits[i] |= b8 << (8u * (r & 3u));
                }
            }
            uint32_t any = 0;
#pragma unroll
            for (uint32_t i = 0; i < 16u; ++i) {
                pw[i * kInitBitStride + lane] = bits[i];
                any |= bits[i];
            }
            const uint64_t m0 = __ballot(any != 0u);
            if (lane == 0u) live[(size_t)col * wpc + chunk] = m0;
            if (m0 == 0ull) continue;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // keys of the live bricks; lane = row (x, y) of the brick, its 8 bits from the brick's lane
            const uint32_t x = bx * kRelaxBrick + (lane >> 3), y = by * kRelaxBrick + (lane & 7u);
            const bool row = x < nx && y < ny;
            const uint64_t rowc = ((uint64_t)x * ny + y) * nz;
            for (uint64_t m = m0; m; m &= m - 1ull) {
                const uint32_t l = (uint32_t)__builtin_ctzll(m);
                const uint32_t zb = (chunk * 64u + l) * kRelaxBrick, zr = min(kRelaxBrick, nz - zb);
                const uint32_t b8 = (pw[(lane >> 2) * kInitBitStride + l] >> (8u * (lane & 3u))) & 0xFFu;
                if (!row) continue;
                const uint64_t c = rowc + zb;
                uint32_t key[8];
#pragma unroll
                for (uint32_t k = 0; k < 8u; ++k) key[k] = kEmptyKey;
                if (b8) {
#pragma unroll
                    for (uint32_t k = 0; k < 8u; ++k)
                        if ((b8 >> k) & 1u) key[k] = kUnreachedKey | flood_weight(noise_seed, (uint32_t)(c + k), amp);
                }
                if (zr == kRelaxBrick && quads) {
                    *(VF_AS_GLOBAL vf_u32x4*)(keys + c) = vf_u32x4{ key[0], key[1], key[2], key[3] };
                    *(VF_AS_GLOBAL vf_u32x4*)(keys + c + 4) = vf_u32x4{ key[4], key[5], key[6], key[7] };
                } else {
                    for (uint32_t k = 0; k < zr; ++k) keys[c + k] = key[k];
                }
            }
            __builtin_amdgcn_wave_barrier();  // the planes are rewritten for the next word
        }
    }
}

__global__ __launch_bounds__(kCellThreads) void k_flood_finish(Launch L)
{
    constexpr uint32_t kRun = 2;  // tiles per step: all key loads are issued before any label is written
    __shared__ JobDev sJ;
    __shared__ uint16_t lab[kMaxDeviceSeeds];  // seed index -> 8-bit fragment label of the staged job
    uint32_t staged = 0xFFFFFFFFu;
    walk_tile_runs<kRun>(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t j, uint32_t local, uint32_t n) {
        if (!J.active || J.algo != VF_FLOOD) return;
        if (j != staged) {  // block-uniform
            for (uint32_t i = threadIdx.x; i < J.S && i < kMaxDeviceSeeds; i += blockDim.x)
                lab[i] = (uint16_t)(J.seeds[i].w & 0xFFu);  // unmaskRightMost(value, 8)
            __syncthreads();
            staged = j;
        }
        const uint64_t V = J.V;
        const uint32_t mask = (1u << J.key_bits) - 1u;
        // a cell's key is EMPTY exactly where its occupancy byte is 0 (the init writes EMPTY there, relax never
        // lowers an EMPTY key, a seed cell gets its occupancy byte set): keys are read only for the 8-cell
        // segments that hold an occupied cell, a small part of a vessel's box
        uint64_t occ[kRun];
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint64_t c0 = (uint64_t)(local + r) * kTileCells + threadIdx.x * kCellsPerThread;
            occ[r] = r < n && c0 < V ? load8b(J.occ, c0, V) : 0ull;
        }
        uint32_t key[kRun][8];
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint64_t c0 = (uint64_t)(local + r) * kTileCells + threadIdx.x * kCellsPerThread;
            if (occ[r] == 0ull) {
#pragma unroll
                for (int k = 0; k < 8; ++k) key[r][k] = kEmptyKey;
            } else if (c0 + 8 <= V) {
                const uint4 a = gload4(J.keys + c0);
                const uint4 b = gload4(J.keys + c0 + 4);
                key[r][0] = a.x; key[r][1] = a.y; key[r][2] = a.z; key[r][3] = a.w;
                key[r][4] = b.x; key[r][5] = b.y; key[r][6] = b.z; key[r][7] = b.w;
            } else {
                for (int k = 0; k < 8; ++k)
                    key[r][k] = r < n && c0 + k < V ? gcast<const uint32_t>(J.keys)[c0 + k] : kEmptyKey;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint64_t c0 = (uint64_t)(local + r) * kTileCells + threadIdx.x * kCellsPerThread;
            if (r >= n || c0 >= V) break;
            uint16_t out[8];
            for (int k = 0; k < 8; ++k) {
                if (((occ[r] >> (8 * k)) & 0xFFull) == 0ull || key[r][k] == kEmptyKey) out[k] = kEmpty;
                else if (key[r][k] >= kUnreachedKey) out[k] = kFree;
                else out[k] = lab[(key[r][k] >> 2) & mask];
            }
            store8(J.labels, c0, V, out);
        }
    });
}

// ---- sub-seeds -------------------------------------------------------------
// Seeds labelled fragment | id << 8 (Seeder::mergeSeeds, Seeder.cpp:115-152).
// While growing, the reference lowers every cell's id to the smallest id met
// across same-fragment neighbours (floodFracturer-comp.glsl:49-62), so each
// connected component of a fragment's cells ends with the smallest id of its
// seeds; cells whose id is not the fragment's smallest become FREE and the
// grid is flooded again from all remaining cells (disjointSet-comp.glsl,
// disjointSetStack-comp.glsl, FloodFracturer.cpp:134-177).
//
// After the first drain every seed's region is connected (each cell's
// shortest path to its seed stays in the region), so a fragment's components
// are unions of the regions of its seeds that touch.  A union-find over the
// seeds joins touching same-fragment regions; a component is kept iff its
// smallest live id equals the fragment's smallest.  Freed cells become
// unreached, kept cells become sources (distance 1, their owner as the tie
// index), and the rounds restart over the bricks holding freed cells.

// Parents only ever move from a root to a smaller index, so chains strictly
// decrease and both loops end; a stale read at worst costs a failed CAS.
__device__ inline uint32_t uf_find(uint32_t* parent, uint32_t x)
{
    for (;;) {
        const uint32_t p = __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == x) return x;
        x = p;
    }
}

__device__ inline void uf_unite(uint32_t* parent, uint32_t a, uint32_t b)
{
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    while (a != b) {
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(parent + b, b, a);
        if (old == b) return;
        b = uf_find(parent, old);  // b was linked meanwhile: continue from its new root
        a = uf_find(parent, a);
    }
}

__device__ inline bool reached_key(uint32_t k) { return k != kEmptyKey && k < kUnreachedKey; }

// One block per job: every seed its own set.
__global__ __launch_bounds__(256) void k_flood_sub_init(Launch L)
{
    const JobDev* J = L.jobs + blockIdx.x;
    if (!J->active || J->algo != VF_FLOOD || !J->sub) return;
    for (uint32_t s = threadIdx.x; s < J->S; s += blockDim.x) J->sub[s] = s;
}

// Join the sets of seeds whose cells touch through the flood's stencil and
// carry the same fragment id.  The relation is symmetric, so each cell looks
// at the forward half of the stencil (3 of 6, 13 of 26 neighbours).
__global__ __launch_bounds__(kCellThreads) void k_flood_sub_touch(Launch L)
{
    __shared__ JobDev sJ;
    __shared__ uint8_t frag[kMaxDeviceSeeds];
    uint32_t staged = 0xFFFFFFFFu;
    walk_tiles(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t j, uint32_t local, uint32_t) {
        if (!J.active || J.algo != VF_FLOOD || !J.sub) return;
        if (j != staged) {  // block-uniform
            for (uint32_t i = threadIdx.x; i < J.S && i < kMaxDeviceSeeds; i += blockDim.x)
                frag[i] = (uint8_t)(J.seeds[i].w & 0xFFu);
            __syncthreads();
            staged = j;
        }
        const uint64_t c0 = (uint64_t)local * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J.V) return;
        const uint32_t mask = (1u << J.key_bits) - 1u;
        const bool n26 = J.neighbourhood != 6;
        uint32_t x, y, z;
        decompose(c0, J, x, y, z);
        uint32_t last_s = 0xFFFFFFFFu, last_t = 0xFFFFFFFFu;
        for (int k = 0; k < kCellsPerThread && c0 + k < J.V; ++k, advance(x, y, z, J)) {
            if (!J.occ[c0 + k]) continue;  // keys exist only in bricks with an occupied cell
            const uint32_t key = J.keys[c0 + k];
            if (!reached_key(key)) continue;
            const uint32_t s = (key >> 2) & mask;
            for (int i = 0; i < (n26 ? 13 : 3); ++i) {
                // forward offsets o: dx = 1 (o 0..8), then dx = 0, dy = 1 (o 9..11), then (0, 0, 1);
                // the 6-neighbour ones are o = 4, 10, 12
                const int o = n26 ? i : (i == 0 ? 4 : i == 1 ? 10 : 12);
                int dx, dy, dz;
                if (o < 9) { dx = 1; dy = o / 3 - 1; dz = o % 3 - 1; }
                else if (o < 12) { dx = 0; dy = 1; dz = o - 10; }
                else { dx = 0; dy = 0; dz = 1; }
                const int xx = (int)x + dx, yy = (int)y + dy, zz = (int)z + dz;
                if (xx >= (int)J.nx || yy < 0 || yy >= (int)J.ny || zz < 0 || zz >= (int)J.nz) continue;
                const size_t cn = ((size_t)xx * J.ny + (uint32_t)yy) * J.nz + (uint32_t)zz;
                if (!J.occ[cn]) continue;
                const uint32_t kn = J.keys[cn];
                if (!reached_key(kn)) continue;
                const uint32_t t = (kn >> 2) & mask;
                if (t == s || frag[t] != frag[s] || (s == last_s && t == last_t)) continue;
                last_s = s;
                last_t = t;
                uf_unite(J.sub, s, t);
            }
        }
    });
}

// One block per job: component minima of the live ids, the fragment minima,
// the keep flags; when a live seed's component is not kept, restart the
// rounds (both flag parities are clear after a drain).
__global__ __launch_bounds__(1024) void k_flood_sub_keep(Launch L)
{
    __shared__ uint32_t cmin[kMaxDeviceSeeds];
    __shared__ uint32_t fmin[256];
    __shared__ uint32_t freed;
    const JobDev* J = L.jobs + blockIdx.x;
    if (!J->active || J->algo != VF_FLOOD || !J->sub) return;
    const uint32_t S = J->S, mask = (1u << J->key_bits) - 1u;
    uint32_t* parent = J->sub;
    uint32_t* keep = J->sub + S;
    for (uint32_t i = threadIdx.x; i < kMaxDeviceSeeds; i += blockDim.x) cmin[i] = 0xFFFFFFFFu;
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) fmin[i] = 0xFFFFFFFFu;
    if (threadIdx.x == 0) freed = 0;
    __syncthreads();
    // a seed is live when it owns its own cell (a later seed on the cell wins)
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        const uint4 q = J->seeds[s];
        const uint32_t k = J->keys[((size_t)q.x * J->ny + q.y) * J->nz + q.z];
        const bool live = reached_key(k) && ((k >> 2) & mask) == s;
        keep[s] = live ? 1u : 0u;
        if (live) atomicMin(&cmin[uf_find(parent, s)], q.w >> 8);
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x)
        if (keep[s]) atomicMin(&fmin[J->seeds[s].w & 0xFFu], cmin[uf_find(parent, s)]);
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        if (!keep[s]) continue;
        if (cmin[uf_find(parent, s)] != fmin[J->seeds[s].w & 0xFFu]) {
            keep[s] = 0u;
            atomicAdd(&freed, 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && freed) {
        J->fctl[0] = 1;
        J->fctl[1] = 0;
        J->fctl[2] = 0;
        J->fctl[3] = 1;
        J->fctl[4] = J->fctl[5] = 0;  // the next compaction starts round 1
        J->stats[kStatFloodDone] = 0;  // running again (multi-job drivers watch this word)
    }
}

// Restarted jobs: kept cells become sources at distance 1, freed cells
// unreached, and the bricks holding freed cells are flagged for round 1.
__global__ __launch_bounds__(kCellThreads) void k_flood_sub_reseed(Launch L)
{
    __shared__ JobDev sJ;
    __shared__ uint8_t keep[kMaxDeviceSeeds];
    uint32_t staged = 0xFFFFFFFFu;
    walk_tiles(L, L.tile_prefix, L.total_tiles, &sJ, [&](const JobDev& J, uint32_t j, uint32_t local, uint32_t) {
        if (!J.active || J.algo != VF_FLOOD || !J.sub || J.fctl[1] != 0u) return;
        if (j != staged) {  // block-uniform
            for (uint32_t i = threadIdx.x; i < J.S && i < kMaxDeviceSeeds; i += blockDim.x)
                keep[i] = (uint8_t)J.sub[J.S + i];
            __syncthreads();
            staged = j;
        }
        const uint64_t c0 = (uint64_t)local * kTileCells + threadIdx.x * kCellsPerThread;
        if (c0 >= J.V) return;
        const uint32_t b = J.key_bits, mask = (1u << b) - 1u;
        const uint32_t nby = bdim(J.ny), nbz = bdim(J.nz);
        unsigned long long* flags1 = J.bflags + relax_words_per_parity(J.nbricks);  // round 1 consumes parity 1
        uint32_t x, y, z, last = 0xFFFFFFFFu;
        decompose(c0, J, x, y, z);
        for (int k = 0; k < kCellsPerThread && c0 + k < J.V; ++k, advance(x, y, z, J)) {
            if (!J.occ[c0 + k]) continue;  // keys exist only in bricks with an occupied cell
            const uint32_t key = J.keys[c0 + k];
            if (!reached_key(key)) continue;
            const uint32_t s = (key >> 2) & mask;
            if (keep[s]) {
                J.keys[c0 + k] = (((1u << b) | s) << 2) | (key & 3u);
            } else {
                J.keys[c0 + k] = kUnreachedKey | (key & 3u);
                const uint32_t id = ((x / kRelaxBrick) * nby + y / kRelaxBrick) * nbz + z / kRelaxBrick;
                if (id != last) {
                    atomicOr(flags1 + (id >> 6), 1ull << (id & 63u));
                    last = id;
                }
            }
        }
    });
}

}  // namespace

void launch_flood_init(const Launch& L, const FloodRoundScratch& W)
{
    TimedLaunch tl("flood_init", L.s);
    if (W.total_tasks == 0) return;
    const uint32_t blocks = std::max(1u, std::min(W.total_tasks * 16u, active_shape().cell_blocks));  // 4 waves each
    k_flood_init<<<blocks, 256, 0, L.s>>>(L, W.task_prefix, W.total_tasks);
    VF_CHECK_LAUNCH();
}

void launch_flood_seed(const Launch& L, const FloodRoundScratch& W)
{
    TimedLaunch tl("flood_seed", L.s);
    k_flood_seed<<<L.njobs, 256, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
    if (W.total_tasks == 0) return;
    const uint32_t blocks = std::max(1u, std::min(W.total_tasks * 16u, active_shape().cell_blocks));  // 4096 bricks a task
    k_flood_nbr<<<blocks, 256, 0, L.s>>>(L, W.task_prefix, W.total_tasks);
    VF_CHECK_LAUNCH();
}

uint32_t flood_list_cap(uint32_t total_tasks) { return (total_tasks + kXcds - 1) / kXcds * 4096u; }

void launch_flood_rounds(const Launch& L, const FloodRoundScratch& W, uint32_t rounds, uint32_t blocks)
{
    if (L.njobs > kFloodMaxJobs) fail(VF_ERR_INVALID_ARG, "flood launch with %u jobs (max %u)", L.njobs, kFloodMaxJobs);
    if (W.list_cap < flood_list_cap(W.total_tasks)) fail(VF_ERR_INVALID_ARG, "flood worklists too small");
    if (rounds == 0) return;
    rounds += rounds & 1u;  // even chunks: a job's round slots (fctl[4..5]) follow the round's index in its chunk
    blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    // one compaction wave per task is enough (tasks are dealt round-robin over the XCDs)
    const uint32_t cblocks = std::max(kXcds, std::min(blocks, (W.total_tasks + kXcds - 1) / kXcds * kXcds));
    unsigned long long* lists[2] = { W.lists, W.lists + (size_t)kXcds * W.list_cap };
    uint32_t* counts[2] = { W.counts, W.counts + kXcds };
    for (uint32_t i = 0; i < rounds; ++i) {
        const uint32_t p = i & 1u;  // round i lists into set p; set p ^ 1 (the previous round's) is cleared
        {
            TimedLaunch tl("flood_compact", L.s);
            k_flood_compact<<<cblocks, 64, 0, L.s>>>(L, W.task_prefix, W.total_tasks, lists[p], W.list_cap, counts[p],
                                                     counts[p ^ 1u], i);
            VF_CHECK_LAUNCH();
        }
        TimedLaunch tl("flood_relax", L.s);
        k_flood_relax<<<blocks, 64, 0, L.s>>>(L, lists[p], W.list_cap, counts[p]);
        VF_CHECK_LAUNCH();
    }
    TimedLaunch tl("flood_plan", L.s);
    k_flood_plan<<<1, 1024, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_flood_sub(const Launch& L)
{
    const uint32_t blocks = std::max(1u, std::min(L.total_tiles, active_shape().cell_blocks));
    {
        TimedLaunch tl("flood_sub_init", L.s);
        k_flood_sub_init<<<L.njobs, 256, 0, L.s>>>(L);
        VF_CHECK_LAUNCH();
    }
    {
        TimedLaunch tl("flood_sub_touch", L.s);
        k_flood_sub_touch<<<blocks, kCellThreads, 0, L.s>>>(L);
        VF_CHECK_LAUNCH();
    }
    {
        TimedLaunch tl("flood_sub_keep", L.s);
        k_flood_sub_keep<<<L.njobs, 1024, 0, L.s>>>(L);
        VF_CHECK_LAUNCH();
    }
    TimedLaunch tl("flood_sub_reseed", L.s);
    k_flood_sub_reseed<<<blocks, kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

void launch_flood_finish(const Launch& L)
{
    TimedLaunch tl("flood_finish", L.s);
    const uint32_t blocks = std::max(1u, std::min(L.total_tiles, active_shape().cell_blocks));
    k_flood_finish<<<blocks, kCellThreads, 0, L.s>>>(L);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
