// Naive Voronoi fracture (NaiveFracturer::buildCPU, NaiveFracturer.cpp:26-68;
// naiveFracturer-comp.glsl:19-42) and the optional isolated-region removal
// (removeIsolatedRegionsCPU, NaiveFracturer.cpp:111-150).
//
// Every non-EMPTY cell takes the label of argmin over seeds of (distance,
// seed index).  Distances are exact integers: squared Euclidean, L1, L-inf
// (the reference's float distances on integer offsets order identically),
// except Euclidean on grids large enough for float roots to collide, which
// ranks by the float root itself.
// The argmin is a single unsigned min over packed keys (distance << b | i),
// so ties go to the lowest index exactly as the reference's strict '<' does.
// Seeds are staged through LDS in chunks; each tile's occupied cells are
// compacted and a thread applies every seed to its few cells in registers.
#include "cellops.hpp"

namespace vf {

namespace {

constexpr uint32_t kSeedChunk = 1024;

// (d << b) | i.  32-bit keys: d < 2^24 (at most 3 * 2047^2) and b <= 17, so it is one 24-bit
// multiply-add, d * 2^b + i.
template <int DIST, typename K>
__device__ inline K pack_key(uint32_t d, uint32_t i, uint32_t b)
{
    if constexpr (sizeof(K) == 4) return (K)(__umul24(d, 1u << b) + i);
    else return ((K)d << b) | (K)i;
}

// |a.lo - b.lo| + |a.hi - b.hi| + acc over the 16-bit halves (v_sad_u16, one instruction); cell and seed
// coordinates are below 2^11 (grid edges of at most 2048 cells), so the halves never wrap.
__device__ inline uint32_t sad16(uint32_t a, uint32_t b, uint32_t acc) { return __builtin_amdgcn_sad_u16(a, b, acc); }

// Distance key of one cell to one seed; xy and sxy hold x | y << 16.  kDistEuclidF32 (64-bit keys
// only) ranks by the bits of the correctly rounded float root of the squared distance, which is
// below 2^24 and so converts exactly: the reference's float sqrt (NaiveFracturer.cpp:12-15).
template <int DIST, typename K>
__device__ inline K seed_key(uint32_t x, uint32_t y, uint32_t z, uint32_t xy, uint32_t sx, uint32_t sy, uint32_t sz,
                             uint32_t sxy, uint32_t id, uint32_t b)
{
    uint32_t d;
    if (DIST == VF_MANHATTAN) {
        d = sad16(xy, sxy, sad16(z, sz, 0u));
    } else {
        const uint32_t ax = sad16(x, sx, 0u), ay = sad16(y, sy, 0u), az = sad16(z, sz, 0u);
        if (DIST == VF_CHEBYSHEV) d = max(ax, max(ay, az));
        else  // offsets < 2^11: 16-bit operands select the full-rate 24-bit multiply
            d = (ax & 0xFFFFu) * (ax & 0xFFFFu) + (ay & 0xFFFFu) * (ay & 0xFFFFu) + (az & 0xFFFFu) * (az & 0xFFFFu);
        if (DIST == kDistEuclidF32) d = sqrt_rn_bits(d);
    }
    return pack_key<DIST, K>(d, id, b);
}

// argmin over seeds [base, base + n) (staged in LDS) for Q compacted cells of
// a thread: each seed is read once (an LDS broadcast) and applied to all Q
// cells held in registers.
template <int DIST, typename K, int Q>
__device__ inline void naive_cells_q(K best[4], const uint32_t cx[4], const uint32_t cy[4], const uint32_t cz[4],
                                     const uint32_t* sx, const uint32_t* sy, const uint32_t* sz, uint32_t base, uint32_t n,
                                     uint32_t b)
{
    K m[Q];
    uint32_t cxy[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        m[q] = best[q];
        cxy[q] = cx[q] | (cy[q] << 16);
    }
    // four seeds per step, one 16-byte LDS read per coordinate array (the arrays are 16-byte aligned)
    uint32_t i = 0;
    for (; i + 4 <= n; i += 4) {
        const uint4 X4 = *reinterpret_cast<const uint4*>(sx + i), Y4 = *reinterpret_cast<const uint4*>(sy + i),
                    Z4 = *reinterpret_cast<const uint4*>(sz + i);
        const uint32_t xs[4] = { X4.x, X4.y, X4.z, X4.w }, ys[4] = { Y4.x, Y4.y, Y4.z, Y4.w },
                       zs[4] = { Z4.x, Z4.y, Z4.z, Z4.w };
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t xy = xs[k] | (ys[k] << 16), id = base + i + k;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const K key = seed_key<DIST, K>(cx[q], cy[q], cz[q], cxy[q], xs[k], ys[k], zs[k], xy, id, b);
                m[q] = key < m[q] ? key : m[q];
            }
        }
    }
    for (; i < n; ++i) {
        const uint32_t x = sx[i], y = sy[i], z = sz[i], xy = x | (y << 16), id = base + i;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const K key = seed_key<DIST, K>(cx[q], cy[q], cz[q], cxy[q], x, y, z, xy, id, b);
            m[q] = key < m[q] ? key : m[q];
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) best[q] = m[q];
}

// qmax (1..4): cells per thread, the same for the whole wave (lanes past the
// tile's last occupied cell compute on a dummy cell)
template <int DIST, typename K>
__device__ inline void naive_cells(K best[4], const uint32_t cx[4], const uint32_t cy[4], const uint32_t cz[4], uint32_t qmax,
                                   const uint32_t* sx, const uint32_t* sy, const uint32_t* sz, uint32_t base, uint32_t n,
                                   uint32_t b)
{
    if (qmax <= 1) naive_cells_q<DIST, K, 1>(best, cx, cy, cz, sx, sy, sz, base, n, b);
    else if (qmax == 2) naive_cells_q<DIST, K, 2>(best, cx, cy, cz, sx, sy, sz, base, n, b);
    else if (qmax == 3) naive_cells_q<DIST, K, 3>(best, cx, cy, cz, sx, sy, sz, base, n, b);
    else naive_cells_q<DIST, K, 4>(best, cx, cy, cz, sx, sy, sz, base, n, b);
}


// One tile per step: the tile's occupied cells are compacted in LDS, so the
// seed loop runs over occupied cells only (shells occupy a few percent of
// their box); labels return through LDS and leave as one 16-byte store per
// thread.  A tile costs several block barriers and dependent LDS steps, so
// the register budget is set for more waves in flight: 5 per SIMD with
// 32-bit keys (95 VGPRs instead of 127, 1.06 -> 0.94 ms on the mixed batch),
// 4 with 64-bit keys, both without spills.  RUN tiles form a walker run whose
// occupancy loads are issued together (a launch shape the tuner picks).
template <typename K, uint32_t RUN>
__global__ __launch_bounds__(kCellThreads) __attribute__((amdgpu_waves_per_eu(sizeof(K) == 4 ? 5 : 4))) void k_naive(Launch L)
{
    __shared__ __attribute__((aligned(16))) uint32_t sx[kSeedChunk], sy[kSeedChunk], sz[kSeedChunk];
    __shared__ uint16_t sw[kSeedChunk];
    __shared__ uint16_t cells[kTileCells];  // compacted offsets of the occupied cells
    __shared__ uint4 lab4[kTileCells / 8];  // the tile's output labels, 8 per thread
    typedef uint16_t __attribute__((may_alias)) label_alias;  // single labels of lab4
    label_alias* lab = reinterpret_cast<label_alias*>(lab4);
    __shared__ uint32_t wsum[kCellThreads / 64];
    __shared__ JobDev sJ;
    __shared__ uint32_t sp[kWalkMaxJobs + 1];
    constexpr int kWant = sizeof(K) == 8 ? 1 : 0;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t staged = 0xFFFFFFFFu;  // job whose seeds (one chunk) sit in LDS
    auto stage = [&](const JobDev& J, uint32_t base, uint32_t n) {
        __syncthreads();
        for (uint32_t i = tid; i < n; i += blockDim.x) {
            const uint4 q = J.seeds[base + i];
            sx[i] = q.x; sy[i] = q.y; sz[i] = q.z; sw[i] = (uint16_t)q.w;
        }
        __syncthreads();
    };
    constexpr uint32_t kRun = RUN;
    walk_tile_runs_strided<kRun>(L, L.tile_prefix, L.total_tiles, &sJ, sp, [&](const JobDev& J, uint32_t j, uint32_t first,
                                                                              uint32_t nrun) {
        if (!J.active || J.algo != VF_NAIVE || J.naive_key64 != kWant || J.S == 0) return;
        const uint32_t S = J.S;
        const bool one_chunk = S <= kSeedChunk;
        if (one_chunk && staged != j) {  // block-uniform
            stage(J, 0, S);
            staged = j;
        }
        uint64_t occs[kRun];
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint64_t c0 = (uint64_t)(first + r) * kTileCells + tid * kCellsPerThread;
            occs[r] = r < nrun && c0 < J.V ? load8b(J.occ, c0, J.V) : 0ull;
        }
        auto tile = [&](uint32_t local, uint64_t occ) {
        const uint64_t t0 = (uint64_t)local * kTileCells;
        const uint64_t c0 = t0 + tid * kCellsPerThread;
        if (!__syncthreads_or(occ != 0ull)) {  // an EMPTY tile (most of a vessel's box): labels stay EMPTY
            if (c0 < J.V) {
                const uint16_t zero[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
                store8(J.labels, c0, J.V, zero);
            }
            return;
        }
        // block-wide exclusive scan of the occupied counts -> compacted positions
        const uint32_t n = (uint32_t)__popcll(occ & 0x0101010101010101ull);
        uint32_t inc = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += t;
        }
        if (lane == 63u) wsum[wave] = inc;
        lab4[tid] = make_uint4(0u, 0u, 0u, 0u);  // EMPTY unless written below
        __syncthreads();
        uint32_t pos = inc - n, total = 0;
        for (uint32_t w = 0; w < kCellThreads / 64; ++w) {
            if (w < wave) pos += wsum[w];
            total += wsum[w];
        }
        for (int k = 0; k < 8; ++k)
            if ((occ >> (8 * k)) & 0xFFu) cells[pos++] = (uint16_t)(tid * 8 + k);
        __syncthreads();
        if (total) {
            // thread t holds compacted cells t, t + 256, ...: a wave's first lane holds the most, so
            // waves past the last cell skip the seed loop.  Cells are taken four per thread at a time
            // (the second group only in tiles with more than 1024 occupied cells), so the registers
            // hold four cells' coordinates and keys, not eight.
            const uint32_t nq = tid < total ? (total - tid + kCellThreads - 1) / kCellThreads : 0u;
            const uint32_t tid0 = wave * 64u;
            const uint32_t qmax = tid0 < total ? (total - tid0 + kCellThreads - 1) / kCellThreads : 0u;  // wave-uniform
            const uint32_t b = J.key_bits;
            const int dist = J.dist;
            const K mask = ((K)1 << b) - 1;
            const uint32_t groups = total > 4u * kCellThreads ? 2u : 1u;  // block-uniform
            for (uint32_t g = 0; g < groups; ++g) {
                const uint32_t q0 = 4u * g;
                const uint32_t gq = nq > q0 ? min(nq - q0, 4u) : 0u;
                const uint32_t gqmax = qmax > q0 ? min(qmax - q0, 4u) : 0u;  // wave-uniform
                uint32_t cx[4], cy[4], cz[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) {
                    cx[q] = cy[q] = cz[q] = 0;
                    if (q < gq) decompose(t0 + cells[tid + (q0 + q) * kCellThreads], J, cx[q], cy[q], cz[q]);
                }
                K best[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) best[q] = ~(K)0;
                for (uint32_t base = 0; base < S; base += kSeedChunk) {
                    const uint32_t m = min(kSeedChunk, S - base);
                    if (!one_chunk) {
                        stage(J, base, m);
                        staged = 0xFFFFFFFFu;
                    }
                    if (gqmax == 0) continue;  // staging above stays block-wide
                    if (dist == VF_MANHATTAN) naive_cells<VF_MANHATTAN, K>(best, cx, cy, cz, gqmax, sx, sy, sz, base, m, b);
                    else if (dist == VF_CHEBYSHEV) naive_cells<VF_CHEBYSHEV, K>(best, cx, cy, cz, gqmax, sx, sy, sz, base, m, b);
                    else if (dist == VF_EUCLIDEAN) naive_cells<VF_EUCLIDEAN, K>(best, cx, cy, cz, gqmax, sx, sy, sz, base, m, b);
                    else if constexpr (sizeof(K) == 8) naive_cells<kDistEuclidF32, K>(best, cx, cy, cz, gqmax, sx, sy, sz, base, m, b);
                }
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) {
                    if (q >= gq) break;
                    const uint32_t id = (uint32_t)(best[q] & mask);
                    lab[cells[tid + (q0 + q) * kCellThreads]] = one_chunk ? sw[id] : (uint16_t)J.seeds[id].w;
                }
            }
        }
        __syncthreads();
        if (c0 < J.V) {
            const uint4 v = lab4[tid];
            const uint16_t out[8] = { (uint16_t)v.x, (uint16_t)(v.x >> 16), (uint16_t)v.y, (uint16_t)(v.y >> 16),
                                      (uint16_t)v.z, (uint16_t)(v.z >> 16), (uint16_t)v.w, (uint16_t)(v.w >> 16) };
            store8(J.labels, c0, J.V, out);
        }
        __syncthreads();  // lab and cells are reused by the next tile
        };
#pragma unroll 1
        for (uint32_t r = 0; r < nrun; ++r) {
            uint64_t occ = occs[0];  // occs[r] by selects (a dynamic index would put occs in scratch)
#pragma unroll
            for (uint32_t i = 1; i < kRun; ++i) occ = r == i ? occs[i] : occ;
            tile(first + r, occ);
        }
    });
}

// ---------------------------------------------------------------------------
// Isolated-region removal: label-propagation BFS from the seeds through cells
// holding the seed's label.  Each cell can be claimed by one label only (its
// own), so the reached set, and thus the result, is independent of order.
__global__ void k_ri_init(const uint16_t* labels, uint16_t* out, uint32_t* visited, uint64_t V, uint64_t words)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        visited[i] = 0;
}

__global__ void k_ri_seed(uint16_t* out, uint32_t* visited, uint32_t* frontier, uint32_t* fl, uint32_t* counters,
                          uint32_t ny, uint32_t nz, const uint4* seeds, uint32_t S)
{
    // seeds stamped in list order: the last seed on a cell keeps it
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        for (uint32_t s = 0; s < S; ++s) {
            uint32_t c = (seeds[s].x * ny + seeds[s].y) * nz + seeds[s].z;
            out[c] = (uint16_t)seeds[s].w;
            visited[c >> 5] |= 1u << (c & 31);
            frontier[s] = c;
            fl[s] = seeds[s].w;
        }
        counters[0] = S;
        counters[1] = 0;
    }
}

__global__ void k_ri_step(const uint16_t* labels, uint16_t* out, uint32_t* visited, const uint32_t* fa, const uint32_t* la,
                          uint32_t* fb, uint32_t* lb, uint32_t* counters, uint32_t cur, uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint32_t n = counters[cur];
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const uint32_t v = fa[e], w = la[e];
        const uint32_t yz = ny * nz;
        const uint32_t x = v / yz, r = v - x * yz, y = r / nz, z = r - y * nz;
        const int dx[6] = { 1, -1, 0, 0, 0, 0 }, dy[6] = { 0, 0, 1, -1, 0, 0 }, dz[6] = { 0, 0, 0, 0, 1, -1 };
        for (int k = 0; k < 6; ++k) {
            const int xx = (int)x + dx[k], yy = (int)y + dy[k], zz = (int)z + dz[k];
            if (xx < 0 || yy < 0 || zz < 0 || xx >= (int)nx || yy >= (int)ny || zz >= (int)nz) continue;
            const uint32_t u = ((uint32_t)xx * ny + (uint32_t)yy) * nz + (uint32_t)zz;
            if ((uint32_t)labels[u] != w) continue;
            const uint32_t bit = 1u << (u & 31);
            if (visited[u >> 5] & bit) continue;
            const uint32_t old = atomicOr(visited + (u >> 5), bit);
            if (old & bit) continue;
            out[u] = (uint16_t)w;
            const uint32_t slot = atomicAdd(counters + (cur ^ 1), 1u);
            fb[slot] = u;
            lb[slot] = w;
        }
    }
}

__global__ void k_ri_reset(uint32_t* counters, uint32_t idx) { counters[idx] = 0; }

}  // namespace

void launch_naive(const Launch& L, bool key64)
{
    if (L.njobs > kWalkMaxJobs) fail(VF_ERR_INVALID_ARG, "naive launch with %u jobs (max %u)", L.njobs, kWalkMaxJobs);
    const uint32_t blocks = std::max(1u, std::min(L.total_tiles, active_shape().cell_blocks));
    TimedLaunch tl(key64 ? "naive64" : "naive", L.s);
    const uint32_t run = active_shape().naive_run;
    if (key64) {
        if (run == 4) k_naive<unsigned long long, 4><<<blocks, kCellThreads, 0, L.s>>>(L);
        else if (run == 16) k_naive<unsigned long long, 16><<<blocks, kCellThreads, 0, L.s>>>(L);
        else k_naive<unsigned long long, 8><<<blocks, kCellThreads, 0, L.s>>>(L);
    } else {
        if (run == 4) k_naive<uint32_t, 4><<<blocks, kCellThreads, 0, L.s>>>(L);
        else if (run == 16) k_naive<uint32_t, 16><<<blocks, kCellThreads, 0, L.s>>>(L);
        else k_naive<uint32_t, 8><<<blocks, kCellThreads, 0, L.s>>>(L);
    }
    VF_CHECK_LAUNCH();
}

void launch_remove_isolated(hipStream_t s, uint16_t* labels, uint16_t* out, uint32_t* fa, uint32_t* fb, uint32_t* counters,
                            uint32_t* visited, const uint32_t dims[3], const uint4* seeds, uint32_t S, uint32_t* host_counter)
{
    // fa/fb each hold 2*V uint32 (cell, label) halves; host_counter is pinned memory
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2];
    uint32_t* la = fa + V;
    uint32_t* lb = fb + V;
    TimedLaunch tl("remove_isolated", s);
    k_ri_init<<<1024, 256, 0, s>>>(labels, out, visited, V, (V + 31) / 32);
    k_ri_seed<<<1, 64, 0, s>>>(out, visited, fa, la, counters, dims[1], dims[2], seeds, S);
    VF_CHECK_LAUNCH();
    uint32_t cur = 0;
    for (;;) {
        k_ri_step<<<1024, 256, 0, s>>>(labels, out, visited, cur ? fb : fa, cur ? lb : la, cur ? fa : fb, cur ? la : lb,
                                       counters, cur, dims[0], dims[1], dims[2]);
        VF_CHECK_LAUNCH();
        k_ri_reset<<<1, 1, 0, s>>>(counters, cur);
        VF_HIP(hipMemcpyAsync(host_counter, counters + (cur ^ 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        VF_HIP(hipStreamSynchronize(s));
        cur ^= 1;
        if (*host_counter == 0) break;
    }
}

}  // namespace vf
