// All-labels marching cubes with an exact weld
// (RegularGrid::toTriangleMesh RegularGrid.cpp:443-486 ->
//  MarchingCubes::triangulateFieldGPU MarchingCubes.cpp:364-433,
//  marchingCubes-comp.glsl:96-156).
//
// The reference runs one full-grid dispatch per label, appends triangles with
// an atomic counter (nondeterministic order) and welds with 30-bit Morton codes.
// Here one pass over the padded lattice serves every label.  For label L the
// field is (cell & 0x7FFF) == L on the grid padded by FREE cells, isolevel 0.5,
// so every vertex is the midpoint of a lattice edge whose two ends disagree
// about L: the weld is exact by edge id (3 * padded index + axis).
// Output is canonical: labels ascending; per label, vertices in edge-id order
// and triangles in (cube, table slot) order, corners emitted (0, 2, 1).
//
// Lattice point p = (px, py, pz) owns its three +axis edges and, when it is a
// valid corner-0, the cube spanning x in [px-1, px], y in [py, py+1],
// z in [pz, pz+1].  Kernels:
//   k_mc_count  per tile: per-label triangle / vertex counts -> (label, tile) matrices
//   scans       label-major exclusive prefix of both matrices (counting sort)
//   k_mc_emit   per tile, per present label: block scan of per-thread counts,
//               write triangle edge ids + flags and vertex edge ids in order
//   k_mc_finish vertex positions (FP32, aabb_min + cell_size * (p - 1)) and
//               triangle indices (binary search of the label's sorted edges)
#include <algorithm>
#include <cstring>
#include <vector>

#include "cellops.hpp"
#include "grid.hpp"
#include "mc_table.hpp"

namespace vf {

namespace {

constexpr uint32_t kMcThreads = 256;
constexpr uint32_t kMcPts = 8;                       // lattice points per thread per sub-chunk
constexpr uint32_t kMcChunk = kMcThreads * kMcPts;   // points per sub-chunk
constexpr uint32_t kMcMaxLabels = 4096;
constexpr uint32_t kMcNoLabel = 0xFFFFu;

__constant__ int c_mc_nb[8][3] = { { 0, 0, 0 }, { 0, 0, 1 }, { -1, 0, 1 }, { -1, 0, 0 },
                                   { 0, 1, 0 }, { 0, 1, 1 }, { -1, 1, 1 }, { -1, 1, 0 } };
__constant__ int c_mc_edge[12][2] = { { 0, 1 }, { 1, 2 }, { 2, 3 }, { 3, 0 }, { 4, 5 }, { 5, 6 },
                                      { 6, 7 }, { 7, 4 }, { 0, 4 }, { 1, 5 }, { 2, 6 }, { 3, 7 } };
__constant__ unsigned long long c_mc_tri[256];

struct McDev {
    const uint16_t* grid;
    uint32_t nx, ny, nz;      // grid dims
    uint32_t Nx, Ny, Nz;      // padded lattice dims
    uint64_t P;               // lattice points
    uint32_t tiles, k;        // tiles, sub-chunks per tile
    uint32_t nl;
    const uint16_t* lindex;   // 65536: label -> dense index or kMcNoLabel
    const uint16_t* labels;   // nl labels ascending
    uint32_t* mtri;           // nl * tiles
    uint32_t* mvert;          // nl * tiles
    uint32_t* tile_total;     // tiles
    const unsigned long long* otri;   // nl * tiles + 1 exclusive prefix
    const unsigned long long* overt;
    unsigned long long* tri_edge;     // 3 per triangle
    uint8_t* tri_flag;
    unsigned long long* vert_edge;
    float mn[3], cs[3];
    float* xyz;
    uint32_t* tri_idx;
    uint32_t nv, nt;
};

__device__ inline uint16_t raw_at(const McDev& M, int px, int py, int pz)
{
    if (px <= 0 || py <= 0 || pz <= 0 || px >= (int)M.Nx - 1 || py >= (int)M.Ny - 1 || pz >= (int)M.Nz - 1) return kFree;
    return M.grid[((size_t)(px - 1) * M.ny + (py - 1)) * M.nz + (pz - 1)];
}

// Everything a lattice point contributes, for all labels.
struct PointInfo {
    bool cube;               // p is a valid corner-0
    uint16_t corner[8];      // masked corner values
    uint16_t c0raw;          // raw value of corner 0 (boundary flag)
    uint16_t own;            // masked value of p
    uint16_t nbr[3];         // masked value at p + e_axis (kMcNoLabel beyond the lattice)
    int px, py, pz;
};

__device__ inline void load_point(const McDev& M, uint64_t p, PointInfo& I)
{
    const uint32_t yz = M.Ny * M.Nz;
    I.px = (int)(p / yz);
    const uint32_t r = (uint32_t)(p - (uint64_t)I.px * yz);
    I.py = (int)(r / M.Nz);
    I.pz = (int)(r - (uint32_t)I.py * M.Nz);
    const uint16_t own_raw = raw_at(M, I.px, I.py, I.pz);
    I.own = own_raw & kMask15;
    I.nbr[0] = I.px + 1 < (int)M.Nx ? (uint16_t)(raw_at(M, I.px + 1, I.py, I.pz) & kMask15) : (uint16_t)kMcNoLabel;
    I.nbr[1] = I.py + 1 < (int)M.Ny ? (uint16_t)(raw_at(M, I.px, I.py + 1, I.pz) & kMask15) : (uint16_t)kMcNoLabel;
    I.nbr[2] = I.pz + 1 < (int)M.Nz ? (uint16_t)(raw_at(M, I.px, I.py, I.pz + 1) & kMask15) : (uint16_t)kMcNoLabel;
    I.cube = I.px >= 1 && I.py <= (int)M.Ny - 2 && I.pz <= (int)M.Nz - 2;
    I.c0raw = own_raw;
    if (I.cube) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            I.corner[i] = raw_at(M, I.px + c_mc_nb[i][0], I.py + c_mc_nb[i][1], I.pz + c_mc_nb[i][2]) & kMask15;
    }
}

__device__ inline uint32_t tri_count_for(const PointInfo& I, uint16_t L)
{
    if (!I.cube) return 0;
    uint32_t conf = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) conf |= (uint32_t)(I.corner[i] != L) << i;
    return (uint32_t)(c_mc_tri[conf] >> 60);
}

__device__ inline uint32_t vert_count_for(const PointInfo& I, uint16_t L)
{
    uint32_t n = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        n += (I.nbr[a] != kMcNoLabel && I.own != I.nbr[a] && (I.own == L || I.nbr[a] == L)) ? 1u : 0u;
    return n;
}

__global__ __launch_bounds__(kMcThreads) void k_mc_count(McDev M)
{
    extern __shared__ uint32_t sh[];
    uint32_t* ct = sh;             // nl
    uint32_t* cv = sh + M.nl;      // nl
    __shared__ uint32_t any;
    for (uint32_t t = blockIdx.x; t < M.tiles; t += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < 2 * M.nl; i += blockDim.x) sh[i] = 0;
        if (threadIdx.x == 0) any = 0;
        __syncthreads();
        for (uint32_t c = 0; c < M.k; ++c) {
            const uint64_t p0 = ((uint64_t)t * M.k + c) * kMcChunk + (uint64_t)threadIdx.x * kMcPts;
            for (uint32_t q = 0; q < kMcPts; ++q) {
                const uint64_t p = p0 + q;
                if (p >= M.P) break;
                PointInfo I;
                load_point(M, p, I);
                // labels touching this point: cube corners and edge ends
                uint16_t cand[11];
                int nc = 0;
                if (I.cube)
                    for (int i = 0; i < 8; ++i) cand[nc++] = I.corner[i];
                cand[nc++] = I.own;
                for (int a = 0; a < 3; ++a)
                    if (I.nbr[a] != kMcNoLabel) cand[nc++] = I.nbr[a];
                for (int i = 0; i < nc; ++i) {
                    const uint16_t L = cand[i];
                    bool seen = false;
                    for (int j = 0; j < i; ++j) seen |= cand[j] == L;
                    if (seen) continue;
                    const uint16_t li = M.lindex[L];
                    if (li == kMcNoLabel) continue;
                    const uint32_t nt = tri_count_for(I, L), nv = vert_count_for(I, L);
                    if (nt) atomicAdd(ct + li, nt);
                    if (nv) atomicAdd(cv + li, nv);
                    if (nt | nv) any = 1;
                }
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < M.nl; i += blockDim.x) {
            M.mtri[(size_t)i * M.tiles + t] = ct[i];
            M.mvert[(size_t)i * M.tiles + t] = cv[i];
        }
        if (threadIdx.x == 0) M.tile_total[t] = any;
        __syncthreads();
    }
}

__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t* wsum, uint32_t* total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
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

__device__ inline uint64_t edge_id(const McDev& M, int px, int py, int pz, int a, int b)
{
    const int ax = px + c_mc_nb[a][0], ay = py + c_mc_nb[a][1], az = pz + c_mc_nb[a][2];
    const int bx = px + c_mc_nb[b][0], by = py + c_mc_nb[b][1], bz = pz + c_mc_nb[b][2];
    const int axis = ax != bx ? 0 : (ay != by ? 1 : 2);
    const uint64_t lo = ((uint64_t)min(ax, bx) * M.Ny + (uint64_t)min(ay, by)) * M.Nz + (uint64_t)min(az, bz);
    return 3ull * lo + (uint64_t)axis;
}

__global__ __launch_bounds__(kMcThreads) void k_mc_emit(McDev M)
{
    __shared__ uint16_t present[256];
    __shared__ uint32_t npresent;
    __shared__ unsigned long long tbase[256], vbase[256];
    __shared__ uint32_t wsum[kMcThreads / 64];
    for (uint32_t t = blockIdx.x; t < M.tiles; t += gridDim.x) {
        if (M.tile_total[t] == 0) continue;
        // labels with output in this tile, in groups of 256
        for (uint32_t g0 = 0; g0 < M.nl; g0 += 256) {
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t n = 0;
                for (uint32_t li = g0; li < min(M.nl, g0 + 256); ++li) {
                    const size_t m = (size_t)li * M.tiles + t;
                    if (M.mtri[m] | M.mvert[m]) {
                        present[n] = (uint16_t)li;
                        tbase[n] = M.otri[m];
                        vbase[n] = M.overt[m];
                        ++n;
                    }
                }
                npresent = n;
            }
            __syncthreads();
            const uint32_t np = npresent;
            if (np == 0) continue;
            for (uint32_t c = 0; c < M.k; ++c) {
                const uint64_t p0 = ((uint64_t)t * M.k + c) * kMcChunk + (uint64_t)threadIdx.x * kMcPts;
                PointInfo I[kMcPts];
                bool valid[kMcPts];
                for (uint32_t q = 0; q < kMcPts; ++q) {
                    valid[q] = p0 + q < M.P;
                    if (valid[q]) load_point(M, p0 + q, I[q]);
                }
                for (uint32_t li_k = 0; li_k < np; ++li_k) {
                    const uint32_t li = present[li_k];
                    const uint16_t L = M.labels[li];
                    uint32_t ntl = 0, nvl = 0;
                    for (uint32_t q = 0; q < kMcPts; ++q)
                        if (valid[q]) {
                            ntl += tri_count_for(I[q], L);
                            nvl += vert_count_for(I[q], L);
                        }
                    uint32_t tot;
                    const uint32_t ex = block_excl_scan_u32((ntl << 16) | nvl, wsum, &tot);
                    unsigned long long to = tbase[li_k] + (ex >> 16), vo = vbase[li_k] + (ex & 0xFFFFu);
                    for (uint32_t q = 0; q < kMcPts; ++q) {
                        if (!valid[q]) continue;
                        const PointInfo& P = I[q];
                        // vertices: +x, +y, +z edges in axis order
                        for (int a = 0; a < 3; ++a) {
                            if (P.nbr[a] != kMcNoLabel && P.own != P.nbr[a] && (P.own == L || P.nbr[a] == L)) {
                                const uint64_t lo = ((uint64_t)P.px * M.Ny + (uint64_t)P.py) * M.Nz + (uint64_t)P.pz;
                                M.vert_edge[vo++] = 3ull * lo + (uint64_t)a;
                            }
                        }
                        if (P.cube) {
                            uint32_t conf = 0;
                            for (int i = 0; i < 8; ++i) conf |= (uint32_t)(P.corner[i] != L) << i;
                            const unsigned long long w = c_mc_tri[conf];
                            const int n = (int)(w >> 60);
                            for (int s = 0; s < n; ++s) {
                                const int e0 = (int)((w >> (12 * s)) & 15), e1 = (int)((w >> (12 * s + 4)) & 15),
                                          e2 = (int)((w >> (12 * s + 8)) & 15);
                                unsigned long long* dst = M.tri_edge + 3 * to;
                                dst[0] = edge_id(M, P.px, P.py, P.pz, c_mc_edge[e0][0], c_mc_edge[e0][1]);
                                dst[1] = edge_id(M, P.px, P.py, P.pz, c_mc_edge[e2][0], c_mc_edge[e2][1]);
                                dst[2] = edge_id(M, P.px, P.py, P.pz, c_mc_edge[e1][0], c_mc_edge[e1][1]);
                                M.tri_flag[to] = (uint8_t)((P.c0raw >> 15) & 1u);
                                ++to;
                            }
                        }
                    }
                    __syncthreads();
                    if (threadIdx.x == 0) {
                        tbase[li_k] += tot >> 16;
                        vbase[li_k] += tot & 0xFFFFu;
                    }
                    __syncthreads();
                }
            }
        }
    }
}

__global__ void k_mc_finish(McDev M, const unsigned long long* voff_l, const unsigned long long* toff_l)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < M.nv; v += stride) {
        const uint64_t e = M.vert_edge[v];
        const int axis = (int)(e % 3);
        const uint64_t p = e / 3;
        const uint32_t yz = M.Ny * M.Nz;
        const uint32_t P[3] = { (uint32_t)(p / yz), (uint32_t)((p % yz) / M.Nz), (uint32_t)(p % M.Nz) };
        for (int q = 0; q < 3; ++q) {
            const float tq = (float)P[q] - 1.0f + (q == axis ? 0.5f : 0.0f);
            M.xyz[3 * v + q] = M.mn[q] + M.cs[q] * tq;
        }
    }
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < M.nt; t += stride) {
        // label of triangle t: largest l with toff_l[l] <= t
        uint32_t lo = 0, hi = M.nl;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (toff_l[mid] <= t) lo = mid; else hi = mid;
        }
        const unsigned long long vb = voff_l[lo], ve = voff_l[lo + 1];
        for (int k = 0; k < 3; ++k) {
            const unsigned long long e = M.tri_edge[3 * t + k];
            unsigned long long a = vb, b = ve;
            while (a < b) {
                const unsigned long long mid = (a + b) >> 1;
                if (M.vert_edge[mid] < e) a = mid + 1; else b = mid;
            }
            M.tri_idx[3 * t + k] = (uint32_t)(a - vb);
        }
    }
}

__global__ void k_mc_offsets(const unsigned long long* o, uint32_t nl, uint32_t tiles, unsigned long long* per_label)
{
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l <= nl; l += gridDim.x * blockDim.x)
        per_label[l] = o[(size_t)l * tiles];
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
        for (int i = 0; i < 3; ++i) {
            const uint32_t g = base + D.tri[3 * (size_t)t + i];
            D.tri_g[3 * (size_t)t + i] = g;
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

__global__ void k_smooth_step(SmoothDev D, const float* src, float* dst, uint32_t target, int check, float weight)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < D.nv; v += gridDim.x * blockDim.x) {
        const float* p = src + 3 * (size_t)v;
        float* o = dst + 3 * (size_t)v;
        float px = p[0], py = p[1], pz = p[2];
        if (D.vflag[v] == target) {
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

void mc_smooth(McScratch& W, hipStream_t s, uint32_t nb_iters, float nb_weight, uint32_t b_iters, float b_weight)
{
    if (!W.valid) fail(VF_ERR_INVALID_ARG, "no marching-cubes mesh to smooth");
    if (!(nb_weight >= 0.0f && nb_weight <= 1.0f && b_weight >= 0.0f && b_weight <= 1.0f))
        fail(VF_ERR_INVALID_ARG, "smoothing weights must lie in [0, 1]");
    if (W.nv == 0 || W.nt == 0 || (nb_iters == 0 && b_iters == 0)) return;
    auto& B = W.buf;
    SmoothDev D{};
    D.nv = W.nv;
    D.nt = W.nt;
    D.nl = W.nl;
    D.tri = B[12].as<uint32_t>();
    D.tri_flag = B[8].as<uint8_t>();
    D.voff = B[11].as<unsigned long long>();
    D.toff = D.voff + (W.nl + 1);
    B[14].ensure((size_t)W.nt * 12 + 64);
    B[15].ensure((size_t)W.nv * 8 + 64);         // vflag, deg
    B[13].ensure(std::max(scan_tmp_bytes(W.nv), (size_t)W.nt * 12) + 64);  // scan scratch, then adjacency
    D.tri_g = B[14].as<uint32_t>();
    D.vflag = B[15].as<uint32_t>();
    D.deg = D.vflag + W.nv;
    DevBuf& offs = W.buf[5];                     // reuse: nv + 1 adjacency offsets
    offs.ensure(((size_t)W.nv + 1) * 8 + 64);
    DevBuf& tmp = W.buf[2];                      // ping-pong positions
    tmp.ensure((size_t)W.nv * 12 + 64);
    VF_HIP(hipMemsetAsync(D.vflag, 0, (size_t)W.nv * 8, s));
    const uint32_t gt = stride_blocks(W.nt, 256, 8192), gv = stride_blocks(W.nv, 256, 8192);
    k_smooth_prep<<<gt, 256, 0, s>>>(D);
    VF_CHECK_LAUNCH();
    scan_exclusive_u32_to_u64(D.deg, offs.as<unsigned long long>(), W.nv, B[13].p, s);
    VF_HIP(hipMemsetAsync(D.deg, 0, (size_t)W.nv * 4, s));
    D.adj_off = offs.as<unsigned long long>();
    D.adj = B[13].as<uint32_t>();                // the scan scratch is free again
    k_smooth_fill<<<gt, 256, 0, s>>>(D);
    VF_CHECK_LAUNCH();
    float* a = B[10].as<float>();
    float* b = tmp.as<float>();
    auto pass = [&](uint32_t iters, float weight, uint32_t target, int check) {
        for (uint32_t i = 0; i < iters; ++i) {
            TimedLaunch tl("mc_smooth", s);
            k_smooth_step<<<gv, 256, 0, s>>>(D, a, b, target, check, weight);
            VF_CHECK_LAUNCH();
            std::swap(a, b);
        }
    };
    pass(nb_iters, nb_weight, 0, 1);
    pass(b_iters, b_weight, 1, 0);
    if (a != B[10].as<float>()) VF_HIP(hipMemcpyAsync(B[10].p, a, (size_t)W.nv * 12, hipMemcpyDeviceToDevice, s));
    VF_HIP(hipStreamSynchronize(s));
}


void mc_run(McScratch& W, hipStream_t s, const uint16_t* grid, const uint32_t dims[3], const float mn[3], const float cs[3],
            const uint32_t* hist_dev, uint32_t want, vf_mesh_info* info)
{
    if (!g_table_loaded) {
        VF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_mc_tri), kMcTriPacked, sizeof(kMcTriPacked)));
        g_table_loaded = true;
    }
    // target labels: countValues (RegularGrid.cpp:601-625) -> labels with a nonzero count
    std::vector<uint32_t> hist(65536);
    VF_HIP(hipMemcpyAsync(hist.data(), hist_dev, 65536 * 4, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    std::vector<uint16_t> labels, lindex(65536, (uint16_t)kMcNoLabel);
    for (uint32_t l = 0; l < 65536; ++l)
        if (hist[l] && (want == VF_ALL_LABELS || want == l)) {
            lindex[l] = (uint16_t)labels.size();
            labels.push_back((uint16_t)l);
        }
    const uint32_t nl = (uint32_t)labels.size();
    if (nl > kMcMaxLabels) fail(VF_ERR_INVALID_ARG, "marching cubes supports at most %u labels (grid has %u)", kMcMaxLabels, nl);
    const uint64_t Nx = dims[0] + 2ull, Ny = dims[1] + 2ull, Nz = dims[2] + 2ull;
    const uint64_t P = Nx * Ny * Nz;
    if (3 * P >= (1ull << 40)) fail(VF_ERR_INVALID_ARG, "grid too large for marching cubes");
    W.nl = nl;
    W.nv = W.nt = 0;
    W.valid = false;
    auto& B = W.buf;
    B[0].ensure(65536 * 2);
    B[1].ensure((size_t)std::max<uint32_t>(nl, 1) * 2);
    VF_HIP(hipMemcpyAsync(B[0].p, lindex.data(), 65536 * 2, hipMemcpyHostToDevice, s));
    if (nl) VF_HIP(hipMemcpyAsync(B[1].p, labels.data(), (size_t)nl * 2, hipMemcpyHostToDevice, s));
    B[11].ensure((size_t)(nl + 1) * 16 + 64);  // per-label offsets (voff | toff), u64
    if (nl == 0) {
        std::vector<unsigned long long> z(2, 0);
        VF_HIP(hipMemcpyAsync(B[11].p, z.data(), 16, hipMemcpyHostToDevice, s));
        VF_HIP(hipStreamSynchronize(s));
        info->num_vertices = info->num_triangles = info->num_labels = 0;
        W.valid = true;
        return;
    }
    // tiles: keep the (label, tile) matrices near 2M entries
    const uint64_t tiles0 = (P + kMcChunk - 1) / kMcChunk;
    uint32_t k = (uint32_t)std::max<uint64_t>(1, (nl * tiles0 + (2u << 20) - 1) / (2u << 20));
    const uint32_t tiles = (uint32_t)((tiles0 + k - 1) / k);
    const size_t mat = (size_t)nl * tiles;
    B[2].ensure(mat * 4 + 64);
    B[3].ensure(mat * 4 + 64);
    B[4].ensure((size_t)tiles * 4 + 64);
    B[5].ensure((mat + 1) * 8 + 64);
    B[6].ensure((mat + 1) * 8 + 64);
    B[13].ensure(scan_tmp_bytes(mat) + 64);

    McDev M{};
    M.grid = grid;
    M.nx = dims[0]; M.ny = dims[1]; M.nz = dims[2];
    M.Nx = (uint32_t)Nx; M.Ny = (uint32_t)Ny; M.Nz = (uint32_t)Nz;
    M.P = P;
    M.tiles = tiles;
    M.k = k;
    M.nl = nl;
    M.lindex = B[0].as<uint16_t>();
    M.labels = B[1].as<uint16_t>();
    M.mtri = B[2].as<uint32_t>();
    M.mvert = B[3].as<uint32_t>();
    M.tile_total = B[4].as<uint32_t>();
    M.otri = B[5].as<unsigned long long>();
    M.overt = B[6].as<unsigned long long>();
    for (int q = 0; q < 3; ++q) {
        M.mn[q] = mn[q];
        M.cs[q] = cs[q];
    }
    {
        TimedLaunch tl("mc_count", s);
        const uint32_t blocks = std::min<uint32_t>(tiles, 4096);
        k_mc_count<<<blocks, kMcThreads, 2 * nl * sizeof(uint32_t), s>>>(M);
        VF_CHECK_LAUNCH();
    }
    scan_exclusive_u32_to_u64(M.mtri, B[5].as<unsigned long long>(), mat, B[13].p, s);
    scan_exclusive_u32_to_u64(M.mvert, B[6].as<unsigned long long>(), mat, B[13].p, s);
    unsigned long long tot[2];
    VF_HIP(hipMemcpyAsync(&tot[0], B[5].as<unsigned long long>() + mat, 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipMemcpyAsync(&tot[1], B[6].as<unsigned long long>() + mat, 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    if (tot[0] > 0xFFFFFFFFull || tot[1] > 0xFFFFFFFFull) fail(VF_ERR_OVERFLOW, "marching-cubes output exceeds 2^32 elements");
    M.nt = (uint32_t)tot[0];
    M.nv = (uint32_t)tot[1];
    B[7].ensure((size_t)M.nt * 24 + 64);
    B[8].ensure((size_t)M.nt + 64);
    B[9].ensure((size_t)M.nv * 8 + 64);
    B[10].ensure((size_t)M.nv * 12 + 64);
    B[12].ensure((size_t)M.nt * 12 + 64);
    M.tri_edge = B[7].as<unsigned long long>();
    M.tri_flag = B[8].as<uint8_t>();
    M.vert_edge = B[9].as<unsigned long long>();
    M.xyz = B[10].as<float>();
    M.tri_idx = B[12].as<uint32_t>();
    {
        TimedLaunch tl("mc_emit", s);
        const uint32_t blocks = std::min<uint32_t>(tiles, 4096);
        k_mc_emit<<<blocks, kMcThreads, 0, s>>>(M);
        VF_CHECK_LAUNCH();
    }
    auto* voff_l = B[11].as<unsigned long long>();
    auto* toff_l = voff_l + (nl + 1);
    k_mc_offsets<<<1, 256, 0, s>>>(M.overt, nl, tiles, voff_l);
    k_mc_offsets<<<1, 256, 0, s>>>(M.otri, nl, tiles, toff_l);
    VF_CHECK_LAUNCH();
    {
        TimedLaunch tl("mc_finish", s);
        const uint64_t n = std::max<uint64_t>(M.nv, M.nt);
        k_mc_finish<<<stride_blocks(n, 256, 8192), 256, 0, s>>>(M, voff_l, toff_l);
        VF_CHECK_LAUNCH();
    }
    VF_HIP(hipStreamSynchronize(s));
    W.nv = M.nv;
    W.nt = M.nt;
    info->num_vertices = M.nv;
    info->num_triangles = M.nt;
    info->num_labels = nl;
    W.valid = true;
}

void mc_read(McScratch& W, hipStream_t s, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* voff,
             uint32_t* toff)
{
    auto& B = W.buf;
    const uint32_t nl = W.nl;
    if (xyz && W.nv) VF_HIP(hipMemcpyAsync(xyz, B[10].p, (size_t)W.nv * 12, hipMemcpyDeviceToHost, s));
    if (tri && W.nt) VF_HIP(hipMemcpyAsync(tri, B[12].p, (size_t)W.nt * 12, hipMemcpyDeviceToHost, s));
    if (tri_flag && W.nt) VF_HIP(hipMemcpyAsync(tri_flag, B[8].p, (size_t)W.nt, hipMemcpyDeviceToHost, s));
    if (labels && nl) VF_HIP(hipMemcpyAsync(labels, B[1].p, (size_t)nl * 2, hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> off(2 * (size_t)(nl + 1));
    VF_HIP(hipMemcpyAsync(off.data(), B[11].p, off.size() * 8, hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    for (uint32_t l = 0; l <= nl; ++l) {
        if (voff) voff[l] = (uint32_t)off[l];
        if (toff) toff[l] = (uint32_t)off[nl + 1 + l];
    }
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
    if (nv) VF_HIP(hipMemcpyAsync(xyz, B[10].as<float>() + 3ull * v0, (size_t)nv * 12, hipMemcpyDeviceToHost, g->stream));
    if (nt) VF_HIP(hipMemcpyAsync(tri, B[12].as<uint32_t>() + 3ull * t0, (size_t)nt * 12, hipMemcpyDeviceToHost, g->stream));
    VF_HIP(hipStreamSynchronize(g->stream));
}

}  // namespace vf
