// Conservative voxelization: FP32 triangle/box SAT shell plus a z-parity
// interior fill, replacing the reference's raster Tetravoxelizer
// (RegularGrid::fill, RegularGrid.cpp:173-212; Tetravoxelizer.cpp:264-315).
//
// Pipeline (all jobs of a launch together, no host synchronisation):
//   1. k_vox_setup   per triangle: SAT candidate cell box, snapped fixed-point
//                    vertices, parity column range, pair counts
//   2. scans         exclusive prefix of the pair counts (load balancing)
//   3. k_vox_shell   one thread per (triangle, column, 32-cell z word): SAT
//                    (Intersections3D.h:204-420 arithmetic) per cell with the
//                    z-free terms once per column, one atomicOr per word
//   4. k_vox_parity  one thread per (triangle, column) pair: top-left covered
//                    column centres toggle the first cell centre above the
//                    crossing (exact int64 edge functions), atomicXor
//   5. k_vox_fill    one lane per 32-cell column word: prefix-XOR in the word,
//                    carry across the column by wave ballot, occupancy bytes
//                    = shell | parity, occupied count
// Compiled with -ffp-contract=off: FP32 results equal the CPU oracle's.
#include "cellops.hpp"

namespace vf {

namespace {

constexpr int kSub = 256;
constexpr float kSnapLim = 524288.0f;
constexpr uint32_t kVoxThreads = 256;
constexpr float kRangeMargin = 0.015625f;   // 1/64 cell
constexpr uint32_t kInlineShell = 8;        // SAT candidates a setup thread tests itself
constexpr uint32_t kInlineCols = 8;         // parity columns a setup thread toggles itself
constexpr uint32_t kColBits = 4;            // column index bits in the setup kernel's owner tables
static_assert(kInlineShell <= (1u << kColBits) && kInlineCols <= (1u << kColBits), "owner table encoding");

__device__ inline bool axis_sep(float pa, float pb, float rad)
{
    float mn = pa < pb ? pa : pb;
    float mx = pa < pb ? pb : pa;
    return mn > rad || mx < -rad;
}

__device__ inline void cell_box(int i, float mn, float cs, float* center, float* half)
{
    float lo = mn + (float)i * cs;
    float hi = mn + (float)(i + 1) * cs;
    *center = (hi + lo) / 2.0f;
    *half = hi - *center;
}

// The triangle/box SAT of Intersections3D::intersect(Triangle3D&, AABB&) (Intersections3D.h:204-420:
// nine edge-cross axes, the three box faces, the triangle plane) for the cells (x, y, z0..z1) of one
// grid column, z0..z1 inside one 32-cell word: bit (z & 31) is set when cell (x, y, z) passes.  For
// each cell every FP32 operation is the reference's (same operands, same order, no contraction), so
// a cell passes exactly when the per-cell test passes; the operations that do not depend on z (x and
// y offsets and edge components, the third axis of each edge, the x and y box faces, the normal's z
// component) are evaluated once per column, and a cell's tests are combined without branches (which
// axis separates first does not change the outcome).
__device__ uint32_t sat_column(const float* a, const float* b, const float* c, float cx, float hx, float cy, float hy,
                               int z0, int z1, float mnz, float csz)
{
    const float v0x = a[0] - cx, v1x = b[0] - cx, v2x = c[0] - cx;
    const float v0y = a[1] - cy, v1y = b[1] - cy, v2y = c[1] - cy;
    const float e0x = v1x - v0x, e1x = v2x - v1x, e2x = v0x - v2x;
    const float e0y = v1y - v0y, e1y = v2y - v1y, e2y = v0y - v2y;
    const float f0x = fabsf(e0x), f0y = fabsf(e0y), f1x = fabsf(e1x), f1y = fabsf(e1y), f2x = fabsf(e2x), f2y = fabsf(e2y);
    bool sep = axis_sep(e0y * v1x - e0x * v1y, e0y * v2x - e0x * v2y, f0y * hx + f0x * hy);
    sep |= axis_sep(e1y * v0x - e1x * v0y, e1y * v1x - e1x * v1y, f1y * hx + f1x * hy);
    sep |= axis_sep(e2y * v1x - e2x * v1y, e2y * v2x - e2x * v2y, f2y * hx + f2x * hy);
    {
        float mn = v0x, mx = v0x;
        if (v1x < mn) mn = v1x;
        if (v1x > mx) mx = v1x;
        if (v2x < mn) mn = v2x;
        if (v2x > mx) mx = v2x;
        sep |= mn > hx || mx < -hx;
        mn = v0y; mx = v0y;
        if (v1y < mn) mn = v1y;
        if (v1y > mx) mx = v1y;
        if (v2y < mn) mn = v2y;
        if (v2y > mx) mx = v2y;
        sep |= mn > hy || mx < -hy;
    }
    if (sep) return 0u;
    const float n2 = e0x * e1y - e1x * e0y;
    const float xlo = -hx - v0x, xhi = hx - v0x, ylo = -hy - v0y, yhi = hy - v0y;
    uint32_t bits = 0;
    for (int z = z0; z <= z1; ++z) {
        float cz, hz;
        cell_box(z, mnz, csz, &cz, &hz);
        const float v0z = a[2] - cz, v1z = b[2] - cz, v2z = c[2] - cz;
        const float e0z = v1z - v0z, e1z = v2z - v1z, e2z = v0z - v2z;
        const float f0z = fabsf(e0z), f1z = fabsf(e1z), f2z = fabsf(e2z);
        bool s = axis_sep(e0z * v0y - e0y * v0z, e0z * v2y - e0y * v2z, f0z * hy + f0y * hz);
        s |= axis_sep(-e0z * v0x + e0x * v0z, -e0z * v2x + e0x * v2z, f0z * hx + f0x * hz);
        s |= axis_sep(e1z * v0y - e1y * v0z, e1z * v2y - e1y * v2z, f1z * hy + f1y * hz);
        s |= axis_sep(-e1z * v0x + e1x * v0z, -e1z * v2x + e1x * v2z, f1z * hx + f1x * hz);
        s |= axis_sep(e2z * v0y - e2y * v0z, e2z * v1y - e2y * v1z, f2z * hy + f2y * hz);
        s |= axis_sep(-e2z * v0x + e2x * v0z, -e2z * v1x + e2x * v1z, f2z * hx + f2x * hz);
        float mn = v0z, mx = v0z;
        if (v1z < mn) mn = v1z;
        if (v1z > mx) mx = v1z;
        if (v2z < mn) mn = v2z;
        if (v2z > mx) mx = v2z;
        s |= mn > hz || mx < -hz;
        const float n0 = e0y * e1z - e1y * e0z;
        const float n1 = e0z * e1x - e1z * e0x;
        const float vmin0 = n0 > 0.0f ? xlo : xhi, vmax0 = n0 > 0.0f ? xhi : xlo;
        const float vmin1 = n1 > 0.0f ? ylo : yhi, vmax1 = n1 > 0.0f ? yhi : ylo;
        const float vmin2 = n2 > 0.0f ? -hz - v0z : hz - v0z, vmax2 = n2 > 0.0f ? hz - v0z : -hz - v0z;
        const float dmin = n0 * vmin0 + n1 * vmin1 + n2 * vmin2;
        const float dmax = n0 * vmax0 + n1 * vmax1 + n2 * vmax2;
        s |= dmin > 0.0f || !(dmax >= 0.0f);
        bits |= s ? 0u : 1u << (z & 31);
    }
    return bits;
}

// (x - mn) / cs in grid cells, the quotient both the candidate range and the snap start from
__device__ inline float grid_coord(float x, float mn, float cs) { return (x - mn) / cs; }

__device__ inline int32_t snap_coord(float q)
{
    q = q * (float)kSub;
    if (!(q >= -kSnapLim)) q = -kSnapLim;
    if (q > kSnapLim) q = kSnapLim;
    return (int32_t)rintf(q);
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// TriSetup records are stored as 8 planes of 16-byte pieces: piece k of record g at plane k, index g.
// A wave's stores of 64 consecutive records are then 8 contiguous 1 KiB runs instead of 16-byte pieces
// 128 bytes apart; readers fetch the pieces they use.
static_assert(sizeof(TriSetup) == 8 * sizeof(uint4), "TriSetup is eight 16-byte pieces");
__device__ inline void store_setup(TriSetup* base, uint32_t total, uint32_t g, const TriSetup& S)
{
    uint4 piece[8];
    __builtin_memcpy(piece, &S, sizeof(TriSetup));
    uint4* dst = reinterpret_cast<uint4*>(base);
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) gstore4(dst + (size_t)k * total + g, piece[k]);
}
__device__ inline TriSetup load_setup(const TriSetup* base, uint32_t total, uint32_t g)
{
    uint4 piece[8];
    const uint4* src = reinterpret_cast<const uint4*>(base);
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) piece[k] = src[(size_t)k * total + g];
    TriSetup S;
    __builtin_memcpy(&S, piece, sizeof(TriSetup));
    return S;
}

__device__ inline bool edge_incl(int64_t e, int64_t dx, int64_t dy)
{
    return e > 0 || (e == 0 && (dy < 0 || (dy == 0 && dx < 0)));
}

// floor(a / b) for b > 0: an FP32 estimate, then exact int64 remainder corrections (the estimate is
// within one of the quotient wherever |a / b| < 2^20; the loops make it exact everywhere)
__device__ inline int64_t floordiv64(int64_t a, int64_t b)
{
    int64_t q = (int64_t)floorf((float)a * __builtin_amdgcn_rcpf((float)b));
    int64_t r = a - q * b;
    while (r < 0) {
        --q;
        r += b;
    }
    while (r >= b) {
        ++q;
        r -= b;
    }
    return q;
}

// Column (i, j): if its centre is covered by the triangle's xy projection
// (top-left rule on exact int64 edge functions), toggle the parity bit of the
// first cell whose centre lies above the crossing.
__device__ inline void parity_column(const JobDev* J, const TriSetup& S, int i, int jj)
{
    const int64_t X0 = S.X[0], X1 = S.X[1], X2 = S.X[2];
    const int64_t Y0 = S.Y[0], Y1 = S.Y[1], Y2 = S.Y[2];
    const int64_t px = (int64_t)i * kSub + kSub / 2, py = (int64_t)jj * kSub + kSub / 2;
    const int64_t d0x = X2 - X1, d0y = Y2 - Y1, d1x = X0 - X2, d1y = Y0 - Y2, d2x = X1 - X0, d2y = Y1 - Y0;
    const int64_t E0 = d0x * (py - Y1) - d0y * (px - X1);
    const int64_t E1 = d1x * (py - Y2) - d1y * (px - X2);
    const int64_t E2 = d2x * (py - Y0) - d2y * (px - X0);
    if (!edge_incl(E0, d0x, d0y) || !edge_incl(E1, d1x, d1y) || !edge_incl(E2, d2x, d2y)) return;
    const int64_t num = E0 * (int64_t)S.Z[0] + E1 * (int64_t)S.Z[1] + E2 * (int64_t)S.Z[2];
    int64_t k = floordiv64(num - (int64_t)(kSub / 2) * S.area, (int64_t)kSub * S.area) + 1;
    if (k < 0) k = 0;
    if (k >= (int64_t)J->nz) return;
    const size_t word = ((size_t)i * J->ny + (size_t)jj) * J->wp + (size_t)(k >> 5);
    atomicXor(J->toggle_bits + word, 1u << (k & 31));
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_setup(const JobDev* jobs, uint32_t njobs, const uint32_t* tri_prefix,
                                                           uint32_t total, TriSetup* setup, uint32_t* cnt_shell, uint32_t* cnt_col)
{
    // Small triangles' candidate columns (shell) and covered columns (parity) are tested in this kernel,
    // shared by the lanes of a wave: each lane stashes its triangle in the wave's slots and lists its
    // columns as (lane << kColBits | column) in an owner table; the wave then takes the listed columns 64 at a
    // time, instead of every lane looping over its own (as many steps as its largest triangle).
    constexpr uint32_t kWaves = kVoxThreads / 64;
    // 16-bit pairs hold cell indices (< 2048) and extents: lo0 | lo1 << 16, lo2 | hi2 << 16, ey | nj << 16, i0 | j0 << 16
    enum { kLo01, kLo2Hi2, kEyNj, kI0J0, kJob, kX, kY = kX + 3, kZ = kY + 3, kAreaLo = kZ + 3, kAreaHi, kFields };
    __shared__ uint32_t sp[kWalkMaxJobs + 1];
    __shared__ float sP[kWaves][9][64];
    __shared__ int32_t sF[kWaves][kFields][64];
    __shared__ uint16_t sOwnS[kWaves][64 * kInlineShell], sOwnP[kWaves][64 * kInlineCols];
    for (uint32_t i = threadIdx.x; i <= njobs; i += blockDim.x) sp[i] = tri_prefix[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t g0 = blockIdx.x * blockDim.x; g0 < total; g0 += gridDim.x * blockDim.x) {  // block-uniform
        const uint32_t g = g0 + threadIdx.x;
        const bool valid = g < total;
        const uint32_t j = valid ? find_job(sp, njobs, g) : 0u;
        const JobDev* J = jobs + j;
        TriSetup S;
        S.job = j;
        S.area = 0;
        S.i0 = S.j0 = 0;
        S.i1 = S.j1 = -1;
        for (int q = 0; q < 3; ++q) S.lo[q] = S.hi[q] = S.X[q] = S.Y[q] = S.Z[q] = 0;
        uint32_t nshell = 0, ncol = 0;
        const uint32_t t = g - sp[j];
        uint32_t vi[3] = { 0, 0, 0 };
        bool in_range = true;
        if (valid && J->active) {
            for (int k = 0; k < 3; ++k) {
                vi[k] = J->tri[3 * (size_t)t + k];
                in_range &= vi[k] < J->nv;
            }
            if (!in_range)  // a device mesh with a bad index: flag the job and skip the triangle
                atomicOr(reinterpret_cast<unsigned int*>(J->stats + kStatError), (unsigned int)kDevErrMesh);
        }
        if (valid && J->active && in_range) {
            const float* p[3] = { J->xyz + 3 * (size_t)vi[0], J->xyz + 3 * (size_t)vi[1], J->xyz + 3 * (size_t)vi[2] };
            float P[3][3];
            bool finite = true;
            for (int k = 0; k < 3; ++k)
                for (int q = 0; q < 3; ++q) {
                    P[k][q] = p[k][q];
                    S.P[k][q] = P[k][q];
                    finite &= __builtin_isfinite(P[k][q]);
                }
            const uint32_t dims[3] = { J->nx, J->ny, J->nz };
            if (finite) {
                // vertex coordinates in cells; (v - mn) / cs is monotone in v, so the extremes
                // of these are the quotients of the extreme coordinates
                float G[3][3];
                for (int k = 0; k < 3; ++k)
                    for (int q = 0; q < 3; ++q) G[k][q] = grid_coord(P[k][q], J->mn[q], J->cs[q]);
                bool empty = false;
                for (int q = 0; q < 3; ++q) {
                    float mn = G[0][q], mx = G[0][q];
                    if (G[1][q] < mn) mn = G[1][q];
                    if (G[1][q] > mx) mx = G[1][q];
                    if (G[2][q] < mn) mn = G[2][q];
                    if (G[2][q] > mx) mx = G[2][q];
                    // candidate cells: those within 1/64 cell of the triangle's extent (the
                    // SAT's FP32 box test can only accept touching cells, and its rounding
                    // is far below the margin)
                    float fl = floorf(mn - kRangeMargin);
                    float fh = floorf(mx + kRangeMargin);
                    const float lim = (float)dims[q] + 1.0f;
                    fl = fl < -2.0f ? -2.0f : (fl > lim ? lim : fl);
                    fh = fh < -2.0f ? -2.0f : (fh > lim ? lim : fh);
                    if ((int)fh < 0 || (int)fl > (int)dims[q] - 1) empty = true;
                    S.lo[q] = clampi((int)fl, 0, (int)dims[q] - 1);
                    S.hi[q] = clampi((int)fh, 0, (int)dims[q] - 1);
                }
                if (!empty)
                    nshell = (uint32_t)(S.hi[0] - S.lo[0] + 1) * (uint32_t)(S.hi[1] - S.lo[1] + 1) *
                             (uint32_t)(S.hi[2] - S.lo[2] + 1);
                int64_t X[3], Y[3], Z[3];
                for (int k = 0; k < 3; ++k) {
                    X[k] = snap_coord(G[k][0]);
                    Y[k] = snap_coord(G[k][1]);
                    Z[k] = snap_coord(G[k][2]);
                }
                int64_t area = (X[1] - X[0]) * (Y[2] - Y[0]) - (Y[1] - Y[0]) * (X[2] - X[0]);
                if (area != 0) {
                    if (area < 0) {
                        int64_t tx = X[1], ty = Y[1], tz = Z[1];
                        X[1] = X[2]; Y[1] = Y[2]; Z[1] = Z[2];
                        X[2] = tx; Y[2] = ty; Z[2] = tz;
                        area = -area;
                    }
                    int32_t mnx = (int32_t)X[0], mxx = (int32_t)X[0], mny = (int32_t)Y[0], mxy = (int32_t)Y[0];
                    for (int k = 1; k < 3; ++k) {
                        mnx = min(mnx, (int32_t)X[k]); mxx = max(mxx, (int32_t)X[k]);
                        mny = min(mny, (int32_t)Y[k]); mxy = max(mxy, (int32_t)Y[k]);
                    }
                    int i0 = (mnx - kSub / 2 + kSub - 1) >> 8, i1 = (mxx - kSub / 2) >> 8;
                    int j0 = (mny - kSub / 2 + kSub - 1) >> 8, j1 = (mxy - kSub / 2) >> 8;
                    i0 = max(i0, 0); j0 = max(j0, 0);
                    i1 = min(i1, (int)dims[0] - 1); j1 = min(j1, (int)dims[1] - 1);
                    for (int k = 0; k < 3; ++k) {
                        S.X[k] = (int32_t)X[k]; S.Y[k] = (int32_t)Y[k]; S.Z[k] = (int32_t)Z[k];
                    }
                    S.area = area;
                    S.i0 = i0; S.i1 = i1; S.j0 = j0; S.j1 = j1;
                    if (i1 >= i0 && j1 >= j0) ncol = (uint32_t)(i1 - i0 + 1) * (uint32_t)(j1 - j0 + 1);
                }
            }
        }
        // small triangles are tested here; the others go to the pair kernels, the shell as segments of
        // a column (x, y) that lie in one 32-cell word of z
        uint32_t nis = 0, nip = 0;  // this lane's inline shell columns and parity columns
        if (nshell && nshell <= kInlineShell) {
            nis = (uint32_t)(S.hi[0] - S.lo[0] + 1) * (uint32_t)(S.hi[1] - S.lo[1] + 1);
            nshell = 0;
        } else if (nshell) {
            nshell = (uint32_t)(S.hi[0] - S.lo[0] + 1) * (uint32_t)(S.hi[1] - S.lo[1] + 1) *
                     (uint32_t)((S.hi[2] >> 5) - (S.lo[2] >> 5) + 1);
        }
        if (ncol && ncol <= kInlineCols) {
            nip = ncol;
            ncol = 0;
        }
        if (valid) {
            // only the pair kernels read a record, and only a triangle with pairs left
            if (nshell | ncol) store_setup(setup, total, g, S);
            cnt_shell[g] = nshell;
            cnt_col[g] = ncol;
        }
        // stash the triangle and list its columns
        if (nis | nip) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int q = 0; q < 3; ++q) sP[wave][3 * k + q][lane] = S.P[k][q];
            sF[wave][kLo01][lane] = S.lo[0] | S.lo[1] << 16;
            sF[wave][kLo2Hi2][lane] = S.lo[2] | S.hi[2] << 16;
            sF[wave][kEyNj][lane] = (S.hi[1] - S.lo[1] + 1) | (S.j1 - S.j0 + 1) << 16;
            sF[wave][kI0J0][lane] = S.i0 | S.j0 << 16;
            sF[wave][kJob][lane] = (int32_t)j;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sF[wave][kX + k][lane] = S.X[k];
                sF[wave][kY + k][lane] = S.Y[k];
                sF[wave][kZ + k][lane] = S.Z[k];
            }
            sF[wave][kAreaLo][lane] = (int32_t)(uint32_t)S.area;
            sF[wave][kAreaHi][lane] = (int32_t)(uint32_t)((uint64_t)S.area >> 32);
        }
        uint32_t es = nis, ep = nip;  // wave inclusive scans
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t us = __shfl_up(es, o, 64), up = __shfl_up(ep, o, 64);
            if (lane >= (uint32_t)o) {
                es += us;
                ep += up;
            }
        }
        for (uint32_t k = 0; k < nis; ++k) sOwnS[wave][es - nis + k] = (uint16_t)(lane << kColBits | k);
        for (uint32_t k = 0; k < nip; ++k) sOwnP[wave][ep - nip + k] = (uint16_t)(lane << kColBits | k);
        const uint32_t ts = __shfl(es, 63, 64), tp = __shfl(ep, 63, 64);
        __syncthreads();
        for (uint32_t it = lane; it < ts; it += 64) {
            const uint32_t e = sOwnS[wave][it], o = e >> kColBits, c = e & ((1u << kColBits) - 1u);
            const int ey = sF[wave][kEyNj][o] & 0xFFFF, lo01 = sF[wave][kLo01][o], z2 = sF[wave][kLo2Hi2][o];
            const int x = (lo01 & 0xFFFF) + (int)c / ey, y = (lo01 >> 16) + (int)c % ey;
            const int lo2 = z2 & 0xFFFF, hi2 = z2 >> 16;
            const JobDev* Jo = jobs + sF[wave][kJob][o];
            float pa[3], pb[3], pc[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                pa[q] = sP[wave][q][o];
                pb[q] = sP[wave][3 + q][o];
                pc[q] = sP[wave][6 + q][o];
            }
            float cx, hx, cy, hy;
            cell_box(x, Jo->mn[0], Jo->cs[0], &cx, &hx);
            cell_box(y, Jo->mn[1], Jo->cs[1], &cy, &hy);
            for (int w = lo2 >> 5; w <= hi2 >> 5; ++w) {
                const uint32_t bits = sat_column(pa, pb, pc, cx, hx, cy, hy, max(lo2, 32 * w), min(hi2, 32 * w + 31),
                                                 Jo->mn[2], Jo->cs[2]);
                if (bits) atomicOr(Jo->shell_bits + ((size_t)x * Jo->ny + (size_t)y) * Jo->wp + (size_t)w, bits);
            }
        }
        for (uint32_t it = lane; it < tp; it += 64) {
            const uint32_t e = sOwnP[wave][it], o = e >> kColBits, c = e & ((1u << kColBits) - 1u);
            TriSetup Q;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Q.X[k] = sF[wave][kX + k][o];
                Q.Y[k] = sF[wave][kY + k][o];
                Q.Z[k] = sF[wave][kZ + k][o];
            }
            Q.area = (int64_t)(((uint64_t)(uint32_t)sF[wave][kAreaHi][o] << 32) | (uint32_t)sF[wave][kAreaLo][o]);
            const int nj = sF[wave][kEyNj][o] >> 16, ij = sF[wave][kI0J0][o];
            parity_column(jobs + sF[wave][kJob][o], Q, (ij & 0xFFFF) + (int)c / nj, (ij >> 16) + (int)c % nj);
        }
        __syncthreads();  // the slots and owner tables are reused by the next 256 triangles
    }
}

// Largest g in [0, n) with off[g] <= p (off[0] == 0 <= p), searched by the
// whole wave with 64 probes per step: four dependent loads cover 16M entries.
__device__ inline uint32_t wave_search(const unsigned long long* off, uint32_t n, unsigned long long p)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t lo = 0, hi = n;
    while (hi - lo > 64u) {
        const uint32_t step = (hi - lo + 63u) / 64u;
        const uint32_t idx = lo + lane * step;
        const bool ok = idx < hi && off[idx] <= p;
        const uint64_t m = __ballot(ok);
        const uint32_t last = 63u - (uint32_t)__clzll((long long)m);
        lo = lo + last * step;
        hi = min(lo + step, hi);
    }
    const uint32_t idx = lo + lane;
    const bool ok = idx < hi && off[idx] <= p;
    const uint64_t m = __ballot(ok);
    return lo + (63u - (uint32_t)__clzll((long long)m));
}

// Pair kernels: a block takes chunks of kPairChunk consecutive pairs.  Two
// waves locate the chunk's first and last triangle, the block caches their
// offsets in LDS and each thread finds its triangle there.
constexpr uint32_t kPairChunk = 4 * kVoxThreads;
constexpr uint32_t kOffCache = 1024;

template <class F>
__device__ inline void for_each_pair(const unsigned long long* off, uint32_t ntris, F&& f)
{
    __shared__ unsigned long long soff[kOffCache + 1];
    __shared__ uint32_t sg[2];
    const unsigned long long start = off[0], total = off[ntris];  // pairs start..total - 1
    const uint32_t wave = threadIdx.x >> 6;
    for (unsigned long long p0 = start + (unsigned long long)blockIdx.x * kPairChunk; p0 < total;
         p0 += (unsigned long long)gridDim.x * kPairChunk) {
        const unsigned long long p1 = min(p0 + kPairChunk, total);
        if (wave < 2) {
            const uint32_t g = wave_search(off, ntris + 1, wave == 0 ? p0 : p1 - 1);
            if ((threadIdx.x & 63u) == 0) sg[wave] = g;
        }
        __syncthreads();
        const uint32_t g0 = sg[0], n = sg[1] - g0 + 1;  // triangles g0 .. g0 + n - 1 hold the chunk
        const bool cached = n <= kOffCache;
        if (cached)
            for (uint32_t i = threadIdx.x; i <= n; i += blockDim.x) soff[i] = off[g0 + i];
        __syncthreads();
        // every lane of a wave calls f in each step (valid == false past the chunk),
        // so f may combine its result across the wave
        for (unsigned long long base = p0; base < p1; base += blockDim.x) {
            const unsigned long long p = base + threadIdx.x;
            const bool valid = p < p1;
            uint32_t lo = 0, hi = n;  // largest i with off[g0 + i] <= p
            while (valid && hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((cached ? soff[mid] : off[g0 + mid]) <= p) lo = mid; else hi = mid;
            }
            const uint32_t g = g0 + lo;
            f(g, valid ? p - (cached ? soff[lo] : off[g]) : 0ull, valid);
        }
        __syncthreads();
    }
}

// One thread per column segment of a triangle's candidate box: the cells (x, y, z) with z in one
// 32-cell word, tested by sat_column and merged into the shell word with one atomicOr.  Held to 64
// VGPRs (8 waves per SIMD instead of 6, no spills: 1.97 -> 1.94 ms on the mixed batch).
__global__ __launch_bounds__(kVoxThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_vox_shell(const JobDev* jobs, const TriSetup* setup,
                                                           const unsigned long long* off, uint32_t ntris)
{
    for_each_pair(off, ntris, [&](uint32_t g, unsigned long long local, bool valid) {
        if (!valid) return;
        const TriSetup Sg = load_setup(setup, ntris, g);
        const TriSetup* S = &Sg;
        const JobDev* J = jobs + S->job;
        const int lo2 = S->lo[2], hi2 = S->hi[2];
        const uint32_t w0 = (uint32_t)lo2 >> 5, nseg = ((uint32_t)hi2 >> 5) - w0 + 1;
        const uint32_t ey = (uint32_t)(S->hi[1] - S->lo[1] + 1);
        uint32_t col = (uint32_t)local, seg = 0;  // < 2^32: columns x segments of one triangle
        if (nseg > 1) {
            seg = col % nseg;
            col /= nseg;
        }
        const int y = S->lo[1] + (int)(col % ey), x = S->lo[0] + (int)(col / ey);
        const uint32_t w = w0 + seg;
        float cx, hx, cy, hy;
        cell_box(x, J->mn[0], J->cs[0], &cx, &hx);
        cell_box(y, J->mn[1], J->cs[1], &cy, &hy);
        const uint32_t bits = sat_column(S->P[0], S->P[1], S->P[2], cx, hx, cy, hy, max(lo2, (int)(32 * w)),
                                         min(hi2, (int)(32 * w + 31)), J->mn[2], J->cs[2]);
        if (bits) atomicOr(J->shell_bits + ((size_t)x * J->ny + (size_t)y) * J->wp + w, bits);
    });
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_parity(const JobDev* jobs, const TriSetup* setup,
                                                            const unsigned long long* off, uint32_t ntris)
{
    for_each_pair(off, ntris, [&](uint32_t g, unsigned long long local, bool valid) {
        if (!valid) return;
        const TriSetup Sg = load_setup(setup, ntris, g);
        const TriSetup* S = &Sg;
        const JobDev* J = jobs + S->job;
        const uint32_t nj = (uint32_t)(S->j1 - S->j0 + 1);
        parity_column(J, *S, S->i0 + (int)(local / nj), S->j0 + (int)(local % nj));
    });
}

// 4 bits -> 4 bytes of 0/1 (bit i -> byte i)
__device__ inline uint32_t spread4(uint32_t b) { return ((b & 0xFu) * 0x00204081u) & 0x01010101u; }

// tile_cnt (or null): per 2048-cell tile (indexed by L.tile_prefix) occupied-cell counts for jobs that seed
// anywhere, added here from the occupancy bits (zeroed by k_vox_clear); a wave's 64 words are whole
// columns, at most 2048 consecutive cells, so they touch at most two tiles
__global__ __launch_bounds__(kVoxThreads) void k_vox_fill(Launch L, const uint32_t* col_prefix, uint32_t total_groups,
                                                          uint32_t* tile_cnt)
{
    __shared__ JobDev sJ;
    __shared__ unsigned long long part[kVoxThreads / 64];
    // a wave's 64 words are whole columns (wp divides 64): at most 2048 consecutive occupancy bytes, staged
    // here and stored as dwords with consecutive lanes on consecutive dwords (a lane's own 32 bytes would
    // put a wave's 16-byte stores 32 bytes apart)
    __shared__ uint32_t stage[kVoxThreads / 64][512];
    unsigned long long acc = 0;
    constexpr uint32_t kRun = 4;  // column groups per step: all their bit loads are issued first
    walk_tile_runs<kRun>(L, col_prefix, total_groups, &sJ, [&](const JobDev& J, uint32_t jidx, uint32_t first, uint32_t nrun) {
        if (!J.active) return;
        const uint64_t nwords = (uint64_t)J.nx * J.ny * J.wp;
        uint32_t tgs[kRun], shs[kRun];
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint64_t w = (uint64_t)(first + r) * kVoxThreads + threadIdx.x;
            const bool valid = r < nrun && w < nwords;
            tgs[r] = valid ? gcast<const uint32_t>(J.toggle_bits)[w] : 0u;
            shs[r] = valid ? gcast<const uint32_t>(J.shell_bits)[w] : 0u;
        }
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            if (r >= nrun) break;
            const uint64_t w = (uint64_t)(first + r) * kVoxThreads + threadIdx.x;
            const bool valid = w < nwords;
            const uint32_t tg = tgs[r];
            const uint32_t sh = shs[r];
            uint32_t p = tg;
            p ^= p << 1; p ^= p << 2; p ^= p << 4; p ^= p << 8; p ^= p << 16;
            const uint64_t par = __ballot(__popc(tg) & 1);
            const uint32_t lane = threadIdx.x & 63;
            const uint32_t seg = lane & ~(J.wp - 1);
            const uint64_t below = (lane ? (~0ull >> (64 - lane)) : 0ull) & ~(seg ? (~0ull >> (64 - seg)) : 0ull);
            if (__popcll(par & below) & 1) p = ~p;
            uint32_t bits = 0;
            uint32_t nb = 0;
            uint64_t base = 0;
            if (valid) {
                const uint32_t lgwp = (uint32_t)__builtin_ctz(J.wp);  // wp is a power of two
                const uint64_t col = w >> lgwp;
                const uint32_t z0 = (uint32_t)(w & (J.wp - 1u)) * 32u;
                if (z0 < J.nz) {
                    nb = min(32u, J.nz - z0);
                    bits = (sh | p) & (nb == 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u));
                    base = col * J.nz + z0;
                }
            }
            if ((J.nz & 3u) == 0) {
                // the wave's region starts at lane 0's cell (a column start when lane 0's word is valid;
                // when it is not, no lane's is) and spans its valid columns
                const uint32_t lgwp = (uint32_t)__builtin_ctz(J.wp);
                const uint64_t w0 = w - lane;
                const uint64_t col0 = w0 >> lgwp, cols = (uint64_t)J.nx * J.ny;
                const uint32_t ncols = col0 < cols ? (uint32_t)min<uint64_t>(64u >> lgwp, cols - col0) : 0u;
                const uint32_t ndw = ncols * (J.nz / 4);  // dwords of the region, <= 512
                uint32_t* st = stage[threadIdx.x >> 6];
                if (nb) {
                    const uint32_t off = (uint32_t)((w & (64u - 1u)) >> lgwp) * (J.nz / 4) + (uint32_t)(w & (J.wp - 1u)) * 8u;
#pragma unroll
                    for (uint32_t q = 0; q < 8; ++q)
                        if (4 * q < nb) st[off + q] = spread4(bits >> (4 * q));
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                auto* dst = gcast<uint32_t>(J.occ + (col0 * J.nz));
                for (uint32_t i = lane; i < ndw; i += 64) dst[i] = st[i];
                __builtin_amdgcn_wave_barrier();  // the stage is rewritten by the next group
            } else if (nb) {
                uint8_t* o = J.occ + base;
                for (uint32_t q = 0; q < nb; ++q) o[q] = (uint8_t)((bits >> q) & 1u);
            }
            acc += (unsigned long long)__popc(bits);
            if (tile_cnt && J.seed_location == VF_SEED_BOTH) {
                const uint64_t b0 = __builtin_amdgcn_readfirstlane((uint32_t)base) |
                                    ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32);
                const uint64_t ta = b0 >> 11;  // the wave's first tile (lane 0 holds a column start)
                const uint64_t split = (ta + 1) << 11;  // first cell of the next tile
                uint32_t na = 0;
                if (nb) {
                    const uint32_t below = base >= split ? 0u : (uint32_t)min<uint64_t>(split - base, 32);
                    na = __popc(bits & (below == 32 ? 0xFFFFFFFFu : ((1u << below) - 1u)));
                }
                // both counts (<= 32 per lane, <= 2048 per wave) packed in one 32-bit wave sum
                uint32_t both = na | ((uint32_t)__popc(bits) - na) << 16;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) both += __shfl_xor(both, o, 64);
                const uint32_t sa = both & 0xFFFFu, sb = both >> 16;
                if ((threadIdx.x & 63u) == 0) {
                    const uint32_t t0 = L.tile_prefix[jidx];
                    if (sa) atomicAdd(tile_cnt + t0 + (uint32_t)ta, sa);
                    if (sb) atomicAdd(tile_cnt + t0 + (uint32_t)ta + 1u, sb);
                }
            }
        }
    }, [&](uint32_t j) {
        const unsigned long long s = wave_sum_u64(acc);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long tot = 0;
            for (uint32_t w = 0; w < kVoxThreads / 64; ++w) tot += part[w];
            if (tot) atomicAdd(L.jobs[j].stats + kStatOccupied, tot);
        }
        __syncthreads();
        acc = 0;
    });
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_clear(Launch L, const uint32_t* col_prefix, uint32_t total_groups,
                                                           uint32_t* tile_cnt)
{
    if (tile_cnt)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < L.total_tiles; i += gridDim.x * blockDim.x) tile_cnt[i] = 0u;
    __shared__ JobDev sJ;
    walk_tiles(L, col_prefix, total_groups, &sJ, [&](const JobDev& J, uint32_t, uint32_t group, uint32_t) {
        if (!J.active) return;
        const uint64_t nwords = (uint64_t)J.nx * J.ny * J.wp;
        const uint64_t w = (uint64_t)group * kVoxThreads + threadIdx.x;
        if (w < nwords) {
            gcast<uint32_t>(J.toggle_bits)[w] = 0u;
            gcast<uint32_t>(J.shell_bits)[w] = 0u;
        }
        if (w == 0) J.stats[kStatOccupied] = 0;
    });
}

}  // namespace


void launch_clear_bits(const Launch& L, const uint32_t* col_prefix, uint32_t total_groups, uint32_t* tile_cnt)
{
    TimedLaunch tl("vox_clear", L.s);
    const uint32_t blocks = std::max(1u, std::min(total_groups, active_shape().cell_blocks));
    k_vox_clear<<<blocks, kVoxThreads, 0, L.s>>>(L, col_prefix, total_groups, tile_cnt);
    VF_CHECK_LAUNCH();
}

// X.setup / X.counts / X.shell_off / X.scan_tmp sized for X.total_tris (shell_off and the scan for twice that).
void launch_voxelize(const Launch& L, const VoxScratch& X, const uint32_t* col_prefix, uint32_t total_groups,
                     uint32_t* tile_cnt)
{
    if (L.njobs > kWalkMaxJobs) fail(VF_ERR_INVALID_ARG, "voxelize launch with %u jobs (max %u)", L.njobs, kWalkMaxJobs);
    launch_clear_bits(L, col_prefix, total_groups, tile_cnt);
    if (X.total_tris) {
        uint32_t* cnt_shell = X.counts;
        uint32_t* cnt_col = cnt_shell + X.total_tris;
        {
            TimedLaunch tl("vox_setup", L.s);
            k_vox_setup<<<stride_blocks(X.total_tris, kVoxThreads), kVoxThreads, 0, L.s>>>(
                L.jobs, L.njobs, X.tri_prefix, X.total_tris, X.setup, cnt_shell, cnt_col);
            VF_CHECK_LAUNCH();
        }
        // one scan over both count arrays (contiguous): the parity offsets continue from the shell total
        scan_exclusive_u32_to_u64(cnt_shell, X.shell_off, 2ull * X.total_tris, X.scan_tmp, L.s);
        const unsigned long long* col_off = X.shell_off + X.total_tris;
        {
            TimedLaunch tl("vox_shell", L.s);
            k_vox_shell<<<active_shape().pair_blocks, kVoxThreads, 0, L.s>>>(L.jobs, X.setup, X.shell_off, X.total_tris);
            VF_CHECK_LAUNCH();
        }
        {
            TimedLaunch tl("vox_parity", L.s);
            k_vox_parity<<<std::max(64u, active_shape().pair_blocks / 2), kVoxThreads, 0, L.s>>>(L.jobs, X.setup, col_off,
                                                                                                X.total_tris);
            VF_CHECK_LAUNCH();
        }
    }
    TimedLaunch tl("vox_fill", L.s);
    const uint32_t blocks = std::max(1u, std::min(total_groups, active_shape().cell_blocks));
    k_vox_fill<<<blocks, kVoxThreads, 0, L.s>>>(L, col_prefix, total_groups, tile_cnt);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
