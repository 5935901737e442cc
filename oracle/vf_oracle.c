/*
 * vf_oracle.c — CPU oracle for the voxelize -> fracture path.
 *
 * TEST INFRASTRUCTURE ONLY (see vf_oracle.h).  Plain C, single threaded except
 * where noted, compiled with -ffp-contract=off so every FP32 expression rounds
 * exactly as written.  Reference file:line citations are relative to the
 * reference repository root (MeshFragments/ is abbreviated MF/).
 */
#include "vf_oracle.h"
#include "vfo_mc_table.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define VFO_EMPTY 0u
#define VFO_FREE 1u
#define VFO_MASK15 0x7FFFu

static inline size_t gidx(size_t x, size_t y, size_t z, const uint32_t* d)
{
    /* RegularGrid::getPositionIndex (MF/Source/DataStructures/RegularGrid.cpp:1031-1034) */
    return (x * d[1] + y) * d[2] + z;
}

/* ========================================================================= */
/* Voxelization                                                              */
/* ========================================================================= */

/*
 * Triangle/box overlap by the separating-axis theorem (Akenine-Moller), the
 * spec named for the voxelizer.  Same tests and the same FP32 arithmetic per
 * test as Intersections3D::intersect(Triangle3D&, AABB&) and its axis helpers
 * (MF/Source/Geometry/3D/Intersections3D.h:204-420): 9 edge-cross-axis tests,
 * 3 box-face tests, then the triangle-plane test.  Rejection uses strict '>'
 * so touching counts as overlap.
 */
static int axis_sep(float pa, float pb, float rad)
{
    float mn = pa < pb ? pa : pb;
    float mx = pa < pb ? pb : pa;
    return (mn > rad || mx < -rad);
}

int vfo_sat_triangle_box(const float a[3], const float b[3], const float c[3],
                         const float center[3], const float half[3])
{
    float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
    for (int q = 0; q < 3; ++q) {
        v0[q] = a[q] - center[q];
        v1[q] = b[q] - center[q];
        v2[q] = c[q] - center[q];
    }
    for (int q = 0; q < 3; ++q) {
        e0[q] = v1[q] - v0[q];
        e1[q] = v2[q] - v1[q];
        e2[q] = v0[q] - v2[q];
    }
    float fx, fy, fz, pa, pb, rad;

    /* edge 0 */
    fx = fabsf(e0[0]); fy = fabsf(e0[1]); fz = fabsf(e0[2]);
    pa = e0[2] * v0[1] - e0[1] * v0[2]; pb = e0[2] * v2[1] - e0[1] * v2[2];
    rad = fz * half[1] + fy * half[2]; if (axis_sep(pa, pb, rad)) return 0;
    pa = -e0[2] * v0[0] + e0[0] * v0[2]; pb = -e0[2] * v2[0] + e0[0] * v2[2];
    rad = fz * half[0] + fx * half[2]; if (axis_sep(pa, pb, rad)) return 0;
    pa = e0[1] * v1[0] - e0[0] * v1[1]; pb = e0[1] * v2[0] - e0[0] * v2[1];
    rad = fy * half[0] + fx * half[1]; if (axis_sep(pa, pb, rad)) return 0;

    /* edge 1 */
    fx = fabsf(e1[0]); fy = fabsf(e1[1]); fz = fabsf(e1[2]);
    pa = e1[2] * v0[1] - e1[1] * v0[2]; pb = e1[2] * v2[1] - e1[1] * v2[2];
    rad = fz * half[1] + fy * half[2]; if (axis_sep(pa, pb, rad)) return 0;
    pa = -e1[2] * v0[0] + e1[0] * v0[2]; pb = -e1[2] * v2[0] + e1[0] * v2[2];
    rad = fz * half[0] + fx * half[2]; if (axis_sep(pa, pb, rad)) return 0;
    pa = e1[1] * v0[0] - e1[0] * v0[1]; pb = e1[1] * v1[0] - e1[0] * v1[1];
    rad = fy * half[0] + fx * half[1]; if (axis_sep(pa, pb, rad)) return 0;

    /* edge 2 */
    fx = fabsf(e2[0]); fy = fabsf(e2[1]); fz = fabsf(e2[2]);
    pa = e2[2] * v0[1] - e2[1] * v0[2]; pb = e2[2] * v1[1] - e2[1] * v1[2];
    rad = fz * half[1] + fy * half[2]; if (axis_sep(pa, pb, rad)) return 0;
    pa = -e2[2] * v0[0] + e2[0] * v0[2]; pb = -e2[2] * v1[0] + e2[0] * v1[2];
    rad = fz * half[0] + fx * half[2]; if (axis_sep(pa, pb, rad)) return 0;
    pa = e2[1] * v1[0] - e2[0] * v1[1]; pb = e2[1] * v2[0] - e2[0] * v2[1];
    rad = fy * half[0] + fx * half[1]; if (axis_sep(pa, pb, rad)) return 0;

    /* box faces: triangle bounds against the box extent */
    for (int q = 0; q < 3; ++q) {
        float mn = v0[q], mx = v0[q];
        if (v1[q] < mn) mn = v1[q];
        if (v1[q] > mx) mx = v1[q];
        if (v2[q] < mn) mn = v2[q];
        if (v2[q] > mx) mx = v2[q];
        if (mn > half[q] || mx < -half[q]) return 0;
    }

    /* triangle plane (planeBoxOverlap, Intersections3D.h:270-294) */
    float n[3];
    n[0] = e0[1] * e1[2] - e1[1] * e0[2];
    n[1] = e0[2] * e1[0] - e1[2] * e0[0];
    n[2] = e0[0] * e1[1] - e1[0] * e0[1];
    float vmin[3], vmax[3];
    for (int q = 0; q < 3; ++q) {
        if (n[q] > 0.0f) { vmin[q] = -half[q] - v0[q]; vmax[q] = half[q] - v0[q]; }
        else             { vmin[q] = half[q] - v0[q];  vmax[q] = -half[q] - v0[q]; }
    }
    float dmin = n[0] * vmin[0] + n[1] * vmin[1] + n[2] * vmin[2];
    if (dmin > 0.0f) return 0;
    float dmax = n[0] * vmax[0] + n[1] * vmax[1] + n[2] * vmax[2];
    return dmax >= 0.0f;
}

/* Cell box of index i along one axis: faces at aabb_min + i*cs, shared exactly
 * by neighbouring cells; centre/extent as AABB::center()/extent()
 * (MF/Source/Geometry/3D/AABB.h:41,51). */
static void cell_box(int i, float mn, float cs, float* center, float* half)
{
    float lo = mn + (float)i * cs;
    float hi = mn + (float)(i + 1) * cs;
    *center = (hi + lo) / 2.0f;
    *half = hi - *center;
}

#define VFO_SUB 256          /* fixed-point sub-cell resolution for the parity scan */
#define VFO_SNAP_LIM 524288.0f

static int32_t snap_coord(float x, float mn, float cs)
{
    float q = (x - mn) / cs;
    q = q * (float)VFO_SUB;
    if (!(q >= -VFO_SNAP_LIM)) q = -VFO_SNAP_LIM;
    if (q > VFO_SNAP_LIM) q = VFO_SNAP_LIM;
    return (int32_t)rintf(q);
}

static int64_t floordiv64(int64_t a, int64_t b) /* b > 0 */
{
    int64_t q = a / b;
    if ((a % b) != 0 && a < 0) q -= 1;
    return q;
}

static inline int32_t floordiv_sub(int32_t a) { return a >> 8; } /* floor(a/256) */

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* top-left tie rule for a CCW edge with direction (dx, dy) */
static inline int edge_incl(int64_t e, int64_t dx, int64_t dy)
{
    return e > 0 || (e == 0 && (dy < 0 || (dy == 0 && dx < 0)));
}

int64_t vfo_voxelize(const float* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt,
                     const float aabb_min[3], const float aabb_max[3], const uint32_t dims[3],
                     uint8_t* occ, uint8_t* shell_out, uint8_t* parity_out)
{
    (void)nv;
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    float cs[3];
    for (int q = 0; q < 3; ++q) cs[q] = (aabb_max[q] - aabb_min[q]) / (float)dims[q];

    uint8_t* shell = (uint8_t*)calloc(V, 1);
    uint8_t* toggle = (uint8_t*)calloc(V, 1);
    if (!shell || !toggle) { free(shell); free(toggle); return -1; }

    for (uint32_t t = 0; t < nt; ++t) {
        const float* p[3] = { xyz + 3 * (size_t)tri[3 * t], xyz + 3 * (size_t)tri[3 * t + 1],
                              xyz + 3 * (size_t)tri[3 * t + 2] };
        int finite = 1;
        for (int k = 0; k < 3; ++k)
            for (int q = 0; q < 3; ++q) finite &= isfinite(p[k][q]) != 0;
        if (!finite) continue; /* non-finite triangles mark nothing */

        /* --- shell: SAT against every candidate cell of the triangle's bbox --- */
        int lo[3], hi[3], empty = 0;
        for (int q = 0; q < 3; ++q) {
            float mn = p[0][q], mx = p[0][q];
            if (p[1][q] < mn) mn = p[1][q];
            if (p[1][q] > mx) mx = p[1][q];
            if (p[2][q] < mn) mn = p[2][q];
            if (p[2][q] > mx) mx = p[2][q];
            /* cells within 1/64 cell of the triangle's extent: a superset of the
             * cells the FP32 SAT can accept (it only adds touching cells) */
            float fl = floorf((mn - aabb_min[q]) / cs[q] - 0.015625f);
            float fh = floorf((mx - aabb_min[q]) / cs[q] + 0.015625f);
            const float lim = (float)dims[q] + 1.0f;
            fl = fl < -2.0f ? -2.0f : (fl > lim ? lim : fl);
            fh = fh < -2.0f ? -2.0f : (fh > lim ? lim : fh);
            if ((int)fh < 0 || (int)fl > (int)dims[q] - 1) empty = 1;
            lo[q] = clampi((int)fl, 0, (int)dims[q] - 1);
            hi[q] = clampi((int)fh, 0, (int)dims[q] - 1);
        }
        if (empty) { lo[0] = 1; hi[0] = 0; }
        for (int x = lo[0]; x <= hi[0]; ++x)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    size_t id = gidx(x, y, z, dims);
                    if (shell[id]) continue;
                    float c[3], h[3];
                    cell_box(x, aabb_min[0], cs[0], &c[0], &h[0]);
                    cell_box(y, aabb_min[1], cs[1], &c[1], &h[1]);
                    cell_box(z, aabb_min[2], cs[2], &c[2], &h[2]);
                    if (vfo_sat_triangle_box(p[0], p[1], p[2], c, h)) shell[id] = 1;
                }

        /* --- z-parity: toggle the first cell centre above each crossing --- */
        int64_t X[3], Y[3], Z[3];
        for (int k = 0; k < 3; ++k) {
            X[k] = snap_coord(p[k][0], aabb_min[0], cs[0]);
            Y[k] = snap_coord(p[k][1], aabb_min[1], cs[1]);
            Z[k] = snap_coord(p[k][2], aabb_min[2], cs[2]);
        }
        int64_t area = (X[1] - X[0]) * (Y[2] - Y[0]) - (Y[1] - Y[0]) * (X[2] - X[0]);
        if (area == 0) continue;
        if (area < 0) {
            int64_t tx = X[1], ty = Y[1], tz = Z[1];
            X[1] = X[2]; Y[1] = Y[2]; Z[1] = Z[2];
            X[2] = tx; Y[2] = ty; Z[2] = tz;
            area = -area;
        }
        int32_t mnx = (int32_t)X[0], mxx = (int32_t)X[0], mny = (int32_t)Y[0], mxy = (int32_t)Y[0];
        for (int k = 1; k < 3; ++k) {
            if (X[k] < mnx) mnx = (int32_t)X[k];
            if (X[k] > mxx) mxx = (int32_t)X[k];
            if (Y[k] < mny) mny = (int32_t)Y[k];
            if (Y[k] > mxy) mxy = (int32_t)Y[k];
        }
        int i0 = floordiv_sub(mnx - VFO_SUB / 2 + VFO_SUB - 1), i1 = floordiv_sub(mxx - VFO_SUB / 2);
        int j0 = floordiv_sub(mny - VFO_SUB / 2 + VFO_SUB - 1), j1 = floordiv_sub(mxy - VFO_SUB / 2);
        if (i0 < 0) i0 = 0;
        if (j0 < 0) j0 = 0;
        if (i1 > (int)dims[0] - 1) i1 = (int)dims[0] - 1;
        if (j1 > (int)dims[1] - 1) j1 = (int)dims[1] - 1;
        const int64_t d0x = X[2] - X[1], d0y = Y[2] - Y[1];
        const int64_t d1x = X[0] - X[2], d1y = Y[0] - Y[2];
        const int64_t d2x = X[1] - X[0], d2y = Y[1] - Y[0];
        for (int i = i0; i <= i1; ++i) {
            const int64_t px = (int64_t)i * VFO_SUB + VFO_SUB / 2;
            for (int j = j0; j <= j1; ++j) {
                const int64_t py = (int64_t)j * VFO_SUB + VFO_SUB / 2;
                int64_t E0 = d0x * (py - Y[1]) - d0y * (px - X[1]);
                int64_t E1 = d1x * (py - Y[2]) - d1y * (px - X[2]);
                int64_t E2 = d2x * (py - Y[0]) - d2y * (px - X[0]);
                if (!edge_incl(E0, d0x, d0y) || !edge_incl(E1, d1x, d1y) || !edge_incl(E2, d2x, d2y))
                    continue;
                int64_t num = E0 * Z[0] + E1 * Z[1] + E2 * Z[2];
                int64_t k = floordiv64(num - (int64_t)(VFO_SUB / 2) * area, (int64_t)VFO_SUB * area) + 1;
                if (k < 0) k = 0;
                if (k >= (int64_t)dims[2]) continue;
                toggle[gidx(i, j, (size_t)k, dims)] ^= 1;
            }
        }
    }

    int64_t count = 0;
    for (uint32_t x = 0; x < dims[0]; ++x)
        for (uint32_t y = 0; y < dims[1]; ++y) {
            uint8_t par = 0;
            for (uint32_t z = 0; z < dims[2]; ++z) {
                size_t id = gidx(x, y, z, dims);
                par ^= toggle[id];
                uint8_t o = (uint8_t)(shell[id] | par);
                occ[id] = o;
                count += o;
                if (shell_out) shell_out[id] = shell[id];
                if (parity_out) parity_out[id] = par;
            }
        }
    free(shell);
    free(toggle);
    return count;
}

/* ========================================================================= */
/* Grid helpers                                                              */
/* ========================================================================= */

void vfo_occ_to_grid(const uint8_t* occ, uint16_t* grid, size_t n)
{
    /* RegularGrid::fill marks occupied cells VOXEL_FREE (RegularGrid.cpp:185-201) */
    for (size_t i = 0; i < n; ++i) grid[i] = occ[i] ? VFO_FREE : VFO_EMPTY;
}

void vfo_homogenize(uint16_t* grid, size_t n)
{
    /* RegularGrid::homogenize (RegularGrid.cpp:533-541) */
    for (size_t i = 0; i < n; ++i)
        if (grid[i] != VFO_EMPTY) grid[i] = VFO_FREE;
}

void vfo_undo_mask(uint16_t* grid, size_t n)
{
    /* undoMask-comp.glsl:19-23 unmaskBit(.., 15) via RegularGrid::undoMask (RegularGrid.cpp:488-503) */
    for (size_t i = 0; i < n; ++i) grid[i] = (uint16_t)(grid[i] & VFO_MASK15);
}

void vfo_unmask_rightmost(uint16_t* grid, size_t n, uint32_t position)
{
    /* undoMask-comp.glsl:25-29 unmaskRightMost */
    for (size_t i = 0; i < n; ++i) grid[i] = (uint16_t)(grid[i] & (uint16_t)((1u << position) - 1u));
}

void vfo_detect_boundaries(const uint16_t* in, uint16_t* out, const uint32_t dims[3], int b)
{
    /* detectBoundaries-comp.glsl:18-43: a labelled cell is boundary when the
     * clamped (2b+1)^3 window holds a value > FREE (bit 15 cleared) that differs
     * from the cell's own raw value. */
    const int nx = (int)dims[0], ny = (int)dims[1], nz = (int)dims[2];
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                size_t id = gidx(x, y, z, dims);
                uint16_t c = in[id];
                out[id] = c;
                if (c <= VFO_FREE) continue;
                int x0 = clampi(x - b, 0, nx - 1), x1 = clampi(x + b, 0, nx - 1);
                int y0 = clampi(y - b, 0, ny - 1), y1 = clampi(y + b, 0, ny - 1);
                int z0 = clampi(z - b, 0, nz - 1), z1 = clampi(z + b, 0, nz - 1);
                int bd = 0;
                for (int xx = x0; xx <= x1 && !bd; ++xx)
                    for (int yy = y0; yy <= y1 && !bd; ++yy)
                        for (int zz = z0; zz <= z1 && !bd; ++zz) {
                            uint16_t v = (uint16_t)(in[gidx(xx, yy, zz, dims)] & VFO_MASK15);
                            if (v > VFO_FREE && v != c) bd = 1;
                        }
                if (bd) out[id] = (uint16_t)(c | 0x8000u);
            }
}

void vfo_histogram(const uint16_t* grid, size_t n, uint32_t* counts)
{
    /* RegularGrid::countValues (RegularGrid.cpp:601-625): count unmask(value) for value > FREE */
    memset(counts, 0, 65536 * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i)
        if (grid[i] > VFO_FREE) counts[grid[i] & VFO_MASK15]++;
}

/* ========================================================================= */
/* Naive Voronoi fracture                                                    */
/* ========================================================================= */

static inline int64_t seed_dist(int64_t dx, int64_t dy, int64_t dz, int dist)
{
    /* NaiveFracturer distance lambdas (MF/Source/Fracturer/NaiveFracturer.cpp:12-23) and
     * distance.glsl:1-20.  L1 and L-infinity are integers (exact as the reference's floats: below
     * 2^24).  Euclidean is the reference's float sqrt(x*x + y*y + z*z): the sum of squares is an
     * exact float (below 3 * 2047^2 < 2^24) and sqrtf rounds it correctly; the returned value is
     * the bit pattern of that float, which orders as the float does (non-negative).  Distinct
     * integers can share a float root (from 4197200 on), and such seeds tie as in the reference. */
    if (dx < 0) dx = -dx;
    if (dy < 0) dy = -dy;
    if (dz < 0) dz = -dz;
    if (dist == VFO_MANHATTAN) return dx + dy + dz;
    if (dist == VFO_CHEBYSHEV) return dx > dy ? (dx > dz ? dx : dz) : (dy > dz ? dy : dz);
    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
    const float r = sqrtf(fx * fx + fy * fy + fz * fz);
    uint32_t bits;
    memcpy(&bits, &r, 4);
    return (int64_t)bits;
}

void vfo_naive(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
               const uint32_t* seeds, uint32_t S, int dist)
{
    /* NaiveFracturer::buildCPU (NaiveFracturer.cpp:26-68) / naiveFracturer-comp.glsl:19-42:
     * every non-EMPTY cell takes the label of the first seed (list order) with the
     * strictly smallest distance. */
    for (uint32_t x = 0; x < dims[0]; ++x)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t z = 0; z < dims[2]; ++z) {
                size_t id = gidx(x, y, z, dims);
                out[id] = in[id];
                if (in[id] == VFO_EMPTY) continue;
                int64_t best = INT64_MAX;
                for (uint32_t s = 0; s < S; ++s) {
                    int64_t d = seed_dist((int64_t)x - seeds[4 * s], (int64_t)y - seeds[4 * s + 1],
                                          (int64_t)z - seeds[4 * s + 2], dist);
                    if (d < best) { best = d; out[id] = (uint16_t)seeds[4 * s + 3]; }
                }
            }
}

/* ========================================================================= */
/* Flood (region-growing) fracture                                           */
/* ========================================================================= */

static inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16; h *= 0x85EBCA6Bu;
    h ^= h >> 13; h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

uint32_t vfo_hash32(uint64_t key, uint32_t counter)
{
    /* counter-based hash: murmur3 finaliser applied twice, keyed by 64 bits */
    uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
    return fmix32(fmix32(counter ^ lo) + hi * 0x9E3779B1u + 0x7F4A7C15u);
}

uint32_t vfo_noise_delta(uint64_t noise_seed, uint32_t idx, uint32_t amp)
{
    return (uint32_t)(((uint64_t)vfo_hash32(noise_seed, idx) * (uint64_t)(amp + 1u)) >> 32);
}

static const int NB26[26][3] = {
    /* FloodFracturer::MOORE order (FloodFracturer.cpp:13-27) */
    {-1,-1,-1},{1,1,1},{-1,-1,0},{1,1,0},{-1,-1,1},{1,1,-1},{0,-1,-1},{0,1,1},{0,-1,0},{0,1,0},
    {0,-1,1},{0,1,-1},{1,-1,-1},{-1,1,1},{1,-1,0},{-1,1,0},{1,-1,1},{-1,1,-1},{-1,0,-1},{1,0,1},
    {-1,0,0},{1,0,0},{-1,0,1},{1,0,-1},{0,0,-1},{0,0,1}};
static const int NB6[6][3] = {
    /* FloodFracturer::VON_NEUMANN order (FloodFracturer.cpp:8-11) */
    {1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};

/* max_label 255: plain fragment ids; 0xFFFF: fragment id in the low 8 bits
 * (at least 2) and a sub-seed id in the high 8 bits (Seeder.h:12). */
static int flood_validate(const uint32_t dims[3], const uint32_t* seeds, uint32_t S, int nb, uint32_t max_label)
{
    if (nb != 6 && nb != 26) return -1;
    for (uint32_t s = 0; s < S; ++s) {
        if (seeds[4 * s] >= dims[0] || seeds[4 * s + 1] >= dims[1] || seeds[4 * s + 2] >= dims[2]) return -1;
        if ((seeds[4 * s + 3] & 0xFFu) < 2 || seeds[4 * s + 3] > max_label) return -1;
    }
    return 0;
}

/* A seed acts as a source unless a later seed stamps the same cell
 * (FloodFracturer.cpp:99-105 writes seeds in list order). */
static uint8_t* flood_active_seeds(const uint32_t dims[3], const uint32_t* seeds, uint32_t S)
{
    uint8_t* act = (uint8_t*)malloc(S ? S : 1);
    for (uint32_t s = 0; s < S; ++s) {
        act[s] = 1;
        for (uint32_t t = s + 1; t < S; ++t)
            if (seeds[4 * t] == seeds[4 * s] && seeds[4 * t + 1] == seeds[4 * s + 1] &&
                seeds[4 * t + 2] == seeds[4 * s + 2]) { act[s] = 0; break; }
    }
    (void)dims;
    return act;
}

typedef struct { uint32_t* v; size_t n, cap; } u32vec;
static void u32push(u32vec* a, uint32_t x)
{
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 64; a->v = (uint32_t*)realloc(a->v, a->cap * 4); }
    a->v[a->n++] = x;
}

/* Dial's bucket algorithm over the lexicographic key (D << 32 | s): key[] holds
 * the sources at D = 0 (listed in src) and UINT64_MAX elsewhere; a path may
 * enter the cells with occ[] set, each adding 1 + noise_delta(cell). */
static void flood_dial(const uint8_t* occ, const uint32_t dims[3], int nb, uint64_t noise_seed, uint32_t amp,
                       uint64_t* key, const u32vec* src)
{
    const int (*off)[3] = nb == 6 ? NB6 : NB26;
    u32vec* bucket = NULL;
    size_t nbuckets = 0;
#define BUCKET(d) (((d) >= nbuckets ? (bucket = (u32vec*)realloc(bucket, ((d) + 64) * sizeof(u32vec)), \
                     memset(bucket + nbuckets, 0, ((d) + 64 - nbuckets) * sizeof(u32vec)), nbuckets = (d) + 64) : 0), &bucket[d])
    for (size_t e = 0; e < src->n; ++e) u32push(BUCKET(0), src->v[e]);
    for (size_t d = 0; d < nbuckets; ++d) {
        for (size_t e = 0; e < bucket[d].n; ++e) {
            uint32_t v = bucket[d].v[e];
            if ((key[v] >> 32) != d) continue;
            uint64_t s = key[v] & 0xFFFFFFFFu;
            int z = (int)(v % dims[2]), y = (int)((v / dims[2]) % dims[1]), x = (int)(v / ((size_t)dims[1] * dims[2]));
            for (int k = 0; k < nb; ++k) {
                int xx = x + off[k][0], yy = y + off[k][1], zz = z + off[k][2];
                if (xx < 0 || yy < 0 || zz < 0 || xx >= (int)dims[0] || yy >= (int)dims[1] || zz >= (int)dims[2]) continue;
                size_t u = gidx(xx, yy, zz, dims);
                if (!occ[u]) continue;
                uint64_t du = d + 1 + vfo_noise_delta(noise_seed, (uint32_t)u, amp);
                uint64_t cand = (du << 32) | s;
                if (cand < key[u]) {
                    if ((key[u] >> 32) != du) u32push(BUCKET(du), (uint32_t)u);
                    key[u] = cand;
                }
            }
        }
    }
#undef BUCKET
    for (size_t d = 0; d < nbuckets; ++d) free(bucket[d].v);
    free(bucket);
}

/*
 * Deterministic flood.  The reference grows all seeds one BFS step at a time
 * and resolves simultaneous claims by a race (floodFracturer-comp.glsl:24-63,
 * FloodFracturer::build FloodFracturer.cpp:98-191).  The deterministic
 * refinement computed here: owner(v) = argmin over seeds s of (D_s(v), s),
 * where D_s is the node-weighted geodesic distance in the 6- or 26-connected
 * graph of occupied cells, weight(v) = 1 + noise_delta(v).  With zero noise
 * this is one of the race's possible outcomes.  Computed by Dial's bucket
 * algorithm over the lexicographic key (D, s).
 *
 * Sub-seeds (labels fragment | id << 8, as Seeder::mergeSeeds makes them,
 * Seeder.cpp:115-152): while growing, the reference lowers the id of every
 * cell to the smallest id met across same-fragment neighbours
 * (floodFracturer-comp.glsl:49-62), so each connected component of a
 * fragment's cells ends with the smallest id of its seeds; then every cell
 * whose id is not the fragment's smallest becomes FREE
 * (disjointSet-comp.glsl, disjointSetStack-comp.glsl) and the grid is flooded
 * again from all remaining cells at once (FloodFracturer.cpp:134-177).  Here:
 * components of {v : fragment(owner(v)) = f} under the flood's stencil, kept
 * iff their smallest id equals the fragment's smallest; the freed cells are
 * re-flooded from every kept cell at distance 0, ties to the smaller owner
 * index.  A second pass frees nothing (each fragment is then one component
 * carrying its smallest id), as in the reference's loop.
 *
 * Unreached occupied cells stay FREE; the final unmaskRightMost keeps the low
 * 8 bits.
 */
int vfo_flood(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
              const uint32_t* seeds, uint32_t S, int nb,
              uint64_t noise_seed, uint32_t amp)
{
    if (flood_validate(dims, seeds, S, nb, 0xFFFF)) return -1;
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    const int (*off)[3] = nb == 6 ? NB6 : NB26;
    uint64_t* key = (uint64_t*)malloc(V * sizeof(uint64_t));
    uint8_t* occ = (uint8_t*)calloc(V ? V : 1, 1);
    for (size_t i = 0; i < V; ++i) { key[i] = UINT64_MAX; occ[i] = in[i] != VFO_EMPTY; }
    uint8_t* act = flood_active_seeds(dims, seeds, S);
    u32vec src = { 0 };
    for (uint32_t s = 0; s < S; ++s) {
        if (!act[s]) continue;
        size_t id = gidx(seeds[4 * s], seeds[4 * s + 1], seeds[4 * s + 2], dims);
        key[id] = (uint64_t)s;
        occ[id] = 1;  /* grid.set makes a seed cell occupied */
        u32push(&src, (uint32_t)id);
    }
    flood_dial(occ, dims, nb, noise_seed, amp, key, &src);

    /* sub-seed merge: components of each fragment's cells, their smallest ids */
#define OWNER(v) ((uint32_t)(key[v] & 0xFFFFFFFFu))
#define FRAG(v) (seeds[4 * OWNER(v) + 3] & 0xFFu)
#define SUBID(v) (seeds[4 * OWNER(v) + 3] >> 8)
    uint32_t* comp = (uint32_t*)malloc(V * sizeof(uint32_t));
    for (size_t i = 0; i < V; ++i) comp[i] = UINT32_MAX;
    u32vec cmin = { 0 }, queue = { 0 };
    uint32_t fmin[256];
    for (int f = 0; f < 256; ++f) fmin[f] = UINT32_MAX;
    for (size_t v0 = 0; v0 < V; ++v0) {
        if (key[v0] == UINT64_MAX || comp[v0] != UINT32_MAX) continue;
        const uint32_t c = (uint32_t)cmin.n, f = FRAG(v0);
        uint32_t m = SUBID(v0);
        queue.n = 0;
        u32push(&queue, (uint32_t)v0);
        comp[v0] = c;
        for (size_t h = 0; h < queue.n; ++h) {
            uint32_t v = queue.v[h];
            if (SUBID(v) < m) m = SUBID(v);
            int z = (int)(v % dims[2]), y = (int)((v / dims[2]) % dims[1]), x = (int)(v / ((size_t)dims[1] * dims[2]));
            for (int k = 0; k < nb; ++k) {
                int xx = x + off[k][0], yy = y + off[k][1], zz = z + off[k][2];
                if (xx < 0 || yy < 0 || zz < 0 || xx >= (int)dims[0] || yy >= (int)dims[1] || zz >= (int)dims[2]) continue;
                size_t u = gidx(xx, yy, zz, dims);
                if (key[u] == UINT64_MAX || comp[u] != UINT32_MAX || FRAG(u) != f) continue;
                comp[u] = c;
                u32push(&queue, (uint32_t)u);
            }
        }
        u32push(&cmin, m);
        if (m < fmin[f]) fmin[f] = m;
    }
    int freed = 0;
    for (size_t v = 0; v < V && !freed; ++v)
        if (key[v] != UINT64_MAX && cmin.v[comp[v]] != fmin[FRAG(v)]) freed = 1;
    if (freed) {
        /* re-flood the freed cells from every kept cell */
        src.n = 0;
        for (size_t v = 0; v < V; ++v) {
            if (key[v] == UINT64_MAX) continue;
            if (cmin.v[comp[v]] == fmin[FRAG(v)]) {
                key[v] = OWNER(v);
                u32push(&src, (uint32_t)v);
            } else {
                key[v] = UINT64_MAX;
            }
        }
        flood_dial(occ, dims, nb, noise_seed, amp, key, &src);
    }
#undef OWNER
#undef FRAG
#undef SUBID
    for (size_t i = 0; i < V; ++i) {
        if (key[i] != UINT64_MAX) out[i] = (uint16_t)(seeds[4 * (key[i] & 0xFFFFFFFFu) + 3] & 0xFFu);
        else out[i] = occ[i] ? VFO_FREE : VFO_EMPTY;
    }
    free(comp); free(cmin.v); free(queue.v); free(src.v);
    free(key);
    free(occ);
    free(act);
    return 0;
}

/* binary heap of (dist, cell) for the brute-force check */
typedef struct { uint64_t d; uint32_t v; } hent;
static void hpush(hent** h, size_t* n, size_t* cap, hent e)
{
    if (*n == *cap) { *cap = *cap ? *cap * 2 : 256; *h = (hent*)realloc(*h, *cap * sizeof(hent)); }
    size_t i = (*n)++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if ((*h)[p].d <= e.d) break;
        (*h)[i] = (*h)[p];
        i = p;
    }
    (*h)[i] = e;
}
static hent hpop(hent* h, size_t* n)
{
    hent top = h[0], last = h[--(*n)];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        uint64_t md = last.d;
        if (l < *n && h[l].d < md) { m = l; md = h[l].d; }
        if (r < *n && h[r].d < md) { m = r; }
        if (m == i) break;
        h[i] = h[m];
        i = m;
    }
    if (*n) h[i] = last;
    return top;
}

int vfo_flood_dijkstra(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
                       const uint32_t* seeds, uint32_t S, int nb,
                       uint64_t noise_seed, uint32_t amp)
{
    if (flood_validate(dims, seeds, S, nb, 255)) return -1;
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    const int (*off)[3] = nb == 6 ? NB6 : NB26;
    uint64_t* best = (uint64_t*)malloc(V * sizeof(uint64_t));
    uint32_t* bests = (uint32_t*)malloc(V * sizeof(uint32_t));
    uint64_t* dist = (uint64_t*)malloc(V * sizeof(uint64_t));
    for (size_t i = 0; i < V; ++i) { best[i] = UINT64_MAX; bests[i] = UINT32_MAX; }
    uint8_t* act = flood_active_seeds(dims, seeds, S);
    hent* heap = NULL;
    size_t hn = 0, hcap = 0;
    for (uint32_t s = 0; s < S; ++s) {
        if (!act[s]) continue;
        for (size_t i = 0; i < V; ++i) dist[i] = UINT64_MAX;
        size_t src = gidx(seeds[4 * s], seeds[4 * s + 1], seeds[4 * s + 2], dims);
        dist[src] = 0;
        hn = 0;
        hpush(&heap, &hn, &hcap, (hent){ 0, (uint32_t)src });
        while (hn) {
            hent e = hpop(heap, &hn);
            if (e.d != dist[e.v]) continue;
            int z = (int)(e.v % dims[2]), y = (int)((e.v / dims[2]) % dims[1]), x = (int)(e.v / ((size_t)dims[1] * dims[2]));
            for (int k = 0; k < nb; ++k) {
                int xx = x + off[k][0], yy = y + off[k][1], zz = z + off[k][2];
                if (xx < 0 || yy < 0 || zz < 0 || xx >= (int)dims[0] || yy >= (int)dims[1] || zz >= (int)dims[2]) continue;
                size_t u = gidx(xx, yy, zz, dims);
                if (in[u] == VFO_EMPTY || u == src) continue;
                uint64_t nd = e.d + 1 + vfo_noise_delta(noise_seed, (uint32_t)u, amp);
                if (nd < dist[u]) { dist[u] = nd; hpush(&heap, &hn, &hcap, (hent){ nd, (uint32_t)u }); }
            }
        }
        for (size_t i = 0; i < V; ++i)
            if (dist[i] < best[i] || (dist[i] == best[i] && dist[i] != UINT64_MAX && s < bests[i])) {
                best[i] = dist[i];
                bests[i] = s;
            }
    }
    for (size_t i = 0; i < V; ++i) {
        if (bests[i] != UINT32_MAX) out[i] = (uint16_t)(seeds[4 * bests[i] + 3] & 0xFFu);
        else out[i] = in[i] != VFO_EMPTY ? VFO_FREE : VFO_EMPTY;
    }
    free(heap); free(best); free(bests); free(dist); free(act);
    return 0;
}

void vfo_remove_isolated(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
                         const uint32_t* seeds, uint32_t S)
{
    /* NaiveFracturer::removeIsolatedRegionsCPU (NaiveFracturer.cpp:111-150): a FIFO
     * BFS from the seeds over 6-neighbours holding the seed's label; cells it
     * does not reach become EMPTY. */
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    memset(out, 0, V * sizeof(uint16_t));
    u32vec q = { 0 }, ql = { 0 };
    for (uint32_t s = 0; s < S; ++s) out[gidx(seeds[4 * s], seeds[4 * s + 1], seeds[4 * s + 2], dims)] = (uint16_t)seeds[4 * s + 3];
    for (uint32_t s = 0; s < S; ++s) {
        u32push(&q, (uint32_t)gidx(seeds[4 * s], seeds[4 * s + 1], seeds[4 * s + 2], dims));
        u32push(&ql, seeds[4 * s + 3]);
    }
    for (size_t h = 0; h < q.n; ++h) {
        uint32_t v = q.v[h], w = ql.v[h];
        int z = (int)(v % dims[2]), y = (int)((v / dims[2]) % dims[1]), x = (int)(v / ((size_t)dims[1] * dims[2]));
        for (int k = 0; k < 6; ++k) {
            int xx = x + NB6[k][0], yy = y + NB6[k][1], zz = z + NB6[k][2];
            if (xx < 0 || yy < 0 || zz < 0 || xx >= (int)dims[0] || yy >= (int)dims[1] || zz >= (int)dims[2]) continue;
            size_t u = gidx(xx, yy, zz, dims);
            if (in[u] == (uint16_t)w && (uint32_t)in[u] == w && out[u] == VFO_EMPTY) {
                out[u] = (uint16_t)w;
                u32push(&q, (uint32_t)u);
                u32push(&ql, w);
            }
        }
    }
    free(q.v);
    free(ql.v);
}

/* ========================================================================= */
/* Erosion                                                                   */
/* ========================================================================= */

int vfo_erosion_mask(int type, uint32_t size, uint8_t* mask, float* activations)
{
    /* RegularGrid::erode (RegularGrid.cpp:82-121): an odd cube of side size
     * (even sizes grow by one); SQUARE all cells, CROSS the three axis lines
     * through the centre (activations = maskSize / 3 as the reference sets
     * it), ELLIPSE the cells whose float distance to the centre is below
     * centre + FLT_EPSILON.  Returns the side; activations = fraction. */
    if (size % 2 == 0) ++size;
    const uint32_t n = size * size * size, c = size / 2;
    float act = 0.0f;
    memset(mask, 0, n);
    if (type == VFO_EROSION_SQUARE) {
        memset(mask, 1, n);
        act = (float)n;
    } else if (type == VFO_EROSION_CROSS) {
        for (uint32_t i = 0; i < size; ++i) {
            mask[i * size * size + c * size + c] = 1;
            mask[c * size * size + i * size + c] = 1;
            mask[c * size * size + c * size + i] = 1;
        }
        act = 1.0f / 3.0f * (float)n;
    } else if (type == VFO_EROSION_ELLIPSE) {
        for (uint32_t x = 0; x < size; ++x)
            for (uint32_t y = 0; y < size; ++y)
                for (uint32_t z = 0; z < size; ++z) {
                    float dx = (float)x - (float)c, dy = (float)y - (float)c, dz = (float)z - (float)c;
                    float d = sqrtf(dx * dx + dy * dy + dz * dz);
                    if (d < (float)c + FLT_EPSILON) {
                        mask[x * size * size + y * size + z] = 1;
                        act += 1.0f;
                    }
                }
    } else {
        return -1;
    }
    *activations = act / (float)n;
    return (int)size;
}

float vfo_erosion_noise(uint64_t seed, uint32_t cell)
{
    /* the reference's uniform float noise buffer (RegularGrid::fillNoiseBuffer,
     * RegularGrid.cpp:238-244) as a counter-based draw: 24 bits to [0, 1) */
    return (float)(vfo_hash32(seed, cell) >> 8) * (1.0f / 16777216.0f);
}

void vfo_erode_step(const uint16_t* in, uint16_t* out, const uint32_t dims[3], const uint8_t* mask, uint32_t size,
                    float limit, float probability, uint64_t noise_seed)
{
    /* erodeGrid-comp.glsl:26-59.  The shader's isBoundary is
     * unmaskedBit(value, 15) != 0, i.e. any non-EMPTY cell, so every cell above
     * FREE passing the noise gate is tested; neighbours count when their raw
     * value (boundary bit included) equals the cell's and the mask is set;
     * the cell empties when count / (in-bounds window cells) < limit
     * (limit = activations * threshold, both FP32). */
    const int nx = (int)dims[0], ny = (int)dims[1], nz = (int)dims[2], h = (int)size / 2;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                size_t id = gidx(x, y, z, dims);
                uint16_t v = in[id];
                out[id] = v;
                if (!(v > VFO_FREE && (v & VFO_MASK15) != 0 && vfo_erosion_noise(noise_seed, (uint32_t)id) < probability))
                    continue;
                uint32_t count = 0, global = 0;
                for (int xx = clampi(x - h, 0, nx - 1); xx <= clampi(x + h, 0, nx - 1); ++xx)
                    for (int yy = clampi(y - h, 0, ny - 1); yy <= clampi(y + h, 0, ny - 1); ++yy)
                        for (int zz = clampi(z - h, 0, nz - 1); zz <= clampi(z + h, 0, nz - 1); ++zz) {
                            if (in[gidx(xx, yy, zz, dims)] == v &&
                                mask[(xx - (x - h)) * size * size + (yy - (y - h)) * size + (zz - (z - h))])
                                ++count;
                            ++global;
                        }
                if ((float)count / (float)global < limit) out[id] = VFO_EMPTY;
            }
}

void vfo_remove_isolated_cells(const uint16_t* in, uint16_t* out, const uint32_t dims[3])
{
    /* removeIsolatedRegionsGrid-comp.glsl:16-38: a cell with fewer than 6
     * other cells of its raw value in the clamped 3x3x3 window becomes EMPTY.
     * The shader updates the grid in place, so its result depends on thread
     * order; this reads the input only. */
    const int nx = (int)dims[0], ny = (int)dims[1], nz = (int)dims[2];
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                size_t id = gidx(x, y, z, dims);
                uint16_t v = in[id];
                int count = -1;
                for (int xx = clampi(x - 1, 0, nx - 1); xx <= clampi(x + 1, 0, nx - 1); ++xx)
                    for (int yy = clampi(y - 1, 0, ny - 1); yy <= clampi(y + 1, 0, ny - 1); ++yy)
                        for (int zz = clampi(z - 1, 0, nz - 1); zz <= clampi(z + 1, 0, nz - 1); ++zz)
                            count += in[gidx(xx, yy, zz, dims)] == v;
                out[id] = count < 6 ? VFO_EMPTY : v;
            }
}

int vfo_erode(uint16_t* grid, const uint32_t dims[3], int type, uint32_t size, uint32_t iterations, float probability,
              float threshold, uint64_t noise_seed)
{
    /* RegularGrid::erode (RegularGrid.cpp:82-159): per iteration
     * detectBoundaries(1) then one erosion step; then removeIsolatedRegions. */
    uint32_t side = size % 2 ? size : size + 1;
    uint8_t* mask = (uint8_t*)malloc((size_t)side * side * side);
    float act;
    if (vfo_erosion_mask(type, size, mask, &act) < 0) { free(mask); return -1; }
    const float limit = act * threshold;
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    uint16_t* tmp = (uint16_t*)malloc(V * sizeof(uint16_t));
    for (uint32_t it = 0; it < iterations; ++it) {
        vfo_detect_boundaries(grid, tmp, dims, 1);
        vfo_erode_step(tmp, grid, dims, mask, side, limit, probability, noise_seed);
    }
    vfo_remove_isolated_cells(grid, tmp, dims);
    memcpy(grid, tmp, V * sizeof(uint16_t));
    free(tmp);
    free(mask);
    return 0;
}

/* ========================================================================= */
/* Seeding                                                                   */
/* ========================================================================= */

static int is_surface(const uint16_t* g, const uint32_t* d, int x, int y, int z)
{
    /* RegularGrid::isBoundary(x, y, z, 1) (RegularGrid.cpp:543-558): an EMPTY cell in the clamped 3^3 window */
    int x0 = clampi(x - 1, 0, (int)d[0] - 1), x1 = clampi(x + 1, 0, (int)d[0] - 1);
    int y0 = clampi(y - 1, 0, (int)d[1] - 1), y1 = clampi(y + 1, 0, (int)d[1] - 1);
    int z0 = clampi(z - 1, 0, (int)d[2] - 1), z1 = clampi(z + 1, 0, (int)d[2] - 1);
    for (int a = x0; a <= x1; ++a)
        for (int b = y0; b <= y1; ++b)
            for (int c = z0; c <= z1; ++c)
                if (g[gidx(a, b, c, d)] == VFO_EMPTY) return 1;
    return 0;
}

static int cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

int vfo_seed_uniform(const uint16_t* grid, const uint32_t dims[3], uint32_t S,
                     uint64_t seed, int location, uint32_t* seeds_out)
{
    /* Seeder::uniform (MF/Source/Fracturer/Seeder.cpp:154-208) with a counter-based
     * draw: S distinct candidate cells (occupied; OUTER/INNER filter by
     * isBoundary), uniform over candidate ranks, emitted in (x,y,z) order with
     * labels 2, 3, ... */
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    uint32_t* cand = (uint32_t*)malloc(V * sizeof(uint32_t));
    size_t O = 0;
    for (uint32_t x = 0; x < dims[0]; ++x)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t z = 0; z < dims[2]; ++z) {
                size_t id = gidx(x, y, z, dims);
                if (grid[id] == VFO_EMPTY) continue;
                if (location != VFO_SEED_BOTH) {
                    int sf = is_surface(grid, dims, (int)x, (int)y, (int)z);
                    if ((location == VFO_SEED_OUTER) != (sf != 0)) continue;
                }
                cand[O++] = (uint32_t)id;
            }
    if (O < S || O > 0xFFFFFFFFull) { free(cand); return -1; }
    uint32_t* ranks = (uint32_t*)malloc((S ? S : 1) * sizeof(uint32_t));
    uint32_t got = 0;
    for (uint32_t j = 0; got < S; ++j) {
        if (j == 1000000u) { free(cand); free(ranks); return -2; }
        uint32_t r = (uint32_t)(((uint64_t)vfo_hash32(seed, j) * (uint64_t)O) >> 32);
        int dup = 0;
        for (uint32_t k = 0; k < got; ++k) if (ranks[k] == r) { dup = 1; break; }
        if (!dup) ranks[got++] = r;
    }
    qsort(ranks, S, sizeof(uint32_t), cmp_u32);
    for (uint32_t k = 0; k < S; ++k) {
        uint32_t id = cand[ranks[k]];
        seeds_out[4 * k + 0] = id / (dims[1] * dims[2]);
        seeds_out[4 * k + 1] = (id / dims[2]) % dims[1];
        seeds_out[4 * k + 2] = id % dims[2];
        seeds_out[4 * k + 3] = 2 + k;
    }
    free(cand);
    free(ranks);
    return 0;
}

static float merge_dist(const uint32_t* a, const uint32_t* b, int dist)
{
    /* Seeder::mergeSeeds' float distances on vec3 coordinates (Seeder.cpp:126-140): glm::distance
     * for Euclidean, sums and maxima of float absolute differences otherwise */
    const float dx = (float)a[0] - (float)b[0], dy = (float)a[1] - (float)b[1], dz = (float)a[2] - (float)b[2];
    if (dist == VFO_MANHATTAN) return fabsf(dx) + fabsf(dy) + fabsf(dz);
    if (dist == VFO_CHEBYSHEV) return fmaxf(fabsf(dx), fmaxf(fabsf(dy), fabsf(dz)));
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

int vfo_seed_merge(const uint32_t* frags, uint32_t F, uint32_t* seeds, uint32_t N, int dist)
{
    /* Seeder::mergeSeeds (Seeder.cpp:115-152): each seed takes the label of its
     * nearest fragment seed (first on ties) with a per-fragment running sub-id
     * in bits 8..15, with the reference's float distances and strict '<'. */
    uint32_t next[256] = { 0 };
    if (N && !F) return -1;
    for (uint32_t f = 0; f < F; ++f)
        if (frags[4 * f + 3] > 255) return -1;
    for (uint32_t i = 0; i < N; ++i) {
        float best = FLT_MAX;
        uint32_t nearest = 0;
        for (uint32_t f = 0; f < F; ++f) {
            float d = merge_dist(seeds + 4 * i, frags + 4 * f, dist);
            if (d < best) { best = d; nearest = f; }
        }
        uint32_t w = frags[4 * nearest + 3];
        if (++next[w] > 255) return -2;
        seeds[4 * i + 3] = w | (next[w] << 8);
    }
    return 0;
}

/* floor(min + (max - min) * h / 2^32): static_cast<int>(min + (max - min) * u)
 * of RandomUtilities::getUniformRandomInt (RandomUtilities.h:141-144) for
 * u = h / 2^32 whenever the result is not negative */
static int64_t uniform_int(int64_t mn, int64_t mx, uint32_t h)
{
    int64_t t = (mx - mn) * (int64_t)h;
    return mn + (t >= 0 ? t >> 32 : -((-t + 0xFFFFFFFFll) >> 32));
}

uint64_t vfo_near_key(uint64_t seed, uint32_t impact)
{
    return seed + (uint64_t)(impact + 1u) * 0x9E3779B97F4A7C15ull;
}

int vfo_seed_near(const uint16_t* grid, const uint32_t dims[3], const uint32_t* frags, uint32_t F,
                  uint32_t impacts, uint32_t num_seeds, uint32_t spreading, uint64_t seed,
                  uint32_t* out, uint32_t* count)
{
    /* Seeder::nearSeeds (Seeder.cpp:49-113) with counter-based draws.  Impact i
     * uses key k_i = vfo_near_key(seed, i): fragment index uniform_int(0, F-1,
     * h(k_i, 0)), seed count min(uniform_int(1, pending, h(k_i, 1)), pending);
     * candidate t draws axis a as n_a/2 - sum_{j<spreading} h(k_i, 2 + (3t+a)
     * spreading + j) mod (n_a / spreading) (getBiasedRandomInt), wraps it
     * around the fragment seed, and is accepted when within min(n)/2 of it,
     * occupied, on the surface (isBoundary) and new.  New seeds follow the
     * fragments in (x, y, z) order with labels after the last fragment's. */
    if (!F || !spreading || spreading > 64) return -1;
    for (int a = 0; a < 3; ++a)
        if (dims[a] < spreading) return -1;
    const uint32_t n[3] = { dims[0], dims[1], dims[2] };
    const uint32_t m[3] = { n[0] / spreading, n[1] / spreading, n[2] / spreading };
    const uint32_t mind = (n[0] < n[1] ? (n[0] < n[2] ? n[0] : n[2]) : (n[1] < n[2] ? n[1] : n[2])) / 2;
    uint32_t* chosen = (uint32_t*)malloc((num_seeds ? num_seeds : 1) * sizeof(uint32_t));
    uint32_t nchosen = 0, pending = num_seeds;
    for (uint32_t i = 0; i < impacts; ++i) {
        const uint64_t k = vfo_near_key(seed, i);
        const uint32_t* fr = frags + 4 * (uint32_t)uniform_int(0, (int64_t)F - 1, vfo_hash32(k, 0));
        int64_t want = uniform_int(1, pending, vfo_hash32(k, 1));
        if (want > pending) want = pending;
        if (want < 0) want = 0;
        uint32_t got = 0;
        for (uint32_t t = 0; got < (uint32_t)want; ++t) {
            if (t == 1000000u) { free(chosen); return -2; }
            uint32_t c[3];
            int64_t d2 = 0;
            for (int a = 0; a < 3; ++a) {
                uint32_t b = 0;
                for (uint32_t j = 0; j < spreading; ++j) b += vfo_hash32(k, 2 + (3 * t + a) * spreading + j) % m[a];
                int64_t o = (int64_t)(n[a] / 2) - b;
                c[a] = (uint32_t)(((int64_t)fr[a] + o + n[a]) % n[a]);
                d2 += ((int64_t)c[a] - fr[a]) * ((int64_t)c[a] - fr[a]);
            }
            if (d2 > (int64_t)mind * mind) continue;
            size_t id = gidx(c[0], c[1], c[2], dims);
            if (grid[id] == VFO_EMPTY || !is_surface(grid, dims, (int)c[0], (int)c[1], (int)c[2])) continue;
            int dup = 0;
            for (uint32_t q = 0; q < nchosen; ++q) if (chosen[q] == id) { dup = 1; break; }
            if (dup) continue;
            chosen[nchosen++] = (uint32_t)id;
            ++got;
        }
        pending -= (uint32_t)want;
    }
    qsort(chosen, nchosen, sizeof(uint32_t), cmp_u32);
    memcpy(out, frags, (size_t)F * 4 * sizeof(uint32_t));
    uint32_t label = frags[4 * (F - 1) + 3];
    for (uint32_t q = 0; q < nchosen; ++q) {
        uint32_t id = chosen[q];
        uint32_t* o = out + 4 * (F + q);
        o[0] = id / (dims[1] * dims[2]);
        o[1] = (id / dims[2]) % dims[1];
        o[2] = id % dims[2];
        o[3] = ++label;
    }
    *count = F + nchosen;
    free(chosen);
    return 0;
}

/* ========================================================================= */
/* Marching cubes                                                            */
/* ========================================================================= */

/* cube corner offsets (marchingCubes-comp.glsl:28-38) and edge corner pairs (:40-54) */
static const int MC_NB[8][3] = { {0,0,0},{0,0,1},{-1,0,1},{-1,0,0},{0,1,0},{0,1,1},{-1,1,1},{-1,1,0} };
static const int MC_EDGE[12][2] = { {0,1},{1,2},{2,3},{3,0},{4,5},{5,6},{6,7},{7,4},{0,4},{1,5},{2,6},{3,7} };

uint64_t vfo_mc_packed(uint32_t config) { return vfo_mc_tri_packed[config & 255u]; }

typedef struct { uint64_t* v; size_t n, cap; } u64vec;
static void u64push(u64vec* a, uint64_t x)
{
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 256; a->v = (uint64_t*)realloc(a->v, a->cap * 8); }
    a->v[a->n++] = x;
}
static int cmp_u64(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/*
 * Per-label marching cubes over the grid padded by one FREE cell
 * (MarchingCubes::setGrid MarchingCubes.cpp:516-533, march() in
 * marchingCubes-comp.glsl:96-156).  Labels are those countValues reports
 * (RegularGrid::toTriangleMesh RegularGrid.cpp:443-486).  Field f = (cell &
 * 0x7FFF) == label, isolevel 0.5, so each vertex is an exact edge midpoint;
 * triangle corners are emitted in table order (0, 2, 1); the flag is bit 15
 * of the cube's corner 0.  Output is canonical: per label, triangles in cube
 * order then table order, and vertices welded exactly by lattice edge, in
 * edge-id order.  World position = aabb_min + cell_size * (p - 1).
 */
int vfo_marching_cubes(const uint16_t* grid, const uint32_t dims[3],
                       const float aabb_min[3], const float aabb_max[3],
                       uint32_t want, vfo_mesh* m)
{
    memset(m, 0, sizeof(*m));
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    uint32_t* counts = (uint32_t*)malloc(65536 * sizeof(uint32_t));
    vfo_histogram(grid, V, counts);
    const int64_t N[3] = { (int64_t)dims[0] + 2, (int64_t)dims[1] + 2, (int64_t)dims[2] + 2 };
    float cs[3];
    double cs64[3];
    for (int q = 0; q < 3; ++q) {
        cs[q] = (aabb_max[q] - aabb_min[q]) / (float)dims[q];
        cs64[q] = ((double)aabb_max[q] - (double)aabb_min[q]) / (double)dims[q];
    }
#define PADV(px, py, pz) (((px) <= 0 || (py) <= 0 || (pz) <= 0 || (px) >= N[0] - 1 || (py) >= N[1] - 1 || (pz) >= N[2] - 1) \
                            ? (uint16_t)VFO_FREE : grid[gidx((px) - 1, (py) - 1, (pz) - 1, dims)])

    /* each label's bounding box in grid coordinates: a cube none of whose corners lies in it has
     * configuration 255, which emits nothing, so the cube loop is limited to the cubes touching it */
    int32_t* bb = (int32_t*)malloc(65536 * 6 * sizeof(int32_t));
    for (uint32_t L = 0; L < 65536; ++L) {
        bb[6 * L + 0] = bb[6 * L + 1] = bb[6 * L + 2] = INT32_MAX;
        bb[6 * L + 3] = bb[6 * L + 4] = bb[6 * L + 5] = -1;
    }
    for (uint32_t x = 0; x < dims[0]; ++x)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t z = 0; z < dims[2]; ++z) {
                int32_t* b = bb + 6 * (grid[gidx(x, y, z, dims)] & VFO_MASK15);
                if ((int32_t)x < b[0]) b[0] = (int32_t)x;
                if ((int32_t)y < b[1]) b[1] = (int32_t)y;
                if ((int32_t)z < b[2]) b[2] = (int32_t)z;
                if ((int32_t)x > b[3]) b[3] = (int32_t)x;
                if ((int32_t)y > b[4]) b[4] = (int32_t)y;
                if ((int32_t)z > b[5]) b[5] = (int32_t)z;
            }

    u64vec tri_edges = { 0 }, tri_meta = { 0 }, verts = { 0 };
    uint32_t nl = 0;
    uint16_t* labels = (uint16_t*)malloc(65536 * sizeof(uint16_t));
    uint32_t* voff = (uint32_t*)malloc(65537 * sizeof(uint32_t));
    uint32_t* toff = (uint32_t*)malloc(65537 * sizeof(uint32_t));
    voff[0] = toff[0] = 0;
    for (uint32_t L = 0; L < 65536; ++L) {
        if (!counts[L]) continue;
        if (want != 0xFFFFFFFFu && want != L) continue;
        labels[nl] = (uint16_t)L;
        size_t tbegin = tri_meta.n;
        /* cube (cx, cy, cz) has corners at grid x in {cx - 1, cx - 2}, y in {cy - 1, cy}, z in {cz - 1, cz} */
        const int32_t* b = bb + 6 * L;
        const int64_t x0 = b[0] + 1 > 1 ? b[0] + 1 : 1, x1 = b[3] + 2 < N[0] - 1 ? b[3] + 2 : N[0] - 1;
        const int64_t y0 = b[1], y1 = b[4] + 1 < N[1] - 2 ? b[4] + 1 : N[1] - 2;
        const int64_t z0 = b[2], z1 = b[5] + 1 < N[2] - 2 ? b[5] + 1 : N[2] - 2;
        for (int64_t cx = x0; cx <= x1; ++cx)
            for (int64_t cy = y0; cy <= y1; ++cy)
                for (int64_t cz = z0; cz <= z1; ++cz) {
                    int conf = 0;
                    uint16_t c0 = 0;
                    for (int i = 0; i < 8; ++i) {
                        uint16_t raw = PADV(cx + MC_NB[i][0], cy + MC_NB[i][1], cz + MC_NB[i][2]);
                        if (i == 0) c0 = raw;
                        if ((uint32_t)(raw & VFO_MASK15) != L) conf |= 1 << i;
                    }
                    uint64_t w = vfo_mc_tri_packed[conf];
                    int ntri = (int)(w >> 60);
                    uint64_t cube = (uint64_t)((cx * N[1] + cy) * N[2] + cz);
                    for (int t = 0; t < ntri; ++t) {
                        int es[3] = { (int)((w >> (12 * t)) & 15), (int)((w >> (12 * t + 8)) & 15),
                                      (int)((w >> (12 * t + 4)) & 15) }; /* corner order 0, 2, 1 */
                        for (int k = 0; k < 3; ++k) {
                            int a = MC_EDGE[es[k]][0], b = MC_EDGE[es[k]][1];
                            int64_t pa[3] = { cx + MC_NB[a][0], cy + MC_NB[a][1], cz + MC_NB[a][2] };
                            int64_t pb[3] = { cx + MC_NB[b][0], cy + MC_NB[b][1], cz + MC_NB[b][2] };
                            int axis = pa[0] != pb[0] ? 0 : (pa[1] != pb[1] ? 1 : 2);
                            int64_t lo[3] = { pa[0] < pb[0] ? pa[0] : pb[0], pa[1] < pb[1] ? pa[1] : pb[1],
                                              pa[2] < pb[2] ? pa[2] : pb[2] };
                            uint64_t eid = 3ull * (uint64_t)((lo[0] * N[1] + lo[1]) * N[2] + lo[2]) + (uint64_t)axis;
                            u64push(&tri_edges, eid);
                            u64push(&verts, eid);
                        }
                        u64push(&tri_meta, (cube << 1) | (uint64_t)((c0 >> 15) & 1u));
                    }
                }
        /* exact weld: the label's vertices are its distinct lattice edges, in id order */
        size_t vb = voff[nl];
        qsort(verts.v + vb, verts.n - vb, sizeof(uint64_t), cmp_u64);
        size_t u = vb;
        for (size_t i = vb; i < verts.n; ++i)
            if (i == vb || verts.v[i] != verts.v[u - 1]) verts.v[u++] = verts.v[i];
        verts.n = u;
        voff[nl + 1] = (uint32_t)verts.n;
        toff[nl + 1] = (uint32_t)tri_meta.n;
        (void)tbegin;
        ++nl;
    }
    m->nl = nl;
    m->nv = (uint32_t)verts.n;
    m->nt = (uint32_t)tri_meta.n;
    m->labels = labels;
    m->voff = voff;
    m->toff = toff;
    m->edge = verts.v;
    m->xyz = (float*)malloc((m->nv ? m->nv : 1) * 3 * sizeof(float));
    m->xyz64 = (double*)malloc((m->nv ? m->nv : 1) * 3 * sizeof(double));
    for (uint32_t i = 0; i < m->nv; ++i) {
        uint64_t eid = verts.v[i];
        int axis = (int)(eid % 3);
        uint64_t p = eid / 3;
        int64_t pz = (int64_t)(p % (uint64_t)N[2]), py = (int64_t)((p / (uint64_t)N[2]) % (uint64_t)N[1]),
                px = (int64_t)(p / ((uint64_t)N[1] * (uint64_t)N[2]));
        int64_t P[3] = { px, py, pz };
        for (int q = 0; q < 3; ++q) {
            float t = (float)P[q] - 1.0f + (q == axis ? 0.5f : 0.0f);
            m->xyz[3 * i + q] = aabb_min[q] + cs[q] * t;
            m->xyz64[3 * i + q] = (double)aabb_min[q] + cs64[q] * ((double)P[q] - 1.0 + (q == axis ? 0.5 : 0.0));
        }
    }
    m->tri = (uint32_t*)malloc((m->nt ? m->nt : 1) * 3 * sizeof(uint32_t));
    m->tri_flag = (uint8_t*)malloc(m->nt ? m->nt : 1);
    m->tri_cube = (uint64_t*)malloc((m->nt ? m->nt : 1) * sizeof(uint64_t));
    for (uint32_t l = 0; l < nl; ++l)
        for (uint32_t t = toff[l]; t < toff[l + 1]; ++t) {
            m->tri_flag[t] = (uint8_t)(tri_meta.v[t] & 1u);
            m->tri_cube[t] = tri_meta.v[t] >> 1;
            for (int k = 0; k < 3; ++k) {
                uint64_t eid = tri_edges.v[3 * (size_t)t + k];
                size_t lo = voff[l], hi = voff[l + 1];
                while (lo < hi) {
                    size_t mid = (lo + hi) / 2;
                    if (verts.v[mid] < eid) lo = mid + 1; else hi = mid;
                }
                m->tri[3 * (size_t)t + k] = (uint32_t)(lo - voff[l]);
            }
        }
    free(tri_edges.v);
    free(tri_meta.v);
    free(counts);
    free(bb);
#undef PADV
    return 0;
}

void vfo_mesh_free(vfo_mesh* m)
{
    free(m->xyz); free(m->xyz64); free(m->edge); free(m->tri); free(m->tri_flag);
    free(m->tri_cube); free(m->labels); free(m->voff); free(m->toff);
    memset(m, 0, sizeof(*m));
}

/* ========================================================================= */
/* RLE grid format                                                           */
/* ========================================================================= */

size_t vfo_rle_encode(const uint16_t* grid, const uint32_t dims[3], uint8_t* out, size_t cap)
{
    /* RegularGrid::exportRLE (RegularGrid.cpp:672-714): uvec3 dims, then maximal
     * runs as packed little-endian {uint16 value, uint32 count}. */
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    size_t pos = 0;
    if (cap >= 12) memcpy(out, dims, 12);
    pos = 12;
    size_t i = 0;
    while (i < V) {
        uint16_t v = grid[i];
        size_t j = i;
        while (j < V && grid[j] == v && j - i < 0xFFFFFFFFu) ++j;
        uint32_t n = (uint32_t)(j - i);
        if (pos + 6 <= cap) { memcpy(out + pos, &v, 2); memcpy(out + pos + 2, &n, 4); }
        pos += 6;
        i = j;
    }
    return pos;
}

int vfo_rle_decode(const uint8_t* in, size_t n, uint32_t dims_out[3], uint16_t* grid, size_t cap)
{
    /* docs/decompress/decompress_grid.py:16-32 */
    if (n < 12 || (n - 12) % 6) return -1;
    memcpy(dims_out, in, 12);
    size_t V = (size_t)dims_out[0] * dims_out[1] * dims_out[2], pos = 0;
    if (V > cap) return -2;
    for (size_t o = 12; o < n; o += 6) {
        uint16_t v; uint32_t c;
        memcpy(&v, in + o, 2);
        memcpy(&c, in + o + 2, 4);
        if (pos + c > V) return -3;
        for (uint32_t k = 0; k < c; ++k) grid[pos++] = v;
    }
    return pos == V ? 0 : -4;
}

/* ------------------------------------------------------------------------- */
/* Laplacian smoothing.  Positions are scaled by 10^4 and truncated to int
 * (ivec3(p * UINT_MULT)), neighbours are summed per face (each face adds its
 * other two vertices to a vertex: laplacianSmoothing-comp.glsl:24-45), and
 * the finish step moves a target vertex to mix(p, sum / count / 10^4, w)
 * (finishLaplacianSmoothing-comp.glsl:18-30).  Integer sums make the result
 * independent of accumulation order. */
static void smooth_pass(vfo_mesh* m, const uint32_t* tri_g, const uint8_t* vflag, uint32_t iters, float weight,
                        uint8_t target, int check)
{
    const uint32_t nv = m->nv, nt = m->nt;
    int32_t* acc = (int32_t*)calloc((size_t)nv * 4 + 4, sizeof(int32_t));
    for (uint32_t it = 0; it < iters; ++it) {
        memset(acc, 0, (size_t)nv * 4 * sizeof(int32_t));
        for (uint32_t t = 0; t < nt; ++t) {
            const uint32_t* f = tri_g + 3 * (size_t)t;
            if (check && !(vflag[f[0]] == target && vflag[f[1]] == target && vflag[f[2]] == target)) continue;
            for (int i = 0; i < 3; ++i) {
                const float* p = m->xyz + 3 * (size_t)f[i];
                const int32_t vx = (int32_t)(p[0] * 10000.0f), vy = (int32_t)(p[1] * 10000.0f),
                              vz = (int32_t)(p[2] * 10000.0f);
                const uint32_t a = f[(i + 1) % 3], b = f[(i + 2) % 3];
                acc[4 * (size_t)a + 0] += vx; acc[4 * (size_t)a + 1] += vy; acc[4 * (size_t)a + 2] += vz; acc[4 * (size_t)a + 3] += 1;
                acc[4 * (size_t)b + 0] += vx; acc[4 * (size_t)b + 1] += vy; acc[4 * (size_t)b + 2] += vz; acc[4 * (size_t)b + 3] += 1;
            }
        }
        for (uint32_t v = 0; v < nv; ++v) {
            const int32_t* a = acc + 4 * (size_t)v;
            if (vflag[v] != target || a[3] <= 0) continue;
            float* p = m->xyz + 3 * (size_t)v;
            for (int q = 0; q < 3; ++q) {
                const float avg = (float)a[q] / (float)a[3] / 10000.0f;
                p[q] = p[q] * (1.0f - weight) + avg * weight;  /* mix(p, avg, weight) */
            }
        }
    }
    free(acc);
}

void vfo_mesh_smooth(vfo_mesh* m, uint32_t nb_iters, float nb_weight, uint32_t b_iters, float b_weight)
{
    const uint32_t nv = m->nv, nt = m->nt;
    if (!nv || !nt) return;
    uint32_t* tri_g = (uint32_t*)malloc((size_t)nt * 3 * sizeof(uint32_t));
    uint8_t* vflag = (uint8_t*)calloc(nv, 1);
    for (uint32_t l = 0; l < m->nl; ++l)
        for (uint32_t t = m->toff[l]; t < m->toff[l + 1]; ++t)
            for (int i = 0; i < 3; ++i) {
                const uint32_t g = m->voff[l] + m->tri[3 * (size_t)t + i];
                tri_g[3 * (size_t)t + i] = g;
                if (m->tri_flag[t]) vflag[g] = 1;
            }
    smooth_pass(m, tri_g, vflag, nb_iters, nb_weight, 0, 1);
    smooth_pass(m, tri_g, vflag, b_iters, b_weight, 1, 0);
    free(tri_g);
    free(vflag);
}
