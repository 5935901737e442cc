/*
 * vf_oracle — CPU restatement of VoxelFragmentML's voxelize -> fracture path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Grid layout follows the reference: linear index x*ny*nz + y*nz + z (z
 * fastest, RegularGrid.cpp:1031-1034), uint16 cell values with 0 = EMPTY,
 * 1 = FREE, >= 2 = fragment label and bit 15 = boundary mask
 * (RegularGrid.h:12-18, voxelMask.glsl:1-32).
 *
 * Each function cites the reference code whose semantics it restates.  The
 * parity pins (golden vectors, closed forms, brute force) live in tests/.
 */
#ifndef VF_ORACLE_H
#define VF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VFO_EUCLIDEAN = 0, VFO_MANHATTAN = 1, VFO_CHEBYSHEV = 2 };

/* ---- voxelization: FP32 triangle/box SAT shell + z-parity interior ------ */
int vfo_sat_triangle_box(const float v0[3], const float v1[3], const float v2[3],
                         const float center[3], const float half[3]);
/* occ, shell, parity: V bytes each (shell/parity may be NULL). Returns #occupied. */
int64_t vfo_voxelize(const float* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt,
                     const float aabb_min[3], const float aabb_max[3], const uint32_t dims[3],
                     uint8_t* occ, uint8_t* shell, uint8_t* parity);

/* ---- grid helpers ------------------------------------------------------- */
void vfo_occ_to_grid(const uint8_t* occ, uint16_t* grid, size_t n);
void vfo_homogenize(uint16_t* grid, size_t n);
void vfo_undo_mask(uint16_t* grid, size_t n);
void vfo_unmask_rightmost(uint16_t* grid, size_t n, uint32_t position);
void vfo_detect_boundaries(const uint16_t* in, uint16_t* out, const uint32_t dims[3], int size);
void vfo_histogram(const uint16_t* grid, size_t n, uint32_t* counts /* 65536 */);

/* ---- fracture ----------------------------------------------------------- */
/* seeds: S x {x, y, z, label} (glm::uvec4 layout) */
void vfo_naive(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
               const uint32_t* seeds, uint32_t S, int dist);
uint32_t vfo_noise_delta(uint64_t noise_seed, uint32_t idx, uint32_t amp);
/* neighbourhood: 6 or 26.  Returns 0 on success. */
int vfo_flood(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
              const uint32_t* seeds, uint32_t S, int neighbourhood,
              uint64_t noise_seed, uint32_t amp);
/* independent brute force of the first growth: one Dijkstra per seed, then
 * lexicographic argmin (labels 2..255, no sub-seed pass) */
int vfo_flood_dijkstra(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
                       const uint32_t* seeds, uint32_t S, int neighbourhood,
                       uint64_t noise_seed, uint32_t amp);
void vfo_remove_isolated(const uint16_t* in, uint16_t* out, const uint32_t dims[3],
                         const uint32_t* seeds, uint32_t S);

/* ---- erosion (RegularGrid::erode) --------------------------------------- */
enum { VFO_EROSION_SQUARE = 0, VFO_EROSION_ELLIPSE = 1, VFO_EROSION_CROSS = 2 };  /* FractureParameters::ErosionType */
/* mask: side^3 bytes (side = size rounded up to odd); returns side or -1 */
int vfo_erosion_mask(int type, uint32_t size, uint8_t* mask, float* activations);
float vfo_erosion_noise(uint64_t seed, uint32_t cell);
void vfo_erode_step(const uint16_t* in, uint16_t* out, const uint32_t dims[3], const uint8_t* mask, uint32_t side,
                    float limit, float probability, uint64_t noise_seed);
void vfo_remove_isolated_cells(const uint16_t* in, uint16_t* out, const uint32_t dims[3]);
int vfo_erode(uint16_t* grid, const uint32_t dims[3], int type, uint32_t size, uint32_t iterations, float probability,
              float threshold, uint64_t noise_seed);

/* ---- seeding ------------------------------------------------------------ */
enum { VFO_SEED_INNER = 0, VFO_SEED_OUTER = 1, VFO_SEED_BOTH = 2 };  /* Seeder::Location order (Seeder.h:24) */
uint32_t vfo_hash32(uint64_t key, uint32_t counter);
int vfo_seed_uniform(const uint16_t* grid, const uint32_t dims[3], uint32_t S,
                     uint64_t seed, int location, uint32_t* seeds_out);
/* Seeder::mergeSeeds: seeds[i].label = nearest fragment's label | running sub-id << 8.
 * 0, -1 bad input, -2 more than 255 sub-ids for one fragment. */
int vfo_seed_merge(const uint32_t* frags, uint32_t F, uint32_t* seeds, uint32_t N, int dist);
/* Seeder::nearSeeds with counter-based draws; out: F + num_seeds entries. 0, -1 bad input, -2 draw cap */
uint64_t vfo_near_key(uint64_t seed, uint32_t impact);
int vfo_seed_near(const uint16_t* grid, const uint32_t dims[3], const uint32_t* frags, uint32_t F,
                  uint32_t impacts, uint32_t num_seeds, uint32_t spreading, uint64_t seed,
                  uint32_t* out, uint32_t* count);

/* ---- marching cubes ----------------------------------------------------- */
typedef struct {
    uint32_t nv, nt, nl;
    float* xyz;          /* 3*nv FP32 world positions */
    double* xyz64;       /* 3*nv FP64 world positions (oracle of the oracle) */
    uint64_t* edge;      /* nv lattice edge ids: 3*padded_index + axis */
    uint32_t* tri;       /* 3*nt vertex indices, local to the label's vertex range */
    uint8_t* tri_flag;   /* nt boundary flags */
    uint64_t* tri_cube;  /* nt padded cube index of each triangle */
    uint16_t* labels;    /* nl target labels, ascending */
    uint32_t* voff;      /* nl+1 vertex offsets per label */
    uint32_t* toff;      /* nl+1 triangle offsets per label */
} vfo_mesh;
int vfo_marching_cubes(const uint16_t* grid, const uint32_t dims[3],
                       const float aabb_min[3], const float aabb_max[3],
                       uint32_t label /* 0xFFFFFFFF = all */, vfo_mesh* out);
void vfo_mesh_free(vfo_mesh* m);
/* Laplacian smoothing of every fragment mesh (MarchingCubes::smoothSurface,
 * MarchingCubes.cpp:395-401, 491-514; laplacianSmoothing-comp.glsl,
 * finishLaplacianSmoothing-comp.glsl): nb_iters Jacobi steps moving the
 * non-boundary vertices (faces with three non-boundary vertices only) with
 * weight nb_weight, then b_iters steps moving the boundary vertices (every
 * face) with b_weight.  A vertex is a boundary vertex when any of its
 * triangles is flagged.  Updates m->xyz in place (FP32). */
void vfo_mesh_smooth(vfo_mesh* m, uint32_t nb_iters, float nb_weight, uint32_t b_iters, float b_weight);
uint64_t vfo_mc_packed(uint32_t config);

/* ---- RLE grid format (RegularGrid::exportRLE) ----------------------------- */
size_t vfo_rle_encode(const uint16_t* grid, const uint32_t dims[3], uint8_t* out, size_t cap);
int vfo_rle_decode(const uint8_t* in, size_t n, uint32_t dims_out[3], uint16_t* grid, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
